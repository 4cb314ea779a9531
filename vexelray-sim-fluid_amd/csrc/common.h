// Internal declarations shared by the runtime and the kernel files of libvexelhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

namespace vxr {

// ------------------------------------------------------------------ device helpers

// The IR's min/max as selects: a signed zero goes the same way on the device and in the
// CPU restatement, and no NaN rule of v_min/v_max enters.
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }

// Hardware float atomic add into global memory (global_atomic_add_f32, no return).
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

constexpr int kWave = 64;

// The least or greatest value over a wave's 64 lanes, all of them active, in every lane, without
// going through LDS: DPP folds each row of 16 lanes (swap neighbours, swap pairs, mirror within 8,
// mirror within 16), then the four rows' results are read into scalar registers.
template <typename Op>
__device__ __forceinline__ int wave_fold(int v, Op op) {
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false));  // row_mirror
    return op(op(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
              op(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

__device__ __forceinline__ int wave_min(int v) {
    return wave_fold(v, [](int a, int b) { return a < b ? a : b; });
}

__device__ __forceinline__ int wave_max(int v) {
    return wave_fold(v, [](int a, int b) { return a > b ? a : b; });
}

// R consecutive floats from i (i a multiple of R, the arrays 256-byte aligned): one load.
template <int R>
__device__ __forceinline__ void load_run(const float* __restrict__ p, int64_t i, float (&v)[R]) {
    if constexpr (R == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p + i);
        v[0] = t.x;
        v[1] = t.y;
        v[2] = t.z;
        v[3] = t.w;
    } else if constexpr (R == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p + i);
        v[0] = t.x;
        v[1] = t.y;
    } else {
        v[0] = p[i];
    }
}

// ------------------------------------------------------------------ launch plumbing

struct Timer;  // event pairs per kernel name, owned by the runtime

// Every launch goes through here so timing can bracket it with events on `stream`.
struct Launch {
    hipStream_t stream;
    Timer* timer;  // null when timing is off
};

void timed_begin(const Launch& l, const char* kernel);
void timed_end(const Launch& l, const char* kernel);

// Returns hipGetLastError() after a launch as a status, with a message.
int check_launch(const char* kernel);

// ------------------------------------------------------------------ shallow water

struct SweArgs {
    int nx, ny;
    int edges[4];  // west, east, south, north: 0 wall, 1 open
    float* out[3];
    const float* in[3];
    const int32_t* params;  // g, courant, dx, dy, dry as f32 bits
    const int32_t* speed_in;
    int32_t* speed_out;
    int32_t* speed_clear;
    const int32_t* budget_in;
    int32_t* budget_out;
    int32_t* clamped;
    // A step of fixed_ticks ticks (> 0) instead of the tick rule: speed_in and budget_in are not
    // read, and speed_out, speed_clear, budget_out and clamped may be null (not written then).
    int32_t fixed_ticks;
    // When set, the step runs only if *gate is nonzero (every wave leaves at once otherwise).
    const int32_t* gate;
};
// One HLL step; `fast` trades the reference's bits for the hardware reciprocal and square root.
int swe_step(const Launch& l, const SweArgs& a, int rows, bool fast = false);

// ------------------------------------------------------------------ fill

int fill_zero(const Launch& l, void* p, int64_t words);

// ------------------------------------------------------------------ MPM

// Particle fields: 3D x y z u v w m j c00..c22 (17); 2D x y u v m j c00 c01 c10 c11 (10).
struct Mpm3Args {
    int nx, ny, nz;
    int64_t count;     // real particles
    int64_t capacity;  // allocated (multiple of the chunk), tail particles have m = 0
    float* p[17];
    float* g[7];  // gm gmu gmv gmw gu gv gw
    const int32_t* params;
    // Bricks of 8x8x8 nodes: P2G sets `touched` for every brick its stencils reach; the grid pass
    // then moves it to `live` (bricks whose sums and velocities may be nonzero) and clears it.
    // Clear and grid visit, in each row of nodes, only the span of bricks flagged in either.
    uint8_t* touched;
    uint8_t* live;
    // The slot each particle was loaded as (permuted with the particles by a sort), the position
    // of each slot (home[slot[p]] = p), and the particles from the first whose key is the
    // inactive one (m <= 0) on: P2G and G2P skip them; equal to count unless the last sort
    // found inactive slots and nothing has given mass to a slot since.
    int32_t* slot;
    int32_t* home;
    int32_t* active;
    // With tension: per brick, bit 0 = within one brick of a touched or live brick at the last
    // region computation, bit 1 = bit 0 then or at the computation before (the blur and stress
    // passes visit the span of bit 1).
    uint8_t* near;
};
constexpr int kBrick = 8;
inline int64_t brick_count3(int nx, int ny, int nz) {
    return (int64_t)((nx + kBrick - 1) / kBrick) * ((ny + kBrick - 1) / kBrick) * ((nz + kBrick - 1) / kBrick);
}
int mpm3_clear(const Launch& l, const Mpm3Args& a);
// P2G schedule 2 (fixed point): the sums of the amounts rounded to 64-bit integers (mass 2^-40,
// momentum 2^-32), the same bits in any order. The clear zeroes them beside the floats; the
// scatter adds into them and then decodes them into gm gmu gmv gmw.
struct Fixed3 {
    unsigned long long* q[4];
};
int mpm3_clear_fixed(const Launch& l, const Mpm3Args& a, const Fixed3& fx);
int mpm3_scatter_fixed(const Launch& l, const Mpm3Args& a, const Fixed3& fx);
int mpm3_scatter_atomic(const Launch& l, const Mpm3Args& a);
// Chunks of 64 * run particles, run in {1, 2, 4}; `ints`: the wave tiles sum in 32-bit integers
// (LDS integer atomics) instead of floats (read-add-write rounds).
int mpm3_scatter_tiled(const Launch& l, const Mpm3Args& a, int run, bool ints);
int mpm3_grid(const Launch& l, const Mpm3Args& a);
int mpm3_advect(const Launch& l, const Mpm3Args& a);
int mpm3_advect_tiled(const Launch& l, const Mpm3Args& a, int run, bool fast);
int mpm3_keys(const Launch& l, const Mpm3Args& a, int32_t* keys, int32_t* index);
// The key of a slot with m <= 0: one past every base cell's, so the sort puts those slots last.
int32_t mpm3_inactive_key(int nx, int ny, int nz);
// After the permute of a sort (`a` over the permuted fields): home from slot, and active from
// the sorted keys.
int mpm3_after_sort(const Launch& l, const Mpm3Args& a, const int32_t* keys);
// Every slot at its own position (slot = home = identity) and active = count.
int mpm3_reset_slots(const Launch& l, const Mpm3Args& a);
// home from slot (after the host wrote slot), and active = count.
int mpm3_rebuild_home(const Launch& l, const Mpm3Args& a);
// After a grid pass that did not go through mpm3_grid: touched bricks become live, touched clears.
int mpm3_flags(const Launch& l, const Mpm3Args& a);
// 3D surface tension (core/particle/Tension3.java): blur (scaled: gm to a colour), stress
// (t = txx tyy tzz txy txz tyz), and the grid update with the capillary force.
int tension3_region(const Launch& l, const Mpm3Args& a, bool all);
int tension3_blur(const Launch& l, const Mpm3Args& a, const float* src, float* dst, bool scaled, bool dense);
int tension3_stress(const Launch& l, const Mpm3Args& a, const float* colour, float* const t[6], bool dense);
int mpm3_grid_tension(const Launch& l, const Mpm3Args& a, float* const t[6], const float* colour);
// Flip3.lod / Flip3.lodTarget over count / 8 groups of 8 consecutive slots.
int mpm3_lod(const Launch& l, const Mpm3Args& a, float* level, const float* target, float* stay);
int mpm3_lod_target(const Launch& l, const Mpm3Args& a, const float* level, float* target, const float* view);
int mpm3_key_bits(int nx, int ny, int nz);

// The non-empty groups (256 cells, half a sort brick: keys key >> 8) of the last sort that laid
// the particles out group by group for P2G schedule 3: per group, its cells ranked by particle
// count in generations of 64, and the particles generation by generation, rank-major.
struct Layers {
    int32_t* layer;   // each non-empty group's number, in key order
    int32_t* start;   // its first particle
    int32_t* count;   // mpm3_layer_cells() words a group: the particles of each of its cells
    uint16_t* cell;   // mpm3_layer_cells() a group: the cell at each place of the count order
    int32_t* meta;    // [0] non-empty groups, [1] end of the active particles, [2..5] work tickets (zero between launches)
    int64_t cap;      // groups the key space holds
};
int mpm3_layer_cells();
int64_t mpm3_layer_capacity(int nx, int ny, int nz);
// Words of scratch mpm3_layers takes: the new index, the new keys (capacity each), block counts.
int64_t mpm3_layer_scratch_words(int64_t capacity);
// After the radix sort (keys sorted, index its permutation): the layer tables, the source index
// of the layer-by-layer order (scratch[0, count)), and the keys in that order (copied into keys).
int mpm3_layers(const Launch& l, const Mpm3Args& a, int32_t* keys, const int32_t* index, int32_t* scratch,
                const Layers& L);
// P2G over the layers (particles in the order mpm3_layers made, possibly moved since).
int mpm3_scatter_layers(const Launch& l, const Mpm3Args& a, const Layers& L);

struct Mpm2Args {
    int nx, ny;
    int64_t count, capacity;
    float* p[10];
    float* g[5];  // gm gmu gmv gu gv
    const int32_t* params;
};
int mpm2_clear(const Launch& l, const Mpm2Args& a);
int mpm2_scatter_atomic(const Launch& l, const Mpm2Args& a);
int mpm2_scatter_runs(const Launch& l, const Mpm2Args& a, int run);
int mpm2_grid(const Launch& l, const Mpm2Args& a);
int mpm2_advect(const Launch& l, const Mpm2Args& a);
int mpm2_keys(const Launch& l, const Mpm2Args& a, int32_t* keys, int32_t* index);
int mpm2_starts(const Launch& l, const int32_t* sorted_keys, int64_t count, int32_t cells, int32_t* starts);
// 2D surface tension (core/particle/Tension.java), dense over the grid: blur (scaled: gm to a
// colour), stress (t = txx tyy txy), and the grid update with the capillary force (reads
// gm gmu gmv, writes gu gv).
int tension2_blur(const Launch& l, const Mpm2Args& a, const float* src, float* dst, bool scaled);
int tension2_stress(const Launch& l, const Mpm2Args& a, const float* colour, float* const t[3]);
int mpm2_grid_tension(const Launch& l, const Mpm2Args& a, float* const t[3], const float* colour);

// Many 2D scenes in one set of buffers, stepped by one launch per pass (a single scene with the
// fixed-point P2G is a batch of one). Scene s: particles [poff, poff + count) padded to
// kBatchAlign, nodes [goff, goff + nx*ny) padded to 256, starts [soff, soff + cells + 1) padded
// to 256, 28 params at params[28 s].
constexpr int kBatchAlign = 1024;
struct Scene2 {
    int nx, ny;
    int32_t cells, unused;
    int64_t poff, count, goff, soff;
};
struct Batch2Args {
    const Scene2* scenes;
    const int32_t* pscene;  // scene of each kBatchAlign particles
    const int32_t* gscene;  // scene of each 256 nodes
    const int32_t* sscene;  // scene of each 256 starts
    int64_t pchunks, gchunks, schunks;
    float* p[10];
    float* g[5];
    unsigned long long* q[3];  // fixed-point sums: mass (2^-40), momentum u, v (2^-32)
    const int32_t* params;
    int cell_bits;
    // A byte for each unit of 64 nodes (four to a 256-node chunk): P2G sets `touched` for every
    // unit a row of its tile box reaches; the grid pass moves it to `live` (units whose sums and
    // velocities may be nonzero) and clears it. Clear and grid visit only units flagged in either.
    uint8_t* touched;
    uint8_t* live;
    // Per starts chunk: its first cell's key, its scene's first particle, how many of its cells
    // the scene has (bit 16 set on the scene's last chunk) and the scene's particles' end, set
    // with the layout; and the first of its scene's sorted particles whose key is at least the
    // first cell's, which the sort writes before the starts.
    const int32_t* sinfo;  // 4 words a chunk
    int32_t* bounds;
};
// `blocks` > 0 caps the grid (each block then strides over chunks, or over windows of 64 unit
// flags for clear, grid and decode); 0 is a block per chunk or window.
int batch2_clear(const Launch& l, const Batch2Args& a, int blocks);
int batch2_scatter(const Launch& l, const Batch2Args& a, int run, int blocks);
// The fixed-point sums decoded into gm gmu gmv (no velocities), and, after a grid update that did
// not go through batch2_grid, the chunk flags moved on (touched -> live, touched cleared).
int batch2_decode(const Launch& l, const Batch2Args& a, int blocks);
int batch2_flags(const Launch& l, const Batch2Args& a);
int batch2_grid(const Launch& l, const Batch2Args& a, int blocks);
int batch2_advect(const Launch& l, const Batch2Args& a, int blocks);
int batch2_keys(const Launch& l, const Batch2Args& a, int32_t* keys, int32_t* index, int blocks);
int batch2_starts(const Launch& l, const Batch2Args& a, const int32_t* keys, int32_t* starts, int blocks);

// ------------------------------------------------------------------ fixed-point bilinear scatter

// core/particle/FixedScatter.java: run <= 0 one particle a thread, else runs of `run` consecutive
// particles folded before their atomics. Adds into the i32 grids (n x n).
int fixed_scatter(const Launch& l, int n, int64_t count, const float* x, const float* y, const float* u,
                  const float* v, const float* m, int mass_bits, int velocity_bits, int run, int* gm, int* gmu,
                  int* gmv);

// ------------------------------------------------------------------ diagnostics

// Diagnostics.of on the device: out = volume, min depth, max depth, max speed, max Froude, max
// Courant, non-finite, negative, wet. Scratch of diag_scratch_words words. Synchronous.
int64_t diag_scratch_words();
int swe_diagnostics(const Launch& l, const float* h, const float* hu, const float* hv, int64_t cells, double g,
                    double dry, double dt, double dx, double dy, int32_t* scratch, double out[9]);
// ParticleDiagnostics.of on the device, sorting the J keys in keys/index (capacity words each)
// with the radix sort's scratch; `small`: 16 words. Synchronous.
int particle_diagnostics(const Launch& l, const float* j, int64_t count, int64_t capacity, int32_t* keys,
                         int32_t* index, int32_t* sort_scratch, int digit_bits, int32_t* small, double speed,
                         double sound, double dt, double out[7]);

// ParticleDiagnostics.of with its sort in the library's fixed device array (no allocation): at
// most particle_diagnostics_static_limit() particles; `why` says what failed. Synchronous.
int64_t particle_diagnostics_static_limit();
int particle_diagnostics_static(const Launch& l, const float* j, int64_t count, double speed, double sound,
                                double dt, double out[7], const char** why);

// ------------------------------------------------------------------ sort

// Stable LSD radix sort of (key, index) pairs, keys in [0, 2^bits). Scratch sized by
// radix_scratch_words. Leaves the sorted pairs in keys/index.
int64_t radix_scratch_words(int64_t n, int digit_bits);
int radix_sort(const Launch& l, int32_t* keys, int32_t* index, int64_t n, int bits, int digit_bits,
               int32_t* scratch);
// out[f][i] = in[f][index[i]] for `fields` float fields.
int permute_fields(const Launch& l, float* const* in, float* const* out, int fields, const int32_t* index,
                   int64_t n);

}  // namespace vxr
