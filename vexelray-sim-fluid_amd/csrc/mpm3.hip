// The 3D MLS-MPM / APIC substep (core/particle/Flip3.java) for gfx950.
//
// Layout in HBM: 17 particle fields as separate f32 arrays (SoA, x y z u v w m j c00..c22),
// padded to a whole number of 2048-particle chunks with m = 0 in the padding; the grid as 7
// f32 node arrays indexed (k*ny + j)*nx + i (gm gmu gmv gmw, then gu gv gw).
//
// P2G comes in the reference's schedule (one thread per particle, 108 float atomics,
// Flip3.java:193-246) and a sorted-segmented one: a block takes a contiguous chunk of
// particles, a thread folds runs of consecutive particles that share a stencil base into 108
// register sums, each run lands in an LDS tile of the chunk's nodes by LDS float atomics, and
// the tile leaves with one global atomic per node and field. Correct in any order; fast when
// the particles are sorted by cell (the sort keys bricks of 8x8x8 cells together).
//
// The grid update and G2P keep the reference's arithmetic order exactly, so from the same
// grid they reproduce the CPU restatement bit for bit; G2P reads its nodes from an LDS tile.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "common.h"

namespace vxr {
namespace {

__device__ __forceinline__ float bits_f(int32_t w) { return __int_as_float(w); }

enum { DT = 0, GX = 1, GY = 2, GZ = 3, BULK = 4, RHO0 = 5 };  // Flip3.java:58-64
constexpr float kWall = 1.0f;                                // Flip.WALL
constexpr float kEmpty = 1e-9f;                              // Flip3.EMPTY
constexpr float kJBound = 10.0f;                             // Flip3.J_BOUND
constexpr int kBlock = 256;

// One axis of Flip3.Stencil3.of, Flip3.java:670-676 and :682-688.
struct Axis {
    int cell;
    float off;
    float w[3];
};

__device__ __forceinline__ Axis axis(float pos, int n) {
    Axis s;
    const float x = clampf(pos, 0.5f, (float)n - 1.5f);
    s.cell = (int)mn(x - 0.5f, (float)(n - 3));  // x - 1/2 >= 0, so truncation is the floor
    s.off = x - (float)s.cell;
    const float a = 1.5f - s.off;
    const float c = s.off - 1.0f;
    const float e = s.off - 0.5f;
    s.w[0] = 0.5f * (a * a);
    s.w[1] = 0.75f - (c * c);
    s.w[2] = 0.5f * (e * e);
    return s;
}

__device__ __forceinline__ int base_cell(float pos, int n) {
    const float x = clampf(pos, 0.5f, (float)n - 1.5f);
    return (int)mn(x - 0.5f, (float)(n - 3));
}

// The 108 amounts of one particle (Flip3.java:216-243), node k's four at 4k .. 4k+3, added
// into acc. The arithmetic of each amount is the reference's.
__device__ __forceinline__ void p2g_amounts(const Axis (&ax)[3], float m, const float (&vel)[3], const float (&c)[9],
                                            float s, float (&acc)[108]) {
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const float wgt = ax[0].w[k % 3] * (ax[1].w[(k / 3) % 3] * ax[2].w[k / 9]);
        const float d[3] = {(float)(k % 3) - ax[0].off, (float)((k / 3) % 3) - ax[1].off,
                            (float)(k / 9) - ax[2].off};
        const float wm = wgt * m;
        acc[4 * k] += wm;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float at3 = vel[a] + (c[3 * a] * d[0] + (c[3 * a + 1] * d[1] + c[3 * a + 2] * d[2]));
            const float push = s * d[a];
            acc[4 * k + 1 + a] += wm * at3 + wgt * push;
        }
    }
}

struct Particle {
    Axis ax[3];
    float m, vel[3], c[9], s;
};

__device__ __forceinline__ void load_particle(const Mpm3Args& a, int64_t q, float dt, float bulk, float rho0,
                                              Particle& p) {
    p.m = a.p[6][q];
    p.ax[0] = axis(a.p[0][q], a.nx);
    p.ax[1] = axis(a.p[1][q], a.ny);
    p.ax[2] = axis(a.p[2][q], a.nz);
#pragma unroll
    for (int i = 0; i < 3; ++i) p.vel[i] = a.p[3 + i][q];
#pragma unroll
    for (int i = 0; i < 9; ++i) p.c[i] = a.p[8 + i][q];
    p.s = (4.0f * (dt * (p.m / rho0))) * (bulk * (1.0f - a.p[7][q]));
}

// ------------------------------------------------------------------ touched bricks, clear

// Marks the bricks holding nodes [lo, hi] (inclusive, per axis) as touched; lanes of a wave
// share the stores.
__device__ __forceinline__ void mark_bricks(const Mpm3Args& a, const int (&lo)[3], const int (&hi)[3], int lane,
                                            int lanes) {
    const int bxn = (a.nx + kBrick - 1) / kBrick, byn = (a.ny + kBrick - 1) / kBrick;
    const int b0[3] = {lo[0] / kBrick, lo[1] / kBrick, lo[2] / kBrick};
    const int e[3] = {hi[0] / kBrick - b0[0] + 1, hi[1] / kBrick - b0[1] + 1, hi[2] / kBrick - b0[2] + 1};
    const int total = e[0] * e[1] * e[2];
    if (e[0] <= 2 && e[1] <= 2 && e[2] <= 2) {
        // every caller's box: a stencil or one brick's stencils, at most 10 nodes an axis, so at
        // most two bricks an axis and eight in all; brick t's x and y offsets are bits of t
        const int sx = e[0] - 1, sy = e[1] - 1;
#pragma unroll 1
        for (int t = lane; t < total; t += lanes) {
            const int bx = b0[0] + (t & sx), by = b0[1] + ((t >> sx) & sy), bz = b0[2] + (t >> (sx + sy));
            a.touched[((int64_t)bz * byn + by) * bxn + bx] = 1;
        }
        return;
    }
#pragma unroll 1
    for (int t = lane; t < total; t += lanes) {
        const int bx = b0[0] + t % e[0], by = b0[1] + (t / e[0]) % e[1], bz = b0[2] + t / (e[0] * e[1]);
        a.touched[((int64_t)bz * byn + by) * bxn + bx] = 1;
    }
}

// A wave a row of nodes (y, z): the x-range from the first to the last brick of the row flagged
// in `touched` or `live` (brick rows of 64 at a time). Returns false when none is.
__device__ __forceinline__ bool row_range(const Mpm3Args& a, int y, int z, int lane, int* x0, int* x1) {
    const int bxn = (a.nx + kBrick - 1) / kBrick, byn = (a.ny + kBrick - 1) / kBrick;
    const int64_t base = ((int64_t)(z / kBrick) * byn + y / kBrick) * bxn;
    int first = INT_MAX, last = -1;
    for (int b0 = 0; b0 < bxn; b0 += kWave) {
        const int b = b0 + lane;
        const bool on = b < bxn && (a.touched[base + b] | a.live[base + b]) != 0;
        const uint64_t m = __ballot(on);
        if (m) {
            first = min(first, b0 + (int)__builtin_ctzll(m));
            last = b0 + 63 - (int)__builtin_clzll(m);
        }
    }
    if (last < 0) return false;
    *x0 = first * kBrick;
    *x1 = min((last + 1) * kBrick, a.nx);
    return true;
}

// Zeroes the sums over the flagged range of every row (rows of 4 nodes in float4 when aligned).
__global__ __launch_bounds__(256) void clear_rows_kernel(Mpm3Args a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= (int64_t)a.ny * a.nz) return;
    const int y = (int)(row % a.ny), z = (int)(row / a.ny);
    int x0, x1;
    if (!row_range(a, y, z, lane, &x0, &x1)) return;
    const int64_t first = row * a.nx;
    if ((a.nx & 3) == 0) {
        const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int x = x0 + 4 * lane; x < x1; x += 4 * kWave)
#pragma unroll
            for (int f = 0; f < 4; ++f) *reinterpret_cast<float4*>(a.g[f] + first + x) = zero;
    } else {
        for (int x = x0 + lane; x < x1; x += kWave)
#pragma unroll
            for (int f = 0; f < 4; ++f) a.g[f][first + x] = 0.0f;
    }
}

// After a grid pass: the velocities are nonzero at most where P2G reached; the sums will be
// zeroed there by the next clear. live := touched, touched := 0.
__global__ void flags_kernel(uint8_t* __restrict__ touched, uint8_t* __restrict__ live, int64_t bricks) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bricks) return;
    live[b] = touched[b];
    touched[b] = 0;
}

// ------------------------------------------------------------------ P2G, the reference schedule

// One particle's 108 amounts straight into the grid by float atomics, in the reference's
// arithmetic (Flip3.java:225-243).
__device__ __forceinline__ void atomic_particle(const Mpm3Args& a, const Axis (&ax)[3], float m, const float (&vel)[3],
                                                const float (&c)[9], float s) {
    const int64_t base = ((int64_t)ax[2].cell * a.ny + ax[1].cell) * a.nx + ax[0].cell;
    const int64_t plane = (int64_t)a.nx * a.ny;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const float wgt = ax[0].w[k % 3] * (ax[1].w[(k / 3) % 3] * ax[2].w[k / 9]);
        const float d[3] = {(float)(k % 3) - ax[0].off, (float)((k / 3) % 3) - ax[1].off, (float)(k / 9) - ax[2].off};
        const float wm = wgt * m;
        const int64_t node = base + (k / 9) * plane + ((k / 3) % 3) * a.nx + k % 3;
        atomic_add(a.g[0] + node, wm);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float at3 = vel[i] + (c[3 * i] * d[0] + (c[3 * i + 1] * d[1] + c[3 * i + 2] * d[2]));
            const float push = s * d[i];
            atomic_add(a.g[1 + i] + node, wm * at3 + wgt * push);
        }
    }
}

__global__ __launch_bounds__(kBlock) void p2g3_atomic_kernel(Mpm3Args a) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= *a.active) return;  // the inactive slots a sort moved to the end
    if (!(a.p[6][q] > 0.0f)) return;
    const float dt = bits_f(a.params[DT]), bulk = bits_f(a.params[BULK]), rho0 = bits_f(a.params[RHO0]);
    Particle p;
    load_particle(a, q, dt, bulk, rho0, p);
    const int lo[3] = {p.ax[0].cell, p.ax[1].cell, p.ax[2].cell};
    const int hi[3] = {lo[0] + 2, lo[1] + 2, lo[2] + 2};
    mark_bricks(a, lo, hi, 0, 1);
    atomic_particle(a, p.ax, p.m, p.vel, p.c, p.s);
}

// ------------------------------------------------------------------ block bounding box

// Min and max over a wave by ds_bpermute exchanges. The DPP folds (wave_min, wave_max) issue
// fewer instructions but, where the wave holds its particles' fields, need more registers:
// G2P went from 122 to 130 VGPRs (four waves a SIMD to three) and P2G spilled into AGPRs.
__device__ __forceinline__ int lanes_min(int v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ int lanes_max(int v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

struct Box {
    int lo[3], hi[3];
    int ex, ey, ez, nodes;
    bool empty;
};

// The block-wide box of base cells from each thread's own lo/hi (empty: lo > hi). Every
// thread returns the same box; nodes = INT_MAX when the box's nodes exceed `cap`. Uses red[24].
__device__ Box reduce_box(int (&lo)[3], int (&hi)[3], int* red, int cap) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        lo[i] = lanes_min(lo[i]);
        hi[i] = lanes_max(hi[i]);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            red[wave * 6 + i] = lo[i];
            red[wave * 6 + 3 + i] = hi[i];
        }
    }
    __syncthreads();
    Box b;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        b.lo[i] = INT_MAX;
        b.hi[i] = INT_MIN;
    }
    for (int w = 0; w < kBlock / kWave; ++w) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            b.lo[i] = min(b.lo[i], red[w * 6 + i]);
            b.hi[i] = max(b.hi[i], red[w * 6 + 3 + i]);
        }
    }
    b.empty = b.lo[0] > b.hi[0];
    b.ex = b.empty ? 0 : b.hi[0] - b.lo[0] + 3;
    b.ey = b.empty ? 0 : b.hi[1] - b.lo[1] + 3;
    b.ez = b.empty ? 0 : b.hi[2] - b.lo[2] + 3;
    const int64_t nodes = (int64_t)b.ex * b.ey * b.ez;
    b.nodes = b.empty ? 0 : (nodes > cap ? INT_MAX : (int)nodes);
    return b;
}

// ------------------------------------------------------------------ P2G, sorted-segmented

// Node (i, j, l) of a particle's stencil, k = i + 3j + 9l, receives mass w m and momentum
// w (E + B n) with n = (i, j, l), B_ab = m C_ab + s [a = b], E_a = m v_a - B_a . off: the
// reference's (w m)(v + C d) + w s d (Flip3.java:225-243) with d = n - off, regrouped so the
// per-node work is one add and one fused multiply-add per field. The weights are the
// reference's products exactly; the regrouping moves each amount by a few ulps.
// `mm` is the mass the mass sums take and `m`, `s` those the momentum sums take: the particle's own
// for the float tiles, each times a power of two (exact) for the integer tiles, whose mass and
// momentum sums are counted in different units.
__device__ __forceinline__ void p2g_fold(const Axis (&ax)[3], float mm, float m, const float (&vel)[3],
                                         const float (&c)[9], float s, float (&acc)[108]) {
    float B[3][3], E[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) B[a][b] = a == b ? fmaf(m, c[3 * a + b], s) : m * c[3 * a + b];
        E[a] = fmaf(m, vel[a], -fmaf(B[a][0], ax[0].off, fmaf(B[a][1], ax[1].off, B[a][2] * ax[2].off)));
    }
    // n B_a for a node offset n of 0, 1 or 2 is nothing, B_a or B_a + B_a: written out, since
    // the compiler must keep a multiplication by 0.0 (it is -0 or NaN for some B).
    auto plus = [](float t, float b, int n) { return n == 0 ? t : (n == 1 ? t + b : t + (b + b)); };
    float Ej[3][3];  // E + j B_1
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int a = 0; a < 3; ++a) Ej[j][a] = plus(E[a], B[a][1], j);
#pragma unroll
    for (int l = 0; l < 3; ++l) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float wyz = ax[1].w[j] * ax[2].w[l];
            float T[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) T[a] = plus(Ej[j][a], B[a][2], l);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int k = i + 3 * j + 9 * l;
                const float wgt = ax[0].w[i] * wyz;
                acc[4 * k] = fmaf(wgt, mm, acc[4 * k]);
#pragma unroll
                for (int a = 0; a < 3; ++a) acc[4 * k + 1 + a] = fmaf(wgt, plus(T[a], B[a][0], i), acc[4 * k + 1 + a]);
            }
        }
    }
}

__device__ __forceinline__ void p2g_fold(const Axis (&ax)[3], float m, const float (&vel)[3], const float (&c)[9],
                                         float s, float (&acc)[108]) {
    p2g_fold(ax, m, m, vel, c, s, acc);
}

// Lanes of one wave exchange values through LDS below without atomics. The hardware keeps a
// wave's LDS operations in program order; this keeps the compiler from moving, merging or
// forwarding LDS accesses across the point (a lane must re-read what other lanes wrote).
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// The wave's box of base cells from each lane's lo/hi (empty: lo > hi); the same in every lane.
__device__ __forceinline__ Box wave_box(int (&lo)[3], int (&hi)[3], int cap) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        lo[i] = lanes_min(lo[i]);
        hi[i] = lanes_max(hi[i]);
    }
    Box b;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        b.lo[i] = lo[i];
        b.hi[i] = hi[i];
    }
    b.empty = b.lo[0] > b.hi[0];
    b.ex = b.empty ? 0 : b.hi[0] - b.lo[0] + 3;
    b.ey = b.empty ? 0 : b.hi[1] - b.lo[1] + 3;
    b.ez = b.empty ? 0 : b.hi[2] - b.lo[2] + 3;
    const int64_t nodes = (int64_t)b.ex * b.ey * b.ez;
    b.nodes = b.empty ? 0 : (nodes > cap ? INT_MAX : (int)nodes);
    return b;
}

constexpr int kWaveTile = 1024;  // nodes of a wave's LDS tile: 1024 x (16 B + 4 B claim) = 20 KB
constexpr int kMaxPasses = 16;   // bricks a wave takes through the tile before going direct

// The tile's nodes: lo .. lo + dims - 1 (a box of base cells and the two nodes past its top).
struct Slab {
    int lo[3], dims[3], nodes;
};

__device__ __forceinline__ Slab tile_of(const Box& box) {
    Slab sl;
    sl.lo[0] = box.lo[0];
    sl.lo[1] = box.lo[1];
    sl.lo[2] = box.lo[2];
    sl.dims[0] = box.ex;
    sl.dims[1] = box.ey;
    sl.dims[2] = box.ez;
    sl.nodes = box.ex * box.ey * box.ez;
    return sl;
}

// The 8x8x8-cell brick of base cell c, numbered as the sort key numbers bricks.
__device__ __forceinline__ int brick_of_cell(const int (&c)[3], int nx, int ny) {
    const int bx = (nx - 2 + 7) / 8, by = (ny - 2 + 7) / 8;
    return ((c[2] >> 3) * by + (c[1] >> 3)) * bx + (c[0] >> 3);
}

// The least of v[] above `after` (INT_MAX if none), in this lane; lanes_min takes it over the wave.
template <int R>
__device__ __forceinline__ int min_of(const int (&v)[R], int after) {
    int m = INT_MAX;
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (v[r] > after) m = min(m, v[r]);
    return m;
}

template <int R>
__device__ __forceinline__ int max_of(const int (&v)[R]) {
    int m = INT_MIN;
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (v[r] != INT_MAX) m = max(m, v[r]);
    return m;
}

__device__ __forceinline__ void clear_tile(float4* tile, const Slab& sl) {
    lds_order();  // whatever the wave read of the tile before is done with
    for (int t = threadIdx.x; t < sl.nodes; t += kWave) tile[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Called by the whole wave; `mine` says whether this lane has a run at tile node `run` to flush.
// Lanes flushing the same node take turns: each writes its lane id into claim[run] and the lane
// it reads back wins the round, so within a round no two lanes touch one node at one offset.
// When every winner's base lies in one z layer (the usual case for cell-sorted particles), the
// three z layers of their stencils cannot meet, so the three read-add-writes of one (dx, dy)
// go together and the round is a chain of 9 instead of 27.
__device__ __forceinline__ void flush_run(float4* tile, int* claim, const Slab& sl, int run, bool mine,
                                          float (&acc)[108]) {
    const int lane = threadIdx.x;
    const int b = mine ? run : 0;  // lanes with nothing to flush never index with their run
    const int ex = sl.dims[0], ey = sl.dims[1];
    const int layer = b / (ex * ey);
    bool pending = mine;
    while (__ballot(pending)) {
        if (pending) claim[b] = lane;
        lds_order();
        const bool won = pending && claim[b] == lane;
        // a round has a winner (the last lane to write each claimed word); do all share its layer?
        const uint64_t winners = __ballot(won);
        const int layer0 = __builtin_amdgcn_readlane(layer, (int)__builtin_ctzll(winners));
        if (__ballot(won && layer != layer0) == 0) {
            if (won) {
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    float4 t[3];
                    int node[3];
#pragma unroll
                    for (int dz = 0; dz < 3; ++dz) {
                        node[dz] = b + (dz * ey + j / 3) * ex + j % 3;
                        t[dz] = tile[node[dz]];
                    }
#pragma unroll
                    for (int dz = 0; dz < 3; ++dz) {
                        const int k = 9 * dz + j;
                        t[dz].x += acc[4 * k];
                        t[dz].y += acc[4 * k + 1];
                        t[dz].z += acc[4 * k + 2];
                        t[dz].w += acc[4 * k + 3];
                        tile[node[dz]] = t[dz];
                    }
                    lds_order();
                }
            }
        } else if (won) {
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const int node = b + ((k / 9) * ey + (k / 3) % 3) * ex + k % 3;
                float4 t = tile[node];
                t.x += acc[4 * k];
                t.y += acc[4 * k + 1];
                t.z += acc[4 * k + 2];
                t.w += acc[4 * k + 3];
                tile[node] = t;
                lds_order();
            }
        }
        pending = pending && !won;
    }
    if (mine) {
#pragma unroll
        for (int i = 0; i < 108; ++i) acc[i] = 0.0f;
    }
}

// Before a wave's last flush: where lanes 2k and 2k+1 hold runs with one base (a cell's
// particles split over the two lanes, the usual case for sorted input), lane 2k takes lane
// 2k+1's sums and lane 2k+1 has nothing left to flush, so the flush needs one claim round where
// it took two. Called by the whole wave (DPP reads every lane).
__device__ __forceinline__ void merge_pairs(int& run, float (&acc)[108]) {
    const int next = __builtin_amdgcn_mov_dpp(run, 0xF5, 0xF, 0xF, false);  // quad_perm [1,1,3,3]
    const int prev = __builtin_amdgcn_mov_dpp(run, 0xA0, 0xF, 0xF, false);  // quad_perm [0,0,2,2]
    const bool even = (threadIdx.x & 1) == 0;
    const bool take = even && run >= 0 && next == run;
    const bool given = !even && run >= 0 && prev == run;
#pragma unroll
    for (int i = 0; i < 108; ++i) {
        const float v = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(acc[i]), 0xF5, 0xF, 0xF, false));
        if (take) acc[i] += v;
        if (given) acc[i] = 0.0f;  // handed over: the sums start from zero again, as after a flush
    }
    if (given) run = -1;
}

// The bricks a lone particle's stencil reaches, marked by this lane alone.
__device__ __forceinline__ void mark_one(const Mpm3Args& a, const Axis (&ax)[3]) {
    const int lo[3] = {ax[0].cell, ax[1].cell, ax[2].cell};
    const int hi[3] = {lo[0] + 2, lo[1] + 2, lo[2] + 2};
    mark_bricks(a, lo, hi, 0, 1);
}

// One particle into this lane's run: its stencil, a flush when its base changes, its amounts.
// `in_pass`: the particle has mass and lies in the brick this pass takes.
__device__ __forceinline__ void take(const Mpm3Args& a, const float (&law)[3], float4* tile, int* claim,
                                     const Slab& sl, bool in_pass, float x, float y, float z, float m, float j,
                                     const float (&vel)[3], const float (&c)[9], int& run_base, float (&acc)[108]) {
    Axis ax[3] = {axis(x, a.nx), axis(y, a.ny), axis(z, a.nz)};
    const bool live = in_pass;
    const int base = ((ax[2].cell - sl.lo[2]) * sl.dims[1] + (ax[1].cell - sl.lo[1])) * sl.dims[0] +
                     (ax[0].cell - sl.lo[0]);
    const bool change = live && base != run_base;
    flush_run(tile, claim, sl, run_base, change && run_base >= 0, acc);
    if (change) run_base = base;
    if (live) {
        const float dt = law[0], bulk = law[1], rho0 = law[2];
        const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - j));
        p2g_fold(ax, m, vel, c, s, acc);
    }
}

// The tile for the base cells [lo, hi] of the wave's lanes (reduced over the wave), its bricks
// marked (the caller clears its nodes).
__device__ __forceinline__ Slab enter_tile(const Mpm3Args& a, int (&lo)[3], int (&hi)[3]) {
    const Box box = wave_box(lo, hi, kWaveTile);
    const int nlo[3] = {box.lo[0], box.lo[1], box.lo[2]};
    const int nhi[3] = {box.hi[0] + 2, box.hi[1] + 2, box.hi[2] + 2};
    mark_bricks(a, nlo, nhi, threadIdx.x, kWave);
    return tile_of(box);
}

// The tile's nodes into the grid: one global float atomic per touched node and field.
__device__ __forceinline__ void leave_tile(const Mpm3Args& a, const float4* tile, const Slab& sl) {
    lds_order();
    const int ex = sl.dims[0], ey = sl.dims[1];
    // a node's (x, y, z) in the tile from its number t by float reciprocals: t < 2^12, so t + 1/2
    // lies at least 1/(2 ex) from a multiple of ex and the rounded product floors exactly
    const float inv_ex = 1.0f / (float)ex, inv_ey = 1.0f / (float)ey;
    for (int t = threadIdx.x; t < sl.nodes; t += kWave) {
        const float4 v = tile[t];
        if (v.x == 0.0f) continue;  // a node no live particle weighs on holds nothing
        const int row = (int)(((float)t + 0.5f) * inv_ex);
        const int lx = t - row * ex;
        const int lz = (int)(((float)row + 0.5f) * inv_ey);
        const int ly = row - lz * ey;
        const int64_t node = ((int64_t)(sl.lo[2] + lz) * a.ny + (sl.lo[1] + ly)) * a.nx + (sl.lo[0] + lx);
        atomic_add(a.g[0] + node, v.x);
        atomic_add(a.g[1] + node, v.y);
        atomic_add(a.g[2] + node, v.z);
        atomic_add(a.g[3] + node, v.w);
    }
}

// atomic_particle as a loop over the 27 nodes, the same amounts in the same arithmetic, for a rare
// path: unrolled, its 108 node addresses would be computed up front, beside the caller's state.
__device__ __forceinline__ void atomic_particle_rolled(const Mpm3Args& a, const Axis (&ax)[3], float m,
                                                       const float (&vel)[3], const float (&c)[9], float s) {
    const int64_t base = ((int64_t)ax[2].cell * a.ny + ax[1].cell) * a.nx + ax[0].cell;
    const int64_t plane = (int64_t)a.nx * a.ny;
    auto pick = [](const Axis& x, int i) { return i == 0 ? x.w[0] : (i == 1 ? x.w[1] : x.w[2]); };
#pragma unroll 1
    for (int k = 0; k < 27; ++k) {
        const int i = k % 3, j = (k / 3) % 3, l = k / 9;
        const float wgt = pick(ax[0], i) * (pick(ax[1], j) * pick(ax[2], l));
        const float d[3] = {(float)i - ax[0].off, (float)j - ax[1].off, (float)l - ax[2].off};
        const float wm = wgt * m;
        const int64_t node = base + l * plane + j * a.nx + i;
        atomic_add(a.g[0] + node, wm);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const float at3 = vel[e] + (c[3 * e] * d[0] + (c[3 * e + 1] * d[1] + c[3 * e + 2] * d[2]));
            const float push = s * d[e];
            atomic_add(a.g[1 + e] + node, wm * at3 + wgt * push);
        }
    }
}

// The float tile: a float4 (m, mu, mv, mw) a node, runs added by read-add-write in claim rounds.
struct FloatTile {
    float4* tile;
    int* claim;
    __device__ void clear(const Slab& sl) const { clear_tile(tile, sl); }
    __device__ void take(const Mpm3Args& a, const float (&law)[3], const Slab& sl, bool in_pass, float x, float y,
                         float z, float m, float j, const float (&vel)[3], const float (&c)[9], int& run_base,
                         float (&acc)[108]) const {
        ::vxr::take(a, law, tile, claim, sl, in_pass, x, y, z, m, j, vel, c, run_base, acc);
    }
    __device__ void last(const Slab& sl, int& run_base, float (&acc)[108]) const {
        merge_pairs(run_base, acc);
        flush_run(tile, claim, sl, run_base, run_base >= 0, acc);
    }
    __device__ void leave(const Mpm3Args& a, const Slab& sl) const { leave_tile(a, tile, sl); }
};

// x rounded to the nearest integer, ties to even; |x| < 2^31. (Halves up in one instruction,
// floor(x + 1/2), would be biased: a run's float sum near 2^23 units has only a bit or two below
// the unit, so halves are common there, and all of them would round up.)
__device__ __forceinline__ int round_even(float x) { return (int)__builtin_rintf(x); }

// The integer tile: four planes of kWaveTile 32-bit sums, mass in units of q[0] and momentum in
// units of q[1], powers of two chosen per wave so that no node's sum can reach 2^31 (see the
// kernel). A run leaves as 108 rounded integers by LDS integer atomics, which need no claim
// rounds (lanes flushing one base meet at the same words in one instruction, and the atomics
// serialise them) and no read before the write; each run's amounts are rounded once, to the
// nearest unit.
struct IntTile {
    int* t;
    float q[2], inv[2];
    __device__ void clear(const Slab& sl) const {
        lds_order();
        for (int n = threadIdx.x; n < sl.nodes; n += kWave)
#pragma unroll
            for (int f = 0; f < 4; ++f) t[f * kWaveTile + n] = 0;
    }
    __device__ void flush(const Slab& sl, int run, bool mine, float (&acc)[108]) const {
        if (mine) {
            const int ex = sl.dims[0], ey = sl.dims[1];
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const int node = run + ((k / 9) * ey + (k / 3) % 3) * ex + k % 3;
#pragma unroll
                for (int f = 0; f < 4; ++f) atomicAdd(t + f * kWaveTile + node, round_even(acc[4 * k + f]));
                // a node at a time: scheduled freely, the 108 conversions would all be held at once
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int i = 0; i < 108; ++i) acc[i] = 0.0f;
        }
    }
    __device__ void take(const Mpm3Args& a, const float (&law)[3], const Slab& sl, bool in_pass, float x, float y,
                         float z, float m, float j, const float (&vel)[3], const float (&c)[9], int& run_base,
                         float (&acc)[108]) const {
        Axis ax[3] = {axis(x, a.nx), axis(y, a.ny), axis(z, a.nz)};
        const int base = ((ax[2].cell - sl.lo[2]) * sl.dims[1] + (ax[1].cell - sl.lo[1])) * sl.dims[0] +
                         (ax[0].cell - sl.lo[0]);
        const bool change = in_pass && base != run_base;
        flush(sl, run_base, change && run_base >= 0, acc);
        if (change) run_base = base;
        if (in_pass) {
            const float s = (4.0f * (law[0] * (m / law[2]))) * (law[1] * (1.0f - j));
            p2g_fold(ax, m * inv[0], m * inv[1], vel, c, s * inv[1], acc);
        }
    }
    __device__ void last(const Slab& sl, int& run_base, float (&acc)[108]) const {
        flush(sl, run_base, run_base >= 0, acc);
    }
    // The tile's nodes into the grid: (float)sum * unit, one global float atomic per nonzero node
    // and field. A wave's LDS operations run in order, so the reads see every atomic before them.
    __device__ void leave(const Mpm3Args& a, const Slab& sl) const {
        lds_order();
        const int ex = sl.dims[0], ey = sl.dims[1];
        const float inv_ex = 1.0f / (float)ex, inv_ey = 1.0f / (float)ey;
        for (int n = threadIdx.x; n < sl.nodes; n += kWave) {
            const int im = t[n];
            if (im == 0) continue;  // no mass: no live particle weighs on the node (or only by < half a unit)
            const float v[4] = {(float)im * q[0], (float)t[kWaveTile + n] * q[1], (float)t[2 * kWaveTile + n] * q[1],
                                (float)t[3 * kWaveTile + n] * q[1]};
            const int row = (int)(((float)n + 0.5f) * inv_ex);
            const int lx = n - row * ex;
            const int lz = (int)(((float)row + 0.5f) * inv_ey);
            const int ly = row - lz * ey;
            const int64_t node = ((int64_t)(sl.lo[2] + lz) * a.ny + (sl.lo[1] + ly)) * a.nx + (sl.lo[0] + lx);
#pragma unroll
            for (int f = 0; f < 4; ++f) atomic_add(a.g[f] + node, v[f]);
        }
    }
};

// The sum over a wave's 64 lanes, in every lane (all lanes active): DPP within rows of 16 (swap
// neighbours, swap pairs, mirror within 8, mirror within 16), then the four rows' sums.
__device__ __forceinline__ float wave_sum(float v) {
    v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
    v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
    v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));
    v = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false));
    auto at = [&](int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
    return (at(0) + at(16)) + (at(32) + at(48));
}

// The unit of an integer tile's sums for a wave whose amounts are bounded by `total` (0.4219 total
// bounds any node's sum): 2^(e - 30) for e = floor(log2 total), so a node's sum stays below 2^30
// units; false when total is not finite.
__device__ __forceinline__ bool tile_unit(float total, float& q, float& inv) {
    int e = ((__float_as_int(total) >> 23) & 0xff) - 127;
    if (e == 128) return false;
    e = min(max(e, -96), 126);
    q = __int_as_float((127 + e - 30) << 23);
    inv = __int_as_float((127 + 30 - e) << 23);
    return true;
}

// A wave takes a chunk of 64 * R particles, lane l the R consecutive ones from l * R, each field
// read with one R-wide load so the wave's loads are whole cache lines, all in flight at once.
// Runs of particles sharing a stencil base fold into 108 register sums. A run leaves through the
// wave's own node tile in LDS, then the tile leaves with one global float atomic per touched node
// and field. Two tiles (kInts):
// - the float tile, a float4 (m, mu, mv, mw) per node, by plain read-modify-write (LDS float
//   atomics run at a small fraction of LDS bandwidth on gfx950): lanes flushing the same base take
//   turns -- each writes its lane id into claim[base], the lane it reads back wins this round -- so
//   within a round no two lanes touch one node at one stencil offset, and the rounds, the offsets
//   and the wave's LDS operations are in program order;
// - the integer tile (IntTile): each run's sums rounded to the wave's units and added by LDS
//   integer atomics, with no rounds and no chains of dependent LDS reads. The units are powers of
//   two from bounds on the wave's amounts: a particle's mass amounts are w m and its momentum
//   amounts w (m v_a + B_a . d) with |d_b| <= 3/2, w <= 0.75^3 = 0.4219, so 0.4219 times the sum over
//   the wave of m, and of U = max_a (m |v_a| + 3/2 sum_b |B_ab|), bound any node's sums.
template <int R, bool kInts>
__global__ __launch_bounds__(kWave) void p2g3_wave_kernel(Mpm3Args a) {
    __shared__ __align__(16) unsigned char lds[kInts ? 16 * kWaveTile : 20 * kWaveTile];
    const int lane = threadIdx.x;
    if ((int64_t)blockIdx.x * kWave * R >= *a.active) return;  // a wave of inactive slots
    // dt, bulk and rho0, read before the kernel stores anything, so once and into scalar registers
    const float law[3] = {bits_f(a.params[DT]), bits_f(a.params[BULK]), bits_f(a.params[RHO0])};
    const int64_t first = ((int64_t)blockIdx.x * kWave + lane) * R;  // first + R <= capacity
    float f[17][R];
#pragma unroll
    for (int k = 0; k < 17; ++k) load_run<R>(a.p[k], first, f[k]);
    const int size[3] = {a.nx, a.ny, a.nz};
    // each particle's brick (the sort's 8x8x8 blocks of base cells); INT_MAX for a massless one
    int bk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int c[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) c[i] = base_cell(f[i][r], size[i]);
        bk[r] = f[6][r] > 0.0f ? brick_of_cell(c, a.nx, a.ny) : INT_MAX;
    }
    typename std::conditional<kInts, IntTile, FloatTile>::type tile;
    if constexpr (kInts) {
        tile.t = reinterpret_cast<int*>(lds);
        float sm = 0.0f, su = 0.0f;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (bk[r] == INT_MAX) continue;
            const float m = f[6][r];
            const float s = (4.0f * (law[0] * (m / law[2]))) * (law[1] * (1.0f - f[7][r]));
            // poison: the plain sum of the bounds, non-finite with any of them (fmaxf would drop a NaN)
            float u = 0.0f, poison = 0.0f;
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                float row = 0.0f;
#pragma unroll
                for (int b = 0; b < 3; ++b) row += fabsf(e == b ? fmaf(m, f[8 + 3 * e + b][r], s) : m * f[8 + 3 * e + b][r]);
                const float t = fmaf(m, fabsf(f[3 + e][r]), 1.5f * row);
                u = fmaxf(u, t);
                poison += t;
            }
            sm += m;
            su += fmaf(0.0f, poison, u);
        }
        // rounding in these sums is far inside the factor 0.4219 x 2 of headroom the units leave
        if (!tile_unit(wave_sum(sm), tile.q[0], tile.inv[0]) | !tile_unit(wave_sum(su), tile.q[1], tile.inv[1])) {
            // a non-finite amount: the wave's particles by the reference's direct atomics, so it
            // reaches the grid as the reference's schedule would carry it
#pragma unroll 1
            for (int r = 0; r < R; ++r) {
                if (bk[r] == INT_MAX) continue;
                Particle p;
                load_particle(a, first + r, law[0], law[1], law[2], p);
                mark_one(a, p.ax);
                atomic_particle_rolled(a, p.ax, p.m, p.vel, p.c, p.s);
            }
            return;
        }
    } else {
        tile.tile = reinterpret_cast<float4*>(lds);
        tile.claim = reinterpret_cast<int*>(lds + 16 * kWaveTile);
    }
    const int first_brick = lanes_min(min_of(bk, INT_MIN));
    if (first_brick == INT_MAX) return;
    int cur = first_brick;
    {
        float acc[108];
#pragma unroll
        for (int i = 0; i < 108; ++i) acc[i] = 0.0f;
        if (lanes_max(max_of(bk)) == first_brick) {
            // the common case: one brick, the particles from the registers they were loaded into
            int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (bk[r] == INT_MAX) continue;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const int c = base_cell(f[i][r], size[i]);
                    lo[i] = min(lo[i], c);
                    hi[i] = max(hi[i], c);
                }
            }
            const Slab sl = enter_tile(a, lo, hi);
            tile.clear(sl);
            int run_base = -1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float vel[3] = {f[3][r], f[4][r], f[5][r]};
                float c[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) c[i] = f[8 + i][r];
                tile.take(a, law, sl, bk[r] != INT_MAX, f[0][r], f[1][r], f[2][r], f[6][r], f[7][r], vel, c, run_base,
                          acc);
            }
            tile.last(sl, run_base, acc);
            tile.leave(a, sl);
            return;
        }
        // Several bricks (a wave straddling a brick boundary, or unsorted input): one pass through
        // the tile per brick, lowest first, each re-reading its particles (from cache); the stencils
        // of one brick's particles cover at most 10 x 10 x 10 nodes, so they always fit. Past
        // kMaxPasses bricks, the rest go by the reference's direct atomics.
        for (int pass = 0; pass < kMaxPasses && cur != INT_MAX; ++pass) {
            int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (bk[r] != cur) continue;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const int c = base_cell(a.p[i][first + r], size[i]);
                    lo[i] = min(lo[i], c);
                    hi[i] = max(hi[i], c);
                }
            }
            const Slab sl = enter_tile(a, lo, hi);
            tile.clear(sl);
            int run_base = -1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t q = first + r;
                const float vel[3] = {a.p[3][q], a.p[4][q], a.p[5][q]};
                float c[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) c[i] = a.p[8 + i][q];
                tile.take(a, law, sl, bk[r] == cur, a.p[0][q], a.p[1][q], a.p[2][q], a.p[6][q], a.p[7][q], vel, c,
                          run_base, acc);
            }
            tile.last(sl, run_base, acc);
            tile.leave(a, sl);
            cur = lanes_min(min_of(bk, cur));
        }
        if (cur == INT_MAX) return;
    }
    const float dt = law[0], bulk = law[1], rho0 = law[2];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t q = first + r;
        const int b = bk[r];
        if (b < cur || b == INT_MAX) continue;
        const float m = a.p[6][q];
        Axis ax[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) ax[i] = axis(a.p[i][q], size[i]);
        const float vel[3] = {a.p[3][q], a.p[4][q], a.p[5][q]};
        float c[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) c[i] = a.p[8 + i][q];
        const float sv = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - a.p[7][q]));
        mark_one(a, ax);
        atomic_particle(a, ax, m, vel, c, sv);
    }
}

// ------------------------------------------------------------------ P2G in fixed point
//
// Schedule 2: each of a particle's 108 amounts, computed as the reference computes it
// (Flip3.java:216-243, the arithmetic of atomic_particle), is rounded to a 64-bit integer -- mass
// in units of 2^-40, momentum in 2^-32 -- and the amounts are summed as integers, which
// associate exactly: the sums are the same bits in any order and on any launch shape, equal to
// the CPU restatement's (orc_mpm3_scatter_fixed), and so is everything the grid update and G2P
// compute from them. core/particle/FixedScatter.java:24-31 gives the reason for integers in 2D;
// these are its 3D counterpart on the MLS-MPM amounts. A wave takes 64 * R consecutive particles
// (a lane R of them in a row); a lane folds each run of its particles that share a stencil base
// into integer sums and adds them into an LDS tile of the box the wave's stencils cover, by 64-bit
// LDS atomics (two of the four fields a sweep): within one instruction two lanes meet at a node
// only where a cell's particles span both. The tile then leaves with one 64-bit global atomic per
// nonzero node and field. A wave whose box exceeds the tile (unsorted particles) adds straight into the
// grid. The pass ends by decoding the sums into gm gmu gmv gmw over the touched or live bricks,
// so the passes after it read floats as before.
constexpr int kFixedMassBits = 40;
constexpr int kFixedMomentumBits = 32;
constexpr int kFixedTile = 512;  // nodes: 4 fields x 8 B x 512 = 16 KB of LDS a wave
constexpr int kFixedRun = 4;

__device__ __forceinline__ float dequantize(unsigned long long q, int bits) {
    return (float)ldexp((double)(long long)q, -bits);
}

// An amount in units of 2^-bits, rounded to the nearest integer (ties to even): the scaling is
// exact, so this is the one rounding, the same as the CPU's llrintf(ldexpf(a, bits)). The integer
// is held in a double, and a run's sums of up to kFixedRun amounts stay exact there while each
// amount is below 2^11 in mass or 2^19 in momentum (a particle of mass 0.125 moving at 10^5 node
// spacings a second is far inside both); they convert to 64 bits once, at the flush.
__device__ __forceinline__ double quantize_d(float a, int bits) { return (double)rintf(ldexpf(a, bits)); }

// Where a fixed-point sweep adds: the wave's LDS tile, or (when its box exceeds the tile) the grid.
struct FixedTarget {
    bool tiled;
    Slab sl;
    int64_t row, lay;
};

// One sweep of a lane's R particles over fields F0 and F0 + 1 ((m, mu) or (mv, mw)): each run of
// particles sharing a stencil base folds into 27 x 2 integer sums, added into the tile (or the
// grid) when the base changes and at the end. The amounts are atomic_particle's (Flip3.java:225-243).
template <int R, int F0>
__device__ __forceinline__ void fixed_sweep(const Mpm3Args& a, const Fixed3& fx, unsigned long long (*tile)[kFixedTile],
                                            const FixedTarget& to, int64_t first, float dt, float bulk, float rho0) {
    const int size[3] = {a.nx, a.ny, a.nz};
    double acc[54];
#pragma unroll
    for (int i = 0; i < 54; ++i) acc[i] = 0.0;
    int64_t run = -1;
#pragma unroll 1
    for (int r = 0; r <= R; ++r) {
        // r == R: a pass that only flushes the last run
        const int64_t q = first + (r < R ? r : 0);
        const float m = r < R ? a.p[6][q] : 0.0f;
        const bool live = m > 0.0f;
        Axis ax[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) ax[i] = axis(a.p[i][q], size[i]);
        const int64_t base = to.tiled ? ((int64_t)(ax[2].cell - to.sl.lo[2]) * to.sl.dims[1] +
                                         (ax[1].cell - to.sl.lo[1])) * to.sl.dims[0] +
                                            (ax[0].cell - to.sl.lo[0])
                                      : ((int64_t)ax[2].cell * a.ny + ax[1].cell) * a.nx + ax[0].cell;
        if (run >= 0 && (r == R || (live && base != run))) {
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const int64_t node = run + (k / 9) * to.lay + ((k / 3) % 3) * to.row + k % 3;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const unsigned long long v = (unsigned long long)(long long)acc[2 * k + j];
                    if (to.tiled)
                        atomicAdd(&tile[F0 + j][node], v);
                    else
                        atomicAdd(fx.q[F0 + j] + node, v);
                    acc[2 * k + j] = 0.0;
                }
            }
            run = -1;
        }
        if (!live) continue;
        if (!to.tiled && F0 == 0) mark_one(a, ax);
        run = base;
        const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - a.p[7][q]));
        // field F0 + j's velocity and row of C: u and c0x for mu, v and c1x for mv, w and c2x for mw
        constexpr int kA0 = F0 == 0 ? -1 : 1, kA1 = F0 == 0 ? 0 : 2;
        float vel[2] = {0.0f, a.p[3 + kA1][q]}, c[2][3] = {};
        if (kA0 >= 0) vel[0] = a.p[3 + kA0][q];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (kA0 >= 0) c[0][e] = a.p[8 + 3 * kA0 + e][q];
            c[1][e] = a.p[8 + 3 * kA1 + e][q];
        }
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const float wgt = ax[0].w[k % 3] * (ax[1].w[(k / 3) % 3] * ax[2].w[k / 9]);
            const float d[3] = {(float)(k % 3) - ax[0].off, (float)((k / 3) % 3) - ax[1].off,
                                (float)(k / 9) - ax[2].off};
            const float wm = wgt * m;
            if (kA0 < 0) {
                acc[2 * k] += quantize_d(wm, kFixedMassBits);
            } else {
                const float at3 = vel[0] + (c[0][0] * d[0] + (c[0][1] * d[1] + c[0][2] * d[2]));
                acc[2 * k] += quantize_d(wm * at3 + wgt * (s * d[kA0]), kFixedMomentumBits);
            }
            const float at3 = vel[1] + (c[1][0] * d[0] + (c[1][1] * d[1] + c[1][2] * d[2]));
            acc[2 * k + 1] += quantize_d(wm * at3 + wgt * (s * d[kA1]), kFixedMomentumBits);
        }
    }
}

template <int R>
__global__ __launch_bounds__(kWave) void p2g3_fixed_kernel(Mpm3Args a, Fixed3 fx) {
    __shared__ unsigned long long tile[4][kFixedTile];
    const int lane = threadIdx.x;
    if ((int64_t)blockIdx.x * kWave * R >= *a.active) return;  // a wave of inactive slots
    const float dt = bits_f(a.params[DT]), bulk = bits_f(a.params[BULK]), rho0 = bits_f(a.params[RHO0]);
    const int64_t first = ((int64_t)blockIdx.x * kWave + lane) * R;  // first + R <= capacity
    const int size[3] = {a.nx, a.ny, a.nz};
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t q = first + r;
        if (!(a.p[6][q] > 0.0f)) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = base_cell(a.p[i][q], size[i]);
            lo[i] = min(lo[i], c);
            hi[i] = max(hi[i], c);
        }
    }
    const Box box = wave_box(lo, hi, kFixedTile);
    if (box.empty) return;  // uniform: every lane holds the same box
    FixedTarget to{};
    to.tiled = box.nodes != INT_MAX;
    if (to.tiled) {
        const int nlo[3] = {box.lo[0], box.lo[1], box.lo[2]};
        const int nhi[3] = {box.hi[0] + 2, box.hi[1] + 2, box.hi[2] + 2};
        mark_bricks(a, nlo, nhi, lane, kWave);
        to.sl = tile_of(box);
        for (int t = lane; t < to.sl.nodes; t += kWave)
#pragma unroll
            for (int f = 0; f < 4; ++f) tile[f][t] = 0ull;
        __syncthreads();  // the block is this one wave
    }
    to.row = to.tiled ? to.sl.dims[0] : a.nx;
    to.lay = to.tiled ? (int64_t)to.sl.dims[0] * to.sl.dims[1] : (int64_t)a.nx * a.ny;
    fixed_sweep<R, 0>(a, fx, tile, to, first, dt, bulk, rho0);
    fixed_sweep<R, 2>(a, fx, tile, to, first, dt, bulk, rho0);
    if (!to.tiled) return;
    __syncthreads();
    const Slab& sl = to.sl;
    const int ex = sl.dims[0], ey = sl.dims[1];
    for (int t = lane; t < sl.nodes; t += kWave) {
        const int lx = t % ex, ly = (t / ex) % ey, lz = t / (ex * ey);
        const int64_t node = ((int64_t)(sl.lo[2] + lz) * a.ny + (sl.lo[1] + ly)) * a.nx + (sl.lo[0] + lx);
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const unsigned long long v = tile[f][t];
            if (v != 0ull) atomicAdd(fx.q[f] + node, v);
        }
    }
}

// Over the span of touched or live bricks of each row, as the clear: the fixed-point sums zeroed
// (`decode` false), or decoded into gm gmu gmv gmw (`decode` true).
__global__ __launch_bounds__(256) void fixed_rows_kernel(Mpm3Args a, Fixed3 fx, int decode) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= (int64_t)a.ny * a.nz) return;
    const int y = (int)(row % a.ny), z = (int)(row / a.ny);
    int x0, x1;
    if (!row_range(a, y, z, lane, &x0, &x1)) return;
    const int64_t first = row * a.nx;
    for (int x = x0 + lane; x < x1; x += kWave) {
        const int64_t i = first + x;
        if (decode) {
            a.g[0][i] = dequantize(fx.q[0][i], kFixedMassBits);
#pragma unroll
            for (int f = 1; f < 4; ++f) a.g[f][i] = dequantize(fx.q[f][i], kFixedMomentumBits);
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f) fx.q[f][i] = 0ull;
        }
    }
}

// ------------------------------------------------------------------ P2G by bricks
//
// P2G schedule 3. The unit is a group: the 256 cells (8 x 8 by 4 layers, half a sort brick) whose
// keys share key >> 8. After the radix sort, each non-empty group's cells are ranked by their
// particle count, most first (ties by cell number), and dealt to lanes in generations of 64:
// generation g holds the cells ranked 64 g .. 64 g + 63, lane l the one ranked 64 g + l. The
// group's particles are then laid out generation by generation, and within a generation
// rank-major: first the first particle of each of its cells that has one, in lane order, then
// every second one, and so on (a cell's particles keep their sorted order). A wave takes a group,
// generation by generation: lane l owns one cell, and in round r every lane whose cell has an
// r-th particle reads it -- consecutive words. Cells of one generation hold similar counts, so
// few lanes idle (76% busy on the two-cubes scene, against 47% for a wave per 64-cell layer). A
// lane folds its cell's particles into 108 register sums; at the end of a generation the 64
// lanes add them into the wave's LDS node tile of the group (10 x 10 x 6 nodes) in 27 steps, one
// stencil offset a step: the lanes' cells differ, so no two lanes touch one node in a step, and
// the steps run in program order. The tile leaves with one global float atomic per touched node
// and field, about 1.5 a particle. A particle whose base cell has changed since the sort (it
// moved) goes by the reference's direct atomics. Per-particle LDS float atomics (ds_add_f32) are
// no option: on gfx950 they run at a few percent of the rate of integer LDS atomics.

constexpr int kGroupShift = 8;                 // half a sort brick: keys key >> 8
constexpr int kGroupCells = 1 << kGroupShift;  // 256: 8 x 8 cells by 4 layers
constexpr int kGroupGens = kGroupCells / kWave;
constexpr int kGroupLayers = kGroupCells / 64;
constexpr int kTileSide = 10;                  // the group's stencil nodes in x and y
constexpr int kTileZ = kGroupLayers + 2;       // and in z
constexpr int kTileNodes = kTileSide * kTileSide * kTileZ;
constexpr int kMovedQueue = 256;               // moved particles a wave queues per group
constexpr int kGroupPer = 4096;                // sorted keys a block of the boundary passes takes
constexpr int kGroupWaves = 4096;              // waves of a brick launch, each taking bricks until none is left
constexpr int kOrderRounds = 256;              // rounds a brick's order tables hold

// The next work item of this wave from a ticket counter t[0] (items 0 .. items - 1), or -1 once
// they are gone. t[1] counts the waves that have run out; the last of the launch's `waves` to do
// so puts both words back to zero, so the next launch on the stream starts from zero again.
__device__ __forceinline__ int take_ticket(int32_t* t, int waves, int items) {
    int got = 0;
    if ((threadIdx.x & (kWave - 1)) == 0) got = atomicAdd(t, 1);
    got = __shfl(got, 0);
    if (got < items) return got;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (atomicAdd(t + 1, 1) == waves - 1) {
            atomicExch(t, 0);
            atomicExch(t + 1, 0);
        }
    }
    return -1;
}

// Counts the brick heads (the first sorted particle of each non-empty brick) of a block's 4096
// sorted keys; an inactive particle keeps its sorted place (and key) in the new order.
__global__ __launch_bounds__(256) void group_count_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                          int32_t inactive, const int32_t* __restrict__ index,
                                                          int32_t* __restrict__ new_index, int32_t* __restrict__ new_keys,
                                                          int32_t* __restrict__ heads) {
    __shared__ int part[256 / kWave];
    const int64_t begin = (int64_t)blockIdx.x * kGroupPer;
    const int64_t end = min(begin + (int64_t)kGroupPer, n);
    int c = 0;
    for (int64_t p = begin + threadIdx.x; p < end; p += 256) {
        const int32_t k = keys[p];
        if (k >= inactive) {
            new_index[p] = index[p];
            new_keys[p] = k;
        } else if (p == 0 || (keys[p - 1] >> kGroupShift) != (k >> kGroupShift)) {
            ++c;
        }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) heads[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of the blocks' head counts in place (one block); meta[0] = the non-empty bricks,
// and meta[1] = 0 until the list pass finds the last active particle.
__global__ __launch_bounds__(256) void group_scan_kernel(int32_t* __restrict__ heads, int blocks,
                                                         int32_t* __restrict__ meta, int64_t cap) {
    __shared__ int part[256];
    const int per = (blocks + 255) / 256;
    const int lo = threadIdx.x * per, hi = min(lo + per, blocks);
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += heads[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - sum;
    for (int i = lo; i < hi; ++i) {
        const int v = heads[i];
        heads[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) {
        meta[0] = (int32_t)min((int64_t)part[255], cap);  // never more than the tables hold
        meta[1] = 0;
    }
}

// The list of non-empty bricks in key order: each head writes its brick and first particle at its
// block's offset plus its rank among the block's heads; the last active particle writes the end
// of the active ones (meta[1]).
__global__ __launch_bounds__(256) void group_list_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                         int32_t inactive, const int32_t* __restrict__ offsets,
                                                         int32_t* __restrict__ group, int32_t* __restrict__ start,
                                                         int32_t* __restrict__ meta, int64_t cap) {
    __shared__ int wsum[256 / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int64_t begin = (int64_t)blockIdx.x * kGroupPer;
    int base = offsets[blockIdx.x];
    for (int r = 0; r < kGroupPer / 256; ++r) {
        const int64_t p = begin + r * 256 + threadIdx.x;
        bool head = false;
        int32_t k = 0;
        if (p < n) {
            k = keys[p];
            head = k < inactive && (p == 0 || (keys[p - 1] >> kGroupShift) != (k >> kGroupShift));
            if (k < inactive && (p + 1 == n || keys[p + 1] >= inactive)) meta[1] = (int32_t)(p + 1);
        }
        const uint64_t b = __ballot(head);
        if (lane == 0) wsum[wave] = __popcll(b);
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (head) {
            const int i = base + before + __popcll(b & ((1ull << lane) - 1ull));
            if (i < cap) {
                group[i] = k >> kGroupShift;
                start[i] = (int32_t)p;
            }
        }
        base += total;
        __syncthreads();
    }
}

// A wave a brick (see the section head): the cells' counts (LDS integer atomics over the brick's
// sorted keys), their ranks by count (a stable counting sort on min(count, 255), most first, in
// eight passes of 64 cells matched by ballots), then the new order -- as the permutation's source
// index, and the keys in that order. Each round's offset and mask of lanes go to LDS tables, so
// every particle then finds its place on its own, with coalesced reads: particle p of cell c, of
// rank r in it, goes to S + off[t] + (the lanes below c's in mask[t]), t the round (generation of
// c, r). A brick of more rounds than the tables hold goes round by round.
__global__ __launch_bounds__(256) void group_order_kernel(const int32_t* __restrict__ keys,
                                                          const int32_t* __restrict__ index,
                                                          int32_t* __restrict__ new_index,
                                                          int32_t* __restrict__ new_keys, Layers L, int waves) {
    struct Lds {
        int cnt[kGroupCells];    // particles of each cell
        int first[kGroupCells];  // each cell's first particle
        int rank[kGroupCells];   // each cell's place in count order
        int at[kGroupCells];     // the cell at each place
        int vbase[256];          // per clamped count: cells with more, then the running count of equal ones
        int toff[kOrderRounds];
        uint64_t tmask[kOrderRounds];
        int gbase[kGroupGens + 1];
    };
    __shared__ Lds lds_all[256 / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    Lds& s = lds_all[wave];
    const uint64_t below = (1ull << lane) - 1ull;
    const int nl = L.meta[0];
    for (;;) {
        const int i = take_ticket(L.meta + 4, waves, nl);
        if (i < 0) return;
        const int S = L.start[i];
        const int E = i + 1 < nl ? L.start[i + 1] : L.meta[1];
        const int group = L.layer[i];
        lds_order();  // the previous brick's reads of the tables are behind us
        for (int c = lane; c < kGroupCells; c += kWave) s.cnt[c] = 0;
        for (int v = lane; v < 256; v += kWave) s.vbase[v] = 0;
        lds_order();
        for (int p = S + lane; p < E; p += kWave) atomicAdd(&s.cnt[keys[p] & (kGroupCells - 1)], 1);
        lds_order();
        // first particle of each cell: lane takes cells 8 lane .. 8 lane + 7
        {
            int n8[kGroupGens], sum = 0;
#pragma unroll
            for (int j = 0; j < kGroupGens; ++j) {
                n8[j] = s.cnt[kGroupGens * lane + j];
                sum += n8[j];
            }
            int incl = sum;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const int v = __shfl_up(incl, off);
                if (lane >= off) incl += v;
            }
            int run = S + incl - sum;
#pragma unroll
            for (int j = 0; j < kGroupGens; ++j) {
                s.first[kGroupGens * lane + j] = run;
                L.count[(int64_t)i * kGroupCells + kGroupGens * lane + j] = n8[j];
                run += n8[j];
                atomicAdd(&s.vbase[min(n8[j], 255)], 1);
            }
        }
        lds_order();
        // vbase[v] := cells whose clamped count is above v (lane takes v = 4 lane .. 4 lane + 3)
        {
            int h4[4], sum = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                h4[j] = s.vbase[4 * lane + j];
                sum += h4[j];
            }
            int incl = sum;  // suffix sums: lanes above this one
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const int v = __shfl_down(incl, off);
                if (lane + off < kWave) incl += v;
            }
            int above = incl - sum;
            lds_order();
#pragma unroll
            for (int j = 3; j >= 0; --j) {
                s.vbase[4 * lane + j] = above;
                above += h4[j];
            }
        }
        lds_order();
        // ranks: cells in number order, 64 at a time; equal clamped counts keep that order
        for (int q = 0; q < kGroupGens; ++q) {
            const int c = kWave * q + lane;
            const int v = min(s.cnt[c], 255);
            uint64_t same = ~0ull;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint64_t ones = __ballot((v >> b) & 1);
                same &= ((v >> b) & 1) ? ones : ~ones;
            }
            const int within = __popcll(same & below);
            const int rank = s.vbase[v] + within;
            lds_order();  // every lane has read its count's base before the leaders move it on
            if (within == 0) s.vbase[v] += __popcll(same);
            s.rank[c] = rank;
            s.at[rank] = c;
            lds_order();
        }
        // the rounds: generation by generation, rank by rank
        int t = 0, off = 0;
        int gens = 0;
        for (int g = 0; g < kGroupGens; ++g) {
            const int n = s.cnt[s.at[kWave * g + lane]];
            const int maxn = wave_max(n);
            if (maxn == 0) break;
            ++gens;
            if (lane == 0) s.gbase[g] = t;
            for (int r = 0; r < maxn; ++r, ++t) {
                const uint64_t b = __ballot(n > r);
                if (lane == 0 && t < kOrderRounds) {
                    s.toff[t] = off;
                    s.tmask[t] = b;
                }
                off += __popcll(b);
            }
        }
        lds_order();
        if (t <= kOrderRounds) {
            for (int p = S + lane; p < E; p += kWave) {
                const int k = keys[p];
                const int c = k & (kGroupCells - 1);
                const int rank = s.rank[c];
                const int tt = s.gbase[rank >> 6] + (p - s.first[c]);
                const int dest = S + s.toff[tt] + __popcll(s.tmask[tt] & ((1ull << (rank & 63)) - 1ull));
                new_index[dest] = index[p];
                new_keys[dest] = k;
            }
        } else {
            int pos = S;
            for (int g = 0; g < gens; ++g) {
                const int c = s.at[kWave * g + lane];
                const int n = s.cnt[c];
                const int maxn = wave_max(n);
                for (int r = 0; r < maxn; ++r) {
                    const uint64_t b = __ballot(n > r);
                    if (n > r) {
                        const int dest = pos + __popcll(b & below);
                        new_index[dest] = index[s.first[c] + r];
                        new_keys[dest] = (group << kGroupShift) | c;
                    }
                    pos += __popcll(b);
                }
            }
        }
        for (int k = lane; k < kGroupCells; k += kWave)
            L.cell[(int64_t)i * kGroupCells + k] = (uint16_t)s.at[k];
    }
}

// P2G over the groups of the last sort (see the section head). Each round's fields load two
// rounds ahead of its fold. A particle that moved out of its cell is queued in LDS and goes by
// direct atomics once the group's sums are spent (all of the group's particles are looked at
// again if the queue overflows).
__global__ __launch_bounds__(256) void p2g3_layer_kernel(Mpm3Args a, Layers L, int waves) {
    __shared__ float4 tiles[256 / kWave][kTileNodes];
    __shared__ int queues[256 / kWave][kMovedQueue];
    float4* tile = tiles[threadIdx.x >> 6];
    int* queue = queues[threadIdx.x >> 6];
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t below = (1ull << lane) - 1ull;
    const float dt = bits_f(a.params[DT]), bulk = bits_f(a.params[BULK]), rho0 = bits_f(a.params[RHO0]);
    const int nl = L.meta[0];
    const int bxn = (a.nx - 2 + 7) / 8, byn = (a.ny - 2 + 7) / 8;
    constexpr int kPerBrick = 512 / kGroupCells;
    const int64_t plane = (int64_t)a.nx * a.ny;
    for (;;) {
        const int i = take_ticket(L.meta + 2, waves, nl);
        if (i < 0) return;
        const uint32_t S = (uint32_t)L.start[i];
        const int group = L.layer[i];
        const int brick = group / kPerBrick;
        const int x0 = (brick % bxn) * 8, y0 = ((brick / bxn) % byn) * 8;
        const int z0 = (brick / (bxn * byn)) * 8 + (group % kPerBrick) * kGroupLayers;
        lds_order();  // the previous group's tile and queue reads are behind us
        for (int t = lane; t < kTileNodes; t += kWave) tile[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t pos = S;  // the first particle of the next round to issue
        int moved = 0;     // the group's moved particles so far, the same in every lane
        for (int g = 0; g < kGroupGens; ++g) {
            const int c = L.cell[(int64_t)i * kGroupCells + kWave * g + lane];
            const int nc = L.count[(int64_t)i * kGroupCells + c];
            const int maxn = wave_max(nc);
            if (maxn == 0) break;
            const int hx = c & 7, hy = (c >> 3) & 7, hz = c >> 6;
            auto issue = [&](int r, float (&dst)[17]) {
                const uint64_t b = __ballot(nc > r);
                const uint32_t q = pos + (uint32_t)__popcll(b & below);
                pos += (uint32_t)__popcll(b);
                if (nc > r) {
#pragma unroll
                    for (int k = 0; k < 17; ++k) dst[k] = a.p[k][q];
                }
                return q;
            };
            float n1[17], n2[17];
            uint32_t q1 = issue(0, n1), q2 = 0;
            if (maxn > 1) q2 = issue(1, n2);
            float acc[108];
#pragma unroll
            for (int k = 0; k < 108; ++k) acc[k] = 0.0f;
            for (int r = 0; r < maxn; ++r) {
                float f[17];
#pragma unroll
                for (int k = 0; k < 17; ++k) {
                    f[k] = n1[k];
                    n1[k] = n2[k];
                }
                const uint32_t q = q1;
                q1 = q2;
                if (r + 2 < maxn) q2 = issue(r + 2, n2);
                // `home` stays true for a lane without a particle this round; the ballot below is
                // taken by every lane, so `moved` stays the same in all of them
                bool home = true;
                const float m = f[6];
                if (nc > r && m > 0.0f) {
                    Axis ax[3] = {axis(f[0], a.nx), axis(f[1], a.ny), axis(f[2], a.nz)};
                    home = ax[0].cell == x0 + hx && ax[1].cell == y0 + hy && ax[2].cell == z0 + hz;
                    if (home) {
                        const float vel[3] = {f[3], f[4], f[5]};
                        float cc[9];
#pragma unroll
                        for (int k = 0; k < 9; ++k) cc[k] = f[8 + k];
                        const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - f[7]));
                        p2g_fold(ax, m, vel, cc, s, acc);
                    }
                }
                const uint64_t away = __ballot(!home);
                if (away) {
                    if (!home) {
                        const int at = moved + __popcll(away & below);
                        if (at < kMovedQueue) queue[at] = (int)q;
                    }
                    moved += __popcll(away);
                }
            }
            // the generation's sums into the tile: one stencil offset a step, the lanes' cells distinct
            lds_order();
            if (nc > 0) {
                const int t0 = (hz * kTileSide + hy) * kTileSide + hx;
#pragma unroll
                for (int k = 0; k < 27; ++k) {
                    const int t = t0 + ((k / 9) * kTileSide + (k / 3) % 3) * kTileSide + k % 3;
                    float4 v = tile[t];
                    v.x += acc[4 * k];
                    v.y += acc[4 * k + 1];
                    v.z += acc[4 * k + 2];
                    v.w += acc[4 * k + 3];
                    tile[t] = v;
                    lds_order();
                }
            }
        }
        {
            const int lo[3] = {x0, y0, z0};
            const int hi[3] = {min(x0 + kTileSide - 1, a.nx - 1), min(y0 + kTileSide - 1, a.ny - 1),
                               min(z0 + kTileZ - 1, a.nz - 1)};
            mark_bricks(a, lo, hi, lane, kWave);
        }
        lds_order();
        for (int t = lane; t < kTileNodes; t += kWave) {
            const float4 v = tile[t];
            if (v.x == 0.0f) continue;  // no live particle weighs on the node
            const int tx = t % kTileSide, ty = (t / kTileSide) % kTileSide, tz = t / (kTileSide * kTileSide);
            const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
            if (gx >= a.nx || gy >= a.ny || gz >= a.nz) continue;
            const int64_t node = (int64_t)gz * plane + (int64_t)gy * a.nx + gx;
            atomic_add(a.g[0] + node, v.x);
            atomic_add(a.g[1] + node, v.y);
            atomic_add(a.g[2] + node, v.z);
            atomic_add(a.g[3] + node, v.w);
        }
        if (moved == 0) continue;
        // the particles that left their cell since the sort, now that the sums are spent
        if (moved <= kMovedQueue) {
            for (int j = lane; j < moved; j += kWave) {
                Particle p;
                load_particle(a, queue[j], dt, bulk, rho0, p);
                mark_one(a, p.ax);
                atomic_particle(a, p.ax, p.m, p.vel, p.c, p.s);
            }
            continue;
        }
        pos = S;
        for (int g = 0; g < kGroupGens; ++g) {
            const int c = L.cell[(int64_t)i * kGroupCells + kWave * g + lane];
            const int nc = L.count[(int64_t)i * kGroupCells + c];
            const int maxn = wave_max(nc);
            if (maxn == 0) break;
            const int hx = x0 + (c & 7), hy = y0 + ((c >> 3) & 7), hz = z0 + (c >> 6);
            for (int r = 0; r < maxn; ++r) {
                const uint64_t b = __ballot(nc > r);
                const uint32_t q = pos + (uint32_t)__popcll(b & below);
                pos += (uint32_t)__popcll(b);
                if (nc <= r || !(a.p[6][q] > 0.0f)) continue;
                if (base_cell(a.p[0][q], a.nx) == hx && base_cell(a.p[1][q], a.ny) == hy &&
                    base_cell(a.p[2][q], a.nz) == hz)
                    continue;
                Particle p;
                load_particle(a, q, dt, bulk, rho0, p);
                mark_one(a, p.ax);
                atomic_particle(a, p.ax, p.m, p.vel, p.c, p.s);
            }
        }
    }
}

// ------------------------------------------------------------------ grid update, Flip3.java:267-311

// Flip3.grid for node `node` at `at` with its sums.
__device__ __forceinline__ void grid_node(const Mpm3Args& a, const int (&at)[3], float mass, const float (&mom)[3],
                                          float dt, const float (&grav)[3], float (&vel)[3]) {
    const int size[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int i = 0; i < 3; ++i) vel[i] = 0.0f;
    if (mass > kEmpty) {
#pragma unroll
        for (int i = 0; i < 3; ++i) vel[i] = mom[i] / mass + dt * grav[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (!(at[i] > 1)) vel[i] = mx(vel[i], 0.0f);
        if (at[i] > size[i] - 3) vel[i] = mn(vel[i], 0.0f);
    }
}

// Flip3.grid over the flagged range of every row. Inside it, a node of a brick P2G did not reach
// holds zero sums (the clear zeroed them, nothing added since), so the formula gives it zero
// velocity -- what the full pass gives it -- and overwrites what it held; outside the range every
// node already holds zero sums and zero velocity.
__global__ __launch_bounds__(256) void grid3_kernel(Mpm3Args a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= (int64_t)a.ny * a.nz) return;
    const int y = (int)(row % a.ny), z = (int)(row / a.ny);
    int x0, x1;
    if (!row_range(a, y, z, lane, &x0, &x1)) return;
    const int64_t first = row * a.nx;
    // read before the pass stores anything (a store to the grid could alias them for the compiler)
    const float dt = bits_f(a.params[DT]);
    const float grav[3] = {bits_f(a.params[GX]), bits_f(a.params[GX + 1]), bits_f(a.params[GX + 2])};
    if ((a.nx & 3) == 0) {
        for (int x = x0 + 4 * lane; x < x1; x += 4 * kWave) {
            float4 v[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) v[f] = *reinterpret_cast<const float4*>(a.g[f] + first + x);
            const float m4[4] = {v[0].x, v[0].y, v[0].z, v[0].w};
            const float u4[4] = {v[1].x, v[1].y, v[1].z, v[1].w};
            const float w4[4] = {v[2].x, v[2].y, v[2].z, v[2].w};
            const float q4[4] = {v[3].x, v[3].y, v[3].z, v[3].w};
            float out[3][4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int at[3] = {x + k, y, z};
                const float mom[3] = {u4[k], w4[k], q4[k]};
                float vel[3];
                grid_node(a, at, m4[k], mom, dt, grav, vel);
#pragma unroll
                for (int i = 0; i < 3; ++i) out[i][k] = vel[i];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
                *reinterpret_cast<float4*>(a.g[4 + i] + first + x) = make_float4(out[i][0], out[i][1], out[i][2], out[i][3]);
        }
    } else {
        for (int x = x0 + lane; x < x1; x += kWave) {
            const int64_t node = first + x;
            const int at[3] = {x, y, z};
            const float mom[3] = {a.g[1][node], a.g[2][node], a.g[3][node]};
            float vel[3];
            grid_node(a, at, a.g[0][node], mom, dt, grav, vel);
#pragma unroll
            for (int i = 0; i < 3; ++i) a.g[4 + i][node] = vel[i];
        }
    }
}

// ------------------------------------------------------------------ G2P, Flip3.java:329-383

// keepInside, Flip3.java:644-655
__device__ __forceinline__ void keep_inside(float& at, float& vel, int size) {
    const float lo = kWall;
    const float hi = (float)size - 1.0f - kWall;
    if (at < lo) {
        at = lo;
        vel = mx(vel, 0.0f);
    }
    if (at > hi) {
        at = hi;
        vel = mn(vel, 0.0f);
    }
}

// The end of Flip3.advect (:361-377): C = 4 sum, the move, the walls, J. `j0` is the particle's
// J, read by the caller with its position (a read here would sit behind the stores above it,
// which the compiler cannot move it past, and wait out a whole memory latency per particle).
__device__ __forceinline__ void g2p_finish(const Mpm3Args& a, int64_t q, const float (&pos0)[3], float j0,
                                           float (&vel)[3], const float (&sum)[9], float dt) {
    float c[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) c[i] = 4.0f * sum[i];
    const int size[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float pos = pos0[i] + dt * vel[i];
        keep_inside(pos, vel[i], size[i]);
        a.p[i][q] = pos;
        a.p[3 + i][q] = vel[i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) a.p[8 + i][q] = c[i];
    const float j = j0 * (1.0f + dt * (c[0] + (c[4] + c[8])));
    a.p[7][q] = clampf(j, 1.0f / kJBound, kJBound);
}

// One particle's G2P in the reference's order (Flip3.java:349-360), node velocities read
// through `node_vel(k, axis)`: the CPU restatement's bits.
template <typename NodeVel>
__device__ __forceinline__ void g2p_one(const Mpm3Args& a, int64_t q, const float (&pos0)[3], float j0,
                                        const Axis (&ax)[3], float dt, NodeVel node_vel) {
    float vel[3] = {0.0f, 0.0f, 0.0f};
    float sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const float wgt = ax[0].w[k % 3] * (ax[1].w[(k / 3) % 3] * ax[2].w[k / 9]);
        const float d[3] = {(float)(k % 3) - ax[0].off, (float)((k / 3) % 3) - ax[1].off, (float)(k / 9) - ax[2].off};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float na = wgt * node_vel(k, i);
            vel[i] = vel[i] + na;
#pragma unroll
            for (int e = 0; e < 3; ++e) sum[3 * i + e] = sum[3 * i + e] + na * d[e];
        }
    }
    g2p_finish(a, q, pos0, j0, vel, sum, dt);
}

// The same sums contracted one axis at a time with fused multiply-adds: x first (the weight
// and the weight times the offset), then y, then z -- a third of the operations, each sum
// within a few ulps of the reference's.
template <typename NodeVel>
__device__ __forceinline__ void g2p_fast(const Mpm3Args& a, int64_t q, const float (&pos0)[3], float j0,
                                         const Axis (&ax)[3], float dt, NodeVel node_vel) {
    float wd[3][3];
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int i = 0; i < 3; ++i) wd[e][i] = ax[e].w[i] * ((float)i - ax[e].off);
    float vel[3], sum[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = 0, s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            float R = 0, S = 0, U = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float P = 0, Q = 0;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float g = node_vel(i + 3 * j + 9 * l, c);
                    P = fmaf(ax[0].w[i], g, P);
                    Q = fmaf(wd[0][i], g, Q);
                }
                R = fmaf(ax[1].w[j], P, R);
                S = fmaf(ax[1].w[j], Q, S);
                U = fmaf(wd[1][j], P, U);
            }
            v = fmaf(ax[2].w[l], R, v);
            s0 = fmaf(ax[2].w[l], S, s0);
            s1 = fmaf(ax[2].w[l], U, s1);
            s2 = fmaf(wd[2][l], R, s2);
        }
        vel[c] = v;
        sum[3 * c] = s0;
        sum[3 * c + 1] = s1;
        sum[3 * c + 2] = s2;
    }
    g2p_finish(a, q, pos0, j0, vel, sum, dt);
}

__global__ __launch_bounds__(kBlock) void g2p3_kernel(Mpm3Args a) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= *a.active || !(a.p[6][q] > 0.0f)) return;
    const float pos[3] = {a.p[0][q], a.p[1][q], a.p[2][q]};
    const float j0 = a.p[7][q];
    Axis ax[3] = {axis(pos[0], a.nx), axis(pos[1], a.ny), axis(pos[2], a.nz)};
    const int64_t base = ((int64_t)ax[2].cell * a.ny + ax[1].cell) * a.nx + ax[0].cell;
    const int64_t plane = (int64_t)a.nx * a.ny;
    const float dt = bits_f(a.params[DT]);
    g2p_one(a, q, pos, j0, ax, dt, [&](int k, int i) {
        return a.g[4 + i][base + (k / 9) * plane + ((k / 3) % 3) * a.nx + k % 3];
    });
}

// The G2P tile: the node velocities u, v, w in three planes at fixed strides, so the 81 stencil
// reads of a particle are one address and constant offsets (the immediates of ds_read_b32). It
// holds boxes of up to 20 x 12 x 10 nodes: a block's 128 cells of sorted particles, also where
// the block straddles two neighbouring bricks; a larger box (unsorted or thinned particles) reads
// the grid.
constexpr int kG2PX = 20, kG2PY = 12, kG2PZ = 10;
constexpr int kG2PNodes = kG2PX * kG2PY * kG2PZ;  // 3 x 2400 x 4 B = 28.8 KB

// A block takes a chunk of kBlock * R particles, lane-consecutive, with the chunk's node
// velocities staged in LDS; every position is loaded once, up front, for the box and the move.
template <int R, bool kFast>
__global__ __launch_bounds__(kBlock) void g2p3_tiled_kernel(Mpm3Args a) {
    __shared__ float tile[3 * kG2PNodes];
    __shared__ int red[6 * (kBlock / kWave)];
    const int64_t chunk = (int64_t)blockIdx.x * kBlock * R;
    if (chunk >= *a.active) return;  // a block of inactive slots
    float pos[R][3], jr[R];
    bool live[R];
    const int size[3] = {a.nx, a.ny, a.nz};
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t q = chunk + (int64_t)r * kBlock + threadIdx.x;  // < capacity
#pragma unroll
        for (int i = 0; i < 3; ++i) pos[r][i] = a.p[i][q];
        jr[r] = a.p[7][q];
        live[r] = q < a.count && a.p[6][q] > 0.0f;
        if (!live[r]) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = base_cell(pos[r][i], size[i]);
            lo[i] = min(lo[i], c);
            hi[i] = max(hi[i], c);
        }
    }
    const Box box = reduce_box(lo, hi, red, INT_MAX);
    if (box.empty) return;
    const bool use_tile = box.ex <= kG2PX && box.ey <= kG2PY && box.ez <= kG2PZ;
    if (use_tile) {
        // every load of the thread's nodes in flight before the first LDS write. A node's
        // (x, y, z) in the box from its number t by float reciprocals: t < 2^12, so t + 1/2
        // lies at least 1/(2 ex) from a multiple of ex and the rounded product floors exactly.
        constexpr int kStage = (kG2PNodes + kBlock - 1) / kBlock;
        const int nodes = box.ex * box.ey * box.ez;
        const float inv_ex = 1.0f / (float)box.ex, inv_ey = 1.0f / (float)box.ey;
        const int64_t plane = (int64_t)a.nx * a.ny;
        float gu[kStage], gv[kStage], gw[kStage];
        int at[kStage];
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int t = threadIdx.x + k * kBlock;
            at[k] = -1;
            if (t < nodes) {
                const int row = (int)(((float)t + 0.5f) * inv_ex);
                const int lx = t - row * box.ex;
                const int lz = (int)(((float)row + 0.5f) * inv_ey);
                const int ly = row - lz * box.ey;
                const int64_t node =
                    (int64_t)(box.lo[2] + lz) * plane + (int64_t)(box.lo[1] + ly) * a.nx + (box.lo[0] + lx);
                at[k] = (lz * kG2PY + ly) * kG2PX + lx;
                gu[k] = a.g[4][node];
                gv[k] = a.g[5][node];
                gw[k] = a.g[6][node];
            }
        }
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            if (at[k] < 0) continue;
            tile[at[k]] = gu[k];
            tile[kG2PNodes + at[k]] = gv[k];
            tile[2 * kG2PNodes + at[k]] = gw[k];
        }
    }
    __syncthreads();
    const float dt = bits_f(a.params[DT]);
    const int64_t plane = (int64_t)a.nx * a.ny;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!live[r]) continue;
        const int64_t q = chunk + (int64_t)r * kBlock + threadIdx.x;
        Axis ax[3] = {axis(pos[r][0], a.nx), axis(pos[r][1], a.ny), axis(pos[r][2], a.nz)};
        if (use_tile) {
            const float* base = tile + ((ax[2].cell - box.lo[2]) * kG2PY + (ax[1].cell - box.lo[1])) * kG2PX +
                                (ax[0].cell - box.lo[0]);
            auto at = [&](int k, int i) { return base[i * kG2PNodes + ((k / 9) * kG2PY + (k / 3) % 3) * kG2PX + k % 3]; };
            if (kFast) g2p_fast(a, q, pos[r], jr[r], ax, dt, at);
            else g2p_one(a, q, pos[r], jr[r], ax, dt, at);
        } else {
            const int64_t base = ((int64_t)ax[2].cell * a.ny + ax[1].cell) * a.nx + ax[0].cell;
            auto at = [&](int k, int i) { return a.g[4 + i][base + (k / 9) * plane + ((k / 3) % 3) * a.nx + k % 3]; };
            if (kFast) g2p_fast(a, q, pos[r], jr[r], ax, dt, at);
            else g2p_one(a, q, pos[r], jr[r], ax, dt, at);
        }
    }
}

// ------------------------------------------------------------------ sort keys

__host__ __device__ __forceinline__ int brick_count(int n) { return (n - 2 + 7) / 8; }

// Stencil3's base cell numbered brick-major: 8x8x8-cell bricks row-major, cells row-major
// inside a brick. Padding particles take the largest key so a stable sort leaves them last.
__global__ __launch_bounds__(kBlock) void keys3_kernel(Mpm3Args a, int32_t* __restrict__ keys,
                                                       int32_t* __restrict__ index, int32_t pad_key,
                                                       int32_t inactive_key) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.capacity) return;
    int32_t key = pad_key;
    if (q < a.count && !(a.p[6][q] > 0.0f)) {
        key = inactive_key;
    } else if (q < a.count) {
        const int cx = base_cell(a.p[0][q], a.nx), cy = base_cell(a.p[1][q], a.ny), cz = base_cell(a.p[2][q], a.nz);
        const int bx = brick_count(a.nx), by = brick_count(a.ny);
        const int brick = ((cz >> 3) * by + (cy >> 3)) * bx + (cx >> 3);
        const int local = ((cz & 7) * 8 + (cy & 7)) * 8 + (cx & 7);
        key = brick * 512 + local;
    }
    keys[q] = key;
    index[q] = (int32_t)q;
}

// home[slot[p]] = p over the sorted particles, and active = the first p whose key is the
// inactive one (count when there is none): the one p where the keys cross it writes it, so
// nothing has to be cleared first.
__global__ __launch_bounds__(kBlock) void after_sort_kernel(Mpm3Args a, const int32_t* __restrict__ keys,
                                                            int32_t inactive_key) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= a.count) return;
    a.home[a.slot[p]] = (int32_t)p;
    const bool here = keys[p] >= inactive_key;
    const bool before = p > 0 && keys[p - 1] >= inactive_key;
    if (here && !before) *a.active = (int32_t)p;
    if (!here && p == a.count - 1) *a.active = (int32_t)a.count;
}

__global__ __launch_bounds__(kBlock) void slots_kernel(Mpm3Args a, int identity) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p == 0) *a.active = (int32_t)a.count;
    if (p >= a.capacity) return;
    if (identity) {
        a.slot[p] = (int32_t)p;
        a.home[p] = (int32_t)p;
    } else if (p < a.count) {
        a.home[a.slot[p]] = (int32_t)p;
    }
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int32_t mpm3_inactive_key(int nx, int ny, int nz) {
    // the key of base cell (n - 3) on every axis, plus one
    const int c[3] = {nx - 3, ny - 3, nz - 3};
    const int bx = (nx - 2 + 7) / 8, by = (ny - 2 + 7) / 8;
    const int brick = ((c[2] >> 3) * by + (c[1] >> 3)) * bx + (c[0] >> 3);
    const int local = ((c[2] & 7) * 8 + (c[1] & 7)) * 8 + (c[0] & 7);
    return brick * 512 + local + 1;
}

int mpm3_key_bits(int nx, int ny, int nz) {
    const int64_t keys = (int64_t)mpm3_inactive_key(nx, ny, nz) + 1;
    int bits = 0;
    while ((1LL << bits) < keys) bits++;
    return bits;
}

int mpm3_after_sort(const Launch& l, const Mpm3Args& a, const int32_t* keys) {
    if (a.count == 0) return 0;
    hipLaunchKernelGGL(after_sort_kernel, dim3(blocks_for(a.count, kBlock)), dim3(kBlock), 0, l.stream, a, keys,
                       mpm3_inactive_key(a.nx, a.ny, a.nz));
    return check_launch("mpm3_after_sort");
}

int mpm3_reset_slots(const Launch& l, const Mpm3Args& a) {
    hipLaunchKernelGGL(slots_kernel, dim3(blocks_for(std::max<int64_t>(a.capacity, 1), kBlock)), dim3(kBlock), 0,
                       l.stream, a, 1);
    return check_launch("mpm3_reset_slots");
}

int mpm3_rebuild_home(const Launch& l, const Mpm3Args& a) {
    hipLaunchKernelGGL(slots_kernel, dim3(blocks_for(std::max<int64_t>(a.capacity, 1), kBlock)), dim3(kBlock), 0,
                       l.stream, a, 0);
    return check_launch("mpm3_rebuild_home");
}

int mpm3_clear(const Launch& l, const Mpm3Args& a) {
    timed_begin(l, "mpm3_clear");
    hipLaunchKernelGGL(clear_rows_kernel, dim3(blocks_for((int64_t)a.ny * a.nz, 4)), dim3(256), 0, l.stream, a);
    timed_end(l, "mpm3_clear");
    return check_launch("mpm3_clear");
}

int mpm3_clear_fixed(const Launch& l, const Mpm3Args& a, const Fixed3& fx) {
    timed_begin(l, "mpm3_clear");
    hipLaunchKernelGGL(clear_rows_kernel, dim3(blocks_for((int64_t)a.ny * a.nz, 4)), dim3(256), 0, l.stream, a);
    hipLaunchKernelGGL(fixed_rows_kernel, dim3(blocks_for((int64_t)a.ny * a.nz, 4)), dim3(256), 0, l.stream, a, fx, 0);
    timed_end(l, "mpm3_clear");
    return check_launch("mpm3_clear");
}

int mpm3_scatter_fixed(const Launch& l, const Mpm3Args& a, const Fixed3& fx) {
    if (a.count == 0) return 0;
    timed_begin(l, "p2g3_fixed");
    hipLaunchKernelGGL(p2g3_fixed_kernel<kFixedRun>, dim3(blocks_for(a.count, kWave * kFixedRun)), dim3(kWave), 0,
                       l.stream, a, fx);
    timed_end(l, "p2g3_fixed");
    if (int rc = check_launch("p2g3_fixed")) return rc;
    timed_begin(l, "fixed3_decode");
    hipLaunchKernelGGL(fixed_rows_kernel, dim3(blocks_for((int64_t)a.ny * a.nz, 4)), dim3(256), 0, l.stream, a, fx, 1);
    timed_end(l, "fixed3_decode");
    return check_launch("fixed3_decode");
}

int mpm3_scatter_atomic(const Launch& l, const Mpm3Args& a) {
    if (a.count == 0) return 0;
    timed_begin(l, "p2g3_atomic");
    hipLaunchKernelGGL(p2g3_atomic_kernel, dim3(blocks_for(a.count, kBlock)), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "p2g3_atomic");
    return check_launch("p2g3_atomic");
}

template <int R>
static void launch_p2g(const Launch& l, const Mpm3Args& a, bool ints) {
    const dim3 grid(blocks_for(a.count, kWave * R));
    if (ints) hipLaunchKernelGGL((p2g3_wave_kernel<R, true>), grid, dim3(kWave), 0, l.stream, a);
    else hipLaunchKernelGGL((p2g3_wave_kernel<R, false>), grid, dim3(kWave), 0, l.stream, a);
}

int mpm3_scatter_tiled(const Launch& l, const Mpm3Args& a, int run, bool ints) {
    if (a.count == 0) return 0;
    timed_begin(l, "p2g3_seg");
    if (run == 4) launch_p2g<4>(l, a, ints);
    else if (run == 2) launch_p2g<2>(l, a, ints);
    else launch_p2g<1>(l, a, ints);
    timed_end(l, "p2g3_seg");
    return check_launch("p2g3_seg");
}

int64_t mpm3_layer_capacity(int nx, int ny, int nz) {
    // groups: every sort brick's 512 cells, kGroupCells at a time
    return (int64_t)((nx - 2 + 7) / 8) * ((ny - 2 + 7) / 8) * ((nz - 2 + 7) / 8) * (512 / kGroupCells);
}

int mpm3_layer_cells() { return kGroupCells; }

int64_t mpm3_layer_scratch_words(int64_t capacity) {
    return 2 * capacity + (capacity + kGroupPer - 1) / kGroupPer + 1;
}

int mpm3_layers(const Launch& l, const Mpm3Args& a, int32_t* keys, const int32_t* index, int32_t* scratch,
                const Layers& L) {
    const int64_t n = a.count;
    const int32_t inactive = mpm3_inactive_key(a.nx, a.ny, a.nz);
    const int blocks = (int)((n + kGroupPer - 1) / kGroupPer);
    int32_t* new_index = scratch;
    int32_t* new_keys = scratch + a.capacity;
    int32_t* heads = scratch + 2 * a.capacity;
    timed_begin(l, "layers3");
    if (blocks > 0)
        hipLaunchKernelGGL(group_count_kernel, dim3(blocks), dim3(256), 0, l.stream, keys, n, inactive, index,
                           new_index, new_keys, heads);
    hipLaunchKernelGGL(group_scan_kernel, dim3(1), dim3(256), 0, l.stream, heads, blocks, L.meta, L.cap);
    if (blocks > 0) {
        hipLaunchKernelGGL(group_list_kernel, dim3(blocks), dim3(256), 0, l.stream, keys, n, inactive, heads, L.layer,
                           L.start, L.meta, L.cap);
        const int waves = (int)std::min<int64_t>(L.cap, kGroupWaves);
        const int wb = (waves + 3) / 4;
        hipLaunchKernelGGL(group_order_kernel, dim3(wb), dim3(256), 0, l.stream, keys, index, new_index, new_keys, L,
                           wb * 4);
        if (hipMemcpyAsync(keys, new_keys, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, l.stream) != hipSuccess)
            return check_launch("layers3 keys");
    }
    timed_end(l, "layers3");
    return check_launch("layers3");
}

int mpm3_scatter_layers(const Launch& l, const Mpm3Args& a, const Layers& L) {
    if (a.count == 0) return 0;
    const int waves = (int)std::min<int64_t>(L.cap, kGroupWaves);
    const int wb = (waves + 3) / 4;
    timed_begin(l, "p2g3_layer");
    hipLaunchKernelGGL(p2g3_layer_kernel, dim3(wb), dim3(256), 0, l.stream, a, L, wb * 4);
    timed_end(l, "p2g3_layer");
    return check_launch("p2g3_layer");
}

int mpm3_flags(const Launch& l, const Mpm3Args& a) {
    const int64_t bricks = brick_count3(a.nx, a.ny, a.nz);
    hipLaunchKernelGGL(flags_kernel, dim3(blocks_for(bricks, 256)), dim3(256), 0, l.stream, a.touched, a.live, bricks);
    return check_launch("grid3_flags");
}

int mpm3_grid(const Launch& l, const Mpm3Args& a) {
    timed_begin(l, "grid3");
    hipLaunchKernelGGL(grid3_kernel, dim3(blocks_for((int64_t)a.ny * a.nz, 4)), dim3(256), 0, l.stream, a);
    const int64_t bricks = brick_count3(a.nx, a.ny, a.nz);
    hipLaunchKernelGGL(flags_kernel, dim3(blocks_for(bricks, 256)), dim3(256), 0, l.stream, a.touched, a.live, bricks);
    timed_end(l, "grid3");
    return check_launch("grid3");
}

int mpm3_advect(const Launch& l, const Mpm3Args& a) {
    if (a.count == 0) return 0;
    timed_begin(l, "g2p3");
    hipLaunchKernelGGL(g2p3_kernel, dim3(blocks_for(a.count, kBlock)), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "g2p3");
    return check_launch("g2p3");
}

template <int R>
static void launch_g2p(const Launch& l, const Mpm3Args& a, bool fast) {
    const unsigned blocks = blocks_for(a.count, kBlock * R);
    if (fast) hipLaunchKernelGGL((g2p3_tiled_kernel<R, true>), dim3(blocks), dim3(kBlock), 0, l.stream, a);
    else hipLaunchKernelGGL((g2p3_tiled_kernel<R, false>), dim3(blocks), dim3(kBlock), 0, l.stream, a);
}

int mpm3_advect_tiled(const Launch& l, const Mpm3Args& a, int run, bool fast) {
    if (a.count == 0) return 0;
    const char* name = fast ? "g2p3_fast" : "g2p3_tiled";
    timed_begin(l, name);
    if (run == 4) launch_g2p<4>(l, a, fast);
    else if (run == 2) launch_g2p<2>(l, a, fast);
    else launch_g2p<1>(l, a, fast);
    timed_end(l, name);
    return check_launch(name);
}

int mpm3_keys(const Launch& l, const Mpm3Args& a, int32_t* keys, int32_t* index) {
    const int32_t pad = (int32_t)((1LL << mpm3_key_bits(a.nx, a.ny, a.nz)) - 1);
    timed_begin(l, "keys3");
    hipLaunchKernelGGL(keys3_kernel, dim3(blocks_for(a.capacity, kBlock)), dim3(kBlock), 0, l.stream, a, keys, index,
                       pad, mpm3_inactive_key(a.nx, a.ny, a.nz));
    timed_end(l, "keys3");
    return check_launch("keys3");
}

}  // namespace vxr
