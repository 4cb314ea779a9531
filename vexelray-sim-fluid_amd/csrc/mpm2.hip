// The 2D MLS-MPM / APIC step (core/particle/Flip.java, FlipStep.java) for gfx950.
//
// Layout: 10 particle fields as f32 arrays (x y u v m j c00 c01 c10 c11), padded with m = 0
// to whole chunks; the grid as 5 f32 node arrays, row-major (gm gmu gmv gu gv).
//
// P2G (float, p2g 0 and 1): the reference sums each run of particles sharing a stencil base across a subgroup's
// lanes (Scatter.segmentedDeposit, Scatter.java:385-439) and the run's last lane takes the
// atomics. Here a thread walks a run of consecutive particles and folds those sharing a base
// into 27 register sums, with one atomic per node and field for each run it holds: the same
// sums regrouped, right in any order and cheap in cell order. The grid update and G2P keep
// the reference's arithmetic order, so from the same grid they are its results bit for bit.
//
// P2G in fixed point (p2g 2, and every batch of scenes): each amount rounded to a 64-bit
// integer (mass in 2^-40, momentum in 2^-32) and summed by integer atomics, which associate
// exactly -- core/particle/FixedScatter.java:24-31's reason for integers, applied to these
// amounts -- so the grid is the same bits in any order, on any launch shape, and equals the
// CPU restatement's (orc_mpm2_scatter_fixed) bit for bit.
#include "common.h"

namespace vxr {
namespace {

__device__ __forceinline__ float bits_f(int32_t w) { return __int_as_float(w); }

enum { DT = 0, GX = 1, GY = 2, BULK = 3, RHO0 = 4 };  // Flip.java:102-106
constexpr float kWall = 1.0f;
constexpr float kEmpty = 1e-9f;
constexpr float kJBound = 10.0f;
constexpr int kBlock = 256;

struct Axis {
    int cell;
    float off;
    float w[3];
};

// One axis of Flip.Stencil.of, Flip.java:572-590.
__device__ __forceinline__ Axis axis(float pos, int n) {
    Axis s;
    const float x = clampf(pos, 0.5f, (float)n - 1.5f);
    s.cell = (int)mn(x - 0.5f, (float)(n - 3));
    s.off = x - (float)s.cell;
    const float a = 1.5f - s.off;
    const float c = s.off - 1.0f;
    const float e = s.off - 0.5f;
    s.w[0] = 0.5f * (a * a);
    s.w[1] = 0.75f - (c * c);
    s.w[2] = 0.5f * (e * e);
    return s;
}

// The 27 amounts of one particle (Flip.java:283-300), node k's three at 3k .. 3k+2, handed to
// `take(k, mass, momentum_u, momentum_v)`.
// `v` holds the particle's ten fields (x, y, u, v, m, J, C00, C01, C10, C11).
template <class Take>
__device__ __forceinline__ void amounts(const float (&v)[10], const Axis& ax, const Axis& ay, float dt, float bulk,
                                        float rho0, Take take) {
    const float m = v[4], u = v[2], w = v[3];
    const float c00 = v[6], c01 = v[7], c10 = v[8], c11 = v[9];
    const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - v[5]));
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float wgt = ax.w[k % 3] * ay.w[k / 3];
        const float wm = wgt * m;
        const float dx = (float)(k % 3) - ax.off;
        const float dy = (float)(k / 3) - ay.off;
        const float vu = u + (c00 * dx + c01 * dy);
        const float vv = w + (c10 * dx + c11 * dy);
        take(k, wm, wm * vu + wgt * (s * dx), wm * vv + wgt * (s * dy));
    }
}

__device__ __forceinline__ void amounts(float* const* p, int64_t q, const Axis& ax, const Axis& ay, float dt,
                                        float bulk, float rho0, float (&acc)[27]) {
    float v[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) v[f] = p[f][q];
    amounts(v, ax, ay, dt, bulk, rho0, [&](int k, float a0, float a1, float a2) {
        acc[3 * k] += a0;
        acc[3 * k + 1] += a1;
        acc[3 * k + 2] += a2;
    });
}

// Fixed point: an amount in units of 2^-bits, rounded to the nearest integer (ties to even).
// Scaling by a power of two is exact, so this is the one rounding, the same on the CPU.
constexpr int kMassBits = 40;
constexpr int kMomentumBits = 32;
// The integer is held in a double: a * scale is exact (a power of two) and rintf rounds it to the
// nearest integer, ties to even. Sums of such integers in double are exact while they stay below
// 2^53 in magnitude (a run's sums of up to 4 particles' amounts, far below it), so a run folds
// in four instructions an amount and converts to a 64-bit integer once, at its flush.
__device__ __forceinline__ double quantize_d(float a, float scale) { return (double)rintf(a * scale); }
__device__ __forceinline__ float dequantize(unsigned long long q, int bits) {
    return (float)ldexp((double)(long long)q, -bits);
}

__global__ void clear3_kernel(float* __restrict__ g0, float* __restrict__ g1, float* __restrict__ g2, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        g0[i] = 0.0f;
        g1[i] = 0.0f;
        g2[i] = 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void p2g2_atomic_kernel(Mpm2Args a) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.count) return;
    const float dt = bits_f(a.params[DT]), bulk = bits_f(a.params[BULK]), rho0 = bits_f(a.params[RHO0]);
    const Axis ax = axis(a.p[0][q], a.nx), ay = axis(a.p[1][q], a.ny);
    float acc[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) acc[i] = 0.0f;
    amounts(a.p, q, ax, ay, dt, bulk, rho0, acc);
    const int64_t base = (int64_t)ay.cell * a.nx + ax.cell;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int64_t node = base + (k / 3) * a.nx + k % 3;
#pragma unroll
        for (int f = 0; f < 3; ++f) atomic_add(a.g[f] + node, acc[3 * k + f]);
    }
}

__global__ __launch_bounds__(kBlock) void p2g2_runs_kernel(Mpm2Args a, int run) {
    const int64_t first = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * run;
    const float dt = bits_f(a.params[DT]), bulk = bits_f(a.params[BULK]), rho0 = bits_f(a.params[RHO0]);
    float acc[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) acc[i] = 0.0f;
    int64_t run_base = -1;
    auto flush = [&](int64_t base) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int64_t node = base + (k / 3) * a.nx + k % 3;
#pragma unroll
            for (int f = 0; f < 3; ++f) atomic_add(a.g[f] + node, acc[3 * k + f]);
        }
#pragma unroll
        for (int i = 0; i < 27; ++i) acc[i] = 0.0f;
    };
    for (int r = 0; r < run; ++r) {
        const int64_t q = first + r;
        if (q >= a.count) break;
        const Axis ax = axis(a.p[0][q], a.nx), ay = axis(a.p[1][q], a.ny);
        const int64_t base = (int64_t)ay.cell * a.nx + ax.cell;
        if (base != run_base) {
            if (run_base >= 0) flush(run_base);
            run_base = base;
        }
        amounts(a.p, q, ax, ay, dt, bulk, rho0, acc);
    }
    if (run_base >= 0) flush(run_base);
}

// dt and the two components of gravity, read from a scene's parameters.
struct Law2 {
    float dt, gx, gy;
};

__device__ __forceinline__ Law2 law2(const int32_t* params) {
    return {bits_f(params[DT]), bits_f(params[GX]), bits_f(params[GY])};
}

// Flip.grid, Flip.java:366-412 (weight 1: no buoyancy), for node `node` of an nx x ny grid.
__device__ __forceinline__ void grid2_node(int64_t node, int nx, const Law2& law, float m, float mu, float mv,
                                           int ny, float* gu, float* gv) {
    const int ni = (int)(node % nx), nj = (int)(node / nx);
    float u = 0.0f, w = 0.0f;
    if (m > kEmpty) {
        u = mu / m + law.dt * (law.gx * 1.0f);
        w = mv / m + law.dt * (law.gy * 1.0f);
    }
    if (!(ni > 1)) u = mx(u, 0.0f);
    if (ni > nx - 3) u = mn(u, 0.0f);
    if (!(nj > 1)) w = mx(w, 0.0f);
    if (nj > ny - 3) w = mn(w, 0.0f);
    gu[node] = u;
    gv[node] = w;
}

__global__ __launch_bounds__(kBlock) void grid2_kernel(Mpm2Args a) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)a.nx * a.ny) return;
    grid2_node(node, a.nx, law2(a.params), a.g[0][node], a.g[1][node], a.g[2][node], a.ny, a.g[3], a.g[4]);
}

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

// What G2P reads of a particle and its grid: the position, J, and the 9 stencil nodes' velocities.
struct G2p2In {
    float x, y, j;
    Axis ax, ay;
    float gu[9], gv[9];
};

// The stencil's node velocities for a particle whose position and J are in `in` already.
__device__ __forceinline__ void g2p2_nodes(int nx, int ny, const float* gu, const float* gv, G2p2In& in) {
    in.ax = axis(in.x, nx);
    in.ay = axis(in.y, ny);
    const int64_t base = (int64_t)in.ay.cell * nx + in.ax.cell;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int64_t node = base + (k / 3) * nx + k % 3;
        in.gu[k] = gu[node];
        in.gv[k] = gv[node];
    }
}

__device__ __forceinline__ void g2p2_gather(float* const* p, int64_t q, int nx, int ny, const float* gu,
                                            const float* gv, G2p2In& in) {
    in.x = p[0][q];
    in.y = p[1][q];
    in.j = p[5][q];
    g2p2_nodes(nx, ny, gu, gv, in);
}

// Flip.advectBody, Flip.java:494-545 (not relaxing), from what g2p2_gather read: the particle's
// new fields in `o`, in the order of kG2pFields.
constexpr int kG2pFields[9] = {0, 1, 2, 3, 6, 7, 8, 9, 5};  // x y u v c00 c01 c10 c11 j
__device__ __forceinline__ void g2p2_compute(int nx, int ny, float dt, const G2p2In& in, float (&o)[9]) {
    const Axis& ax = in.ax;
    const Axis& ay = in.ay;
    float u = 0, w = 0, b00 = 0, b01 = 0, b10 = 0, b11 = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float wgt = ax.w[k % 3] * ay.w[k / 3];
        const float dx = (float)(k % 3) - ax.off;
        const float dy = (float)(k / 3) - ay.off;
        const float ui = wgt * in.gu[k];
        const float vi = wgt * in.gv[k];
        u = u + ui;
        w = w + vi;
        b00 = b00 + ui * dx;
        b01 = b01 + ui * dy;
        b10 = b10 + vi * dx;
        b11 = b11 + vi * dy;
    }
    const float c00 = 4.0f * b00, c01 = 4.0f * b01, c10 = 4.0f * b10, c11 = 4.0f * b11;
    float x = in.x + dt * u;
    float y = in.y + dt * w;
    keep_inside(x, u, nx);
    keep_inside(y, w, ny);
    const float j = in.j * (1.0f + dt * (c00 + c11));
    o[0] = x;
    o[1] = y;
    o[2] = u;
    o[3] = w;
    o[4] = c00;
    o[5] = c01;
    o[6] = c10;
    o[7] = c11;
    o[8] = clampf(j, 1.0f / kJBound, kJBound);
}

// The same for particle q, stored. Reads and stores are split so a thread can have several
// particles' reads in flight before its first store (a read cannot be moved past a store that
// may alias it).
__device__ __forceinline__ void g2p2_finish(float* const* p, int64_t q, int nx, int ny, float dt, const G2p2In& in) {
    float o[9];
    g2p2_compute(nx, ny, dt, in, o);
#pragma unroll
    for (int f = 0; f < 9; ++f) p[kG2pFields[f]][q] = o[f];
}

__device__ __forceinline__ void g2p2_particle(float* const* p, int64_t q, int nx, int ny, const float* gu,
                                              const float* gv, const int32_t* params) {
    G2p2In in;
    const float dt = bits_f(params[DT]);
    g2p2_gather(p, q, nx, ny, gu, gv, in);
    g2p2_finish(p, q, nx, ny, dt, in);
}

__global__ __launch_bounds__(kBlock) void g2p2_kernel(Mpm2Args a) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.count) return;
    g2p2_particle(a.p, q, a.nx, a.ny, a.g[3], a.g[4], a.params);
}

// Scatter.cell as a function (Scatter.java:473-477).
__device__ __forceinline__ int32_t cell2(float x, float y, int nx, int ny) {
    const float cx = clampf(x, 0.0f, (float)nx - 1.0f);
    const float cy = clampf(y, 0.0f, (float)ny - 1.0f);
    return (int32_t)mn(cy, (float)ny - 2.0f) * (nx - 1) + (int32_t)mn(cx, (float)nx - 2.0f);
}

// Scatter.cell, Scatter.java:473-477
__global__ __launch_bounds__(kBlock) void keys2_kernel(Mpm2Args a, int32_t* __restrict__ keys,
                                                       int32_t* __restrict__ index, int32_t pad_key) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.capacity) return;
    const int32_t key = q < a.count ? cell2(a.p[0][q], a.p[1][q], a.nx, a.ny) : pad_key;
    keys[q] = key;
    index[q] = (int32_t)q;
}

// starts[c] = the first sorted particle in cell c or after it; starts[cells] = count
__global__ __launch_bounds__(kBlock) void starts_kernel(const int32_t* __restrict__ keys, int64_t count,
                                                        int32_t cells, int32_t* __restrict__ starts) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c > cells) return;
    int64_t lo = 0, hi = count;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    starts[c] = (int32_t)lo;
}


// ------------------------------------------------------------------ scenes in one launch
//
// A batch holds many 2D scenes in one set of buffers: scene s's particles at
// [poff, poff + count) of every field (its range padded to kBatchAlign with m = 0), its nodes
// at [goff, goff + nx*ny) of every grid field, its 28 params at params[28 s], its starts at
// [soff, soff + cells + 1). Each kernel walks fixed chunks -- kBatchAlign particles, 256 nodes,
// 256 starts -- and finds a chunk's scene in a table, so one launch steps every scene, and
// each scene gets exactly the arithmetic a single-scene run gives it: the P2G is the
// fixed-point one, whose integer sums do not depend on the order or grouping of the atomics.

// An integer held exactly in a double (|v| < 2^53) as a 64-bit integer: the high word
// floor(v / 2^32) and the low word v - 2^32 high, both exact. No rounding step is needed, since
// v is an integer already (a general conversion rounds first).
__device__ __forceinline__ long long exact_ll(double v) {
    const double hi = floor(v * 0x1p-32);
    const double lo = fma(hi, -0x1p32, v);  // in [0, 2^32)
    return (long long)(((unsigned long long)(unsigned int)(int)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ void add_fixed(unsigned long long* p, long long v) {
    atomicAdd(p, (unsigned long long)v);
}

// P2G in fixed point. A block takes a chunk of kBatchAlign particles of one scene; a thread
// loads four consecutive ones, which give the chunk's box and then fold, kRun at a time, while
// they share a stencil base, into 27 integer sums. When the
// chunk's stencils fit kTile2 nodes (sorted particles: a few rows of cells) the sums go to an
// LDS tile by 64-bit LDS atomics, and the tile's touched nodes to the grid by one 64-bit global
// atomic each; otherwise straight to the grid. Integer sums: the same bits either way.
constexpr int kTile2 = 2048;  // 3 x 8 B a node: 48 KB
// Nodes a touched or live flag covers: four flags to a 256-node chunk, one byte each.
constexpr int kUnit2 = 64;

// A rounded amount r (an integer in binary32, |r| < 2^40) as the pair h = floor(r / 2^20),
// l = r - 2^20 h in [0, 2^20), h in the high word and l in the low one, every step exact (r
// 2^-20 is a power-of-two scaling, and l, an integer below 2^20, is a binary32). Added as one
// 64-bit integer, the low words' sum stays below 2^30 for up to 1,024 amounts, so no carry
// reaches the high word, whose sum of |h| <= 2^20 stays in int32: the LDS tile sums a block's
// amounts with the 64-bit LDS atomic as before, for five 32-bit instructions an amount instead
// of a fold and split in double precision. unpair20 gives back the exact integer sum.
constexpr float kPairLimit = 0x1.fep39f;  // below 2^40 by a margin for the bound's own rounding
__device__ __forceinline__ unsigned long long pair20(float r) {
    const float h = floorf(r * 0x1p-20f);
    const float l = fmaf(h, -0x1p20f, r);
    return ((unsigned long long)(unsigned)(int)h << 32) | (unsigned long long)(unsigned)(int)l;
}
__device__ __forceinline__ unsigned long long unpair20(unsigned long long t) {
    return (unsigned long long)((long long)(int)(unsigned)(t >> 32) * (1ll << 20) + (long long)(t & 0xffffffffull));
}


// kOwn = 4: a block a chunk, a thread four consecutive particles. kOwn = 1: a block a quarter
// chunk, a thread one particle, for scenes too small to fill the GPU with whole chunks.
template <int kRun, int kOwn>
__global__ __launch_bounds__(kBlock) void batch2_p2g_kernel(Batch2Args a) {
    __shared__ unsigned long long tile[3][kTile2];
    __shared__ int box[5];
    const float mscale = 0x1p40f, pscale = 0x1p32f;
    constexpr int kParts = kBatchAlign / (kBlock * kOwn);  // a chunk's parts, a block each
    static_assert(kBatchAlign % (kBlock * kOwn) == 0 && kOwn % kRun == 0, "a thread's particles split into whole runs");
    for (int64_t bc = blockIdx.x; bc < a.pchunks * kParts; bc += gridDim.x) {
        // the thread's particles, every field with one load, all in flight at once and beside the
        // scene's number (the chunk lies inside its scene's range, padded to whole chunks with
        // massless slots, so these reads need nothing from the scene)
        const int64_t c = bc / kParts;
        const int64_t first = c * kBatchAlign + (bc % kParts) * (kBlock * kOwn) + (int64_t)threadIdx.x * kOwn;
        float f[10][kOwn];
#pragma unroll
        for (int k = 0; k < 10; ++k) load_run<kOwn>(a.p[k], first, f[k]);
        const int s = a.pscene[c];
        const Scene2 sc = a.scenes[s];
        const int32_t* params = a.params + 28 * (int64_t)s;
        const float dt = bits_f(params[DT]), bulk = bits_f(params[BULK]), rho0 = bits_f(params[RHO0]);
        const int64_t live = sc.poff + sc.count;  // the chunk's live particles are a prefix of it

        // the chunk's box of stencil bases
        if (threadIdx.x == 0) {
            box[0] = box[1] = 0x7fffffff;
            box[2] = box[3] = -1;
            box[4] = 0;
        }
        __syncthreads();
        int x0 = 0x7fffffff, y0 = 0x7fffffff, x1 = -1, y1 = -1;
        // a bound on the magnitude of every amount a particle rounds (Flip.java:283-300's terms:
        // weights <= 0.5625, offsets <= 1.5), as a sum so a NaN carries through to "no bound"
        const float inv_rho0 = 1.0f / fabsf(rho0);
        float bound = 0.0f;
#pragma unroll
        for (int r = 0; r < kOwn; ++r) {
            if (first + r >= live) break;
            const int bx = axis(f[0][r], sc.nx).cell, by = axis(f[1][r], sc.ny).cell;
            x0 = min(x0, bx);
            y0 = min(y0, by);
            x1 = max(x1, bx);
            y1 = max(y1, by);
            const float mp = fabsf(f[4][r] * pscale);
            const float sb = 4.0f * fabsf(dt) * (mp * inv_rho0) * fabsf(bulk) * fabsf(1.0f - f[5][r]);
            const float b = 0.5625f * (mp * (256.0f + fabsf(f[2][r]) + fabsf(f[3][r]) +
                                             1.5f * (fabsf(f[6][r]) + fabsf(f[7][r]) + fabsf(f[8][r]) + fabsf(f[9][r]))) +
                                       1.5f * sb);
            bound = b < kPairLimit ? fmaxf(bound, b) : __builtin_inff();
        }
        x0 = wave_min(x0);
        y0 = wave_min(y0);
        x1 = wave_max(x1);
        y1 = wave_max(y1);
        const int bound_bits = wave_max(__float_as_int(bound));  // non-negative: int order is float order
        if ((threadIdx.x & (kWave - 1)) == 0) {
            atomicMin(&box[0], x0);
            atomicMin(&box[1], y0);
            atomicMax(&box[2], x1);
            atomicMax(&box[3], y1);
            atomicMax(&box[4], bound_bits);
        }
        __syncthreads();
        const int bx0 = box[0], by0 = box[1];
        const int w = box[2] - bx0 + 3, h = box[3] - by0 + 3;
        const bool tiled = box[2] >= 0 && w * h <= kTile2;
        const bool pairs = tiled && __int_as_float(box[4]) < kPairLimit;
        if (box[2] >= 0) {
            // the units each node row of the box reaches (the box lies inside the scene's grid)
            for (int r = threadIdx.x; r < h; r += kBlock) {
                const int64_t row0 = sc.goff + (int64_t)(by0 + r) * sc.nx + bx0;
                for (int64_t u = row0 / kUnit2; u <= (row0 + w - 1) / kUnit2; ++u) a.touched[u] = 1;
            }
        }
        if (tiled) {
            for (int i = threadIdx.x; i < w * h; i += kBlock) tile[0][i] = tile[1][i] = tile[2][i] = 0ull;
        }
        __syncthreads();

        unsigned long long* q0 = a.q[0] + sc.goff;
        unsigned long long* q1 = a.q[1] + sc.goff;
        unsigned long long* q2 = a.q[2] + sc.goff;
        if (pairs) {
            // every amount of the block below 2^40: a particle at a time, each amount rounded and
            // split into its pair, added to the tile
#pragma unroll
            for (int r = 0; r < kOwn; ++r) {
                if (first + r >= live) break;
                float v[10];
#pragma unroll
                for (int k = 0; k < 10; ++k) v[k] = f[k][r];
                const Axis ax = axis(v[0], sc.nx), ay = axis(v[1], sc.ny);
                const int at0 = (ay.cell - by0) * w + (ax.cell - bx0);
                v[4] = v[4] * pscale;
                amounts(v, ax, ay, dt, bulk, rho0, [&](int k, float m, float mu, float mv) {
                    const int at = at0 + (k / 3) * w + k % 3;
                    atomicAdd(&tile[0][at], pair20(rintf(m * (mscale / pscale))));
                    atomicAdd(&tile[1][at], pair20(rintf(mu)));
                    atomicAdd(&tile[2][at], pair20(rintf(mv)));
                });
            }
        } else {
#pragma unroll
        for (int pass = 0; pass < kOwn / kRun; ++pass) {
            double acc[27];  // the run's fixed-point sums, exact integers in double
#pragma unroll
            for (int i = 0; i < 27; ++i) acc[i] = 0.0;
            int rx = -1, ry = -1;
            auto flush = [&]() {
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const long long s0 = exact_ll(acc[3 * k]), s1 = exact_ll(acc[3 * k + 1]),
                                    s2 = exact_ll(acc[3 * k + 2]);
                    if (tiled) {
                        const int at = (ry - by0 + k / 3) * w + (rx - bx0 + k % 3);
                        atomicAdd(&tile[0][at], (unsigned long long)s0);
                        atomicAdd(&tile[1][at], (unsigned long long)s1);
                        atomicAdd(&tile[2][at], (unsigned long long)s2);
                    } else {
                        const int64_t node = (int64_t)(ry + k / 3) * sc.nx + rx + k % 3;
                        add_fixed(q0 + node, s0);
                        add_fixed(q1 + node, s1);
                        add_fixed(q2 + node, s2);
                    }
                }
#pragma unroll
                for (int i = 0; i < 27; ++i) acc[i] = 0.0;
            };
#pragma unroll
            for (int r = pass * kRun; r < (pass + 1) * kRun; ++r) {
                if (first + r >= live) break;
                float v[10];
#pragma unroll
                for (int k = 0; k < 10; ++k) v[k] = f[k][r];
                const Axis ax = axis(v[0], sc.nx), ay = axis(v[1], sc.ny);
                if (ax.cell != rx || ay.cell != ry) {
                    if (rx >= 0) flush();
                    rx = ax.cell;
                    ry = ay.cell;
                }
                // The amounts with the mass scaled by pscale = 2^32: every amount is a product,
                // quotient or sum of binary32 values with the mass as a factor, so each comes out
                // scaled by 2^32 exactly, and the momentum's quantisation needs no multiply (the
                // mass's, 2^40, is 2^8 more). Exact while nothing leaves the normal range; a
                // product small enough to lose bits there quantises to zero either way.
                v[4] = v[4] * pscale;
                amounts(v, ax, ay, dt, bulk, rho0, [&](int k, float m, float mu, float mv) {
                    acc[3 * k] += quantize_d(m, mscale / pscale);
                    acc[3 * k + 1] += (double)rintf(mu);
                    acc[3 * k + 2] += (double)rintf(mv);
                });
            }
            if (rx >= 0) flush();
        }
        }
        __syncthreads();
        if (tiled) {
            for (int i = threadIdx.x; i < w * h; i += kBlock) {
                unsigned long long v0 = tile[0][i], v1 = tile[1][i], v2 = tile[2][i];
                if (pairs) {
                    v0 = unpair20(v0);
                    v1 = unpair20(v1);
                    v2 = unpair20(v2);
                }
                if ((v0 | v1 | v2) == 0ull) continue;
                const int64_t node = (int64_t)(by0 + i / w) * sc.nx + bx0 + i % w;
                atomicAdd(q0 + node, v0);
                atomicAdd(q1 + node, v1);
                atomicAdd(q2 + node, v2);
            }
        }
        __syncthreads();  // the tile and box serve the next chunk
    }
}

// The visits of clear, grid and decode. A block takes a window of kW units (kW = 64: 16 chunks,
// 4,096 nodes): wave 0 reads the window's flags with one load a lane and lists its units flagged
// touched or live in LDS, and wave 1 reads the scenes of the window's chunks (when any of its
// flags is set) into an LDS table. Then the block's waves split the listed units, a node a lane,
// four units a wave in flight at once. A window with no flag costs one round trip, and no
// chunk's scene chain (chunk -> scene -> record -> parameters) stands between the flags and the
// sums' reads. Grids of fewer than kSmallWindows such windows take windows of 16 units instead,
// four times the blocks, so that a small scene's few windows are not each a long serial walk.
constexpr int kWin2 = 64;
constexpr int kWinChunks = kWin2 * kUnit2 / kBlock;
constexpr int kWaves2 = kBlock / kWave;
constexpr int kSmallWindows = 2048;
constexpr int kSmallChunks = 256;  // particle chunks below which P2G takes a quarter chunk a block
static_assert(kWin2 == kWave, "a window's flags are one wave's lanes");

struct WinScenes {
    int64_t goff[kWinChunks];
    int nx[kWinChunks], ny[kWinChunks];
    Law2 law[kWinChunks];
};

__device__ __forceinline__ int64_t units2(const Batch2Args& a) { return a.gchunks * (kBlock / kUnit2); }

// Wave 0: the window's units flagged touched or live, as unit numbers within the window in order,
// into list[0, *count). Returns the lane's unit's touched flag.
template <int kW>
__device__ __forceinline__ uint8_t window_units(const Batch2Args& a, int64_t win, int lane, int* list, int* count) {
    const int64_t u = win * kW + lane;
    uint8_t t = 0, l = 0;
    if (lane < kW && u < units2(a)) {
        t = a.touched[u];
        l = a.live[u];
    }
    const bool on = (t | l) != 0;
    const unsigned long long mask = __ballot(on);
    if (on) list[__popcll(mask & ((1ull << lane) - 1ull))] = lane;
    if (lane == 0) *count = __popcll(mask);
    return t;
}

// Wave 1: the scenes of the window's chunks, when any of the window's flags is set.
template <int kW>
__device__ __forceinline__ void window_scenes(const Batch2Args& a, int64_t win, int lane, WinScenes& ws) {
    constexpr int kChunks = kW * kUnit2 / kBlock;
    const int64_t c = win * kChunks + lane;
    const bool in = lane < kChunks && c < a.gchunks;
    uint32_t f = 0;
    if (in) f = ((const uint32_t*)a.touched)[c] | ((const uint32_t*)a.live)[c];  // a chunk's 4 flags
    if (__ballot(f != 0) == 0ull || !in) return;
    const int s = a.gscene[c];
    const Scene2 sc = a.scenes[s];
    ws.goff[lane] = sc.goff;
    ws.nx[lane] = sc.nx;
    ws.ny[lane] = sc.ny;
    ws.law[lane] = law2(a.params + 28 * (int64_t)s);
}

// Zeroes the sums of every unit flagged touched or live; the others hold zero sums already.
template <int kW>
__global__ __launch_bounds__(kBlock) void batch2_clear_kernel(Batch2Args a) {
    __shared__ int list[kWin2];
    __shared__ int count;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t windows = (units2(a) + kW - 1) / kW;
    for (int64_t win = blockIdx.x; win < windows; win += gridDim.x) {
        if (wave == 0) window_units<kW>(a, win, lane, list, &count);
        __syncthreads();
        const int n = count;
        for (int i = wave; i < n; i += kWaves2) {
            const int64_t at = (win * kW + list[i]) * kUnit2 + lane;
            a.q[0][at] = 0ull;
            a.q[1][at] = 0ull;
            a.q[2][at] = 0ull;
        }
        __syncthreads();  // list and count serve the next window
    }
}

// After a grid pass: live := touched, touched := 0, a chunk's four unit flags as one word.
__global__ void batch2_flags_kernel(uint32_t* __restrict__ touched, uint32_t* __restrict__ live, int64_t chunks) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= chunks) return;
    live[c] = touched[c];
    touched[c] = 0;
}

// kVelocity: the sums decoded into gm gmu gmv, then Flip.grid on them, over the units flagged
// touched or live. A live unit P2G did not reach again holds zero sums, so it gets the zero
// velocity the full pass would give it; a unit in neither set already holds zero sums and
// velocities. Each window's flags then move on (live := touched, touched := 0), by the block that
// read them. Without kVelocity, the decoding alone, flags unchanged: for passes that read the mass
// before the grid update.
template <bool kVelocity, int kW>
__global__ __launch_bounds__(kBlock) void batch2_grid_kernel(Batch2Args a) {
    __shared__ int list[kWin2];
    __shared__ int count;
    __shared__ WinScenes ws;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t windows = (units2(a) + kW - 1) / kW;
    for (int64_t win = blockIdx.x; win < windows; win += gridDim.x) {
        uint8_t touched = 0;
        if (wave == 0) touched = window_units<kW>(a, win, lane, list, &count);
        else if (wave == 1) window_scenes<kW>(a, win, lane, ws);
        __syncthreads();
        if (kVelocity && wave == 0) {
            // the flags move on once both waves have read them: live := touched, touched := 0
            const int64_t u = win * kW + lane;
            if (lane < kW && u < units2(a)) {
                a.live[u] = touched;
                a.touched[u] = 0;
            }
        }
        const int n = count;
        // wave w takes listed units w, w + 4, w + 8, w + 12, then the next 16: the sums of all four
        // read before the first store (a short last group reads its last unit again)
        for (int i0 = wave; i0 < n; i0 += 4 * kWaves2) {
            unsigned long long v[4][3];
            int64_t at[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = min(i0 + k * kWaves2, n - 1);
                at[k] = (win * kW + list[i]) * kUnit2 + lane;
#pragma unroll
                for (int f = 0; f < 3; ++f) v[k][f] = a.q[f][at[k]];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + k * kWaves2;
                if (i >= n) break;
                const int w = list[i] / (kBlock / kUnit2);  // the unit's chunk within the window
                const int64_t goff = ws.goff[w];
                const int nx = ws.nx[w], ny = ws.ny[w];
                const int64_t node = at[k] - goff;
                if (node >= (int64_t)nx * ny) continue;  // the padding past the scene's last node
                const float m = dequantize(v[k][0], kMassBits);
                const float mu = dequantize(v[k][1], kMomentumBits);
                const float mv = dequantize(v[k][2], kMomentumBits);
                a.g[0][at[k]] = m;
                a.g[1][at[k]] = mu;
                a.g[2][at[k]] = mv;
                if (kVelocity) grid2_node(node, nx, ws.law[w], m, mu, mv, ny, a.g[3] + goff, a.g[4] + goff);
            }
        }
        __syncthreads();  // list, count and the scene table serve the next window
    }
}

// G2P over a batch. A block takes a chunk of kBatchAlign particles (one scene's), a thread every
// kBlock-th of them. The box of nodes their stencils reach is staged from gu and gv into LDS,
// every load of it in flight before the first LDS store; each particle then reads its 9 nodes
// from the box (from the grid when the box does not fit: unsorted particles) and stores its
// fields, a particle at a time.
constexpr int kTileG2 = 2048;  // gu and gv of a chunk's box: 16 KB
__global__ __launch_bounds__(kBlock) void batch2_g2p_kernel(Batch2Args a) {
    __shared__ float tile[2][kTileG2];
    __shared__ int box[4];
    constexpr int kPer = kBatchAlign / kBlock;
    constexpr int kLoads = kTileG2 / kBlock;
    for (int64_t pc = blockIdx.x; pc < a.pchunks; pc += gridDim.x) {
        // positions and J beside the scene's number: the chunk lies inside its scene's range
        // (padded to whole chunks), so these reads need nothing from the scene
        float px[kPer], py[kPer], pj[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int64_t q = pc * kBatchAlign + k * kBlock + threadIdx.x;
            px[k] = a.p[0][q];
            py[k] = a.p[1][q];
            pj[k] = a.p[5][q];
        }
        const int s = a.pscene[pc];
        const Scene2 sc = a.scenes[s];
        const float dt = bits_f(a.params[28 * (int64_t)s + DT]);
        const int64_t count = sc.poff + sc.count - pc * kBatchAlign;  // the chunk's live particles
        const float* gu = a.g[3] + sc.goff;
        const float* gv = a.g[4] + sc.goff;

        if (threadIdx.x == 0) {
            box[0] = box[1] = 0x7fffffff;
            box[2] = box[3] = -1;
        }
        __syncthreads();
        int x0 = 0x7fffffff, y0 = 0x7fffffff, x1 = -1, y1 = -1;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (k * kBlock + threadIdx.x >= count) break;
            const int cx = axis(px[k], sc.nx).cell, cy = axis(py[k], sc.ny).cell;
            x0 = min(x0, cx);
            y0 = min(y0, cy);
            x1 = max(x1, cx);
            y1 = max(y1, cy);
        }
        x0 = wave_min(x0);
        y0 = wave_min(y0);
        x1 = wave_max(x1);
        y1 = wave_max(y1);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            atomicMin(&box[0], x0);
            atomicMin(&box[1], y0);
            atomicMax(&box[2], x1);
            atomicMax(&box[3], y1);
        }
        __syncthreads();
        const int bx0 = box[0], by0 = box[1];
        const int w = box[2] - bx0 + 3, h = box[3] - by0 + 3;
        const bool tiled = box[2] >= 0 && w * h <= kTileG2;
        if (tiled) {
            // node i of the box: row floor((i + 1/2) / w) by a float reciprocal, exact for i < 2^12
            const float inv_w = 1.0f / (float)w;
            float vu[kLoads], vv[kLoads];
#pragma unroll
            for (int n = 0; n < kLoads; ++n) {
                const int i = threadIdx.x + n * kBlock;
                if (i >= w * h) break;
                const int row = (int)(((float)i + 0.5f) * inv_w);
                const int64_t node = (int64_t)(by0 + row) * sc.nx + bx0 + (i - row * w);
                vu[n] = gu[node];
                vv[n] = gv[node];
            }
#pragma unroll
            for (int n = 0; n < kLoads; ++n) {
                const int i = threadIdx.x + n * kBlock;
                if (i >= w * h) break;
                tile[0][i] = vu[n];
                tile[1][i] = vv[n];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (k * kBlock + threadIdx.x >= count) break;
            const int64_t q = pc * kBatchAlign + k * kBlock + threadIdx.x;
            G2p2In in;
            in.x = px[k];
            in.y = py[k];
            in.j = pj[k];
            if (tiled) {
                in.ax = axis(in.x, sc.nx);
                in.ay = axis(in.y, sc.ny);
                const int at = (in.ay.cell - by0) * w + (in.ax.cell - bx0);
#pragma unroll
                for (int n = 0; n < 9; ++n) {
                    in.gu[n] = tile[0][at + (n / 3) * w + n % 3];
                    in.gv[n] = tile[1][at + (n / 3) * w + n % 3];
                }
            } else {
                g2p2_nodes(sc.nx, sc.ny, gu, gv, in);
            }
            g2p2_finish(a.p, q, sc.nx, sc.ny, dt, in);
        }
        __syncthreads();  // the tile and box serve the next chunk
    }
}

// key = scene << cell_bits | Scatter.cell; a pad particle takes its scene's cell count, so it
// sorts after the scene's particles and before the next scene's.
// A block a chunk of kBatchAlign particles (one scene's), a thread kBatchAlign / kBlock of them,
// every position read before the first key is written.
__global__ __launch_bounds__(kBlock) void batch2_keys_kernel(Batch2Args a, int32_t* __restrict__ keys,
                                                             int32_t* __restrict__ index) {
    constexpr int kPer = kBatchAlign / kBlock;
    for (int64_t pc = blockIdx.x; pc < a.pchunks; pc += gridDim.x) {
        // positions beside the scene's number: the chunk lies inside its scene's padded range
        float x[kPer], y[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int64_t q = pc * kBatchAlign + k * kBlock + threadIdx.x;
            x[k] = a.p[0][q];
            y[k] = a.p[1][q];
        }
        const int s = a.pscene[pc];
        const Scene2 sc = a.scenes[s];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int64_t q = pc * kBatchAlign + k * kBlock + threadIdx.x;
            const int32_t cell = q - sc.poff < sc.count ? cell2(x[k], y[k], sc.nx, sc.ny) : sc.cells;
            keys[q] = (int32_t)(((uint32_t)s << a.cell_bits) | (uint32_t)cell);
            index[q] = (int32_t)q;
        }
    }
}

// starts[soff + c] = the first of the scene's sorted particles in cell c or after it, counted
// from the scene's first particle (Scatter.cellStarts, Scatter.java:484-502). Two passes over
// chunks of kBlock cells. The first writes each chunk's bound (its scene's first particle whose
// key is at least the chunk's first cell's): a particle whose chunk differs from its
// predecessor's bounds every chunk from the one past the predecessor's to its own, and the
// scene's last particle bounds the chunks after its own at the scene's end. The second takes
// kStartsGroup chunks to a block: the first particle of each run of equal keys between a chunk's
// bound and the next chunk's records its index in an LDS table by cell, and a cell's start is
// the least index recorded at that cell or after it (the run count past the last), a suffix
// minimum across the block. Both passes are chains of dependent loads, so each wave takes
// several independent items to keep more loads in flight.
constexpr int kStartsGroup = 4;

__global__ __launch_bounds__(kBlock) void batch2_bounds_kernel(Batch2Args a, const int32_t* __restrict__ keys) {
    constexpr int kPer = kBatchAlign / kBlock;
    const uint32_t mask = (1u << a.cell_bits) - 1u;
    for (int64_t pc = blockIdx.x; pc < a.pchunks; pc += gridDim.x) {
        // the keys beside the scene lookup: the chunk lies inside its scene's padded range
        const int64_t base = pc * kBatchAlign + threadIdx.x;
        int32_t key[kPer], before[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int64_t q = base + j * kBlock;
            key[j] = keys[q];
            before[j] = q > 0 ? keys[q - 1] : 0;  // used for a scene's particles after its first
        }
        const Scene2 sc = a.scenes[a.pscene[pc]];  // a particle chunk lies in one scene
        const int64_t c0 = sc.soff / kBlock;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int64_t q = base + j * kBlock;
            const int64_t i = q - sc.poff;
            if (i >= sc.count) continue;
            const int64_t mine = (int64_t)(((uint32_t)key[j] & mask) / kBlock);
            const int64_t prev = i > 0 ? (int64_t)(((uint32_t)before[j] & mask) / kBlock) : -1;
            for (int64_t k = prev + 1; k <= mine; ++k) a.bounds[c0 + k] = (int32_t)q;
            if (i == sc.count - 1) {
                const int64_t chunks = ((int64_t)sc.cells + kBlock) / kBlock;  // cells + 1 starts
                for (int64_t k = mine + 1; k < chunks; ++k) a.bounds[c0 + k] = (int32_t)(sc.poff + sc.count);
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void batch2_starts_kernel(Batch2Args a, const int32_t* __restrict__ keys,
                                                               int32_t* __restrict__ starts) {
    constexpr int kWaves = kBlock / 64;
    __shared__ int32_t head[kStartsGroup][kBlock];
    __shared__ int32_t wave_min[kStartsGroup][kWaves];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t groups = (a.schunks + kStartsGroup - 1) / kStartsGroup;
    for (int64_t g = blockIdx.x; g < groups; g += gridDim.x) {
        int lo[kStartsGroup], at[kStartsGroup + 1], rel[kStartsGroup], key0[kStartsGroup], valid[kStartsGroup];
        at[0] = 0;
#pragma unroll
        for (int j = 0; j < kStartsGroup; ++j) {
            const int64_t c = g * kStartsGroup + j;
            int n = 0;
            if (c < a.schunks) {
                const int4 info = reinterpret_cast<const int4*>(a.sinfo)[c];
                const bool last = (info.z >> 16) & 1;
                const bool none = info.w == info.y;  // a scene without particles has no bounds written
                lo[j] = none ? info.y : a.bounds[c];
                n = (last || none ? info.w : a.bounds[c + 1]) - lo[j];
                rel[j] = lo[j] - info.y;
                key0[j] = info.x;
                valid[j] = info.z & 0xffff;
            } else {
                lo[j] = rel[j] = key0[j] = valid[j] = 0;
            }
            at[j + 1] = at[j] + n;
            head[j][t] = n;
        }
        __syncthreads();
        // the group's keys as one run of at[kStartsGroup]: index idx lies in the last chunk whose
        // offset is at most idx (empty chunks share their successor's offset)
        for (int idx = t; idx < at[kStartsGroup]; idx += kBlock) {
            int j = 0, from = lo[0], off = 0, k0 = key0[0];
#pragma unroll
            for (int jj = 1; jj < kStartsGroup; ++jj)
                if (idx >= at[jj]) j = jj, from = lo[jj], off = at[jj], k0 = key0[jj];
            const int i = idx - off;
            const int32_t key = keys[from + i];
            if (i == 0 || keys[from + i - 1] != key) head[j][key - k0] = i;
        }
        __syncthreads();
        int v[kStartsGroup];
#pragma unroll
        for (int j = 0; j < kStartsGroup; ++j) {
            v[j] = head[j][t];
            // suffix minimum over the wave: a lane past the last reads its own value
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) v[j] = min(v[j], __shfl_down(v[j], d));
            if (lane == 0) wave_min[j][wave] = v[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kStartsGroup; ++j) {
#pragma unroll
            for (int w = 1; w < kWaves; ++w)
                if (w > wave) v[j] = min(v[j], wave_min[j][w]);
            if (t < valid[j]) starts[(g * kStartsGroup + j) * kBlock + t] = rel[j] + v[j];
        }
        __syncthreads();  // the next group rewrites head and wave_min
    }
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int mpm2_clear(const Launch& l, const Mpm2Args& a) {
    const int64_t nodes = (int64_t)a.nx * a.ny;
    int64_t blocks = blocks_for(nodes, kBlock);
    if (blocks > 8192) blocks = 8192;
    timed_begin(l, "mpm2_clear");
    hipLaunchKernelGGL(clear3_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, l.stream, a.g[0], a.g[1], a.g[2], nodes);
    timed_end(l, "mpm2_clear");
    return check_launch("mpm2_clear");
}

int mpm2_scatter_atomic(const Launch& l, const Mpm2Args& a) {
    if (a.count == 0) return 0;
    timed_begin(l, "p2g2_atomic");
    hipLaunchKernelGGL(p2g2_atomic_kernel, dim3(blocks_for(a.count, kBlock)), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "p2g2_atomic");
    return check_launch("p2g2_atomic");
}

int mpm2_scatter_runs(const Launch& l, const Mpm2Args& a, int run) {
    if (a.count == 0) return 0;
    timed_begin(l, "p2g2_runs");
    hipLaunchKernelGGL(p2g2_runs_kernel, dim3(blocks_for(a.count, (int64_t)kBlock * run)), dim3(kBlock), 0, l.stream,
                       a, run);
    timed_end(l, "p2g2_runs");
    return check_launch("p2g2_runs");
}

int mpm2_grid(const Launch& l, const Mpm2Args& a) {
    timed_begin(l, "grid2");
    hipLaunchKernelGGL(grid2_kernel, dim3(blocks_for((int64_t)a.nx * a.ny, kBlock)), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "grid2");
    return check_launch("grid2");
}

int mpm2_advect(const Launch& l, const Mpm2Args& a) {
    if (a.count == 0) return 0;
    timed_begin(l, "g2p2");
    hipLaunchKernelGGL(g2p2_kernel, dim3(blocks_for(a.count, kBlock)), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "g2p2");
    return check_launch("g2p2");
}

int mpm2_keys(const Launch& l, const Mpm2Args& a, int32_t* keys, int32_t* index) {
    const int32_t pad = (a.nx - 1) * (a.ny - 1);  // one past the last cell
    timed_begin(l, "keys2");
    hipLaunchKernelGGL(keys2_kernel, dim3(blocks_for(a.capacity, kBlock)), dim3(kBlock), 0, l.stream, a, keys, index,
                       pad);
    timed_end(l, "keys2");
    return check_launch("keys2");
}

int mpm2_starts(const Launch& l, const int32_t* sorted_keys, int64_t count, int32_t cells, int32_t* starts) {
    timed_begin(l, "starts2");
    hipLaunchKernelGGL(starts_kernel, dim3(blocks_for((int64_t)cells + 1, kBlock)), dim3(kBlock), 0, l.stream,
                       sorted_keys, count, cells, starts);
    timed_end(l, "starts2");
    return check_launch("starts2");
}

// ------------------------------------------------------------------ batch launchers

namespace {
unsigned grid_of(int64_t chunks, int blocks) {
    if (chunks <= 0) return 0;
    return (unsigned)(blocks > 0 && blocks < chunks ? blocks : chunks);
}
// Windows of kW flags over the grid's units, which the clear, grid and decode blocks stride over:
// 64 a window, or 16 for grids of fewer than kSmallWindows windows of 64.
int window2(const Batch2Args& a) {
    return (a.gchunks * (kBlock / kUnit2) + kWin2 - 1) / kWin2 < kSmallWindows ? 16 : kWin2;
}
int64_t windows2(const Batch2Args& a) {
    const int w = window2(a);
    return (a.gchunks * (kBlock / kUnit2) + w - 1) / w;
}
}  // namespace

int batch2_clear(const Launch& l, const Batch2Args& a, int blocks) {
    const unsigned g = grid_of(windows2(a), blocks);
    if (!g) return 0;
    timed_begin(l, "batch2_clear");
    if (window2(a) == 16) hipLaunchKernelGGL(batch2_clear_kernel<16>, dim3(g), dim3(kBlock), 0, l.stream, a);
    else hipLaunchKernelGGL(batch2_clear_kernel<kWin2>, dim3(g), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "batch2_clear");
    return check_launch("batch2_clear");
}

int batch2_scatter(const Launch& l, const Batch2Args& a, int run, int blocks) {
    const unsigned g = grid_of(a.pchunks, blocks);
    if (!g) return 0;
    timed_begin(l, "batch2_p2g");
    if (a.pchunks < kSmallChunks)
        hipLaunchKernelGGL((batch2_p2g_kernel<1, 1>), dim3(grid_of(a.pchunks * 4, blocks)), dim3(kBlock), 0,
                           l.stream, a);
    else if (run == 4)
        hipLaunchKernelGGL((batch2_p2g_kernel<4, 4>), dim3(g), dim3(kBlock), 0, l.stream, a);
    else if (run == 2)
        hipLaunchKernelGGL((batch2_p2g_kernel<2, 4>), dim3(g), dim3(kBlock), 0, l.stream, a);
    else
        hipLaunchKernelGGL((batch2_p2g_kernel<1, 4>), dim3(g), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "batch2_p2g");
    return check_launch("batch2_p2g");
}

int batch2_grid(const Launch& l, const Batch2Args& a, int blocks) {
    const unsigned g = grid_of(windows2(a), blocks);
    if (!g) return 0;
    timed_begin(l, "batch2_grid");
    if (window2(a) == 16) hipLaunchKernelGGL((batch2_grid_kernel<true, 16>), dim3(g), dim3(kBlock), 0, l.stream, a);
    else hipLaunchKernelGGL((batch2_grid_kernel<true, kWin2>), dim3(g), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "batch2_grid");
    return check_launch("batch2_grid");
}

int batch2_decode(const Launch& l, const Batch2Args& a, int blocks) {
    const unsigned g = grid_of(windows2(a), blocks);
    if (!g) return 0;
    timed_begin(l, "batch2_decode");
    if (window2(a) == 16) hipLaunchKernelGGL((batch2_grid_kernel<false, 16>), dim3(g), dim3(kBlock), 0, l.stream, a);
    else hipLaunchKernelGGL((batch2_grid_kernel<false, kWin2>), dim3(g), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "batch2_decode");
    return check_launch("batch2_decode");
}

int batch2_flags(const Launch& l, const Batch2Args& a) {
    if (a.gchunks <= 0) return 0;
    hipLaunchKernelGGL(batch2_flags_kernel, dim3((unsigned)((a.gchunks + 255) / 256)), dim3(256), 0, l.stream,
                       (uint32_t*)a.touched, (uint32_t*)a.live, a.gchunks);
    return check_launch("batch2_flags");
}

int batch2_advect(const Launch& l, const Batch2Args& a, int blocks) {
    const unsigned g = grid_of(a.pchunks, blocks);
    if (!g) return 0;
    timed_begin(l, "batch2_g2p");
    hipLaunchKernelGGL(batch2_g2p_kernel, dim3(g), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "batch2_g2p");
    return check_launch("batch2_g2p");
}

int batch2_keys(const Launch& l, const Batch2Args& a, int32_t* keys, int32_t* index, int blocks) {
    const unsigned g = grid_of(a.pchunks, blocks);
    if (!g) return 0;
    timed_begin(l, "batch2_keys");
    hipLaunchKernelGGL(batch2_keys_kernel, dim3(g), dim3(kBlock), 0, l.stream, a, keys, index);
    timed_end(l, "batch2_keys");
    return check_launch("batch2_keys");
}

int batch2_starts(const Launch& l, const Batch2Args& a, const int32_t* keys, int32_t* starts, int blocks) {
    const unsigned g = grid_of(a.schunks, blocks);
    if (!g) return 0;
    timed_begin(l, "batch2_starts");
    hipLaunchKernelGGL(batch2_bounds_kernel, dim3(grid_of(a.pchunks, blocks)), dim3(kBlock), 0, l.stream, a, keys);
    hipLaunchKernelGGL(batch2_starts_kernel, dim3(grid_of((a.schunks + kStartsGroup - 1) / kStartsGroup, blocks)),
                       dim3(kBlock), 0, l.stream, a, keys, starts);
    timed_end(l, "batch2_starts");
    return check_launch("batch2_starts");
}

}  // namespace vxr
