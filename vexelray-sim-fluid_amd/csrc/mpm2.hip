// The 2D MLS-MPM / APIC step (core/particle/Flip.java, FlipStep.java) for gfx950.
//
// Layout: 10 particle fields as f32 arrays (x y u v m j c00 c01 c10 c11), padded with m = 0
// to whole chunks; the grid as 5 f32 node arrays, row-major (gm gmu gmv gu gv).
//
// P2G: the reference sums each run of particles sharing a stencil base across a subgroup's
// lanes (Scatter.segmentedDeposit, Scatter.java:385-439) and the run's last lane takes the
// atomics. Here a thread walks a run of consecutive particles and folds those sharing a base
// into 27 register sums, with one atomic per node and field for each run it holds: the same
// sums regrouped, right in any order and cheap in cell order. The grid update and G2P keep
// the reference's arithmetic order, so from the same grid they are its results bit for bit.
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

// The 27 amounts of one particle (Flip.java:283-300), node k's three at 3k .. 3k+2.
__device__ __forceinline__ void amounts(const Mpm2Args& a, int64_t q, const Axis& ax, const Axis& ay, float dt,
                                        float bulk, float rho0, float (&acc)[27]) {
    const float m = a.p[4][q], u = a.p[2][q], w = a.p[3][q];
    const float c00 = a.p[6][q], c01 = a.p[7][q], c10 = a.p[8][q], c11 = a.p[9][q];
    const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - a.p[5][q]));
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float wgt = ax.w[k % 3] * ay.w[k / 3];
        const float wm = wgt * m;
        const float dx = (float)(k % 3) - ax.off;
        const float dy = (float)(k / 3) - ay.off;
        const float vu = u + (c00 * dx + c01 * dy);
        const float vv = w + (c10 * dx + c11 * dy);
        acc[3 * k] += wm;
        acc[3 * k + 1] += wm * vu + wgt * (s * dx);
        acc[3 * k + 2] += wm * vv + wgt * (s * dy);
    }
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
    amounts(a, q, ax, ay, dt, bulk, rho0, acc);
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
        amounts(a, q, ax, ay, dt, bulk, rho0, acc);
    }
    if (run_base >= 0) flush(run_base);
}

// Flip.grid, Flip.java:366-412 (weight 1: no buoyancy)
__global__ __launch_bounds__(kBlock) void grid2_kernel(Mpm2Args a) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)a.nx * a.ny) return;
    const int ni = (int)(node % a.nx), nj = (int)(node / a.nx);
    const float m = a.g[0][node];
    float u = 0.0f, w = 0.0f;
    if (m > kEmpty) {
        const float dt = bits_f(a.params[DT]);
        u = a.g[1][node] / m + dt * (bits_f(a.params[GX]) * 1.0f);
        w = a.g[2][node] / m + dt * (bits_f(a.params[GY]) * 1.0f);
    }
    if (!(ni > 1)) u = mx(u, 0.0f);
    if (ni > a.nx - 3) u = mn(u, 0.0f);
    if (!(nj > 1)) w = mx(w, 0.0f);
    if (nj > a.ny - 3) w = mn(w, 0.0f);
    a.g[3][node] = u;
    a.g[4][node] = w;
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

// Flip.advectBody, Flip.java:494-545 (not relaxing)
__global__ __launch_bounds__(kBlock) void g2p2_kernel(Mpm2Args a) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.count) return;
    const Axis ax = axis(a.p[0][q], a.nx), ay = axis(a.p[1][q], a.ny);
    const int64_t base = (int64_t)ay.cell * a.nx + ax.cell;
    float u = 0, w = 0, b00 = 0, b01 = 0, b10 = 0, b11 = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float wgt = ax.w[k % 3] * ay.w[k / 3];
        const float dx = (float)(k % 3) - ax.off;
        const float dy = (float)(k / 3) - ay.off;
        const int64_t node = base + (k / 3) * a.nx + k % 3;
        const float ui = wgt * a.g[3][node];
        const float vi = wgt * a.g[4][node];
        u = u + ui;
        w = w + vi;
        b00 = b00 + ui * dx;
        b01 = b01 + ui * dy;
        b10 = b10 + vi * dx;
        b11 = b11 + vi * dy;
    }
    const float c00 = 4.0f * b00, c01 = 4.0f * b01, c10 = 4.0f * b10, c11 = 4.0f * b11;
    const float dt = bits_f(a.params[DT]);
    float x = a.p[0][q] + dt * u;
    float y = a.p[1][q] + dt * w;
    keep_inside(x, u, a.nx);
    keep_inside(y, w, a.ny);
    a.p[0][q] = x;
    a.p[1][q] = y;
    a.p[2][q] = u;
    a.p[3][q] = w;
    a.p[6][q] = c00;
    a.p[7][q] = c01;
    a.p[8][q] = c10;
    a.p[9][q] = c11;
    const float j = a.p[5][q] * (1.0f + dt * (c00 + c11));
    a.p[5][q] = clampf(j, 1.0f / kJBound, kJBound);
}

// Scatter.cell, Scatter.java:473-477
__global__ __launch_bounds__(kBlock) void keys2_kernel(Mpm2Args a, int32_t* __restrict__ keys,
                                                       int32_t* __restrict__ index, int32_t pad_key) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.capacity) return;
    int32_t key = pad_key;
    if (q < a.count) {
        const float cx = clampf(a.p[0][q], 0.0f, (float)a.nx - 1.0f);
        const float cy = clampf(a.p[1][q], 0.0f, (float)a.ny - 1.0f);
        key = (int32_t)mn(cy, (float)a.ny - 2.0f) * (a.nx - 1) + (int32_t)mn(cx, (float)a.nx - 2.0f);
    }
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

}  // namespace vxr
