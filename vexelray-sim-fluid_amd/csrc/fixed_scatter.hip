// The fixed-point bilinear particle-to-grid scatter, core/particle/FixedScatter.java:151-228, for
// gfx950: a particle's mass and velocity in quanta, its offset in its cell in 1/1024ths, four
// integer corner shares summing exactly to its mass, momentum as share times velocity, all added
// by 32-bit integer atomics that wrap. Integer sums associate, so every schedule and every order
// gives the grid bit for bit, and that grid is the host's (test/particle/FixedScatterTest.java:
// 402-432). Two schedules: one particle a thread with 12 atomics (FixedScatter.kernel, :151-162),
// and runs of consecutive particles sharing a cell folded in registers first, one atomic per run,
// corner and field (the sums FixedScatter.segmented takes across a subgroup, :169-183).
#include "common.h"

namespace vxr {
namespace {

constexpr int kBlock = 256;
constexpr int kOffsetBits = 10;  // FixedScatter.OFFSET_BITS
constexpr int kOne = 1 << kOffsetBits;

// The IR's toInt as Java's (int): toward zero, NaN to 0, saturating.
__device__ __forceinline__ int java_f2i(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 0x7fffffff;
    if (f <= -2147483648.0f) return (int)0x80000000;
    return (int)f;
}
__device__ __forceinline__ int wmul(int a, int b) { return (int)((unsigned)a * (unsigned)b); }
__device__ __forceinline__ int wadd(int a, int b) { return (int)((unsigned)a + (unsigned)b); }

struct Quanta {
    int corner;     // the cell's lower-left node
    int share[4];   // corners: bit 0 east, bit 1 north
    int uq, vq;
};

// Scatter.Particle.load (Scatter.java:547-556) and FixedScatter.Quanta.of (FixedScatter.java:196-216).
__device__ __forceinline__ Quanta quanta(int n, float x, float y, float u, float v, float m, float mscale,
                                         float vscale) {
    Quanta qt;
    const float cx = mn(mx(x, 0.0f), (float)n - 1.0f);
    const float cy = mn(mx(y, 0.0f), (float)n - 1.0f);
    const int col = java_f2i(mn(cx, (float)n - 2.0f));
    const int row = java_f2i(mn(cy, (float)n - 2.0f));
    const float fx = cx - (float)col, fy = cy - (float)row;
    const int mass = java_f2i(m * mscale);
    qt.uq = java_f2i(u * vscale);
    qt.vq = java_f2i(v * vscale);
    const int ox = java_f2i(fx * (float)kOne), oy = java_f2i(fy * (float)kOne);
    const int high = mass / kOne, low = mass % kOne;
    int left = mass;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int q = wmul((k & 1) ? ox : kOne - ox, (k >> 1) ? oy : kOne - oy);
        const int share = wadd(wmul(q, high), wmul(q, low) / kOne) / kOne;
        qt.share[k] = share;
        left = (int)((unsigned)left - (unsigned)share);
    }
    qt.share[3] = left;
    qt.corner = row * n + col;
    return qt;
}

__device__ __forceinline__ void add(int* p, int v) { atomicAdd(p, v); }

__global__ __launch_bounds__(kBlock) void fixed_direct_kernel(int n, int64_t count, const float* __restrict__ x,
                                                              const float* __restrict__ y, const float* __restrict__ u,
                                                              const float* __restrict__ v, const float* __restrict__ m,
                                                              float mscale, float vscale, int* __restrict__ gm,
                                                              int* __restrict__ gmu, int* __restrict__ gmv) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= count) return;
    const Quanta qt = quanta(n, x[p], y[p], u[p], v[p], m[p], mscale, vscale);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int node = qt.corner + (k >> 1) * n + (k & 1);
        add(gm + node, qt.share[k]);
        add(gmu + node, wmul(qt.share[k], qt.uq));
        add(gmv + node, wmul(qt.share[k], qt.vq));
    }
}

__global__ __launch_bounds__(kBlock) void fixed_runs_kernel(int n, int64_t count, const float* __restrict__ x,
                                                            const float* __restrict__ y, const float* __restrict__ u,
                                                            const float* __restrict__ v, const float* __restrict__ m,
                                                            float mscale, float vscale, int run,
                                                            int* __restrict__ gm, int* __restrict__ gmu,
                                                            int* __restrict__ gmv) {
    const int64_t first = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * run;
    int acc[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) acc[i] = 0;
    int corner = -1;
    auto flush = [&]() {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int node = corner + (k >> 1) * n + (k & 1);
            add(gm + node, acc[3 * k]);
            add(gmu + node, acc[3 * k + 1]);
            add(gmv + node, acc[3 * k + 2]);
        }
#pragma unroll
        for (int i = 0; i < 12; ++i) acc[i] = 0;
    };
    for (int r = 0; r < run; ++r) {
        const int64_t p = first + r;
        if (p >= count) break;
        const Quanta qt = quanta(n, x[p], y[p], u[p], v[p], m[p], mscale, vscale);
        if (qt.corner != corner) {
            if (corner >= 0) flush();
            corner = qt.corner;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc[3 * k] = wadd(acc[3 * k], qt.share[k]);
            acc[3 * k + 1] = wadd(acc[3 * k + 1], wmul(qt.share[k], qt.uq));
            acc[3 * k + 2] = wadd(acc[3 * k + 2], wmul(qt.share[k], qt.vq));
        }
    }
    if (corner >= 0) flush();
}

}  // namespace

int fixed_scatter(const Launch& l, int n, int64_t count, const float* x, const float* y, const float* u,
                  const float* v, const float* m, int mass_bits, int velocity_bits, int run, int* gm, int* gmu,
                  int* gmv) {
    if (count <= 0) return 0;
    const float mscale = (float)ldexp(1.0, mass_bits), vscale = (float)ldexp(1.0, velocity_bits);
    if (run <= 0) {
        timed_begin(l, "fixed_direct");
        hipLaunchKernelGGL(fixed_direct_kernel, dim3((unsigned)((count + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           l.stream, n, count, x, y, u, v, m, mscale, vscale, gm, gmu, gmv);
        timed_end(l, "fixed_direct");
        return check_launch("fixed_direct");
    }
    const int64_t per = (int64_t)kBlock * run;
    timed_begin(l, "fixed_runs");
    hipLaunchKernelGGL(fixed_runs_kernel, dim3((unsigned)((count + per - 1) / per)), dim3(kBlock), 0, l.stream, n,
                       count, x, y, u, v, m, mscale, vscale, run, gm, gmu, gmv);
    timed_end(l, "fixed_runs");
    return check_launch("fixed_runs");
}

}  // namespace vxr
