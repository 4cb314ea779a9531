// 3D surface tension for gfx950: Tension3.blur, Tension3.stress, Tension3.force
// (core/particle/Tension3.java:69-196) and Flip3.gridWithTension (core/particle/Flip3.java:263-311),
// one node a thread, neighbour reads held to the wall ring, in the reference's binary32 operations
// and order (the CPU restatement's results bit for bit). These passes run over the whole grid: the
// blurred colour reaches past the nodes P2G touched.
#include "common.h"

namespace vxr {
namespace {

constexpr int kBlock = 256;
constexpr float kWall = 1.0f;
constexpr int kLow = 1;
constexpr float kMinColour = 0.3f, kMassFloor = 0.3f, kEmpty = 1e-9f;
enum { DT = 0, GX = 1, GY = 2, GZ = 3, BULK = 4, RHO0 = 5, SIGMA = 6 };

__device__ __forceinline__ float bits_f(int32_t w) { return __int_as_float(w); }

__device__ __forceinline__ int held(int c, int off, int size) {
    return (int)clampf((float)c + (float)off, kWall, (float)(size - 1) - kWall);
}

__device__ __forceinline__ int64_t held_at(const Mpm3Args& a, const int (&c)[3], int di, int dj, int dk) {
    const int x = held(c[0], di, a.nx), y = held(c[1], dj, a.ny), z = held(c[2], dk, a.nz);
    return ((int64_t)z * a.ny + y) * a.nx + x;
}

__device__ __forceinline__ void coords(const Mpm3Args& a, int64_t node, int (&c)[3]) {
    c[0] = (int)(node % a.nx);
    c[1] = (int)(node / a.nx % a.ny);
    c[2] = (int)(node / ((int64_t)a.nx * a.ny));
}

__global__ __launch_bounds__(kBlock) void blur_kernel(Mpm3Args a, const float* __restrict__ src,
                                                      float* __restrict__ dst, int scaled) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)a.nx * a.ny * a.nz) return;
    int c[3];
    coords(a, node, c);
    const int size[3] = {a.nx, a.ny, a.nz};
    const float rho0 = bits_f(a.params[RHO0]);
    float sum = 0.0f;
#pragma unroll
    for (int dk = -1; dk <= 1; dk++)
#pragma unroll
        for (int dj = -1; dj <= 1; dj++)
#pragma unroll
            for (int di = -1; di <= 1; di++) {
                const float weight = (float)((2 - abs(di)) * (2 - abs(dj)) * (2 - abs(dk)) / 64.0);
                float value = src[held_at(a, c, di, dj, dk)];
                if (scaled) {
                    float full = rho0;
                    const int d[3] = {di, dj, dk};
#pragma unroll
                    for (int ax = 0; ax < 3; ax++) {
                        const int h = held(c[ax], d[ax], size[ax]);
                        if (h == kLow) full = 0.5f * full;
                        if (h == size[ax] - 1 - kLow) full = 0.5f * full;
                    }
                    value = mn(value / full, 1.0f);
                }
                sum = sum + weight * value;
            }
    dst[node] = sum;
}

__global__ __launch_bounds__(kBlock) void stress_kernel(Mpm3Args a, const float* __restrict__ colour,
                                                        float* __restrict__ t0, float* __restrict__ t1,
                                                        float* __restrict__ t2, float* __restrict__ t3,
                                                        float* __restrict__ t4, float* __restrict__ t5) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)a.nx * a.ny * a.nz) return;
    int c[3];
    coords(a, node, c);
    const float gx = 0.5f * (colour[held_at(a, c, 1, 0, 0)] - colour[held_at(a, c, -1, 0, 0)]);
    const float gy = 0.5f * (colour[held_at(a, c, 0, 1, 0)] - colour[held_at(a, c, 0, -1, 0)]);
    const float gz = 0.5f * (colour[held_at(a, c, 0, 0, 1)] - colour[held_at(a, c, 0, 0, -1)]);
    const float xx = gx * gx, yy = gy * gy, zz = gz * gz;
    const float length = sqrtf((xx + (yy + zz)) + (float)(1e-4 * 1e-4));
    t0[node] = (yy + zz) / length;
    t1[node] = (xx + zz) / length;
    t2[node] = (xx + yy) / length;
    t3[node] = -((gx * gy) / length);
    t4[node] = -((gx * gz) / length);
    t5[node] = -((gy * gz) / length);
}

struct Stress {
    const float* t[6];  // txx tyy tzz txy txz tyz
};

__device__ __forceinline__ const float* component(const Stress& s, int a, int k) {
    // [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]
    const int idx = a == k ? a : (a + k == 1 ? 3 : (a + k == 2 ? 4 : 5));
    return s.t[idx];
}

__global__ __launch_bounds__(kBlock) void grid_tension_kernel(Mpm3Args a, Stress st,
                                                              const float* __restrict__ colour) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)a.nx * a.ny * a.nz) return;
    int c[3];
    coords(a, node, c);
    const int size[3] = {a.nx, a.ny, a.nz};
    const float mass = a.g[0][node];
    float vel[3] = {0.0f, 0.0f, 0.0f};
    if (mass > kEmpty) {
        const float dt = bits_f(a.params[DT]);
#pragma unroll
        for (int ax = 0; ax < 3; ax++) vel[ax] = a.g[1 + ax][node] / mass + dt * bits_f(a.params[GX + ax]);
        // Tension3.force: the flux of the stress across the node's six faces
        float force[3] = {0.0f, 0.0f, 0.0f};
        bool ring = colour[node] > kMinColour;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) ring = ring && !(c[ax] < kLow) && !(c[ax] > size[ax] - 1 - kLow);
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
#pragma unroll
            for (int sign = -1; sign <= 1; sign += 2) {
                int d[3] = {0, 0, 0};
                d[ax] = sign;
                const int64_t there = held_at(a, c, d[0], d[1], d[2]);
                const bool open = ring && colour[there] > kMinColour;
                const bool wall = sign < 0 ? !(c[ax] > kLow) : !(c[ax] < size[ax] - 1 - kLow);
                float shear = 1.0f;
                if (wall) shear = 0.0f;
                if (open) {
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        const float* comp = component(st, ax, k);
                        const float flux = (float)(0.5 * sign) * (comp[node] + comp[there]);
                        force[k] = force[k] + (k == ax ? flux : shear * flux);
                    }
                }
            }
        }
        const float push = (dt * bits_f(a.params[SIGMA])) / mx(mass, kMassFloor * bits_f(a.params[RHO0]));
#pragma unroll
        for (int ax = 0; ax < 3; ax++) vel[ax] = vel[ax] + push * force[ax];
    }
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        if (!(c[ax] > 1)) vel[ax] = mx(vel[ax], 0.0f);
        if (c[ax] > size[ax] - 3) vel[ax] = mn(vel[ax], 0.0f);
        a.g[4 + ax][node] = vel[ax];
    }
}

inline unsigned blocks(const Mpm3Args& a) {
    return (unsigned)(((int64_t)a.nx * a.ny * a.nz + kBlock - 1) / kBlock);
}

}  // namespace

int tension3_blur(const Launch& l, const Mpm3Args& a, const float* src, float* dst, bool scaled) {
    timed_begin(l, "tension3_blur");
    hipLaunchKernelGGL(blur_kernel, dim3(blocks(a)), dim3(kBlock), 0, l.stream, a, src, dst, scaled ? 1 : 0);
    timed_end(l, "tension3_blur");
    return check_launch("tension3_blur");
}

int tension3_stress(const Launch& l, const Mpm3Args& a, const float* colour, float* const t[6]) {
    timed_begin(l, "tension3_stress");
    hipLaunchKernelGGL(stress_kernel, dim3(blocks(a)), dim3(kBlock), 0, l.stream, a, colour, t[0], t[1], t[2], t[3],
                       t[4], t[5]);
    timed_end(l, "tension3_stress");
    return check_launch("tension3_stress");
}

int mpm3_grid_tension(const Launch& l, const Mpm3Args& a, float* const t[6], const float* colour) {
    Stress st;
    for (int k = 0; k < 6; ++k) st.t[k] = t[k];
    timed_begin(l, "grid3_tension");
    hipLaunchKernelGGL(grid_tension_kernel, dim3(blocks(a)), dim3(kBlock), 0, l.stream, a, st, colour);
    timed_end(l, "grid3_tension");
    return check_launch("grid3_tension");
}

}  // namespace vxr
