// Surface tension for gfx950, in the reference's binary32 operations and order (the CPU
// restatement's results bit for bit), neighbour reads held to the wall ring:
// - 3D: Tension3.blur, Tension3.stress, Tension3.force (core/particle/Tension3.java:69-196) and
//   Flip3.gridWithTension (core/particle/Flip3.java:263-311), below;
// - 2D: Tension.blur, Tension.stress, Tension.force (core/particle/Tension.java:98-185) and
//   Flip.gridWithTension (core/particle/Flip.java:342-412), at the end of the file: dense passes
//   over the grid, a node a thread (2D grids are small, and a pass is one launch).
//
// Where they run. The reference sweeps the whole grid each pass. Here the blur and stress passes
// visit, in each row of nodes, the span of bricks in the region: the 3x3x3-brick neighbourhood
// of every brick P2G touched or the grid pass left live, now or at the previous computation. The
// colour spreads one node a blur from the nodes with mass, which lie in touched or live bricks,
// and a node outside the region is at least 9 nodes from every such brick. The host counts the
// spread: while the colour has spread at most 8 nodes (at most 7 for the stress, which reads one
// node further), every node outside the region holds what the full pass would give it: +0 colour
// and diagonal stress, -0 off-diagonal stress (the buffers start at those values). Past that, or
// after the host wrote these buffers, a pass runs dense. The grid pass with tension visits the
// span of touched or live bricks, as the plain grid pass does.
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

// The x-range of row (y, z) from the first to the last brick whose flags, under `mask`, are
// nonzero (brick rows 64 at a time, by ballot). Returns false when there is none.
__device__ __forceinline__ bool span(const Mpm3Args& a, const uint8_t* __restrict__ f0,
                                     const uint8_t* __restrict__ f1, uint8_t mask, int y, int z, int lane, int* x0,
                                     int* x1) {
    const int bxn = (a.nx + kBrick - 1) / kBrick, byn = (a.ny + kBrick - 1) / kBrick;
    const int64_t base = ((int64_t)(z / kBrick) * byn + y / kBrick) * bxn;
    int first = INT_MAX, last = -1;
    for (int b0 = 0; b0 < bxn; b0 += kWave) {
        const int b = b0 + lane;
        const bool on = b < bxn && ((f0[base + b] | (f1 ? f1[base + b] : 0)) & mask) != 0;
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

// The region of the blur and stress passes, per brick: bit 0 = some brick in its 3x3x3
// neighbourhood is touched or live (the colour of this step reaches at most 7 nodes past those
// bricks), bit 1 = bit 0 now or at the previous computation (what was written then and must
// now be rewritten). With `all`, every brick.
__global__ void region_kernel(Mpm3Args a, int all) {
    const int bxn = (a.nx + kBrick - 1) / kBrick, byn = (a.ny + kBrick - 1) / kBrick,
              bzn = (a.nz + kBrick - 1) / kBrick;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= (int64_t)bxn * byn * bzn) return;
    const int bx = (int)(b % bxn), by = (int)(b / bxn % byn), bz = (int)(b / ((int64_t)bxn * byn));
    bool near = all != 0;
    for (int dz = -1; dz <= 1 && !near; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = bx + dx, y = by + dy, z = bz + dz;
                if (x < 0 || y < 0 || z < 0 || x >= bxn || y >= byn || z >= bzn) continue;
                const int64_t nb = ((int64_t)z * byn + y) * bxn + x;
                near = near || (a.touched[nb] | a.live[nb]) != 0;
            }
    const uint8_t old = a.near[b];
    a.near[b] = (uint8_t)((near ? 1 : 0) | ((near || (old & 1)) ? 2 : 0));
}

// Row bases of the nine held (y, z) neighbour rows of row (y, z): base[k][j] for dz = k - 1,
// dy = j - 1.
__device__ __forceinline__ void row_bases(const Mpm3Args& a, int y, int z, int64_t (&base)[3][3],
                                          int (&ys)[3], int (&zs)[3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) ys[j] = held(y, j - 1, a.ny);
#pragma unroll
    for (int k = 0; k < 3; ++k) zs[k] = held(z, k - 1, a.nz);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) base[k][j] = ((int64_t)zs[k] * a.ny + ys[j]) * a.nx;
}

// Tension3.blur. A wave takes one row y of one brick layer of z (8 values of z) and, for each
// 64 values of x in the span of that brick row in the region (or the whole row when dense),
// loads the 3x3 taps of the layer's ten planes of z at once (scaled on the way when `kScaled`);
// each plane serves as the dk = +1, 0 and -1 planes of three successive nodes. Each node sums
// its 27 taps in the reference's order, sum += weight * value, no contraction. Held reads make
// the planes of node z exactly held(z - 1), held(z), held(z + 1).
template <bool kScaled>
__device__ __forceinline__ void load_plane(const Mpm3Args& a, const float* __restrict__ src, int zp,
                                           const int (&ys)[3], const int (&xs)[3], float rho0, float (&p)[9]) {
    const int size[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        // the row's start (the same for the whole wave) and the lane's byte offset in it
        const char* row = reinterpret_cast<const char*>(src + ((int64_t)zp * a.ny + ys[j]) * a.nx);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float value = *reinterpret_cast<const float*>(row + (uint32_t)xs[i] * 4u);
            if (kScaled) {
                // full at a rest mass, halved for each axis on which the node read lies on the wall ring
                const int h[3] = {xs[i], ys[j], zp};
                float full = rho0;
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    if (h[ax] == kLow) full = 0.5f * full;
                    if (h[ax] == size[ax] - 1 - kLow) full = 0.5f * full;
                }
                value = mn(value / full, 1.0f);
            }
            p[3 * j + i] = value;
        }
    }
}

template <bool kScaled>
__global__ __launch_bounds__(kBlock) void blur_march_kernel(Mpm3Args a, const float* __restrict__ src,
                                                            float* __restrict__ dst, int dense) {
    const int lane = threadIdx.x & (kWave - 1);
    const int layers = (a.nz + kBrick - 1) / kBrick;
    // the wave's job, row and span held in scalar registers: the rows' addresses are then scalar
    const int job = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6)));
    if (job >= a.ny * layers) return;
    const int y = job % a.ny, z0 = (job / a.ny) * kBrick;
    const int z1 = min(z0 + kBrick, a.nz);
    int x0 = 0, x1 = a.nx;
    if (!dense && !span(a, a.near, nullptr, 2, y, z0, lane, &x0, &x1)) return;
    x0 = __builtin_amdgcn_readfirstlane(x0);
    x1 = __builtin_amdgcn_readfirstlane(x1);
    const int ys[3] = {held(y, -1, a.ny), held(y, 0, a.ny), held(y, 1, a.ny)};
    const float rho0 = bits_f(a.params[RHO0]);
    for (int xb = x0; xb < x1; xb += kWave) {
        // lanes past the span's end recompute its last node (held x) and do not store
        const int x = min(xb + lane, x1 - 1);
        const int xs[3] = {held(x, -1, a.nx), held(x, 0, a.nx), held(x, 1, a.nx)};
        // every plane the layer reads, in flight at once: node z0 + d reads planes d, d + 1 and
        // d + 2, plane k being held(z0 + k - 1) (in the grid for every k, so the loads past a
        // short last layer are harmless)
        float plane[kBrick + 2][9];
#pragma unroll
        for (int k = 0; k < kBrick + 2; ++k) load_plane<kScaled>(a, src, held(z0 + k, -1, a.nz), ys, xs, rho0, plane[k]);
#pragma unroll
        for (int dz = 0; dz < kBrick; ++dz) {
            const int z = z0 + dz;
            if (z >= z1) break;
            float sum = 0.0f;
#pragma unroll
            for (int dk = -1; dk <= 1; dk++)
#pragma unroll
                for (int dj = -1; dj <= 1; dj++)
#pragma unroll
                    for (int di = -1; di <= 1; di++) {
                        const float weight = (float)((2 - abs(di)) * (2 - abs(dj)) * (2 - abs(dk)) / 64.0);
                        const int t = 3 * (dj + 1) + (di + 1);
                        sum = sum + weight * plane[dz + dk + 1][t];
                    }
            if (xb + lane < x1) dst[((int64_t)z * a.ny + y) * a.nx + x] = sum;
        }
    }
}

// Tension3.stress over one row a wave: central differences of the colour (held reads), then
// the six components over |g|.
__global__ __launch_bounds__(kBlock) void stress_rows_kernel(Mpm3Args a, const float* __restrict__ colour,
                                                             float* __restrict__ t0, float* __restrict__ t1,
                                                             float* __restrict__ t2, float* __restrict__ t3,
                                                             float* __restrict__ t4, float* __restrict__ t5,
                                                             int dense) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t row = (int64_t)blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    if (row >= (int64_t)a.ny * a.nz) return;
    const int y = (int)(row % a.ny), z = (int)(row / a.ny);
    int x0 = 0, x1 = a.nx;
    if (!dense && !span(a, a.near, nullptr, 2, y, z, lane, &x0, &x1)) return;
    int64_t base[3][3];
    int ys[3], zs[3];
    row_bases(a, y, z, base, ys, zs);
    for (int x = x0 + lane; x < x1; x += kWave) {
        const int xm = held(x, -1, a.nx), xh = held(x, 0, a.nx), xp = held(x, 1, a.nx);
        const float gx = 0.5f * (colour[base[1][1] + xp] - colour[base[1][1] + xm]);
        const float gy = 0.5f * (colour[base[1][2] + xh] - colour[base[1][0] + xh]);
        const float gz = 0.5f * (colour[base[2][1] + xh] - colour[base[0][1] + xh]);
        const float xx = gx * gx, yy = gy * gy, zz = gz * gz;
        const float length = sqrtf((xx + (yy + zz)) + (float)(1e-4 * 1e-4));
        const int64_t node = row * a.nx + x;
        t0[node] = (yy + zz) / length;
        t1[node] = (xx + zz) / length;
        t2[node] = (xx + yy) / length;
        t3[node] = -((gx * gy) / length);
        t4[node] = -((gx * gz) / length);
        t5[node] = -((gy * gz) / length);
    }
}

struct Stress {
    const float* t[6];  // txx tyy tzz txy txz tyz
};

// [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]
__device__ __forceinline__ int component(int a, int k) { return a == k ? a : (a + k == 1 ? 3 : (a + k == 2 ? 4 : 5)); }

// Flip3.gridWithTension over the span of touched or live bricks of each row, as the plain grid
// pass: a node there without mass gets zero velocity, and outside the span every node already
// holds zero sums and zero velocity. Tension3.force: the flux of the stress across the node's
// six faces, open where the colour on both sides exceeds MIN_COLOUR; across a face in the wall
// only the normal stress crosses.
__global__ __launch_bounds__(kBlock) void grid_tension_rows_kernel(Mpm3Args a, Stress st,
                                                                   const float* __restrict__ colour) {
    const int lane = threadIdx.x & (kWave - 1);
    // the wave's row and span held in scalar registers: the rows' addresses are then scalar
    const int64_t row = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6)));
    if (row >= (int64_t)a.ny * a.nz) return;
    const int y = (int)(row % a.ny), z = (int)(row / a.ny);
    int x0, x1;
    if (!span(a, a.touched, a.live, 0xff, y, z, lane, &x0, &x1)) return;
    x0 = __builtin_amdgcn_readfirstlane(x0);
    x1 = __builtin_amdgcn_readfirstlane(x1);
    int64_t base[3][3];
    int ys[3], zs[3];
    row_bases(a, y, z, base, ys, zs);
    const int size[3] = {a.nx, a.ny, a.nz};
    // the scene's constants, read before the loop's stores (which could alias them for the compiler)
    const float dt = bits_f(a.params[DT]), sigma = bits_f(a.params[SIGMA]), rho0 = bits_f(a.params[RHO0]);
    const float grav[3] = {bits_f(a.params[GX]), bits_f(a.params[GX + 1]), bits_f(a.params[GX + 2])};
    for (int x = x0 + lane; x < x1; x += kWave) {
        const int64_t node = row * a.nx + x;
        const int c[3] = {x, y, z};
        const int xm = held(x, -1, a.nx), xh = held(x, 0, a.nx), xp = held(x, 1, a.nx);
        // the held neighbour across each face: [axis][sign < 0 ? 0 : 1]
        const int64_t there[3][2] = {{base[1][1] + xm, base[1][1] + xp},
                                     {base[1][0] + xh, base[1][2] + xh},
                                     {base[0][1] + xh, base[2][1] + xh}};
        // every value the node may use, read at once (all in the grid) rather than behind the
        // tests that decide whether it is used: the sums, the colours, the node's six stress
        // components and each face neighbour's three
        const float mass = a.g[0][node];
        const float sums[3] = {a.g[1][node], a.g[2][node], a.g[3][node]};
        const float here = colour[node];
        float across[3][2], mine[6], theirs[3][2][3];
#pragma unroll
        for (int k = 0; k < 6; k++) mine[k] = st.t[k][node];
#pragma unroll
        for (int ax = 0; ax < 3; ax++)
#pragma unroll
            for (int side = 0; side < 2; side++) {
                across[ax][side] = colour[there[ax][side]];
#pragma unroll
                for (int k = 0; k < 3; k++) theirs[ax][side][k] = st.t[component(ax, k)][there[ax][side]];
            }
        float vel[3] = {0.0f, 0.0f, 0.0f};
        if (mass > kEmpty) {
#pragma unroll
            for (int ax = 0; ax < 3; ax++) vel[ax] = sums[ax] / mass + dt * grav[ax];
            float force[3] = {0.0f, 0.0f, 0.0f};
            bool ring = here > kMinColour;
#pragma unroll
            for (int ax = 0; ax < 3; ax++) ring = ring && !(c[ax] < kLow) && !(c[ax] > size[ax] - 1 - kLow);
#pragma unroll
            for (int ax = 0; ax < 3; ax++) {
#pragma unroll
                for (int side = 0; side < 2; side++) {
                    const int sign = side == 0 ? -1 : 1;
                    const bool open = ring && across[ax][side] > kMinColour;
                    const bool wall = sign < 0 ? !(c[ax] > kLow) : !(c[ax] < size[ax] - 1 - kLow);
                    float shear = 1.0f;
                    if (wall) shear = 0.0f;
                    if (open) {
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            const float flux = (float)(0.5 * sign) * (mine[component(ax, k)] + theirs[ax][side][k]);
                            force[k] = force[k] + (k == ax ? flux : shear * flux);
                        }
                    }
                }
            }
            const float push = (dt * sigma) / mx(mass, kMassFloor * rho0);
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
}

// ------------------------------------------------------------------ 2D

namespace t2 {
enum { DT = 0, GX = 1, GY = 2, BULK = 3, RHO0 = 4, SIGMA = 5 };  // Flip.java:102-107
}

// Tension.blur (:98-122): the [1 2 1] x [1 2 1] / 16 blur, taps row by row (dj outer, di inner),
// sum += weight * value; kScaled takes each tap to a colour, min(m / rho0, 1).
template <bool kScaled>
__global__ __launch_bounds__(kBlock) void blur2_kernel(int nx, int ny, const float* __restrict__ src,
                                                       float* __restrict__ dst, const int32_t* __restrict__ params) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)nx * ny) return;
    const int ni = (int)(node % nx), nj = (int)(node / nx);
    const float rho0 = kScaled ? bits_f(params[t2::RHO0]) : 1.0f;
    float sum = 0.0f;
#pragma unroll
    for (int dj = -1; dj <= 1; dj++) {
        const int64_t row = (int64_t)held(nj, dj, ny) * nx;
#pragma unroll
        for (int di = -1; di <= 1; di++) {
            const float weight = (float)((2 - abs(di)) * (2 - abs(dj))) / 16.0f;
            float value = src[row + held(ni, di, nx)];
            if (kScaled) value = mn(value / rho0, 1.0f);
            sum = sum + weight * value;
        }
    }
    dst[node] = sum;
}

// Tension.stress (:135-150): (c_y^2, c_x^2, -c_x c_y) / |grad c|, central differences.
__global__ __launch_bounds__(kBlock) void stress2_kernel(int nx, int ny, const float* __restrict__ colour,
                                                         float* __restrict__ txx, float* __restrict__ tyy,
                                                         float* __restrict__ txy) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)nx * ny) return;
    const int ni = (int)(node % nx), nj = (int)(node / nx);
    const int64_t row = (int64_t)held(nj, 0, ny) * nx;
    const float cx = 0.5f * (colour[row + held(ni, 1, nx)] - colour[row + held(ni, -1, nx)]);
    const int col = held(ni, 0, nx);
    const float cy = 0.5f * (colour[(int64_t)held(nj, 1, ny) * nx + col] - colour[(int64_t)held(nj, -1, ny) * nx + col]);
    const float g = sqrtf((cx * cx + cy * cy) + (float)(1e-4 * 1e-4));
    txx[node] = (cy * cy) / g;
    tyy[node] = (cx * cx) / g;
    txy[node] = -((cx * cy) / g);
}

struct Stress2 {
    const float* txx;
    const float* tyy;
    const float* txy;
};

// Flip.gridWithTension (Flip.java:342-412, weight 1): Flip.grid, plus on a node with mass
// dt sigma F / max(m, MASS_FLOOR rho0), F being Tension.force (:160-185): the mean stress across
// the faces east, west, north and south in that order, each where the colour on both of its
// sides exceeds MIN_COLOUR; then the wall ring.
__global__ __launch_bounds__(kBlock) void grid2_tension_kernel(int nx, int ny, const float* __restrict__ gm,
                                                               const float* __restrict__ gmu,
                                                               const float* __restrict__ gmv, float* __restrict__ gu,
                                                               float* __restrict__ gv, const int32_t* __restrict__ params,
                                                               Stress2 t, const float* __restrict__ colour) {
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= (int64_t)nx * ny) return;
    const int ni = (int)(node % nx), nj = (int)(node / nx);
    const float m = gm[node];
    float u = 0.0f, w = 0.0f;
    if (m > kEmpty) {
        const float dt = bits_f(params[t2::DT]);
        u = gmu[node] / m + dt * (bits_f(params[t2::GX]) * 1.0f);
        w = gmv[node] / m + dt * (bits_f(params[t2::GY]) * 1.0f);
        const float here = colour[node];
        float fx = 0.0f, fy = 0.0f;
#pragma unroll
        for (int f = 0; f < 4; f++) {
            const int di = f == 0 ? 1 : (f == 1 ? -1 : 0);
            const int dj = f == 2 ? 1 : (f == 3 ? -1 : 0);
            const int64_t there = (int64_t)held(nj, dj, ny) * nx + held(ni, di, nx);
            if (here > kMinColour && colour[there] > kMinColour) {
                const float half = (float)(di + dj) * 0.5f;
                const float* along_x = di != 0 ? t.txx : t.txy;
                const float* along_y = di != 0 ? t.txy : t.tyy;
                fx = fx + half * (along_x[node] + along_x[there]);
                fy = fy + half * (along_y[node] + along_y[there]);
            }
        }
        const float push = (dt * bits_f(params[t2::SIGMA])) / mx(m, kMassFloor * bits_f(params[t2::RHO0]));
        u = u + push * fx;
        w = w + push * fy;
    }
    if (!(ni > 1)) u = mx(u, 0.0f);
    if (ni > nx - 3) u = mn(u, 0.0f);
    if (!(nj > 1)) w = mx(w, 0.0f);
    if (nj > ny - 3) w = mn(w, 0.0f);
    gu[node] = u;
    gv[node] = w;
}

inline unsigned node_blocks(int nx, int ny) { return (unsigned)(((int64_t)nx * ny + kBlock - 1) / kBlock); }

inline unsigned row_blocks(const Mpm3Args& a) {
    return (unsigned)(((int64_t)a.ny * a.nz + kBlock / kWave - 1) / (kBlock / kWave));
}

}  // namespace

int tension3_region(const Launch& l, const Mpm3Args& a, bool all) {
    const int64_t bricks = brick_count3(a.nx, a.ny, a.nz);
    hipLaunchKernelGGL(region_kernel, dim3((unsigned)((bricks + 255) / 256)), dim3(256), 0, l.stream, a, all ? 1 : 0);
    return check_launch("tension3_region");
}

int tension3_blur(const Launch& l, const Mpm3Args& a, const float* src, float* dst, bool scaled, bool dense) {
    const int64_t jobs = (int64_t)a.ny * ((a.nz + kBrick - 1) / kBrick);
    const dim3 grid((unsigned)((jobs + kBlock / kWave - 1) / (kBlock / kWave)));
    timed_begin(l, "tension3_blur");
    if (scaled)
        hipLaunchKernelGGL(blur_march_kernel<true>, grid, dim3(kBlock), 0, l.stream, a, src, dst, dense ? 1 : 0);
    else
        hipLaunchKernelGGL(blur_march_kernel<false>, grid, dim3(kBlock), 0, l.stream, a, src, dst, dense ? 1 : 0);
    timed_end(l, "tension3_blur");
    return check_launch("tension3_blur");
}

int tension3_stress(const Launch& l, const Mpm3Args& a, const float* colour, float* const t[6], bool dense) {
    timed_begin(l, "tension3_stress");
    hipLaunchKernelGGL(stress_rows_kernel, dim3(row_blocks(a)), dim3(kBlock), 0, l.stream, a, colour, t[0], t[1],
                       t[2], t[3], t[4], t[5], dense ? 1 : 0);
    timed_end(l, "tension3_stress");
    return check_launch("tension3_stress");
}

int mpm3_grid_tension(const Launch& l, const Mpm3Args& a, float* const t[6], const float* colour) {
    Stress st;
    for (int k = 0; k < 6; ++k) st.t[k] = t[k];
    timed_begin(l, "grid3_tension");
    hipLaunchKernelGGL(grid_tension_rows_kernel, dim3(row_blocks(a)), dim3(kBlock), 0, l.stream, a, st, colour);
    timed_end(l, "grid3_tension");
    return check_launch("grid3_tension");
}

int tension2_blur(const Launch& l, const Mpm2Args& a, const float* src, float* dst, bool scaled) {
    timed_begin(l, "tension2_blur");
    if (scaled)
        hipLaunchKernelGGL(blur2_kernel<true>, dim3(node_blocks(a.nx, a.ny)), dim3(kBlock), 0, l.stream, a.nx, a.ny,
                           src, dst, a.params);
    else
        hipLaunchKernelGGL(blur2_kernel<false>, dim3(node_blocks(a.nx, a.ny)), dim3(kBlock), 0, l.stream, a.nx, a.ny,
                           src, dst, a.params);
    timed_end(l, "tension2_blur");
    return check_launch("tension2_blur");
}

int tension2_stress(const Launch& l, const Mpm2Args& a, const float* colour, float* const t[3]) {
    timed_begin(l, "tension2_stress");
    hipLaunchKernelGGL(stress2_kernel, dim3(node_blocks(a.nx, a.ny)), dim3(kBlock), 0, l.stream, a.nx, a.ny, colour,
                       t[0], t[1], t[2]);
    timed_end(l, "tension2_stress");
    return check_launch("tension2_stress");
}

int mpm2_grid_tension(const Launch& l, const Mpm2Args& a, float* const t[3], const float* colour) {
    timed_begin(l, "grid2_tension");
    hipLaunchKernelGGL(grid2_tension_kernel, dim3(node_blocks(a.nx, a.ny)), dim3(kBlock), 0, l.stream, a.nx, a.ny,
                       a.g[0], a.g[1], a.g[2], a.g[3], a.g[4], a.params, Stress2{t[0], t[1], t[2]}, colour);
    timed_end(l, "grid2_tension");
    return check_launch("grid2_tension");
}

}  // namespace vxr
