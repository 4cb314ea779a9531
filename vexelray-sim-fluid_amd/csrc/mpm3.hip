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
#include <climits>

#include "common.h"

namespace vxr {
namespace {

__device__ __forceinline__ float bits_f(int32_t w) { return __int_as_float(w); }

enum { DT = 0, GX = 1, GY = 2, GZ = 3, BULK = 4, RHO0 = 5 };  // Flip3.java:58-64
constexpr float kWall = 1.0f;                                // Flip.WALL
constexpr float kEmpty = 1e-9f;                              // Flip3.EMPTY
constexpr float kJBound = 10.0f;                             // Flip3.J_BOUND
constexpr int kBlock = 256;
constexpr int kTileNodes = 2560;  // LDS tile: 4 x 2560 x 4 B = 40 KB for P2G, 30 KB for G2P

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

// ------------------------------------------------------------------ clear

__global__ void clear4_kernel(float* __restrict__ g0, float* __restrict__ g1, float* __restrict__ g2,
                              float* __restrict__ g3, int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const float4 z = make_float4(0, 0, 0, 0);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        reinterpret_cast<float4*>(g0)[i] = z;
        reinterpret_cast<float4*>(g1)[i] = z;
        reinterpret_cast<float4*>(g2)[i] = z;
        reinterpret_cast<float4*>(g3)[i] = z;
    }
}

// ------------------------------------------------------------------ P2G, the reference schedule

__global__ __launch_bounds__(kBlock) void p2g3_atomic_kernel(Mpm3Args a) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.count) return;
    if (!(a.p[6][q] > 0.0f)) return;
    const float dt = bits_f(a.params[DT]), bulk = bits_f(a.params[BULK]), rho0 = bits_f(a.params[RHO0]);
    Particle p;
    load_particle(a, q, dt, bulk, rho0, p);
    const int64_t base = ((int64_t)p.ax[2].cell * a.ny + p.ax[1].cell) * a.nx + p.ax[0].cell;
    const int64_t plane = (int64_t)a.nx * a.ny;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        const float wgt = p.ax[0].w[k % 3] * (p.ax[1].w[(k / 3) % 3] * p.ax[2].w[k / 9]);
        const float d[3] = {(float)(k % 3) - p.ax[0].off, (float)((k / 3) % 3) - p.ax[1].off,
                            (float)(k / 9) - p.ax[2].off};
        const float wm = wgt * p.m;
        const int64_t node = base + (k / 9) * plane + ((k / 3) % 3) * a.nx + k % 3;
        atomic_add(a.g[0] + node, wm);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const float at3 = p.vel[ax] + (p.c[3 * ax] * d[0] + (p.c[3 * ax + 1] * d[1] + p.c[3 * ax + 2] * d[2]));
            const float push = p.s * d[ax];
            atomic_add(a.g[1 + ax] + node, wm * at3 + wgt * push);
        }
    }
}

// ------------------------------------------------------------------ block bounding box

struct Box {
    int lo[3], hi[3];
    int ex, ey, ez, nodes;
    bool empty;
};

// The box of base cells of the live particles q in [first, first + n), over the block.
// Every thread returns the same box. Uses red[24] of LDS.
__device__ Box block_box(const Mpm3Args& a, int64_t first, int n, int stride, int* red) {
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    const int sizes[3] = {a.nx, a.ny, a.nz};
    for (int r = 0; r < n; ++r) {
        const int64_t q = first + (int64_t)r * stride;
        if (q >= a.capacity || !(a.p[6][q] > 0.0f)) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int c = base_cell(a.p[i][q], sizes[i]);
            lo[i] = min(lo[i], c);
            hi[i] = max(hi[i], c);
        }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            lo[i] = min(lo[i], __shfl_xor(lo[i], off));
            hi[i] = max(hi[i], __shfl_xor(hi[i], off));
        }
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
    b.nodes = b.empty ? 0 : (nodes > kTileNodes ? INT_MAX : (int)nodes);
    return b;
}

// ------------------------------------------------------------------ P2G, sorted-segmented

__global__ __launch_bounds__(kBlock) void p2g3_tiled_kernel(Mpm3Args a, int run) {
    __shared__ float tile[4 * kTileNodes];
    __shared__ int red[6 * (kBlock / kWave)];
    const int64_t first = (int64_t)blockIdx.x * kBlock * run + (int64_t)threadIdx.x * run;
    const Box box = block_box(a, first, run, 1, red);
    if (box.empty) return;  // the same for every thread of the block
    const bool use_tile = box.nodes <= kTileNodes;
    if (use_tile) {
        for (int t = threadIdx.x; t < 4 * box.nodes; t += kBlock) tile[t] = 0.0f;
    }
    __syncthreads();

    const float dt = bits_f(a.params[DT]), bulk = bits_f(a.params[BULK]), rho0 = bits_f(a.params[RHO0]);
    const int64_t plane = (int64_t)a.nx * a.ny;
    float acc[108];
#pragma unroll
    for (int i = 0; i < 108; ++i) acc[i] = 0.0f;
    int64_t run_base = -1;

    auto flush = [&](int64_t base) {
        if (use_tile) {
            const int b = (int)base;
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const int node = b + ((k / 9) * box.ey + (k / 3) % 3) * box.ex + k % 3;
#pragma unroll
                for (int f = 0; f < 4; ++f) atomicAdd(&tile[f * box.nodes + node], acc[4 * k + f]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 27; ++k) {
                const int64_t node = base + (k / 9) * plane + ((k / 3) % 3) * a.nx + k % 3;
#pragma unroll
                for (int f = 0; f < 4; ++f) atomic_add(a.g[f] + node, acc[4 * k + f]);
            }
        }
#pragma unroll
        for (int i = 0; i < 108; ++i) acc[i] = 0.0f;
    };

    for (int r = 0; r < run; ++r) {
        const int64_t q = first + r;
        if (q >= a.capacity || !(a.p[6][q] > 0.0f)) continue;
        Particle p;
        load_particle(a, q, dt, bulk, rho0, p);
        const int64_t base = use_tile ? (int64_t)(((p.ax[2].cell - box.lo[2]) * box.ey + (p.ax[1].cell - box.lo[1])) *
                                                      box.ex +
                                                  (p.ax[0].cell - box.lo[0]))
                                      : ((int64_t)p.ax[2].cell * a.ny + p.ax[1].cell) * a.nx + p.ax[0].cell;
        if (base != run_base) {
            if (run_base >= 0) flush(run_base);
            run_base = base;
        }
        p2g_amounts(p.ax, p.m, p.vel, p.c, p.s, acc);
    }
    if (run_base >= 0) flush(run_base);

    if (use_tile) {
        __syncthreads();
        for (int t = threadIdx.x; t < box.nodes; t += kBlock) {
            const float m = tile[t];
            if (m == 0.0f) continue;  // a node no live particle weighs on holds nothing
            const int lx = t % box.ex, ly = (t / box.ex) % box.ey, lz = t / (box.ex * box.ey);
            const int64_t node = ((int64_t)(box.lo[2] + lz) * a.ny + (box.lo[1] + ly)) * a.nx + (box.lo[0] + lx);
            atomic_add(a.g[0] + node, m);
            atomic_add(a.g[1] + node, tile[box.nodes + t]);
            atomic_add(a.g[2] + node, tile[2 * box.nodes + t]);
            atomic_add(a.g[3] + node, tile[3 * box.nodes + t]);
        }
    }
}

// ------------------------------------------------------------------ grid update, Flip3.java:267-311

__global__ __launch_bounds__(kBlock) void grid3_kernel(Mpm3Args a) {
    const int64_t nodes = (int64_t)a.nx * a.ny * a.nz;
    const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (node >= nodes) return;
    const int at[3] = {(int)(node % a.nx), (int)((node / a.nx) % a.ny), (int)(node / ((int64_t)a.nx * a.ny))};
    const int size[3] = {a.nx, a.ny, a.nz};
    const float mass = a.g[0][node];
    float vel[3] = {0.0f, 0.0f, 0.0f};
    if (mass > kEmpty) {
        const float dt = bits_f(a.params[DT]);
#pragma unroll
        for (int i = 0; i < 3; ++i) vel[i] = a.g[1 + i][node] / mass + dt * bits_f(a.params[GX + i]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (!(at[i] > 1)) vel[i] = mx(vel[i], 0.0f);
        if (at[i] > size[i] - 3) vel[i] = mn(vel[i], 0.0f);
        a.g[4 + i][node] = vel[i];
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

// One particle's G2P, node velocities read through `node_vel(k, axis)`.
template <typename NodeVel>
__device__ __forceinline__ void g2p_one(const Mpm3Args& a, int64_t q, const Axis (&ax)[3], float dt,
                                        NodeVel node_vel) {
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
    float c[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) c[i] = 4.0f * sum[i];
    const int size[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float pos = a.p[i][q] + dt * vel[i];
        keep_inside(pos, vel[i], size[i]);
        a.p[i][q] = pos;
        a.p[3 + i][q] = vel[i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) a.p[8 + i][q] = c[i];
    const float j = a.p[7][q] * (1.0f + dt * (c[0] + (c[4] + c[8])));
    a.p[7][q] = clampf(j, 1.0f / kJBound, kJBound);
}

__global__ __launch_bounds__(kBlock) void g2p3_kernel(Mpm3Args a) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.count || !(a.p[6][q] > 0.0f)) return;
    Axis ax[3] = {axis(a.p[0][q], a.nx), axis(a.p[1][q], a.ny), axis(a.p[2][q], a.nz)};
    const int64_t base = ((int64_t)ax[2].cell * a.ny + ax[1].cell) * a.nx + ax[0].cell;
    const int64_t plane = (int64_t)a.nx * a.ny;
    const float dt = bits_f(a.params[DT]);
    g2p_one(a, q, ax, dt, [&](int k, int i) {
        return a.g[4 + i][base + (k / 9) * plane + ((k / 3) % 3) * a.nx + k % 3];
    });
}

// The chunk's node velocities staged in LDS; particles of the chunk are taken lane-consecutive.
__global__ __launch_bounds__(kBlock) void g2p3_tiled_kernel(Mpm3Args a, int run) {
    __shared__ float tile[3 * kTileNodes];
    __shared__ int red[6 * (kBlock / kWave)];
    const int64_t chunk = (int64_t)blockIdx.x * kBlock * run;
    const Box box = block_box(a, chunk + threadIdx.x, run, kBlock, red);
    if (box.empty) return;
    const bool use_tile = box.nodes <= kTileNodes;
    if (use_tile) {
        for (int t = threadIdx.x; t < box.nodes; t += kBlock) {
            const int lx = t % box.ex, ly = (t / box.ex) % box.ey, lz = t / (box.ex * box.ey);
            const int64_t node = ((int64_t)(box.lo[2] + lz) * a.ny + (box.lo[1] + ly)) * a.nx + (box.lo[0] + lx);
            tile[t] = a.g[4][node];
            tile[box.nodes + t] = a.g[5][node];
            tile[2 * box.nodes + t] = a.g[6][node];
        }
    }
    __syncthreads();
    const float dt = bits_f(a.params[DT]);
    const int64_t plane = (int64_t)a.nx * a.ny;
    for (int r = 0; r < run; ++r) {
        const int64_t q = chunk + (int64_t)r * kBlock + threadIdx.x;
        if (q >= a.count || !(a.p[6][q] > 0.0f)) continue;
        Axis ax[3] = {axis(a.p[0][q], a.nx), axis(a.p[1][q], a.ny), axis(a.p[2][q], a.nz)};
        if (use_tile) {
            const int base = ((ax[2].cell - box.lo[2]) * box.ey + (ax[1].cell - box.lo[1])) * box.ex +
                             (ax[0].cell - box.lo[0]);
            g2p_one(a, q, ax, dt, [&](int k, int i) {
                return tile[i * box.nodes + base + ((k / 9) * box.ey + (k / 3) % 3) * box.ex + k % 3];
            });
        } else {
            const int64_t base = ((int64_t)ax[2].cell * a.ny + ax[1].cell) * a.nx + ax[0].cell;
            g2p_one(a, q, ax, dt, [&](int k, int i) {
                return a.g[4 + i][base + (k / 9) * plane + ((k / 3) % 3) * a.nx + k % 3];
            });
        }
    }
}

// ------------------------------------------------------------------ sort keys

__device__ __forceinline__ int brick_count(int n) { return (n - 2 + 7) / 8; }

// Stencil3's base cell numbered brick-major: 8x8x8-cell bricks row-major, cells row-major
// inside a brick. Padding particles take the largest key so a stable sort leaves them last.
__global__ __launch_bounds__(kBlock) void keys3_kernel(Mpm3Args a, int32_t* __restrict__ keys,
                                                       int32_t* __restrict__ index, int32_t pad_key) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.capacity) return;
    int32_t key = pad_key;
    if (q < a.count) {
        const int cx = base_cell(a.p[0][q], a.nx), cy = base_cell(a.p[1][q], a.ny), cz = base_cell(a.p[2][q], a.nz);
        const int bx = brick_count(a.nx), by = brick_count(a.ny);
        const int brick = ((cz >> 3) * by + (cy >> 3)) * bx + (cx >> 3);
        const int local = ((cz & 7) * 8 + (cy & 7)) * 8 + (cx & 7);
        key = brick * 512 + local;
    }
    keys[q] = key;
    index[q] = (int32_t)q;
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int mpm3_key_bits(int nx, int ny, int nz) {
    const int64_t keys = (int64_t)((nx - 2 + 7) / 8) * ((ny - 2 + 7) / 8) * ((nz - 2 + 7) / 8) * 512;
    int bits = 0;
    while ((1LL << bits) < keys) bits++;
    return bits;
}

int mpm3_clear(const Launch& l, const Mpm3Args& a) {
    const int64_t nodes = (int64_t)a.nx * a.ny * a.nz;
    const int64_t n4 = (nodes + 3) / 4;  // node arrays are allocated in whole float4s
    int64_t blocks = blocks_for(n4, kBlock);
    if (blocks > 16384) blocks = 16384;
    timed_begin(l, "mpm3_clear");
    hipLaunchKernelGGL(clear4_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, l.stream, a.g[0], a.g[1], a.g[2],
                       a.g[3], n4);
    timed_end(l, "mpm3_clear");
    return check_launch("mpm3_clear");
}

int mpm3_scatter_atomic(const Launch& l, const Mpm3Args& a) {
    if (a.count == 0) return 0;
    timed_begin(l, "p2g3_atomic");
    hipLaunchKernelGGL(p2g3_atomic_kernel, dim3(blocks_for(a.count, kBlock)), dim3(kBlock), 0, l.stream, a);
    timed_end(l, "p2g3_atomic");
    return check_launch("p2g3_atomic");
}

int mpm3_scatter_tiled(const Launch& l, const Mpm3Args& a, int run) {
    if (a.count == 0) return 0;
    timed_begin(l, "p2g3_tiled");
    hipLaunchKernelGGL(p2g3_tiled_kernel, dim3(blocks_for(a.count, kBlock * run)), dim3(kBlock), 0, l.stream, a,
                       run);
    timed_end(l, "p2g3_tiled");
    return check_launch("p2g3_tiled");
}

int mpm3_grid(const Launch& l, const Mpm3Args& a) {
    const int64_t nodes = (int64_t)a.nx * a.ny * a.nz;
    timed_begin(l, "grid3");
    hipLaunchKernelGGL(grid3_kernel, dim3(blocks_for(nodes, kBlock)), dim3(kBlock), 0, l.stream, a);
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

int mpm3_advect_tiled(const Launch& l, const Mpm3Args& a, int run) {
    if (a.count == 0) return 0;
    timed_begin(l, "g2p3_tiled");
    hipLaunchKernelGGL(g2p3_tiled_kernel, dim3(blocks_for(a.count, kBlock * run)), dim3(kBlock), 0, l.stream, a,
                       run);
    timed_end(l, "g2p3_tiled");
    return check_launch("g2p3_tiled");
}

int mpm3_keys(const Launch& l, const Mpm3Args& a, int32_t* keys, int32_t* index) {
    const int32_t pad = (int32_t)((1LL << mpm3_key_bits(a.nx, a.ny, a.nz)) - 1);
    timed_begin(l, "keys3");
    hipLaunchKernelGGL(keys3_kernel, dim3(blocks_for(a.capacity, kBlock)), dim3(kBlock), 0, l.stream, a, keys, index,
                       pad);
    timed_end(l, "keys3");
    return check_launch("keys3");
}

}  // namespace vxr
