// MLS-MPM / APIC passes restated from core/particle/Flip3.java (3D) and core/particle/Flip.java
// (2D), binary32, no contraction, one particle (or node) at a time in index order. The
// float atomics of the reference become plain sums in particle order here: any order is
// the reference's (Scatter.java:48-53), and this one is fixed. Test infrastructure only.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

inline float mn(float a, float b) { return b < a ? b : a; }
inline float mx(float a, float b) { return a < b ? b : a; }
inline float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
inline float bits_to_float(int32_t w) {
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

const float kWall = 1.0f;     // Flip.WALL, core/particle/Flip.java:128
const float kEmpty = 1e-9f;   // Flip.EMPTY / Flip3.EMPTY
const float kJBound = 10.0f;  // Flip3.J_BOUND

// Flip3.Stencil3.weights / Flip.Stencil.weights: 1/2(3/2-f)^2, 3/4-(f-1)^2, 1/2(f-1/2)^2
inline void weights(float f, float w[3]) {
    float a = 1.5f - f;
    float c = f - 1.0f;
    float e = f - 0.5f;
    w[0] = 0.5f * (a * a);
    w[1] = 0.75f - (c * c);
    w[2] = 0.5f * (e * e);
}

// One axis of Flip3.Stencil3.of (Flip3.java:670-676) / Flip.Stencil.of (Flip.java:573-579).
inline void axis(float pos, int n, int32_t* cell, float* off, float w[3]) {
    float x = clampf(pos, 0.5f, (float)n - 1.5f);
    *cell = (int32_t)mn(x - 0.5f, (float)(n - 3));
    *off = x - (float)*cell;
    weights(*off, w);
}

// keepInside, Flip3.java:644-655 and Flip.java:548-559
inline void keep_inside(float* at, float* vel, int size) {
    float lo = kWall;
    float hi = (float)size - 1.0f - kWall;
    if (*at < lo) {
        *at = lo;
        *vel = mx(*vel, 0.0f);
    }
    if (*at > hi) {
        *at = hi;
        *vel = mn(*vel, 0.0f);
    }
}

enum { DT3 = 0, GX3 = 1, GY3 = 2, GZ3 = 3, BULK3 = 4, RHO03 = 5 };     // Flip3.java:58-64
enum { DT2 = 0, GX2 = 1, GY2 = 2, BULK2 = 3, RHO02 = 4 };              // Flip.java:102-106

// One particle's 108 P2G amounts (Flip3.java:217-243), handed to add(node, field, amount) in the
// reference's order: node = base_of(base cell) + dz plane + dy row + dx in the caller's layout.
// Nothing for a particle without mass.
template <typename Base, typename Add>
inline void scatter3_one(int nx, int ny, int nz, float* const p[17], int64_t q, float dt, float bulk, float rho0,
                         Base base_of, int64_t row, int64_t plane, Add add) {
    const float m = p[6][q];
    if (!(m > 0.0f)) return;
    const int size[3] = {nx, ny, nz};
    int32_t cell[3];
    float off[3], w[3][3];
    for (int a = 0; a < 3; a++) axis(p[a][q], size[a], &cell[a], &off[a], w[a]);
    const float vel[3] = {p[3][q], p[4][q], p[5][q]};
    float c[9];
    for (int a = 0; a < 9; a++) c[a] = p[8 + a][q];
    const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - p[7][q]));
    const int64_t base = base_of(cell);
    for (int k = 0; k < 27; k++) {
        const float wgt = w[0][k % 3] * (w[1][(k / 3) % 3] * w[2][k / 9]);
        const float d[3] = {(float)(k % 3) - off[0], (float)((k / 3) % 3) - off[1], (float)(k / 9) - off[2]};
        const float wm = wgt * m;
        const int64_t node = base + (k / 9) * plane + ((k / 3) % 3) * row + k % 3;
        add(node, 0, wm);
        for (int a = 0; a < 3; a++) {
            const float at3 = vel[a] + (c[3 * a] * d[0] + (c[3 * a + 1] * d[1] + c[3 * a + 2] * d[2]));
            const float push = s * d[a];
            add(node, 1 + a, wm * at3 + wgt * push);
        }
    }
}

// Flip3.grid at one node (Flip3.java:267-311, the plain branch).
inline void grid3_node(int nx, int ny, int nz, int64_t node, float dt, const float (&g)[3], const float* gm,
                       const float* const mom[3], float* const out[3]) {
    const int size[3] = {nx, ny, nz};
    const int64_t at[3] = {node % nx, (node / nx) % ny, node / ((int64_t)nx * ny)};
    const float mass = gm[node];
    float vel[3] = {0.0f, 0.0f, 0.0f};
    if (mass > kEmpty) {
        for (int a = 0; a < 3; a++) vel[a] = mom[a][node] / mass + dt * g[a];
    }
    for (int a = 0; a < 3; a++) {
        if (!(at[a] > 1)) vel[a] = mx(vel[a], 0.0f);
        if (at[a] > size[a] - 3) vel[a] = mn(vel[a], 0.0f);
        out[a][node] = vel[a];
    }
}

// Flip3.advect for one particle (Flip3.java:329-383, the plain branch).
inline void advect3_one(int nx, int ny, int nz, float* const p[17], int64_t q, float dt, const float* const vel_in[3]) {
    if (!(p[6][q] > 0.0f)) return;
    // local copies: the particle stores below cannot be taken to change them
    const float* const grid[3] = {vel_in[0], vel_in[1], vel_in[2]};
    const int size[3] = {nx, ny, nz};
    int32_t cell[3];
    float off[3], w[3][3];
    for (int a = 0; a < 3; a++) axis(p[a][q], size[a], &cell[a], &off[a], w[a]);
    const int64_t base = ((int64_t)cell[2] * ny + cell[1]) * nx + cell[0];
    float vel[3] = {0.0f, 0.0f, 0.0f};
    float sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 27; k++) {
        const float wgt = w[0][k % 3] * (w[1][(k / 3) % 3] * w[2][k / 9]);
        const float d[3] = {(float)(k % 3) - off[0], (float)((k / 3) % 3) - off[1], (float)(k / 9) - off[2]};
        const int64_t node = base + (int64_t)(k / 9) * nx * ny + ((k / 3) % 3) * nx + k % 3;
        for (int a = 0; a < 3; a++) {
            const float na = wgt * grid[a][node];
            vel[a] = vel[a] + na;
            for (int e = 0; e < 3; e++) sum[3 * a + e] = sum[3 * a + e] + na * d[e];
        }
    }
    float c[9];
    for (int a = 0; a < 9; a++) c[a] = 4.0f * sum[a];
    for (int a = 0; a < 3; a++) {
        float pos = p[a][q] + dt * vel[a];
        keep_inside(&pos, &vel[a], size[a]);
        p[a][q] = pos;
        p[3 + a][q] = vel[a];
    }
    for (int a = 0; a < 9; a++) p[8 + a][q] = c[a];
    const float j = p[7][q] * (1.0f + dt * (c[0] + (c[4] + c[8])));
    p[7][q] = clampf(j, 1.0f / kJBound, kJBound);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- 3D

void orc_mpm3_clear(int nx, int ny, int nz, float* gm, float* gmu, float* gmv, float* gmw) {
    size_t n = (size_t)nx * ny * nz;
    std::fill(gm, gm + n, 0.0f);
    std::fill(gmu, gmu + n, 0.0f);
    std::fill(gmv, gmv + n, 0.0f);
    std::fill(gmw, gmw + n, 0.0f);
}

// Flip3.scatter, Flip3.java:193-246 (the plain branch)
void orc_mpm3_scatter(int nx, int ny, int nz, int64_t count, float* const p[17], const int32_t params[7],
                      float* gm, float* gmu, float* gmv, float* gmw) {
    float* grid[4] = {gm, gmu, gmv, gmw};
    const float dt = bits_to_float(params[DT3]);
    const float bulk = bits_to_float(params[BULK3]);
    const float rho0 = bits_to_float(params[RHO03]);
    const auto base_of = [&](const int32_t (&cell)[3]) { return ((int64_t)cell[2] * ny + cell[1]) * nx + cell[0]; };
    for (int64_t q = 0; q < count; q++)
        scatter3_one(nx, ny, nz, p, q, dt, bulk, rho0, base_of, nx, (int64_t)nx * ny,
                     [&](int64_t node, int f, float v) { grid[f][node] += v; });
}

// Fixed-point amounts: a * 2^bits is exact in binary32 (a power-of-two scaling), and llrintf
// rounds it to the nearest integer, ties to even, under the default rounding mode.
static inline int64_t quantize(float a, int bits) { return (int64_t)llrintf(std::ldexp(a, bits)); }

// The 3D fixed-point scatter: each amount scatter3_one computes (the reference's arithmetic,
// Flip3.java:216-243) rounded to a 64-bit integer, mass in 2^-40 and momentum in 2^-32, and summed
// with two's-complement wrap-around as the device's 64-bit atomics sum it.
void orc_mpm3_scatter_fixed(int nx, int ny, int nz, int64_t count, float* const p[17], const int32_t params[7],
                            int64_t* qm, int64_t* qmu, int64_t* qmv, int64_t* qmw) {
    int64_t* q[4] = {qm, qmu, qmv, qmw};
    const int bits[4] = {ORC_FIXED_MASS_BITS, ORC_FIXED_MOMENTUM_BITS, ORC_FIXED_MOMENTUM_BITS,
                         ORC_FIXED_MOMENTUM_BITS};
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int f = 0; f < 4; f++) std::fill(q[f], q[f] + nodes, 0);
    const float dt = bits_to_float(params[DT3]);
    const float bulk = bits_to_float(params[BULK3]);
    const float rho0 = bits_to_float(params[RHO03]);
    const auto base_of = [&](const int32_t (&cell)[3]) { return ((int64_t)cell[2] * ny + cell[1]) * nx + cell[0]; };
    for (int64_t k = 0; k < count; k++)
        scatter3_one(nx, ny, nz, p, k, dt, bulk, rho0, base_of, nx, (int64_t)nx * ny, [&](int64_t node, int f, float v) {
            q[f][node] = (int64_t)((uint64_t)q[f][node] + (uint64_t)quantize(v, bits[f]));
        });
}

void orc_mpm3_decode_fixed(int64_t nodes, const int64_t* qm, const int64_t* qmu, const int64_t* qmv,
                           const int64_t* qmw, float* gm, float* gmu, float* gmv, float* gmw) {
    for (int64_t i = 0; i < nodes; i++) {
        gm[i] = (float)std::ldexp((double)qm[i], -ORC_FIXED_MASS_BITS);
        gmu[i] = (float)std::ldexp((double)qmu[i], -ORC_FIXED_MOMENTUM_BITS);
        gmv[i] = (float)std::ldexp((double)qmv[i], -ORC_FIXED_MOMENTUM_BITS);
        gmw[i] = (float)std::ldexp((double)qmw[i], -ORC_FIXED_MOMENTUM_BITS);
    }
}

// Flip3.grid, Flip3.java:267-311 (the plain branch)
void orc_mpm3_grid(int nx, int ny, int nz, const int32_t params[7], const float* gm, const float* gmu,
                   const float* gmv, const float* gmw, float* gu, float* gv, float* gw) {
    const float* mom[3] = {gmu, gmv, gmw};
    float* out[3] = {gu, gv, gw};
    const float dt = bits_to_float(params[DT3]);
    const float g[3] = {bits_to_float(params[GX3]), bits_to_float(params[GY3]), bits_to_float(params[GZ3])};
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t node = 0; node < nodes; node++) grid3_node(nx, ny, nz, node, dt, g, gm, mom, out);
}

// Flip3.advect, Flip3.java:329-383 (the plain branch)
void orc_mpm3_advect(int nx, int ny, int nz, int64_t count, float* const p[17], const int32_t params[7],
                     const float* gu, const float* gv, const float* gw) {
    const float* grid[3] = {gu, gv, gw};
    const float dt = bits_to_float(params[DT3]);
    for (int64_t q = 0; q < count; q++) advect3_one(nx, ny, nz, p, q, dt, grid);
}

void orc_mpm3_step(int nx, int ny, int nz, int64_t count, float* const p[17], const int32_t params[7],
                   float* const grid[7]) {
    orc_mpm3_clear(nx, ny, nz, grid[0], grid[1], grid[2], grid[3]);
    orc_mpm3_scatter(nx, ny, nz, count, p, params, grid[0], grid[1], grid[2], grid[3]);
    orc_mpm3_grid(nx, ny, nz, params, grid[0], grid[1], grid[2], grid[3], grid[4], grid[5], grid[6]);
    orc_mpm3_advect(nx, ny, nz, count, p, params, grid[4], grid[5], grid[6]);
}

// The same substep on `threads` host threads, for the CPU baseline only (not a restatement to
// check against: its P2G sums in another order). Clear, grid update and G2P split the nodes and
// particles evenly; P2G gives each thread a contiguous run of particles, scattered into a private
// box of nodes around them and then added into the grid one thread at a time.
void orc_mpm3_step_omp(int nx, int ny, int nz, int64_t count, float* const p[17], const int32_t params[7],
                       float* const grid[7], int threads) {
    const int64_t nodes = (int64_t)nx * ny * nz;
    const float dt = bits_to_float(params[DT3]);
    const float bulk = bits_to_float(params[BULK3]);
    const float rho0 = bits_to_float(params[RHO03]);
    const float g[3] = {bits_to_float(params[GX3]), bits_to_float(params[GY3]), bits_to_float(params[GZ3])};
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < nodes; i++) grid[0][i] = grid[1][i] = grid[2][i] = grid[3][i] = 0.0f;
#pragma omp parallel num_threads(threads)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        const int64_t q0 = count * t / nt, q1 = count * (t + 1) / nt;
        const int size[3] = {nx, ny, nz};
        int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
        for (int64_t q = q0; q < q1; q++) {
            if (!(p[6][q] > 0.0f)) continue;
            for (int a = 0; a < 3; a++) {
                int32_t cell;
                float off, w[3];
                axis(p[a][q], size[a], &cell, &off, w);
                lo[a] = std::min(lo[a], (int)cell);
                hi[a] = std::max(hi[a], (int)cell);
            }
        }
        if (lo[0] <= hi[0]) {
            const int ex = hi[0] - lo[0] + 3, ey = hi[1] - lo[1] + 3, ez = hi[2] - lo[2] + 3;
            std::vector<float> box((size_t)ex * ey * ez * 4, 0.0f);
            const auto base_of = [&](const int32_t (&cell)[3]) {
                return ((int64_t)(cell[2] - lo[2]) * ey + cell[1] - lo[1]) * ex + cell[0] - lo[0];
            };
            for (int64_t q = q0; q < q1; q++)
                scatter3_one(nx, ny, nz, p, q, dt, bulk, rho0, base_of, ex, (int64_t)ex * ey,
                             [&](int64_t at, int f, float v) { box[4 * at + f] += v; });
#pragma omp critical
            for (int z = 0; z < ez; z++)
                for (int y = 0; y < ey; y++)
                    for (int x = 0; x < ex; x++) {
                        const int64_t at = ((int64_t)z * ey + y) * ex + x;
                        const int64_t node = ((int64_t)(lo[2] + z) * ny + lo[1] + y) * nx + lo[0] + x;
                        for (int f = 0; f < 4; f++) grid[f][node] += box[4 * at + f];
                    }
        }
    }
    const float* mom[3] = {grid[1], grid[2], grid[3]};
    float* out[3] = {grid[4], grid[5], grid[6]};
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t node = 0; node < nodes; node++) grid3_node(nx, ny, nz, node, dt, g, grid[0], mom, out);
    const float* vel[3] = {grid[4], grid[5], grid[6]};
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t q = 0; q < count; q++) advect3_one(nx, ny, nz, p, q, dt, vel);
}

// The 3D sort key: Stencil3's base cell, numbered brick-major (8x8x8-cell bricks, bricks
// row-major, cells row-major within a brick). A bijection of the base cell, so a stable
// sort by it groups each cell's particles and keeps a brick's particles together. A slot without
// mass (m <= 0: a tail slot, or one the level of detail emptied) takes the key one past the
// largest base-cell key, so the sort moves every active slot ahead of every inactive one.
static inline int brick_count(int n) { return (n - 2 + 7) / 8; }

static int32_t base_key(int nx, int ny, int nz, const int32_t (&cell)[3]) {
    const int bx = brick_count(nx), by = brick_count(ny);
    (void)nz;
    int32_t brick = ((cell[2] >> 3) * by + (cell[1] >> 3)) * bx + (cell[0] >> 3);
    int32_t local = ((cell[2] & 7) * 8 + (cell[1] & 7)) * 8 + (cell[0] & 7);
    return brick * 512 + local;
}

int32_t orc_mpm3_inactive_key(int nx, int ny, int nz) {
    const int32_t last[3] = {nx - 3, ny - 3, nz - 3};
    return base_key(nx, ny, nz, last) + 1;
}

int orc_mpm3_key_bits(int nx, int ny, int nz) {
    const int64_t keys = (int64_t)orc_mpm3_inactive_key(nx, ny, nz) + 1;
    int bits = 0;
    while ((1LL << bits) < keys) bits++;
    return bits;
}

void orc_mpm3_keys(int nx, int ny, int nz, int64_t count, const float* x, const float* y, const float* z,
                   const float* m, int32_t* keys) {
    const float* at[3] = {x, y, z};
    const int size[3] = {nx, ny, nz};
    const int32_t inactive = orc_mpm3_inactive_key(nx, ny, nz);
    for (int64_t q = 0; q < count; q++) {
        if (!(m[q] > 0.0f)) {
            keys[q] = inactive;
            continue;
        }
        int32_t cell[3];
        for (int a = 0; a < 3; a++) {
            float px = clampf(at[a][q], 0.5f, (float)size[a] - 1.5f);
            cell[a] = (int32_t)mn(px - 0.5f, (float)(size[a] - 3));
        }
        keys[q] = base_key(nx, ny, nz, cell);
    }
}

// ---------------------------------------------------------------- 2D

void orc_mpm2_clear(int nx, int ny, float* gm, float* gmu, float* gmv) {
    size_t n = (size_t)nx * ny;
    std::fill(gm, gm + n, 0.0f);
    std::fill(gmu, gmu + n, 0.0f);
    std::fill(gmv, gmv + n, 0.0f);
}

// Flip.scatter, Flip.java:269-303, with Scatter.segmentedDeposit's sums (Scatter.java:385-439)
// taken in particle order: a segmented sum is the same sum regrouped.
void orc_mpm2_scatter(int nx, int ny, int64_t count, float* const p[10], const int32_t* params, float* gm,
                      float* gmu, float* gmv) {
    const float dt = bits_to_float(params[DT2]);
    const float bulk = bits_to_float(params[BULK2]);
    const float rho0 = bits_to_float(params[RHO02]);
    for (int64_t q = 0; q < count; q++) {
        int32_t col, row;
        float fx, fy, wx[3], wy[3];
        axis(p[0][q], nx, &col, &fx, wx);
        axis(p[1][q], ny, &row, &fy, wy);
        const int64_t base = (int64_t)row * nx + col;
        const float m = p[4][q], u = p[2][q], w = p[3][q];
        const float c00 = p[6][q], c01 = p[7][q], c10 = p[8][q], c11 = p[9][q];
        const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - p[5][q]));
        for (int k = 0; k < 9; k++) {
            const float wgt = wx[k % 3] * wy[k / 3];
            const float wm = wgt * m;
            const float dx = (float)(k % 3) - fx;
            const float dy = (float)(k / 3) - fy;
            const float vu = u + (c00 * dx + c01 * dy);
            const float vv = w + (c10 * dx + c11 * dy);
            const int64_t node = base + (int64_t)(k / 3) * nx + k % 3;
            gm[node] += wm;
            gmu[node] += wm * vu + wgt * (s * dx);
            gmv[node] += wm * vv + wgt * (s * dy);
        }
    }
}

// The 2D fixed-point scatter: each amount orc_mpm2_scatter computes, rounded by quantize.
void orc_mpm2_scatter_fixed(int nx, int ny, int64_t count, float* const p[10], const int32_t* params,
                            int64_t* qm, int64_t* qmu, int64_t* qmv) {
    const float dt = bits_to_float(params[DT2]);
    const float bulk = bits_to_float(params[BULK2]);
    const float rho0 = bits_to_float(params[RHO02]);
    const int64_t nodes = (int64_t)nx * ny;
    std::fill(qm, qm + nodes, 0);
    std::fill(qmu, qmu + nodes, 0);
    std::fill(qmv, qmv + nodes, 0);
    for (int64_t q = 0; q < count; q++) {
        int32_t col, row;
        float fx, fy, wx[3], wy[3];
        axis(p[0][q], nx, &col, &fx, wx);
        axis(p[1][q], ny, &row, &fy, wy);
        const int64_t base = (int64_t)row * nx + col;
        const float m = p[4][q], u = p[2][q], w = p[3][q];
        const float c00 = p[6][q], c01 = p[7][q], c10 = p[8][q], c11 = p[9][q];
        const float s = (4.0f * (dt * (m / rho0))) * (bulk * (1.0f - p[5][q]));
        for (int k = 0; k < 9; k++) {
            const float wgt = wx[k % 3] * wy[k / 3];
            const float wm = wgt * m;
            const float dx = (float)(k % 3) - fx;
            const float dy = (float)(k / 3) - fy;
            const float vu = u + (c00 * dx + c01 * dy);
            const float vv = w + (c10 * dx + c11 * dy);
            const int64_t node = base + (int64_t)(k / 3) * nx + k % 3;
            // two's-complement wrap-around, as the device's 64-bit atomics
            qm[node] = (int64_t)((uint64_t)qm[node] + (uint64_t)quantize(wm, ORC_FIXED_MASS_BITS));
            qmu[node] = (int64_t)((uint64_t)qmu[node] +
                                  (uint64_t)quantize(wm * vu + wgt * (s * dx), ORC_FIXED_MOMENTUM_BITS));
            qmv[node] = (int64_t)((uint64_t)qmv[node] +
                                  (uint64_t)quantize(wm * vv + wgt * (s * dy), ORC_FIXED_MOMENTUM_BITS));
        }
    }
}

void orc_mpm2_decode_fixed(int64_t nodes, const int64_t* qm, const int64_t* qmu, const int64_t* qmv, float* gm,
                           float* gmu, float* gmv) {
    for (int64_t i = 0; i < nodes; i++) {
        gm[i] = (float)std::ldexp((double)qm[i], -ORC_FIXED_MASS_BITS);
        gmu[i] = (float)std::ldexp((double)qmu[i], -ORC_FIXED_MOMENTUM_BITS);
        gmv[i] = (float)std::ldexp((double)qmv[i], -ORC_FIXED_MOMENTUM_BITS);
    }
}

// Flip.grid, Flip.java:366-412 (the plain branch: weight 1)
void orc_mpm2_grid(int nx, int ny, const int32_t* params, const float* gm, const float* gmu, const float* gmv,
                   float* gu, float* gv) {
    const float dt = bits_to_float(params[DT2]);
    const float gx = bits_to_float(params[GX2]), gy = bits_to_float(params[GY2]);
    const int64_t nodes = (int64_t)nx * ny;
    for (int64_t node = 0; node < nodes; node++) {
        const int64_t ni = node % nx, nj = node / nx;
        const float m = gm[node];
        float u = 0.0f, w = 0.0f;
        if (m > kEmpty) {
            u = gmu[node] / m + dt * (gx * 1.0f);
            w = gmv[node] / m + dt * (gy * 1.0f);
        }
        if (!(ni > 1)) u = mx(u, 0.0f);
        if (ni > nx - 3) u = mn(u, 0.0f);
        if (!(nj > 1)) w = mx(w, 0.0f);
        if (nj > ny - 3) w = mn(w, 0.0f);
        gu[node] = u;
        gv[node] = w;
    }
}

// Flip.advectBody, Flip.java:494-545 (not relaxing)
void orc_mpm2_advect(int nx, int ny, int64_t count, float* const p[10], const int32_t* params, const float* gu,
                     const float* gv) {
    const float dt = bits_to_float(params[DT2]);
    for (int64_t q = 0; q < count; q++) {
        int32_t col, row;
        float fx, fy, wx[3], wy[3];
        axis(p[0][q], nx, &col, &fx, wx);
        axis(p[1][q], ny, &row, &fy, wy);
        const int64_t base = (int64_t)row * nx + col;
        float u = 0, w = 0, b00 = 0, b01 = 0, b10 = 0, b11 = 0;
        for (int k = 0; k < 9; k++) {
            const float wgt = wx[k % 3] * wy[k / 3];
            const float dx = (float)(k % 3) - fx;
            const float dy = (float)(k / 3) - fy;
            const int64_t node = base + (int64_t)(k / 3) * nx + k % 3;
            const float ui = wgt * gu[node];
            const float vi = wgt * gv[node];
            u = u + ui;
            w = w + vi;
            b00 = b00 + ui * dx;
            b01 = b01 + ui * dy;
            b10 = b10 + vi * dx;
            b11 = b11 + vi * dy;
        }
        const float c00 = 4.0f * b00, c01 = 4.0f * b01, c10 = 4.0f * b10, c11 = 4.0f * b11;
        float x = p[0][q] + dt * u;
        float y = p[1][q] + dt * w;
        keep_inside(&x, &u, nx);
        keep_inside(&y, &w, ny);
        p[0][q] = x;
        p[1][q] = y;
        p[2][q] = u;
        p[3][q] = w;
        p[6][q] = c00;
        p[7][q] = c01;
        p[8][q] = c10;
        p[9][q] = c11;
        const float j = p[5][q] * (1.0f + dt * (c00 + c11));
        p[5][q] = clampf(j, 1.0f / kJBound, kJBound);
    }
}

void orc_mpm2_step(int nx, int ny, int64_t count, float* const p[10], const int32_t* params, float* const grid[5]) {
    orc_mpm2_clear(nx, ny, grid[0], grid[1], grid[2]);
    orc_mpm2_scatter(nx, ny, count, p, params, grid[0], grid[1], grid[2]);
    orc_mpm2_grid(nx, ny, params, grid[0], grid[1], grid[2], grid[3], grid[4]);
    orc_mpm2_advect(nx, ny, count, p, params, grid[3], grid[4]);
}

// Scatter.cell, Scatter.java:473-477
int32_t orc_mpm2_cell(float x, float y, int nx, int ny) {
    float cx = clampf(x, 0.0f, (float)nx - 1.0f);
    float cy = clampf(y, 0.0f, (float)ny - 1.0f);
    return (int32_t)mn(cy, (float)ny - 2.0f) * (nx - 1) + (int32_t)mn(cx, (float)nx - 2.0f);
}

void orc_mpm2_keys(int nx, int ny, int64_t count, const float* x, const float* y, int32_t* keys) {
    for (int64_t q = 0; q < count; q++) keys[q] = orc_mpm2_cell(x[q], y[q], nx, ny);
}

// Scatter.cellStarts, Scatter.java:484-502
int64_t orc_mpm2_cell_starts(int nx, int ny, int64_t count, const float* x, const float* y, int32_t* starts) {
    const int64_t cells = (int64_t)(nx - 1) * (ny - 1);
    std::fill(starts, starts + cells + 1, 0);
    int32_t previous = 0;
    for (int64_t q = 0; q < count; q++) {
        int32_t c = orc_mpm2_cell(x[q], y[q], nx, ny);
        if (c < previous) return -1 - q;
        previous = c;
        starts[c + 1]++;
    }
    for (int64_t c = 0; c < cells; c++) starts[c + 1] += starts[c];
    return 0;
}

void orc_stable_order(int64_t count, const int32_t* keys, int32_t* order) {
    std::iota(order, order + count, 0);
    std::stable_sort(order, order + count, [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
}

void orc_bucket_starts(int64_t count, const int32_t* keys, int32_t cells, int32_t* starts) {
    std::fill(starts, starts + (int64_t)cells + 1, 0);
    for (int64_t q = 0; q < count; q++) starts[keys[q] + 1]++;
    for (int64_t c = 0; c < cells; c++) starts[c + 1] += starts[c];
}

// FlipStep.sort(): count by Scatter.cell, scan, permute every field, copy back. Stable here.
void orc_mpm2_sort(int nx, int ny, int64_t count, float* const p[10]) {
    std::vector<int32_t> keys((size_t)count), order((size_t)count);
    orc_mpm2_keys(nx, ny, count, p[0], p[1], keys.data());
    orc_stable_order(count, keys.data(), order.data());
    std::vector<float> tmp((size_t)count);
    for (int f = 0; f < 10; f++) {
        for (int64_t q = 0; q < count; q++) tmp[q] = p[f][order[q]];
        std::memcpy(p[f], tmp.data(), sizeof(float) * count);
    }
}

}  // extern "C"
