// Surface tension restated, binary32 in the IR's order, neighbour reads held to the wall ring.
// Test infrastructure only.
// 3D: Tension3.blur, Tension3.stress and Tension3.force (core/particle/Tension3.java:69-196) and
// Flip3.gridWithTension (core/particle/Flip3.java:263-311).
// 2D: Tension.blur, Tension.stress and Tension.force (core/particle/Tension.java:98-185) and
// Flip.gridWithTension (core/particle/Flip.java:342-412).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oracle.h"

namespace {

inline float mn(float a, float b) { return b < a ? b : a; }
inline float mx(float a, float b) { return a < b ? b : a; }
inline float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
inline float bits_to_float(int32_t w) {
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

const float kWall = 1.0f;
const int kLow = 1;                // (int) Flip3.WALL
const float kMinColour = 0.3f;     // Tension.MIN_COLOUR
const float kMassFloor = 0.3f;     // Tension.MASS_FLOOR
const float kEps2 = (float)(1e-4 * 1e-4);
const float kEmpty = 1e-9f;
enum { DT = 0, GX = 1, GY = 2, GZ = 3, BULK = 4, RHO0 = 5, SIGMA = 6 };  // Flip3.java:58-64

// Tension3.held: the coordinate moved by `off`, held to the wall ring.
inline int held(int c, int off, int size) {
    return (int)clampf((float)c + (float)off, kWall, (float)(size - 1) - kWall);
}

struct Grid3 {
    int nx, ny, nz;
    int64_t at(const int (&c)[3], int di, int dj, int dk) const {
        const int x = held(c[0], di, nx), y = held(c[1], dj, ny), z = held(c[2], dk, nz);
        return ((int64_t)z * ny + y) * nx + x;
    }
    void coords(int64_t node, int (&c)[3]) const {
        c[0] = (int)(node % nx);
        c[1] = (int)(node / nx % ny);
        c[2] = (int)(node / ((int64_t)nx * ny));
    }
};

}  // namespace

extern "C" {

void orc_tension3_blur(int nx, int ny, int nz, const float* src, float* dst, const int32_t params[7], int scaled) {
    const Grid3 g{nx, ny, nz};
    const int size[3] = {nx, ny, nz};
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t node = 0; node < nodes; node++) {
        int c[3];
        g.coords(node, c);
        float sum = 0.0f;
        for (int dk = -1; dk <= 1; dk++)
            for (int dj = -1; dj <= 1; dj++)
                for (int di = -1; di <= 1; di++) {
                    const float weight = (float)((2 - std::abs(di)) * (2 - std::abs(dj)) * (2 - std::abs(dk)) / 64.0);
                    float value = src[g.at(c, di, dj, dk)];
                    if (scaled) {
                        float full = bits_to_float(params[RHO0]);
                        const int d[3] = {di, dj, dk};
                        for (int a = 0; a < 3; a++) {
                            const int h = held(c[a], d[a], size[a]);
                            if (h == kLow) full = 0.5f * full;
                            if (h == size[a] - 1 - kLow) full = 0.5f * full;
                        }
                        value = mn(value / full, 1.0f);
                    }
                    sum = sum + weight * value;
                }
        dst[node] = sum;
    }
}

void orc_tension3_stress(int nx, int ny, int nz, const float* colour, float* const t[6]) {
    const Grid3 g{nx, ny, nz};
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t node = 0; node < nodes; node++) {
        int c[3];
        g.coords(node, c);
        float gr[3];
        for (int a = 0; a < 3; a++) {
            int up[3] = {0, 0, 0}, down[3] = {0, 0, 0};
            up[a] = 1;
            down[a] = -1;
            gr[a] = 0.5f * (colour[g.at(c, up[0], up[1], up[2])] - colour[g.at(c, down[0], down[1], down[2])]);
        }
        const float xx = gr[0] * gr[0], yy = gr[1] * gr[1], zz = gr[2] * gr[2];
        const float length = std::sqrt((xx + (yy + zz)) + kEps2);
        t[0][node] = (yy + zz) / length;
        t[1][node] = (xx + zz) / length;
        t[2][node] = (xx + yy) / length;
        t[3][node] = -((gr[0] * gr[1]) / length);
        t[4][node] = -((gr[0] * gr[2]) / length);
        t[5][node] = -((gr[1] * gr[2]) / length);
    }
}

// Tension3.force at one node (:156-196): the flux of the stress across its six faces.
void orc_tension3_force(int nx, int ny, int nz, const float* colour, const float* const t[6], int64_t node,
                        float out[3]) {
    const Grid3 g{nx, ny, nz};
    const int size[3] = {nx, ny, nz};
    int c[3];
    g.coords(node, c);
    const float* tensor[3][3] = {{t[0], t[3], t[4]}, {t[3], t[1], t[5]}, {t[4], t[5], t[2]}};
    float force[3] = {0.0f, 0.0f, 0.0f};
    bool ring = colour[node] > kMinColour;
    for (int a = 0; a < 3; a++) ring = ring && !(c[a] < kLow) && !(c[a] > size[a] - 1 - kLow);
    for (int a = 0; a < 3; a++) {
        for (int sign = -1; sign <= 1; sign += 2) {
            int d[3] = {0, 0, 0};
            d[a] = sign;
            const int64_t there = g.at(c, d[0], d[1], d[2]);
            const bool open = ring && colour[there] > kMinColour;
            const bool wall = sign < 0 ? !(c[a] > kLow) : !(c[a] < size[a] - 1 - kLow);
            float shear = 1.0f;
            if (wall) shear = 0.0f;
            if (open) {
                for (int k = 0; k < 3; k++) {
                    const float flux = (float)(0.5 * sign) * (tensor[a][k][node] + tensor[a][k][there]);
                    force[k] = force[k] + (k == a ? flux : shear * flux);
                }
            }
        }
    }
    out[0] = force[0];
    out[1] = force[1];
    out[2] = force[2];
}

void orc_mpm3_grid_tension(int nx, int ny, int nz, const int32_t params[7], const float* gm, const float* gmu,
                           const float* gmv, const float* gmw, float* gu, float* gv, float* gw,
                           const float* const t[6], const float* colour) {
    const Grid3 g{nx, ny, nz};
    const int size[3] = {nx, ny, nz};
    const float* mom[3] = {gmu, gmv, gmw};
    float* out[3] = {gu, gv, gw};
    const int grav[3] = {GX, GY, GZ};
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (int64_t node = 0; node < nodes; node++) {
        int c[3];
        g.coords(node, c);
        const float mass = gm[node];
        float vel[3] = {0.0f, 0.0f, 0.0f};
        if (mass > kEmpty) {
            const float dt = bits_to_float(params[DT]);
            for (int a = 0; a < 3; a++) vel[a] = mom[a][node] / mass + dt * bits_to_float(params[grav[a]]);
            float force[3];
            orc_tension3_force(nx, ny, nz, colour, t, node, force);
            const float push = (dt * bits_to_float(params[SIGMA])) / mx(mass, kMassFloor * bits_to_float(params[RHO0]));
            for (int a = 0; a < 3; a++) vel[a] = vel[a] + push * force[a];
        }
        for (int a = 0; a < 3; a++) {
            if (!(c[a] > 1)) vel[a] = mx(vel[a], 0.0f);
            if (c[a] > size[a] - 3) vel[a] = mn(vel[a], 0.0f);
            out[a][node] = vel[a];
        }
    }
}

// ---- 2D (core/particle/Tension.java)

// Tension.blur (:98-122): the [1 2 1] x [1 2 1] / 16 blur of src, rows of taps outer; with
// `scaled` each tap is first taken to a colour, min(m / rho0, 1). params: Flip's 28 words.
void orc_tension2_blur(int nx, int ny, const float* src, float* dst, const int32_t* params, int scaled) {
    const int64_t nodes = (int64_t)nx * ny;
    const float rho0 = bits_to_float(params[4]);  // Flip.RHO0
    for (int64_t node = 0; node < nodes; node++) {
        const int ni = (int)(node % nx), nj = (int)(node / nx);
        float sum = 0.0f;
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const float weight = (float)((2 - std::abs(di)) * (2 - std::abs(dj)) / 16.0);
                float value = src[(int64_t)held(nj, dj, ny) * nx + held(ni, di, nx)];
                if (scaled) value = mn(value / rho0, 1.0f);
                sum = sum + weight * value;
            }
        dst[node] = sum;
    }
}

// Tension.stress (:135-150): (c_y^2, c_x^2, -c_x c_y) / |grad c| from central differences.
void orc_tension2_stress(int nx, int ny, const float* colour, float* txx, float* tyy, float* txy) {
    const int64_t nodes = (int64_t)nx * ny;
    for (int64_t node = 0; node < nodes; node++) {
        const int ni = (int)(node % nx), nj = (int)(node / nx);
        const float cx = 0.5f * (colour[(int64_t)held(nj, 0, ny) * nx + held(ni, 1, nx)] -
                                 colour[(int64_t)held(nj, 0, ny) * nx + held(ni, -1, nx)]);
        const float cy = 0.5f * (colour[(int64_t)held(nj, 1, ny) * nx + held(ni, 0, nx)] -
                                 colour[(int64_t)held(nj, -1, ny) * nx + held(ni, 0, nx)]);
        const float g = std::sqrt((cx * cx + cy * cy) + kEps2);
        txx[node] = (cy * cy) / g;
        tyy[node] = (cx * cx) / g;
        txy[node] = -((cx * cy) / g);
    }
}

// Tension.force (:160-185) at one node: the mean stress across each of its faces east, west,
// north, south in that order, where the colour on both sides exceeds MIN_COLOUR.
void orc_tension2_force(int nx, int ny, const float* colour, const float* txx, const float* tyy, const float* txy,
                        int64_t node, float out[2]) {
    const int ni = (int)(node % nx), nj = (int)(node / nx);
    const int faces[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    float fx = 0.0f, fy = 0.0f;
    const float here = colour[node];
    for (const auto& face : faces) {
        const int64_t there = (int64_t)held(nj, face[1], ny) * nx + held(ni, face[0], nx);
        if (!(here > kMinColour && colour[there] > kMinColour)) continue;
        const float half = (float)(0.5 * (face[0] + face[1]));
        const float* along_x = face[0] != 0 ? txx : txy;
        const float* along_y = face[0] != 0 ? txy : tyy;
        fx = fx + half * (along_x[node] + along_x[there]);
        fy = fy + half * (along_y[node] + along_y[there]);
    }
    out[0] = fx;
    out[1] = fy;
}

// Flip.gridWithTension (Flip.java:342-412 with tension, weight 1): Flip.grid plus
// dt sigma F / max(m, MASS_FLOOR rho0) on every node with mass, before the wall ring.
void orc_mpm2_grid_tension(int nx, int ny, const int32_t* params, const float* gm, const float* gmu, const float* gmv,
                           float* gu, float* gv, const float* txx, const float* tyy, const float* txy,
                           const float* colour) {
    const float dt = bits_to_float(params[0]);
    const float gx = bits_to_float(params[1]), gy = bits_to_float(params[2]);
    const float rho0 = bits_to_float(params[4]), sigma = bits_to_float(params[5]);
    const int64_t nodes = (int64_t)nx * ny;
    for (int64_t node = 0; node < nodes; node++) {
        const int64_t ni = node % nx, nj = node / nx;
        const float m = gm[node];
        float u = 0.0f, w = 0.0f;
        if (m > kEmpty) {
            u = gmu[node] / m + dt * (gx * 1.0f);
            w = gmv[node] / m + dt * (gy * 1.0f);
            float force[2];
            orc_tension2_force(nx, ny, colour, txx, tyy, txy, node, force);
            const float push = (dt * sigma) / mx(m, kMassFloor * rho0);
            u = u + push * force[0];
            w = w + push * force[1];
        }
        if (!(ni > 1)) u = mx(u, 0.0f);
        if (ni > nx - 3) u = mn(u, 0.0f);
        if (!(nj > 1)) w = mx(w, 0.0f);
        if (nj > ny - 3) w = mn(w, 0.0f);
        gu[node] = u;
        gv[node] = w;
    }
}

}  // extern "C"
