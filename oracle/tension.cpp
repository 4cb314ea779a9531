// 3D surface tension restated: Tension3.blur, Tension3.stress and Tension3.force
// (core/particle/Tension3.java:69-196) and Flip3.gridWithTension (core/particle/Flip3.java:263-311),
// binary32 in the IR's order, neighbour reads held to the wall ring. Test infrastructure only.
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

}  // extern "C"
