// HLL shallow-water step restated from core/stencil/ShallowWater.java:166-384, one cell
// at a time in cell order, binary32 throughout, no contraction. The min/max of the IR
// are written as the selects below so that a signed zero is kept the same way the HIP
// kernel keeps it. Test infrastructure only.
#include "oracle.h"

#include <algorithm>
#include <vector>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include <omp.h>

namespace {

inline float mn(float a, float b) { return b < a ? b : a; }
inline float mx(float a, float b) { return a < b ? b : a; }

inline float bits_to_float(int32_t w) {
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}
inline int32_t float_to_bits(float f) {
    int32_t w;
    std::memcpy(&w, &f, 4);
    return w;
}

struct State {
    float h, hu, hv;
};
struct Flux {
    float mass, normal, tangent;
};

// ShallowWater.velocity, :329-333
inline float velocity(float h, float q, float dry) { return h > dry ? q / h : 0.0f; }

// ShallowWater.combine, :381-384
inline float combine(float hi, float lo, float lohi, float span, float fl, float fr, float ul, float ur) {
    return ((hi * fl - lo * fr) + lohi * (ur - ul)) / span;
}

// ShallowWater.hll, :336-378
Flux hll(float g, float dry, const State& l, const State& r, bool normal_is_x) {
    float qnL = normal_is_x ? l.hu : l.hv;
    float qtL = normal_is_x ? l.hv : l.hu;
    float qnR = normal_is_x ? r.hu : r.hv;
    float qtR = normal_is_x ? r.hv : r.hu;
    float unL = velocity(l.h, qnL, dry);
    float utL = velocity(l.h, qtL, dry);
    float unR = velocity(r.h, qnR, dry);
    float utR = velocity(r.h, qtR, dry);
    float cL = std::sqrt(g * l.h);
    float cR = std::sqrt(g * r.h);
    float sL = mn(unL - cL, unR - cR);
    float sR = mx(unL + cL, unR + cR);
    if (!(r.h > dry)) {
        sL = unL - cL;
        sR = unL + 2.0f * cL;
    }
    if (!(l.h > dry)) {
        sL = unR - 2.0f * cR;
        sR = unR + cR;
    }
    float lo = mn(sL, 0.0f);
    float hi = mx(sR, 0.0f);
    float span = mx(hi - lo, 1e-30f);
    float lohi = lo * hi;
    float pL = (0.5f * g) * (l.h * l.h);
    float pR = (0.5f * g) * (r.h * r.h);
    Flux f;
    f.mass = combine(hi, lo, lohi, span, qnL, qnR, l.h, r.h);
    f.normal = combine(hi, lo, lohi, span, qnL * unL + pL, qnR * unR + pR, qnL, qnR);
    f.tangent = combine(hi, lo, lohi, span, qnL * utL, qnR * utR, qtL, qtR);
    return f;
}

// ShallowWater.neighbour, :316-326; edge 0 = WALL (normal momentum reversed), 1 = OPEN.
State neighbour(const float* h, const float* hu, const float* hv, int64_t cell, bool at_edge, int64_t offset,
                int edge, bool normal_is_x) {
    int64_t index = at_edge ? cell : cell + offset;
    State s{h[index], hu[index], hv[index]};
    if (edge == 0 && at_edge) {
        if (normal_is_x) {
            s.hu = -s.hu;
        } else {
            s.hv = -s.hv;
        }
    }
    return s;
}

}  // namespace

extern "C" {

namespace {

// One step's cells [first, last) in cell order; returns the largest speed bits and adds the clamps.
int32_t swe_cells(int nx, int ny, const int edges[4], float* out_h, float* out_hu, float* out_hv, const float* in_h,
                  const float* in_hu, const float* in_hv, float g, float dry, float rx, float ry, int64_t first,
                  int64_t last, int32_t speed_max, int64_t* clamps) {
    for (int64_t cell = first; cell < last; cell++) {
        const int64_t x = cell % nx;
        const int64_t y = cell / nx;
        State here{in_h[cell], in_hu[cell], in_hv[cell]};
        State west = neighbour(in_h, in_hu, in_hv, cell, x == 0, -1, edges[0], true);
        State east = neighbour(in_h, in_hu, in_hv, cell, x == nx - 1, 1, edges[1], true);
        State south = neighbour(in_h, in_hu, in_hv, cell, y == 0, -(int64_t)nx, edges[2], false);
        State north = neighbour(in_h, in_hu, in_hv, cell, y == ny - 1, nx, edges[3], false);
        Flux fw = hll(g, dry, west, here, true);
        Flux fe = hll(g, dry, here, east, true);
        Flux fs = hll(g, dry, south, here, false);
        Flux fn = hll(g, dry, here, north, false);
        // :205-211
        float raw_h = (here.h - rx * (fe.mass - fw.mass)) - ry * (fn.mass - fs.mass);
        float new_hu = (here.hu - rx * (fe.normal - fw.normal)) - ry * (fn.tangent - fs.tangent);
        float new_hv = (here.hv - rx * (fe.tangent - fw.tangent)) - ry * (fn.normal - fs.normal);
        // :218-225
        if (raw_h < 0.0f) *clamps += 1;
        float new_h = raw_h < 0.0f ? 0.0f : raw_h;
        out_h[cell] = new_h;
        out_hu[cell] = new_hu;
        out_hv[cell] = new_hv;
        // :230-235, a signed integer max over the bits of a non-negative float.
        float u = velocity(new_h, new_hu, dry);
        float w = velocity(new_h, new_hv, dry);
        float speed = mx(std::fabs(u), std::fabs(w)) + std::sqrt(g * new_h);
        int32_t bits = float_to_bits(speed);
        if (bits > speed_max) speed_max = bits;
    }
    return speed_max;
}

// The step size of :184-192 (whole ticks, capped by the budget) and the step over the cells on
// `threads` threads; each cell is computed alone, and the speed maximum and the clamp count do not
// depend on the order, so every thread count gives the same bits.
void swe_step(int nx, int ny, const int edges[4], float* out_h, float* out_hu, float* out_hv, const float* in_h,
              const float* in_hu, const float* in_hv, const int32_t params[5], const int32_t* speed_in,
              int32_t* speed_out, int32_t* speed_clear, const int32_t* budget_in, int32_t* budget_out,
              int32_t* clamped, int threads) {
    const float g = bits_to_float(params[0]);
    const float courant = bits_to_float(params[1]);
    const float dx = bits_to_float(params[2]);
    const float dy = bits_to_float(params[3]);
    const float dry = bits_to_float(params[4]);
    const float fastest = bits_to_float(speed_in[0]);
    const float allowed = mn((courant * mn(dx, dy)) / mx(fastest, 1e-30f), 2000.0f);
    int32_t ticks = (int32_t)(allowed * 1000000.0f);
    const int32_t budget = budget_in[0];
    if (budget < ticks) ticks = budget;
    const float dt = (float)ticks * (float)(1.0 / 1000000.0);
    const float rx = dt / dx;
    const float ry = dt / dy;
    const int64_t cells = (int64_t)nx * ny;
    int32_t speed = speed_out[0];
    int64_t clamps = 0;
    if (threads <= 1) {
        speed = swe_cells(nx, ny, edges, out_h, out_hu, out_hv, in_h, in_hu, in_hv, g, dry, rx, ry, 0, cells, speed,
                          &clamps);
    } else {
#pragma omp parallel num_threads(threads) reduction(max : speed) reduction(+ : clamps)
        {
            const int t = omp_get_thread_num(), n = omp_get_num_threads();
            const int64_t first = cells * t / n, last = cells * (t + 1) / n;
            speed = swe_cells(nx, ny, edges, out_h, out_hu, out_hv, in_h, in_hu, in_hv, g, dry, rx, ry, first, last,
                              speed, &clamps);
        }
    }
    speed_out[0] = speed;
    clamped[0] += (int32_t)clamps;
    // :237-240
    speed_clear[0] = 0;
    budget_out[0] = budget - ticks;
}

}  // namespace

void orc_swe_step(int nx, int ny, const int edges[4], float* out_h, float* out_hu, float* out_hv,
                  const float* in_h, const float* in_hu, const float* in_hv, const int32_t params[5],
                  const int32_t* speed_in, int32_t* speed_out, int32_t* speed_clear, const int32_t* budget_in,
                  int32_t* budget_out, int32_t* clamped) {
    swe_step(nx, ny, edges, out_h, out_hu, out_hv, in_h, in_hu, in_hv, params, speed_in, speed_out, speed_clear,
             budget_in, budget_out, clamped, 1);
}

// The same step on `threads` host threads: bit-identical to orc_swe_step.
void orc_swe_step_omp(int nx, int ny, const int edges[4], float* out_h, float* out_hu, float* out_hv,
                      const float* in_h, const float* in_hu, const float* in_hv, const int32_t params[5],
                      const int32_t* speed_in, int32_t* speed_out, int32_t* speed_clear, const int32_t* budget_in,
                      int32_t* budget_out, int32_t* clamped, int threads) {
    swe_step(nx, ny, edges, out_h, out_hu, out_hv, in_h, in_hu, in_hv, params, speed_in, speed_out, speed_clear,
             budget_in, budget_out, clamped, threads);
}

// ShallowWater.fastestWave, :287-295 (float arithmetic with a double threshold and sqrt).
int32_t orc_swe_fastest_wave(int64_t cells, const float* h, const float* hu, const float* hv, double g,
                             double dry) {
    float fastest = 0;
    for (int64_t k = 0; k < cells; k++) {
        float u = (double)h[k] > dry ? hu[k] / h[k] : 0.0f;
        float w = (double)h[k] > dry ? hv[k] / h[k] : 0.0f;
        float s = std::max(std::fabs(u), std::fabs(w)) + (float)std::sqrt(g * (double)h[k]);
        fastest = std::max(fastest, s);
    }
    return float_to_bits(fastest);
}

// Diagnostics.of, core/stencil/Diagnostics.java:84-122.
void orc_swe_diagnostics(int64_t cells, const float* h, const float* hu, const float* hv, double g, double dry,
                         double dt, double dx, double dy, double out[9]) {
    const double shallow = 1e-3;
    double volume = 0, min_depth = std::numeric_limits<double>::infinity(), max_depth = 0;
    double max_speed = 0, max_froude = 0, max_wave = 0;
    int64_t non_finite = 0, negative = 0, wet = 0;
    for (int64_t k = 0; k < cells; k++) {
        if (!std::isfinite(h[k]) || !std::isfinite(hu[k]) || !std::isfinite(hv[k])) {
            non_finite++;
            continue;
        }
        if (h[k] < 0) negative++;
        volume += h[k];
        max_depth = std::max(max_depth, (double)h[k]);
        if ((double)h[k] <= dry) continue;
        wet++;
        min_depth = std::min(min_depth, (double)h[k]);
        // float quotients widened, as Java evaluates hu[k] / h[k] on two floats
        double u = (double)(hu[k] / h[k]);
        double v = (double)(hv[k] / h[k]);
        double c = std::sqrt(g * h[k]);
        double speed = std::sqrt(u * u + v * v);
        max_speed = std::max(max_speed, speed);
        if (h[k] >= shallow) max_froude = std::max(max_froude, speed / c);
        max_wave = std::max(max_wave, std::max(std::fabs(u), std::fabs(v)) + c);
    }
    out[0] = volume * dx * dy;
    out[1] = wet == 0 ? 0 : min_depth;
    out[2] = max_depth;
    out[3] = max_speed;
    out[4] = max_froude;
    out[5] = max_wave * dt / std::min(dx, dy);
    out[6] = (double)non_finite;
    out[7] = (double)negative;
    out[8] = (double)wet;
}

// The same HLL rule in binary64 with a given dt: the error judge for reduced-precision runs.
void orc_swe_step_f64(int nx, int ny, const int edges[4], double* out_h, double* out_hu, double* out_hv,
                      const double* in_h, const double* in_hu, const double* in_hv, double g, double dx, double dy,
                      double dry, double dt) {
    struct S {
        double h, hu, hv;
    };
    auto vel = [&](double h, double q) { return h > dry ? q / h : 0.0; };
    auto flux = [&](const S& l, const S& r, bool nx_, double f[3]) {
        double qnL = nx_ ? l.hu : l.hv, qtL = nx_ ? l.hv : l.hu;
        double qnR = nx_ ? r.hu : r.hv, qtR = nx_ ? r.hv : r.hu;
        double unL = vel(l.h, qnL), utL = vel(l.h, qtL), unR = vel(r.h, qnR), utR = vel(r.h, qtR);
        double cL = std::sqrt(g * l.h), cR = std::sqrt(g * r.h);
        double sL = std::min(unL - cL, unR - cR), sR = std::max(unL + cL, unR + cR);
        if (!(r.h > dry)) {
            sL = unL - cL;
            sR = unL + 2 * cL;
        }
        if (!(l.h > dry)) {
            sL = unR - 2 * cR;
            sR = unR + cR;
        }
        double lo = std::min(sL, 0.0), hi = std::max(sR, 0.0), span = std::max(hi - lo, 1e-300), lohi = lo * hi;
        double pL = 0.5 * g * l.h * l.h, pR = 0.5 * g * r.h * r.h;
        auto comb = [&](double fl, double fr, double ul, double ur) {
            return ((hi * fl - lo * fr) + lohi * (ur - ul)) / span;
        };
        f[0] = comb(qnL, qnR, l.h, r.h);
        f[1] = comb(qnL * unL + pL, qnR * unR + pR, qnL, qnR);
        f[2] = comb(qnL * utL, qnR * utR, qtL, qtR);
    };
    auto nb = [&](int64_t cell, bool edge, int64_t off, int kind, bool nx_) {
        int64_t i = edge ? cell : cell + off;
        S s{in_h[i], in_hu[i], in_hv[i]};
        if (kind == 0 && edge) {
            if (nx_) s.hu = -s.hu;
            else s.hv = -s.hv;
        }
        return s;
    };
    const double rx = dt / dx, ry = dt / dy;
    const int64_t cells = (int64_t)nx * ny;
    for (int64_t cell = 0; cell < cells; cell++) {
        int64_t x = cell % nx, y = cell / nx;
        S c{in_h[cell], in_hu[cell], in_hv[cell]};
        S w = nb(cell, x == 0, -1, edges[0], true), e = nb(cell, x == nx - 1, 1, edges[1], true);
        S s = nb(cell, y == 0, -(int64_t)nx, edges[2], false), n = nb(cell, y == ny - 1, nx, edges[3], false);
        double fw[3], fe[3], fs[3], fn[3];
        flux(w, c, true, fw);
        flux(c, e, true, fe);
        flux(s, c, false, fs);
        flux(c, n, false, fn);
        double h = (c.h - rx * (fe[0] - fw[0])) - ry * (fn[0] - fs[0]);
        out_h[cell] = h < 0 ? 0 : h;
        out_hu[cell] = (c.hu - rx * (fe[1] - fw[1])) - ry * (fn[2] - fs[2]);
        out_hv[cell] = (c.hv - rx * (fe[2] - fw[2])) - ry * (fn[1] - fs[1]);
    }
}

}  // extern "C"

extern "C" {

// ParticleDiagnostics.of (core/particle/ParticleDiagnostics.java:45-65): the valid J (finite and
// above zero) sorted; density 1/J at the extremes and at the 1st and 99th percentile.
void orc_particle_diagnostics(int64_t count, const float* j, double speed, double sound, double dt, double out[7]) {
    std::vector<float> valid;
    valid.reserve((size_t)count);
    for (int64_t k = 0; k < count; k++)
        if (std::isfinite(j[k]) && j[k] > 0) valid.push_back(j[k]);
    const int64_t n = (int64_t)valid.size();
    out[4] = speed / sound;
    out[5] = (speed + sound) * dt;
    out[6] = (double)(count - n);
    if (n == 0) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    std::sort(valid.begin(), valid.end());
    auto percentile = [&](double p) { return std::min<int64_t>(n - 1, (int64_t)(p * (double)n)); };
    // 1 / valid[i] on a float[]: a binary32 quotient, widened (ParticleDiagnostics.java:50-51)
    out[0] = (double)(1.0f / valid[n - 1]);
    out[1] = (double)(1.0f / valid[0]);
    out[2] = (double)(1.0f / valid[percentile(0.99)]);
    out[3] = (double)(1.0f / valid[percentile(0.01)]);
}

}  // extern "C"
