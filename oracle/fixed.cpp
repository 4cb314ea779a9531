// The reference's fixed-point bilinear scatter, restated: core/particle/FixedScatter.java:151-228
// (the direct and segmented kernels and their Quanta), as its own tests restate it on the host
// (test/particle/FixedScatterTest.java:402-432), and the FP64 bilinear deposit those tests judge
// it by (test/particle/ScatterTest.java:429-450). Java int arithmetic wraps; here it runs in
// uint32_t so that it wraps too. Test infrastructure only.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oracle.h"

namespace {

inline float mn(float a, float b) { return b < a ? b : a; }
inline float mx(float a, float b) { return a < b ? b : a; }

// Java's (int) of a float: toward zero, NaN to 0, saturating at the int range.
inline int32_t java_f2i(float f) {
    if (std::isnan(f)) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t wrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

}  // namespace

extern "C" {

int32_t orc_fixed_split(int32_t q, int32_t mass) {
    const int32_t one = 1 << ORC_FIXED_OFFSET_BITS;
    return wrap_add(wrap_mul(q, mass / one), wrap_mul(q, mass % one) / one) / one;
}

void orc_fixed_scatter(int n, int64_t count, const float* x, const float* y, const float* u, const float* v,
                       const float* m, int mass_bits, int velocity_bits, int32_t* gm, int32_t* gmu, int32_t* gmv) {
    const float mass_scale = (float)std::ldexp(1.0, mass_bits);
    const float velocity_scale = (float)std::ldexp(1.0, velocity_bits);
    const int32_t one = 1 << ORC_FIXED_OFFSET_BITS;
    for (int64_t p = 0; p < count; p++) {
        const float px = mn(mx(x[p], 0.0f), (float)n - 1.0f);
        const float py = mn(mx(y[p], 0.0f), (float)n - 1.0f);
        const int32_t col = java_f2i(mn(px, (float)n - 2.0f));
        const int32_t row = java_f2i(mn(py, (float)n - 2.0f));
        const float fx = px - (float)col;
        const float fy = py - (float)row;
        const int32_t mass = java_f2i(m[p] * mass_scale);
        const int32_t uq = java_f2i(u[p] * velocity_scale);
        const int32_t vq = java_f2i(v[p] * velocity_scale);
        const int32_t ox = java_f2i(fx * (float)one);
        const int32_t oy = java_f2i(fy * (float)one);
        int32_t left = mass;
        for (int k = 0; k < 4; k++) {
            int32_t share = left;
            if (k < 3) {
                const int32_t q = wrap_mul((k & 1) ? ox : one - ox, (k >> 1) ? oy : one - oy);
                share = orc_fixed_split(q, mass);
                left = (int32_t)((uint32_t)left - (uint32_t)share);
            }
            const int64_t node = (int64_t)(row + (k >> 1)) * n + col + (k & 1);
            gm[node] = wrap_add(gm[node], share);
            gmu[node] = wrap_add(gmu[node], wrap_mul(share, uq));
            gmv[node] = wrap_add(gmv[node], wrap_mul(share, vq));
        }
    }
}

void orc_scatter_f64(int n, int64_t count, const float* x, const float* y, const float* u, const float* v,
                     const float* m, double* gm, double* gmu, double* gmv) {
    for (int64_t p = 0; p < count; p++) {
        const double px = std::fmin(std::fmax((double)x[p], 0.0), n - 1.0);
        const double py = std::fmin(std::fmax((double)y[p], 0.0), n - 1.0);
        const int32_t col = (int32_t)std::fmin(px, n - 2.0);
        const int32_t row = (int32_t)std::fmin(py, n - 2.0);
        const double fx = px - col, fy = py - row;
        const double mass = m[p];
        for (int k = 0; k < 4; k++) {
            const int east = k & 1, north = k >> 1;
            const double w = (east ? fx : 1 - fx) * (north ? fy : 1 - fy);
            const int64_t node = (int64_t)(row + north) * n + col + east;
            gm[node] += w * mass;
            gmu[node] += w * mass * u[p];
            gmv[node] += w * mass * v[p];
        }
    }
}

}  // extern "C"
