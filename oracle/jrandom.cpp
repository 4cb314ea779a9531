// java.util.Random restated: the 48-bit linear congruential generator of the Java SE
// specification (class java.util.Random: setSeed, next(bits), nextInt, nextInt(bound),
// nextFloat, nextDouble, nextGaussian's polar method). The reference seeds every test
// input with it, e.g. test/particle/FlipTest.java:134-150 (Random(71)),
// test/particle/ScatterTest.java:460-493 (jittered, shuffled), demo FlatRun.fill (Random(1)).
// Test infrastructure only.
#include "oracle.h"

#include <cmath>
#include <cstring>

struct orc_rng {
    int64_t seed;
    bool have_next_gaussian;
    double next_gaussian;
};

static const int64_t kMultiplier = 0x5DEECE66DLL;
static const int64_t kAddend = 0xBLL;
static const int64_t kMask = (1LL << 48) - 1;

static int32_t next_bits(orc_rng* r, int bits) {
    r->seed = (int64_t)(((uint64_t)r->seed * (uint64_t)kMultiplier + (uint64_t)kAddend) & (uint64_t)kMask);
    return (int32_t)(uint32_t)((uint64_t)r->seed >> (48 - bits));
}

extern "C" {

orc_rng* orc_rng_new(int64_t seed) {
    orc_rng* r = new orc_rng;
    r->seed = (seed ^ kMultiplier) & kMask;
    r->have_next_gaussian = false;
    r->next_gaussian = 0;
    return r;
}

void orc_rng_free(orc_rng* r) { delete r; }

int32_t orc_rng_next_int(orc_rng* r) { return next_bits(r, 32); }

int32_t orc_rng_next_int_bound(orc_rng* r, int32_t bound) {
    int32_t v = next_bits(r, 31);
    int32_t m = bound - 1;
    if ((bound & m) == 0) {
        return (int32_t)(((int64_t)bound * (int64_t)v) >> 31);
    }
    for (int32_t u = v; (int32_t)((uint32_t)u - (uint32_t)(v = u % bound) + (uint32_t)m) < 0; u = next_bits(r, 31)) {
    }
    return v;
}

float orc_rng_next_float(orc_rng* r) { return (float)next_bits(r, 24) * 0x1.0p-24f; }

double orc_rng_next_double(orc_rng* r) {
    int64_t hi = (int64_t)next_bits(r, 26);
    int64_t lo = (int64_t)next_bits(r, 27);
    return (double)((hi << 27) + lo) * 0x1.0p-53;
}

double orc_rng_next_gaussian(orc_rng* r) {
    if (r->have_next_gaussian) {
        r->have_next_gaussian = false;
        return r->next_gaussian;
    }
    double v1, v2, s;
    do {
        v1 = 2 * orc_rng_next_double(r) - 1;
        v2 = 2 * orc_rng_next_double(r) - 1;
        s = v1 * v1 + v2 * v2;
    } while (s >= 1 || s == 0);
    // StrictMath.log is fdlibm's; libm's log agrees to the last bit on all but rare inputs.
    double multiplier = std::sqrt(-2 * std::log(s) / s);
    r->next_gaussian = v2 * multiplier;
    r->have_next_gaussian = true;
    return v1 * multiplier;
}

void orc_rng_pattern(orc_rng* r, const char* pattern, int64_t count, float* out) {
    size_t len = std::strlen(pattern);
    int64_t k = 0;
    for (int64_t n = 0; n < count; n++) {
        for (size_t c = 0; c < len; c++) {
            switch (pattern[c]) {
                case 'f': out[k++] = orc_rng_next_float(r); break;
                case 'g': out[k++] = (float)orc_rng_next_gaussian(r); break;
                case 'd': out[k++] = (float)orc_rng_next_double(r); break;
                default: out[k++] = 0; break;
            }
        }
    }
}

void orc_rng_shuffle_order(orc_rng* r, int32_t n, int32_t* order) {
    for (int32_t k = 0; k < n; k++) order[k] = k;
    for (int32_t k = n - 1; k > 0; k--) {
        int32_t j = orc_rng_next_int_bound(r, k + 1);
        int32_t t = order[k];
        order[k] = order[j];
        order[j] = t;
    }
}

}  // extern "C"
