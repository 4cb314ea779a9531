// The 3D level of detail restated: Flip3.lod (core/particle/Flip3.java:420-586), one group of
// GROUP = 8 slots at a time, and Flip3.lodTarget (:600-641). binary32 in the IR's order, min/max
// as selects, no contraction. Test infrastructure only.
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "oracle.h"

namespace {

inline float mn(float a, float b) { return b < a ? b : a; }
inline float mx(float a, float b) { return a < b ? b : a; }
inline float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }

const int kGroup = 8;           // Flip3.GROUP
const float kBlend = 0.15f;     // Flip3.BLEND
const float kSnap = 0.002f;     // Flip3.SNAP
const float kSettled = 1e-6f;   // Flip3.SETTLED
const float kJoin = 0.25f;      // Flip3.JOIN_OFFSET
const float kWall = 1.0f;       // Flip.WALL
const float kHysteresis = 1.25f;
const int kMass = 6, kVolume = 7;

}  // namespace

extern "C" {

void orc_mpm3_lod(int nx, int ny, int nz, int64_t groups, float* const p[17], float* level, const float* target,
                  float* stay) {
    const int size[3] = {nx, ny, nz};
    for (int64_t g = 0; g < groups; g++) {
        const int64_t base = g * kGroup;
        float m[kGroup], keep[kGroup];
        float total = 0.0f, count = 0.0f;
        for (int s = 0; s < kGroup; s++) {
            m[s] = p[kMass][base + s];
            keep[s] = stay[base + s];
            total = total + m[s];
            count = count + keep[s];
        }
        const float want = clampf(target[g], 1.0f, (float)kGroup);
        const float share = total / mx(count, 1.0f);
        float spare = 0.0f;
        for (int s = 0; s < kGroup; s++) spare = spare + mx(m[s] - keep[s] * share, 0.0f);
        float work = 0.0f;
        if (spare > total * kSettled) work = 1.0f;
        if (std::fabs(count - want) > 0.5f) work = 1.0f;
        if (!(total > 0.0f) || !(work > 0.5f)) continue;

        const float inv = 1.0f / total;
        float mean[17] = {};
        for (int f = 0; f < 17; f++) {
            if (f == kMass) continue;
            float acc = 0.0f;
            for (int s = 0; s < kGroup; s++) acc = acc + m[s] * p[f][base + s];
            mean[f] = acc * inv;
        }
        // too many meant to stay: the ones nearest the mean position stop, one at a time
        float remove = mx(count - want, 0.0f);
        float near[kGroup];
        for (int s = 0; s < kGroup; s++) {
            near[s] = 0.0f;
            for (int a = 0; a < 3; a++) {
                const float d = p[a][base + s] - mean[a];
                near[s] = near[s] + d * d;
            }
        }
        for (int round = 0; round < kGroup - 1; round++) {
            if (!(remove > 0.5f)) continue;
            float best = -1.0f, best_at = FLT_MAX;
            for (int s = 0; s < kGroup; s++) {
                if (keep[s] > 0.5f && near[s] < best_at) {
                    best_at = near[s];
                    best = (float)s;
                }
            }
            for (int s = 0; s < kGroup; s++)
                if (best == (float)s) keep[s] = 0.0f;
            remove = remove - 1.0f;
        }
        // too few: slots not staying stay again, an empty one set down at the group's mean first
        float join = mx(want - count, 0.0f);
        for (int s = 0; s < kGroup; s++) {
            if (!(keep[s] < 0.5f) || !(join > 0.5f)) continue;
            keep[s] = 1.0f;
            join = join - 1.0f;
            if (m[s] > 0.0f) continue;
            float offset[3];
            for (int a = 0; a < 3; a++) offset[a] = ((float)((s >> a) & 1) - 0.5f) * 2 * kJoin;
            const int64_t at = base + s;
            for (int a = 0; a < 3; a++) {
                p[a][at] = clampf(mean[a] + offset[a], kWall, (float)(size[a] - 1) - kWall);
                float vel = mean[3 + a];
                for (int e = 0; e < 3; e++) vel = vel + mean[8 + 3 * a + e] * offset[e];
                p[3 + a][at] = vel;
            }
            p[kVolume][at] = mean[kVolume];
            for (int e = 0; e < 9; e++) p[8 + e][at] = mean[8 + e];
        }
        // the shares those meant to stay now have, who has too much and who too little
        float stayers = 0.0f;
        for (int s = 0; s < kGroup; s++) stayers = stayers + keep[s];
        const float equal = total / mx(stayers, 1.0f);
        float extra[kGroup], lack[kGroup], movable = 0.0f;
        for (int s = 0; s < kGroup; s++) {
            extra[s] = mx(m[s] - keep[s] * equal, 0.0f);
            lack[s] = mx(keep[s] * equal - m[s], 0.0f);
            movable = movable + extra[s];
        }
        float rate = kBlend;
        if (movable < total * kSnap) rate = 1.0f;
        if (movable > 0.0f) {
            float gain[kGroup];
            for (int s = 0; s < kGroup; s++) gain[s] = rate * lack[s];
            const float inverse = 1.0f / movable;
            for (int f = 3; f < 17; f++) {
                if (f == kMass) continue;
                float given = 0.0f;
                for (int s = 0; s < kGroup; s++) given = given + extra[s] * p[f][base + s];
                given = given * inverse;
                for (int s = 0; s < kGroup; s++)
                    if (gain[s] > 0.0f)
                        p[f][base + s] = (m[s] * p[f][base + s] + gain[s] * given) / (m[s] + gain[s]);
            }
            for (int s = 0; s < kGroup; s++) p[kMass][base + s] = (m[s] + gain[s]) - rate * extra[s];
        }
        for (int s = 0; s < kGroup; s++) stay[base + s] = keep[s];
        level[g] = stayers;
    }
}

void orc_mpm3_lod_target(int64_t groups, const float* x, const float* y, const float* z, const float* level,
                         float* target, const float view[5]) {
    const float* at[3] = {x, y, z};
    for (int64_t g = 0; g < groups; g++) {
        const int64_t slot = g * kGroup;
        float d2 = 0.0f;
        for (int a = 0; a < 3; a++) {
            const float d = at[a][slot] - view[a];
            d2 = d2 + d * d;
        }
        const float near = view[3];
        const float n2 = near * near;
        const float floor_ = view[4];
        const float n = level[g];
        float want = (float)kGroup;
        for (int k = 0; k < 3; k++) {
            const float ratio = (float)std::pow(4.0, k);
            if (!(d2 < n2 * ratio)) want = (float)(kGroup >> (k + 1));
        }
        want = mx(want, floor_);
        float keep = 0.0f;
        for (int lv = kGroup, k = 0; lv >= 1; lv /= 2, k++) {
            const float lo = k == 0 ? -1.0f : (float)std::pow(4.0, k - 1) / (kHysteresis * kHysteresis);
            const float hi = lv == 1 ? FLT_MAX : (float)std::pow(4.0, k) * kHysteresis * kHysteresis;
            if (n == (float)lv && !(d2 < n2 * lo) && d2 < n2 * hi) keep = 1.0f;
        }
        if (keep > 0.5f && !(n < floor_)) want = n;
        target[g] = want;
    }
}

}  // extern "C"
