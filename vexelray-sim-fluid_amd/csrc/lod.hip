// The 3D level of detail for gfx950: Flip3.lod (core/particle/Flip3.java:420-586) and
// Flip3.lodTarget (:600-641), one thread a group of GROUP = 8 consecutive slots, in the
// reference's binary32 operations and order (so the CPU restatement's results, bit for bit).
// A group is the 8 slots loaded consecutively; a sort moves slots, so a slot's fields are read and
// written at its position home[slot] (the identity until the first sort).
#include "common.h"

namespace vxr {
namespace {

constexpr int kBlock = 256;
constexpr int kGroup = 8;
constexpr float kBlend = 0.15f, kSnap = 0.002f, kSettled = 1e-6f, kJoin = 0.25f, kWall = 1.0f;
constexpr float kHysteresis = 1.25f;
constexpr int kMass = 6, kVolume = 7;

__global__ __launch_bounds__(kBlock) void lod_kernel(Mpm3Args a, int64_t groups, float* __restrict__ level,
                                                     const float* __restrict__ target, float* __restrict__ stay) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    // slots may gain mass: until the next sort, P2G and G2P visit every particle
    if (g == 0) *a.active = (int32_t)a.count;
    if (g >= groups) return;
    const int size[3] = {a.nx, a.ny, a.nz};
    const int64_t base = g * kGroup;
    int64_t at[kGroup];
#pragma unroll
    for (int s = 0; s < kGroup; s++) at[s] = a.home[base + s];
    float m[kGroup], keep[kGroup];
    float total = 0.0f, count = 0.0f;
#pragma unroll
    for (int s = 0; s < kGroup; s++) {
        m[s] = a.p[kMass][at[s]];
        keep[s] = stay[base + s];
        total = total + m[s];
        count = count + keep[s];
    }
    const float want = clampf(target[g], 1.0f, (float)kGroup);
    const float share = total / mx(count, 1.0f);
    float spare = 0.0f;
#pragma unroll
    for (int s = 0; s < kGroup; s++) spare = spare + mx(m[s] - keep[s] * share, 0.0f);
    float work = 0.0f;
    if (spare > total * kSettled) work = 1.0f;
    if (fabsf(count - want) > 0.5f) work = 1.0f;
    if (!(total > 0.0f) || !(work > 0.5f)) return;

    const float inv = 1.0f / total;
    float mean[17];
#pragma unroll
    for (int f = 0; f < 17; f++) {
        mean[f] = 0.0f;
        if (f == kMass) continue;
        float acc = 0.0f;
#pragma unroll
        for (int s = 0; s < kGroup; s++) acc = acc + m[s] * a.p[f][at[s]];
        mean[f] = acc * inv;
    }
    // too many meant to stay: the ones nearest the mean position stop, one at a time
    float remove = mx(count - want, 0.0f);
    float near[kGroup];
#pragma unroll
    for (int s = 0; s < kGroup; s++) {
        near[s] = 0.0f;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const float d = a.p[ax][at[s]] - mean[ax];
            near[s] = near[s] + d * d;
        }
    }
#pragma unroll
    for (int round = 0; round < kGroup - 1; round++) {
        if (!(remove > 0.5f)) continue;
        float best = -1.0f, best_at = 3.40282347e38f;
#pragma unroll
        for (int s = 0; s < kGroup; s++) {
            if (keep[s] > 0.5f && near[s] < best_at) {
                best_at = near[s];
                best = (float)s;
            }
        }
#pragma unroll
        for (int s = 0; s < kGroup; s++)
            if (best == (float)s) keep[s] = 0.0f;
        remove = remove - 1.0f;
    }
    // too few: slots not staying stay again, an empty one set down at the group's mean first
    float join = mx(want - count, 0.0f);
#pragma unroll
    for (int s = 0; s < kGroup; s++) {
        if (!(keep[s] < 0.5f) || !(join > 0.5f)) continue;
        keep[s] = 1.0f;
        join = join - 1.0f;
        if (m[s] > 0.0f) continue;
        float offset[3];
#pragma unroll
        for (int ax = 0; ax < 3; ax++) offset[ax] = ((float)((s >> ax) & 1) - 0.5f) * 2.0f * kJoin;
        const int64_t q = at[s];
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            a.p[ax][q] = clampf(mean[ax] + offset[ax], kWall, (float)(size[ax] - 1) - kWall);
            float vel = mean[3 + ax];
#pragma unroll
            for (int e = 0; e < 3; e++) vel = vel + mean[8 + 3 * ax + e] * offset[e];
            a.p[3 + ax][q] = vel;
        }
        a.p[kVolume][q] = mean[kVolume];
#pragma unroll
        for (int e = 0; e < 9; e++) a.p[8 + e][q] = mean[8 + e];
    }
    // the shares those meant to stay now have, who has too much and who too little
    float stayers = 0.0f;
#pragma unroll
    for (int s = 0; s < kGroup; s++) stayers = stayers + keep[s];
    const float equal = total / mx(stayers, 1.0f);
    float extra[kGroup], lack[kGroup], movable = 0.0f;
#pragma unroll
    for (int s = 0; s < kGroup; s++) {
        extra[s] = mx(m[s] - keep[s] * equal, 0.0f);
        lack[s] = mx(keep[s] * equal - m[s], 0.0f);
        movable = movable + extra[s];
    }
    float rate = kBlend;
    if (movable < total * kSnap) rate = 1.0f;
    if (movable > 0.0f) {
        float gain[kGroup];
#pragma unroll
        for (int s = 0; s < kGroup; s++) gain[s] = rate * lack[s];
        const float inverse = 1.0f / movable;
#pragma unroll
        for (int f = 3; f < 17; f++) {
            if (f == kMass) continue;
            float given = 0.0f;
#pragma unroll
            for (int s = 0; s < kGroup; s++) given = given + extra[s] * a.p[f][at[s]];
            given = given * inverse;
#pragma unroll
            for (int s = 0; s < kGroup; s++)
                if (gain[s] > 0.0f) a.p[f][at[s]] = (m[s] * a.p[f][at[s]] + gain[s] * given) / (m[s] + gain[s]);
        }
#pragma unroll
        for (int s = 0; s < kGroup; s++) a.p[kMass][at[s]] = (m[s] + gain[s]) - rate * extra[s];
    }
#pragma unroll
    for (int s = 0; s < kGroup; s++) stay[base + s] = keep[s];
    level[g] = stayers;
}

__global__ __launch_bounds__(kBlock) void lod_target_kernel(Mpm3Args a, int64_t groups,
                                                            const float* __restrict__ level,
                                                            float* __restrict__ target,
                                                            const float* __restrict__ view) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= groups) return;
    const int64_t first = a.home[g * kGroup];  // the group's first slot
    float d2 = 0.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const float d = a.p[ax][first] - view[ax];
        d2 = d2 + d * d;
    }
    const float near = view[3];
    const float n2 = near * near;
    const float floor_ = view[4];
    const float n = level[g];
    float want = (float)kGroup;
    const float ratio[3] = {1.0f, 4.0f, 16.0f};
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (!(d2 < n2 * ratio[k])) want = (float)(kGroup >> (k + 1));
    want = mx(want, floor_);
    float keep = 0.0f;
    // [lo, hi) of each level's squared distances in units of near^2, widened by the hysteresis
    const float h2 = kHysteresis * kHysteresis;
    const float lo[4] = {-1.0f, 1.0f / h2, 4.0f / h2, 16.0f / h2};
    const float hi[4] = {1.0f * kHysteresis * kHysteresis, 4.0f * kHysteresis * kHysteresis,
                         16.0f * kHysteresis * kHysteresis, 3.40282347e38f};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float lv = (float)(kGroup >> k);
        if (n == lv && !(d2 < n2 * lo[k]) && d2 < n2 * hi[k]) keep = 1.0f;
    }
    if (keep > 0.5f && !(n < floor_)) want = n;
    target[g] = want;
}

}  // namespace

int mpm3_lod(const Launch& l, const Mpm3Args& a, float* level, const float* target, float* stay) {
    const int64_t groups = a.count / kGroup;
    if (groups == 0) return 0;
    timed_begin(l, "lod3");
    hipLaunchKernelGGL(lod_kernel, dim3((unsigned)((groups + kBlock - 1) / kBlock)), dim3(kBlock), 0, l.stream, a,
                       groups, level, target, stay);
    timed_end(l, "lod3");
    return check_launch("lod3");
}

int mpm3_lod_target(const Launch& l, const Mpm3Args& a, const float* level, float* target, const float* view) {
    const int64_t groups = a.count / kGroup;
    if (groups == 0) return 0;
    timed_begin(l, "lod3_target");
    hipLaunchKernelGGL(lod_target_kernel, dim3((unsigned)((groups + kBlock - 1) / kBlock)), dim3(kBlock), 0, l.stream,
                       a, groups, level, target, view);
    timed_end(l, "lod3_target");
    return check_launch("lod3_target");
}

}  // namespace vxr
