// The first-order HLL shallow-water step (core/stencil/ShallowWater.java:166-384) for gfx950.
//
// Layout: h, hu, hv are row-major f32 planes (SoA). A wave owns a 62-column strip and marches
// `rows` rows of it north: lane L holds column x0 + L - 1, so lanes 1..62 own cells and lanes
// 0 and 63 are the strip's west and east halo. Per row each lane computes one east face and
// one north face; the west face comes from the lane below by a cross-lane shuffle and the
// south face is the previous row's north face kept in registers. Each face flux is computed
// once, from the same two states by the same code the reference runs twice, so the result is
// the reference's bit for bit (built with -ffp-contract=off and correctly rounded div/sqrt).
// No LDS and no barriers: a wave needs nothing from another wave.
//
// The fused step-size reduction (:227-235) is a wave max of the new state's wave speed and
// one atomic max per wave; the clamp count (:218-219) is a wave sum and one atomic add.
//
// The step is bound by its VALU work, not by HBM: about 357 vector instructions a cell, most of
// them in its 10 correctly rounded divisions and 2 square roots. kFast = true trades the
// reference's bits for fewer: one reciprocal (the hardware's, refined by a Newton step) per
// face for its three flux quotients and one per cell for its two velocities, and the hardware
// square root. Its results differ from the reference's by a few ulps a step.
//
// Strips are dealt to workgroups so that each of the 8 XCDs takes a contiguous run of them (the
// hardware dispatches workgroups to XCDs in turn): the halo columns and rows a strip shares with
// its neighbours then sit in the same XCD's L2.
#include "common.h"

namespace vxr {
namespace {

struct Cell {
    float h, hu, hv;  // conserved state
    float u, v, c;    // velocity (zero where dry) and sqrt(g h), as the flux takes them
};

struct Flux {
    float mass, normal, tangent;
};

__device__ __forceinline__ float bits_f(int32_t w) { return __int_as_float(w); }

// ShallowWater.velocity, :329-333
__device__ __forceinline__ float velocity(float h, float q, float dry) { return h > dry ? q / h : 0.0f; }

// 1/d from the hardware reciprocal (about 1 ulp) and one Newton step.
__device__ __forceinline__ float recip(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    return fmaf(fmaf(-d, r, 1.0f), r, r);
}

template <bool kFast>
__device__ __forceinline__ float root(float v) {
    if constexpr (kFast) return __builtin_amdgcn_sqrtf(v);
    return sqrtf(v);
}

template <bool kFast>
__device__ __forceinline__ Cell make_cell(float h, float hu, float hv, float g, float dry) {
    Cell s;
    s.h = h;
    s.hu = hu;
    s.hv = hv;
    if constexpr (kFast) {
        const float inv = recip(h);
        s.u = h > dry ? hu * inv : 0.0f;
        s.v = h > dry ? hv * inv : 0.0f;
    } else {
        s.u = velocity(h, hu, dry);
        s.v = velocity(h, hv, dry);
    }
    s.c = root<kFast>(g * h);
    return s;
}

// The wall ghost of ShallowWater.neighbour (:316-326): the edge cell with its normal momentum
// negated. Its velocity is velocity(h, -q), which is -(q/h) exactly when wet (IEEE division is
// odd in its numerator) and +0 when dry.
__device__ __forceinline__ Cell mirror(Cell s, bool normal_is_x, float dry) {
    if (normal_is_x) {
        s.hu = -s.hu;
        s.u = s.h > dry ? -s.u : 0.0f;
    } else {
        s.hv = -s.hv;
        s.v = s.h > dry ? -s.v : 0.0f;
    }
    return s;
}

// ShallowWater.combine, :381-384; with kFast, `span` is 1/span and multiplies
template <bool kFast>
__device__ __forceinline__ float combine(float hi, float lo, float lohi, float span, float fl, float fr, float ul,
                                         float ur) {
    if constexpr (kFast) return ((hi * fl - lo * fr) + lohi * (ur - ul)) * span;
    return ((hi * fl - lo * fr) + lohi * (ur - ul)) / span;
}

// ShallowWater.hll, :336-378
template <bool kFast>
__device__ __forceinline__ Flux hll(const Cell& l, const Cell& r, bool normal_is_x, float g, float dry) {
    const float qnL = normal_is_x ? l.hu : l.hv;
    const float qtL = normal_is_x ? l.hv : l.hu;
    const float qnR = normal_is_x ? r.hu : r.hv;
    const float qtR = normal_is_x ? r.hv : r.hu;
    const float unL = normal_is_x ? l.u : l.v;
    const float utL = normal_is_x ? l.v : l.u;
    const float unR = normal_is_x ? r.u : r.v;
    const float utR = normal_is_x ? r.v : r.u;
    const float cL = l.c;
    const float cR = r.c;
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
    const float lo = mn(sL, 0.0f);
    const float hi = mx(sR, 0.0f);
    float span = mx(hi - lo, 1e-30f);
    if constexpr (kFast) span = recip(span);
    const float lohi = lo * hi;
    const float pL = (0.5f * g) * (l.h * l.h);
    const float pR = (0.5f * g) * (r.h * r.h);
    Flux f;
    f.mass = combine<kFast>(hi, lo, lohi, span, qnL, qnR, l.h, r.h);
    f.normal = combine<kFast>(hi, lo, lohi, span, qnL * unL + pL, qnR * unR + pR, qnL, qnR);
    f.tangent = combine<kFast>(hi, lo, lohi, span, qnL * utL, qnR * utR, qtL, qtR);
    return f;
}

__device__ __forceinline__ Cell shfl_down_cell(const Cell& s) {
    Cell r;
    r.h = __shfl_down(s.h, 1);
    r.hu = __shfl_down(s.hu, 1);
    r.hv = __shfl_down(s.hv, 1);
    r.u = __shfl_down(s.u, 1);
    r.v = __shfl_down(s.v, 1);
    r.c = __shfl_down(s.c, 1);
    return r;
}

constexpr int kStrip = kWave - 2;  // cells a wave owns across
constexpr int kWavesPerBlock = 4;
constexpr int kXcds = 8;

// Workgroup `id` of a launch of `per` * kXcds: XCD id % kXcds takes the run of tiles
// [(id % kXcds) * per, (id % kXcds + 1) * per), strips across first, then bands.
__device__ __forceinline__ int64_t tile_of_workgroup(int64_t id, int64_t per) {
    return (id % kXcds) * per + id / kXcds;
}

template <bool kFast>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void hll_step_kernel(SweArgs a, int rows, int strips,
                                                                          int64_t tiles, int64_t per) {
    const int nx = a.nx, ny = a.ny;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t tile = tile_of_workgroup(blockIdx.x, per);
    if (tile >= tiles) return;  // the last XCD's run may be short; uniform across the block
    const int strip = (int)(tile % strips);
    const int band = (int)(tile / strips) * kWavesPerBlock + (threadIdx.x >> 6);
    const int y_begin = band * rows;
    if (y_begin >= ny) return;  // uniform across the wave; no barrier follows
    if (a.gate && *a.gate == 0) return;  // a gated step not called for
    const int y_end = min(y_begin + rows, ny);
    const int x = strip * kStrip + lane - 1;
    const int xc = min(max(x, 0), nx - 1);
    const bool owner = lane >= 1 && lane <= kStrip && x < nx;

    const float g = bits_f(a.params[0]);
    const float courant = bits_f(a.params[1]);
    const float dx = bits_f(a.params[2]);
    const float dy = bits_f(a.params[3]);
    const float dry = bits_f(a.params[4]);
    // :184-192, the step in whole ticks, capped by the budget (or a fixed number of ticks)
    int32_t ticks = a.fixed_ticks, budget = 0;
    if (ticks <= 0) {
        const float fastest = bits_f(a.speed_in[0]);
        const float allowed = mn((courant * mn(dx, dy)) / mx(fastest, 1e-30f), 2000.0f);
        ticks = (int32_t)(allowed * 1000000.0f);
        budget = a.budget_in[0];
        if (budget < ticks) ticks = budget;
    }
    const float dt = (float)ticks * (float)(1.0 / 1000000.0);
    const float rx = dt / dx;
    const float ry = dt / dy;

    const bool wall_w = a.edges[0] == 0, wall_e = a.edges[1] == 0;
    const bool wall_s = a.edges[2] == 0, wall_n = a.edges[3] == 0;
    const float* __restrict__ ih = a.in[0];
    const float* __restrict__ ihu = a.in[1];
    const float* __restrict__ ihv = a.in[2];

    auto load = [&](int y) {
        const size_t i = (size_t)y * nx + xc;
        return make_cell<kFast>(ih[i], ihu[i], ihv[i], g, dry);
    };

    Cell cur = load(y_begin);
    Cell south = y_begin == 0 ? (wall_s ? mirror(cur, false, dry) : cur) : load(y_begin - 1);
    Flux fs = hll<kFast>(south, cur, false, g, dry);
    Cell nrt = y_begin + 1 < ny ? load(y_begin + 1) : cur;

    int32_t speed_max = 0;
    int clamps = 0;
    for (int y = y_begin; y < y_end; ++y) {
        // the row after next, in flight while this row computes
        const bool prefetch = y + 1 < y_end && y + 2 < ny;
        float ph = 0, phu = 0, phv = 0;
        if (prefetch) {
            const size_t i = (size_t)(y + 2) * nx + xc;
            ph = ih[i];
            phu = ihu[i];
            phv = ihv[i];
        }
        const Cell north = y + 1 < ny ? nrt : (wall_n ? mirror(cur, false, dry) : cur);
        const Flux fn = hll<kFast>(cur, north, false, g, dry);

        Cell right = shfl_down_cell(cur);
        if (x == nx - 1) right = wall_e ? mirror(cur, true, dry) : cur;
        const Flux fe = hll<kFast>(cur, right, true, g, dry);
        Flux fw;
        fw.mass = __shfl_up(fe.mass, 1);
        fw.normal = __shfl_up(fe.normal, 1);
        fw.tangent = __shfl_up(fe.tangent, 1);
        if (x == 0) fw = hll<kFast>(wall_w ? mirror(cur, true, dry) : cur, cur, true, g, dry);

        if (owner) {
            // :205-225; x-faces carry (mass, normal, tangent) = (h, hu, hv), y-faces (h, hv, hu)
            const float raw_h = (cur.h - rx * (fe.mass - fw.mass)) - ry * (fn.mass - fs.mass);
            const float new_hu = (cur.hu - rx * (fe.normal - fw.normal)) - ry * (fn.tangent - fs.tangent);
            const float new_hv = (cur.hv - rx * (fe.tangent - fw.tangent)) - ry * (fn.normal - fs.normal);
            clamps += raw_h < 0.0f ? 1 : 0;
            const float new_h = raw_h < 0.0f ? 0.0f : raw_h;
            const size_t i = (size_t)y * nx + x;
            a.out[0][i] = new_h;
            a.out[1][i] = new_hu;
            a.out[2][i] = new_hv;
            // :230-235
            float u, w;
            if constexpr (kFast) {
                const float inv = recip(new_h);
                u = new_h > dry ? new_hu * inv : 0.0f;
                w = new_h > dry ? new_hv * inv : 0.0f;
            } else {
                u = velocity(new_h, new_hu, dry);
                w = velocity(new_h, new_hv, dry);
            }
            const float speed = mx(fabsf(u), fabsf(w)) + root<kFast>(g * new_h);
            const int32_t bits = __float_as_int(speed);
            speed_max = bits > speed_max ? bits : speed_max;
            if (x == 0 && y == 0 && a.budget_out) {
                // :237-240, invocation zero
                a.speed_clear[0] = 0;
                a.budget_out[0] = budget - ticks;
            }
        }
        fs = fn;
        cur = nrt;
        if (prefetch) nrt = make_cell<kFast>(ph, phu, phv, g, dry);
    }

    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int32_t o = __shfl_xor(speed_max, off);
        speed_max = o > speed_max ? o : speed_max;
        clamps += __shfl_xor(clamps, off);
    }
    if (lane == 0 && a.speed_out) {
        atomicMax(a.speed_out, speed_max);
        if (clamps) atomicAdd(a.clamped, clamps);
    }
}

__global__ void fill_zero_kernel(uint32_t* __restrict__ p, int64_t words) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0;
}

}  // namespace

int swe_step(const Launch& l, const SweArgs& a, int rows, bool fast) {
    const int bands = (a.ny + rows - 1) / rows;
    const int strips = (a.nx + kStrip - 1) / kStrip;
    const int64_t tiles = (int64_t)strips * ((bands + kWavesPerBlock - 1) / kWavesPerBlock);
    const int64_t per = (tiles + kXcds - 1) / kXcds;
    const dim3 grid((unsigned)(per * kXcds));
    timed_begin(l, "hll_step");
    if (fast)
        hipLaunchKernelGGL(hll_step_kernel<true>, grid, dim3(kWave * kWavesPerBlock), 0, l.stream, a, rows, strips,
                           tiles, per);
    else
        hipLaunchKernelGGL(hll_step_kernel<false>, grid, dim3(kWave * kWavesPerBlock), 0, l.stream, a, rows, strips,
                           tiles, per);
    timed_end(l, "hll_step");
    return check_launch("hll_step");
}

int fill_zero(const Launch& l, void* p, int64_t words) {
    if (words <= 0) return 0;
    int64_t blocks = (words + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, l.stream, (uint32_t*)p, words);
    return check_launch("fill_zero");
}

}  // namespace vxr
