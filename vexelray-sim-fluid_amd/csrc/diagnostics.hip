// The reference's health readings as device reductions, so a caller reads nine (or seven) numbers
// instead of the whole state: core/stencil/Diagnostics.java:48-81 over a shallow-water state and
// core/particle/ParticleDiagnostics.java:45-65 over the particles' J.
//
// Shallow water: each thread folds a strided share of the cells in the reference's per-cell
// arithmetic (double, the float quotients widened as Java widens them), then a block tree and a
// final single-block pass combine the partials in a fixed order -- the minima, maxima and counts
// exactly as the sequential loop has them, the volume to double rounding of another summation
// order. Particles: the valid J (finite, above zero) sorted by their bits (positive floats order
// as integers), the extremes and the 1st and 99th percentile picked from the sorted keys.
#include <mutex>

#include "common.h"

namespace vxr {
namespace {

constexpr int kBlock = 256;
constexpr int kParts = 1024;  // blocks of the first pass: fixed, so the summation order is too
constexpr double kShallow = 1e-3;

struct Part {
    double volume, min_depth, max_depth, max_speed, max_froude, max_wave;
    long long non_finite, negative, wet;
};

__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }

__device__ __forceinline__ Part empty_part() {
    Part p;
    p.volume = 0;
    p.min_depth = __longlong_as_double(0x7ff0000000000000LL);  // +inf
    p.max_depth = p.max_speed = p.max_froude = p.max_wave = 0;
    p.non_finite = p.negative = p.wet = 0;
    return p;
}

__device__ __forceinline__ void combine(Part& a, const Part& b) {
    a.volume += b.volume;
    a.min_depth = dmin(a.min_depth, b.min_depth);
    a.max_depth = dmax(a.max_depth, b.max_depth);
    a.max_speed = dmax(a.max_speed, b.max_speed);
    a.max_froude = dmax(a.max_froude, b.max_froude);
    a.max_wave = dmax(a.max_wave, b.max_wave);
    a.non_finite += b.non_finite;
    a.negative += b.negative;
    a.wet += b.wet;
}

// Tree over the block's threads through LDS, in a fixed order; thread 0 ends with the total.
__device__ __forceinline__ Part block_combine(Part mine, Part* lds) {
    lds[threadIdx.x] = mine;
    __syncthreads();
    for (int half = kBlock / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half) {
            Part a = lds[threadIdx.x];
            combine(a, lds[threadIdx.x + half]);
            lds[threadIdx.x] = a;
        }
        __syncthreads();
    }
    return lds[0];
}

__global__ __launch_bounds__(kBlock) void swe_diag_kernel(const float* __restrict__ h, const float* __restrict__ hu,
                                                          const float* __restrict__ hv, int64_t cells, double g,
                                                          double dry, Part* __restrict__ parts) {
    __shared__ Part lds[kBlock];
    Part p = empty_part();
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < cells; k += (int64_t)gridDim.x * kBlock) {
        const float hk = h[k], huk = hu[k], hvk = hv[k];
        if (!isfinite(hk) || !isfinite(huk) || !isfinite(hvk)) {
            p.non_finite++;
            continue;
        }
        if (hk < 0) p.negative++;
        p.volume += hk;
        p.max_depth = dmax(p.max_depth, (double)hk);
        if ((double)hk <= dry) continue;
        p.wet++;
        p.min_depth = dmin(p.min_depth, (double)hk);
        const double u = (double)(huk / hk), v = (double)(hvk / hk);
        const double c = sqrt(g * (double)hk);
        const double speed = sqrt(u * u + v * v);
        p.max_speed = dmax(p.max_speed, speed);
        if ((double)hk >= kShallow) p.max_froude = dmax(p.max_froude, speed / c);
        p.max_wave = dmax(p.max_wave, dmax(fabs(u), fabs(v)) + c);
    }
    const Part total = block_combine(p, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void swe_diag_final_kernel(const Part* __restrict__ parts, int n, double dt,
                                                                double dx, double dy, double* __restrict__ out) {
    __shared__ Part lds[kBlock];
    Part p = empty_part();
    for (int i = threadIdx.x; i < n; i += kBlock) combine(p, parts[i]);
    const Part t = block_combine(p, lds);
    if (threadIdx.x == 0) {
        out[0] = t.volume * dx * dy;
        out[1] = t.wet == 0 ? 0.0 : t.min_depth;
        out[2] = t.max_depth;
        out[3] = t.max_speed;
        out[4] = t.max_froude;
        out[5] = t.max_wave * dt / dmin(dx, dy);
        out[6] = (double)t.non_finite;
        out[7] = (double)t.negative;
        out[8] = (double)t.wet;
    }
}

// key = J's bits for a valid J (finite, above zero), else past every valid key; counts the invalid.
__global__ __launch_bounds__(kBlock) void jkeys_kernel(const float* __restrict__ j, int64_t count, int64_t n,
                                                       int32_t* __restrict__ keys, int32_t* __restrict__ index,
                                                       int* __restrict__ invalid) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= n) return;
    bool valid = false;
    float v = 0.0f;
    if (q < count) {
        v = j[q];
        valid = isfinite(v) && v > 0.0f;
        if (!valid) atomicAdd(invalid, 1);
    }
    keys[q] = valid ? __float_as_int(v) : 0x7fffffff;
    index[q] = (int32_t)q;
}

__global__ void pick_kernel(const int32_t* __restrict__ keys, const int* __restrict__ invalid, int64_t count,
                            int32_t* __restrict__ out) {
    const int64_t n = count - *invalid;
    if (threadIdx.x != 0) return;
    out[4] = (int32_t)n;
    if (n <= 0) return;
    auto percentile = [&](double p) { return min(n - 1, (int64_t)(p * (double)n)); };
    out[0] = keys[n - 1];
    out[1] = keys[0];
    out[2] = keys[percentile(0.99)];
    out[3] = keys[percentile(0.01)];
}

}  // namespace

int64_t diag_scratch_words() { return (int64_t)kParts * (int64_t)(sizeof(Part) / 4) + 32; }

int swe_diagnostics(const Launch& l, const float* h, const float* hu, const float* hv, int64_t cells, double g,
                    double dry, double dt, double dx, double dy, int32_t* scratch, double out[9]) {
    Part* parts = (Part*)scratch;
    double* dev_out = (double*)(scratch + (int64_t)kParts * (sizeof(Part) / 4));
    timed_begin(l, "swe_diagnostics");
    hipLaunchKernelGGL(swe_diag_kernel, dim3(kParts), dim3(kBlock), 0, l.stream, h, hu, hv, cells, g, dry, parts);
    hipLaunchKernelGGL(swe_diag_final_kernel, dim3(1), dim3(kBlock), 0, l.stream, parts, kParts, dt, dx, dy, dev_out);
    timed_end(l, "swe_diagnostics");
    if (int rc = check_launch("swe_diagnostics")) return rc;
    if (hipMemcpyAsync(out, dev_out, 9 * sizeof(double), hipMemcpyDeviceToHost, l.stream) != hipSuccess ||
        hipStreamSynchronize(l.stream) != hipSuccess)
        return check_launch("swe_diagnostics readback");
    return 0;
}

int particle_diagnostics(const Launch& l, const float* j, int64_t count, int64_t capacity, int32_t* keys,
                         int32_t* index, int32_t* sort_scratch, int digit_bits, int32_t* small, double speed,
                         double sound, double dt, double out[7]) {
    int* invalid = (int*)small;
    int32_t* picks = small + 8;
    if (int rc = fill_zero(l, invalid, 1)) return rc;  // a kernel, not a memset: safe to capture
    timed_begin(l, "particle_diagnostics");
    hipLaunchKernelGGL(jkeys_kernel, dim3((unsigned)((capacity + kBlock - 1) / kBlock)), dim3(kBlock), 0, l.stream, j,
                       count, capacity, keys, index, invalid);
    if (int rc = radix_sort(l, keys, index, capacity, 31, digit_bits, sort_scratch)) return rc;
    hipLaunchKernelGGL(pick_kernel, dim3(1), dim3(64), 0, l.stream, keys, invalid, count, picks);
    timed_end(l, "particle_diagnostics");
    if (int rc = check_launch("particle_diagnostics")) return rc;
    int32_t host[5] = {0, 0, 0, 0, 0};
    if (hipMemcpyAsync(host, picks, sizeof host, hipMemcpyDeviceToHost, l.stream) != hipSuccess ||
        hipStreamSynchronize(l.stream) != hipSuccess)
        return check_launch("particle_diagnostics readback");
    const int64_t n = host[4];
    out[4] = speed / sound;
    out[5] = (speed + sound) * dt;
    out[6] = (double)(count - n);
    if (n <= 0) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return 0;
    }
    for (int i = 0; i < 4; ++i) {
        float v;
        memcpy(&v, &host[i], 4);
        out[i] = (double)(1.0f / v);  // the reference divides on a float[] and widens the quotient
    }
    return 0;
}

// ------------------------------------------------------------------ without allocation

// ParticleDiagnostics.of for a caller that cannot allocate (code under graph capture, serving
// loops, harnesses that forbid hipMalloc): the J keys, their index and the radix sort's scratch
// live in `scratch`, a file-scope device array of fixed size that each GPU holds once, for the
// life of the process. At most kStaticParticles particles a call. Every call on one GPU shares
// that GPU's array, so calls that overlap on one GPU (two streams, two host threads) race.
constexpr int64_t kStaticParticles = 1 << 20;
constexpr int kStaticDigits = 8;
constexpr int64_t kStaticSortWords =
    2 * kStaticParticles + (1 << kStaticDigits) * ((kStaticParticles + 4095) / 4096) + (1 << kStaticDigits);
constexpr int64_t kStaticSmallWords = 16;
constexpr int64_t kStaticWords = 2 * kStaticParticles + kStaticSortWords + kStaticSmallWords;  // 16.3 MB

static __device__ __align__(256) int32_t scratch[kStaticWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch[kMaxDevices] = {};  // each device's address of `scratch`, taken once
}  // namespace

int64_t particle_diagnostics_static_limit() { return kStaticParticles; }

// The current device's copy of `scratch`. In host code the array's name stands for no device
// address: hipGetSymbolAddress gives this device's, once, with the device current.
int static_scratch(int32_t** out, const char** why) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *why = "no current device, or a device number past 63";
        return 3;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *why = "hipGetSymbolAddress of the diagnostics scratch failed";
            return 3;
        }
        g_scratch[dev] = (int32_t*)p;
    }
    *out = g_scratch[dev];
    return 0;
}

int particle_diagnostics_static(const Launch& l, const float* j, int64_t count, double speed, double sound,
                                double dt, double out[7], const char** why) {
    if (radix_scratch_words(kStaticParticles, kStaticDigits) > kStaticSortWords) {
        *why = "the diagnostics scratch is smaller than the sort needs";
        return 3;
    }
    int32_t* base = nullptr;
    if (int rc = static_scratch(&base, why)) return rc;
    int32_t* keys = base;
    int32_t* index = keys + kStaticParticles;
    int32_t* sort = index + kStaticParticles;
    int32_t* small = sort + kStaticSortWords;
    return particle_diagnostics(l, j, count, count, keys, index, sort, kStaticDigits, small, speed, sound, dt, out);
}

}  // namespace vxr
