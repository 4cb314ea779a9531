// Stable LSD radix sort of (key, index) pairs and the permutation of particle fields by it.
//
// Replaces the sort by cell of core/particle/FlipStep.java:183-192 (Scatter.sortCount, then
// CountingSort's scan and permute, then Flip.copy). The reference's counting sort takes each
// particle's rank by an atomic, so its order inside a cell is whatever the atomics made; this
// sort is stable, so its order is the host's stable sort by the same key, bit for bit.
//
// A pass over one digit is three kernels: a per-block histogram in LDS, a scan of the
// histograms digit-major, and a scatter in which each block re-reads its elements in order and
// ranks them by digit with wave ballots (lanes whose digits agree on every bit), so equal
// digits keep their input order within a wave, across the block's waves and across blocks.
#include "common.h"

namespace vxr {
namespace {

constexpr int kBlock = 256;
constexpr int kPerBlock = 4096;  // elements a block histograms and scatters per pass
constexpr int kMaxDigits = 256;

__global__ __launch_bounds__(kBlock) void hist_kernel(const int32_t* __restrict__ keys, int64_t n, int shift,
                                                      int mask, int32_t* __restrict__ hist, int blocks) {
    __shared__ int h[kMaxDigits];
    for (int d = threadIdx.x; d <= mask; d += kBlock) h[d] = 0;
    __syncthreads();
    const int64_t begin = (int64_t)blockIdx.x * kPerBlock;
    const int64_t end = min(begin + kPerBlock, n);
    for (int64_t i = begin + threadIdx.x; i < end; i += kBlock) atomicAdd(&h[(keys[i] >> shift) & mask], 1);
    __syncthreads();
    for (int d = threadIdx.x; d <= mask; d += kBlock) hist[(int64_t)d * blocks + blockIdx.x] = h[d];
}

// Block d scans row d of the histogram (exclusive, in place) and writes the row's total.
__global__ __launch_bounds__(kBlock) void scan_rows_kernel(int32_t* __restrict__ hist, int blocks,
                                                           int32_t* __restrict__ totals) {
    __shared__ int part[kBlock];
    int32_t* row = hist + (int64_t)blockIdx.x * blocks;
    const int per = (blocks + kBlock - 1) / kBlock;
    const int lo = threadIdx.x * per, hi = min(lo + per, blocks);
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += row[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - sum;  // exclusive prefix of this thread's slice
    for (int i = lo; i < hi; ++i) {
        const int v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == kBlock - 1) totals[blockIdx.x] = part[kBlock - 1];
}

// Each wave takes a contiguous quarter of the block's elements, kPerWave of them, and ranks them
// stably with ballots (the lanes whose digits agree on every bit) and a running count per digit
// in LDS that only this wave touches, so the waves need no barrier between their rounds. The
// block's elements then go through LDS in digit order -- a digit's elements in input order:
// earlier waves first, then by rank in the wave -- and leave in that order, each to its digit's
// place past all elements of smaller digits and the digit's elements in earlier blocks (the
// scanned histogram). Neighbouring threads write neighbouring addresses within a digit.
constexpr int kWaves = kBlock / kWave;
constexpr int kPerWave = kPerBlock / kWaves;
constexpr int kRounds = kPerWave / kWave;

__global__ __launch_bounds__(kBlock) void scatter_kernel(const int32_t* __restrict__ keys_in,
                                                         const int32_t* __restrict__ vals_in,
                                                         int32_t* __restrict__ keys_out,
                                                         int32_t* __restrict__ vals_out, int64_t n, int shift,
                                                         int mask, int digit_bits, const int32_t* __restrict__ hist,
                                                         const int32_t* __restrict__ totals, int blocks) {
    __shared__ int count[kWaves][kMaxDigits];  // a wave's running count per digit, then its local base
    __shared__ int2 scan[kBlock];              // (all blocks, this block) per digit, scanned
    __shared__ int out_of[kMaxDigits];         // global place of a digit's first element, less its local one
    __shared__ int32_t sorted_keys[kPerBlock], sorted_vals[kPerBlock];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid >> 6;
    const int digits = mask + 1;
    for (int d = tid; d < kWaves * kMaxDigits; d += kBlock) (&count[0][0])[d] = 0;
    __syncthreads();

    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t begin = (int64_t)blockIdx.x * kPerBlock;
    const int64_t first = begin + (int64_t)wave * kPerWave;
    int32_t key[kRounds], val[kRounds];
    int pos[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int64_t i = first + r * kWave + lane;
        key[r] = i < n ? keys_in[i] : 0;
        val[r] = i < n ? vals_in[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const bool valid = first + r * kWave + lane < n;
        const int d = (key[r] >> shift) & mask;
        uint64_t same = __ballot(valid);
        for (int b = 0; b < digit_bits; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t ones = __ballot(bit);
            same &= bit ? ones : ~ones;
        }
        const int rank = __popcll(same & below);
        // every lane of the digit reads the count before the first of them moves it on: one
        // wave's LDS operations happen in program order
        pos[r] = (valid ? count[wave][d] : 0) + rank;
        asm volatile("" ::: "memory");  // the reads above stay before the write below
        if (valid && rank == 0) count[wave][d] += __popcll(same);
        asm volatile("" ::: "memory");
    }
    __syncthreads();
    // per digit: the count over all blocks (totals) and over this block, both scanned
    int2 own = make_int2(0, 0);
    if (tid < digits) {
        own.x = totals[tid];
#pragma unroll
        for (int w = 0; w < kWaves; ++w) own.y += count[w][tid];
    }
    scan[tid] = own;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const int2 v = tid >= off ? scan[tid - off] : make_int2(0, 0);
        __syncthreads();
        scan[tid].x += v.x;
        scan[tid].y += v.y;
        __syncthreads();
    }
    if (tid < digits) {
        const int local = scan[tid].y - own.y;
        out_of[tid] = scan[tid].x - own.x + hist[(int64_t)tid * blocks + blockIdx.x] - local;
        int run = local;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const int c = count[w][tid];
            count[w][tid] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        if (first + r * kWave + lane >= n) continue;
        const int at = count[wave][(key[r] >> shift) & mask] + pos[r];
        sorted_keys[at] = key[r];
        sorted_vals[at] = val[r];
    }
    __syncthreads();
    const int here = (int)min((int64_t)kPerBlock, n - begin);
    for (int i = tid; i < here; i += kBlock) {
        const int32_t k = sorted_keys[i];
        const int at = out_of[(k >> shift) & mask] + i;
        keys_out[at] = k;
        vals_out[at] = sorted_vals[i];
    }
}

template <int kFields>
struct FieldPtrs {
    const float* in[kFields];
    float* out[kFields];
};

template <int kFields>
__global__ __launch_bounds__(kBlock) void permute_kernel(FieldPtrs<kFields> f, const int32_t* __restrict__ index,
                                                         int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t from = index[i];
#pragma unroll
    for (int k = 0; k < kFields; ++k) f.out[k][i] = f.in[k][from];
}

}  // namespace

int64_t radix_scratch_words(int64_t n, int digit_bits) {
    const int64_t blocks = (n + kPerBlock - 1) / kPerBlock;
    const int64_t digits = 1LL << digit_bits;
    return 2 * n + digits * (blocks > 0 ? blocks : 1) + digits;
}

// scratch: keys_tmp[n], vals_tmp[n], hist[digits * blocks], totals[digits]
int radix_sort(const Launch& l, int32_t* keys, int32_t* index, int64_t n, int bits, int digit_bits,
               int32_t* scratch) {
    if (n <= 1 || bits <= 0) return 0;
    const int blocks = (int)((n + kPerBlock - 1) / kPerBlock);
    int32_t* keys_tmp = scratch;
    int32_t* vals_tmp = scratch + n;
    int32_t* hist = scratch + 2 * n;
    int32_t* totals = hist + ((int64_t)1 << digit_bits) * blocks;
    int32_t *kin = keys, *vin = index, *kout = keys_tmp, *vout = vals_tmp;
    for (int shift = 0; shift < bits; shift += digit_bits) {
        const int width = bits - shift < digit_bits ? bits - shift : digit_bits;
        const int mask = (1 << width) - 1;
        timed_begin(l, "radix_hist");
        hipLaunchKernelGGL(hist_kernel, dim3(blocks), dim3(kBlock), 0, l.stream, kin, n, shift, mask, hist, blocks);
        timed_end(l, "radix_hist");
        timed_begin(l, "radix_scan");
        hipLaunchKernelGGL(scan_rows_kernel, dim3(mask + 1), dim3(kBlock), 0, l.stream, hist, blocks, totals);
        timed_end(l, "radix_scan");
        timed_begin(l, "radix_scatter");
        hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(kBlock), 0, l.stream, kin, vin, kout, vout, n, shift,
                           mask, width, hist, totals, blocks);
        timed_end(l, "radix_scatter");
        if (int rc = check_launch("radix_sort")) return rc;
        int32_t* t = kin;
        kin = kout;
        kout = t;
        t = vin;
        vin = vout;
        vout = t;
    }
    if (kin != keys) {
        if (hipMemcpyAsync(keys, kin, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, l.stream) != hipSuccess ||
            hipMemcpyAsync(index, vin, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, l.stream) != hipSuccess)
            return check_launch("radix_sort copy");
    }
    return 0;
}

int permute_fields(const Launch& l, float* const* in, float* const* out, int fields, const int32_t* index,
                   int64_t n) {
    if (n <= 0) return 0;
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    timed_begin(l, "permute");
    if (fields == 18) {
        FieldPtrs<18> f;
        for (int k = 0; k < 18; ++k) {
            f.in[k] = in[k];
            f.out[k] = out[k];
        }
        hipLaunchKernelGGL(permute_kernel<18>, dim3(blocks), dim3(kBlock), 0, l.stream, f, index, n);
    } else if (fields == 10) {
        FieldPtrs<10> f;
        for (int k = 0; k < 10; ++k) {
            f.in[k] = in[k];
            f.out[k] = out[k];
        }
        hipLaunchKernelGGL(permute_kernel<10>, dim3(blocks), dim3(kBlock), 0, l.stream, f, index, n);
    } else {
        timed_end(l, "permute");
        return 1;
    }
    timed_end(l, "permute");
    return check_launch("permute");
}

}  // namespace vxr
