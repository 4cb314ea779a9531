// The convolutional surrogate of the shallow-water step, and the FP64 HLL step that judges it.
//
// No reference counterpart: the reference has only the HLL solver (core/stencil/ShallowWater.java).
// The surrogate maps a state (h, hu, hv) to the state a fixed time later with a residual stack of
// 3x3 convolutions: 16 -> 64, four 64 -> 64, 64 -> 16 channels (the 3 state channels padded to
// 16, the padding held at zero), ReLU between, the last layer's output added to the state.
//
// Activations are NHWC bf16; weights bf16 [cout][tap][cin] with an FP32 master copy; every
// product accumulates in FP32 on the matrix cores (v_mfma_f32_16x16x32_bf16). A 3x3 convolution
// is an implicit GEMM over K = 9 taps x input channels, each lane loading 8 consecutive channels
// of one tap (16 bytes); the staged kernel takes the weights as the row operand (output channels)
// and the pixels as the column operand, the unstaged one the other way round. Zero padding at the image edges, so the
// input gradient is the same kernel run with the taps flipped and the channel roles swapped. The
// weight gradient is a GEMM over pixels, staged through LDS transposed.
#include <climits>

#include "../../include/vexel/vxs.h"
#include "common.h"

namespace vxr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kTileX = 64;  // output pixels of a block: 4 waves of 16

enum { kRelu = 1, kMask = 2, kOutF32 = 4, kAccumulate = 8 };

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
    return z;
}

// out[n][y][x][co] (co < out_ch) = act(bias[co] + sum_{tap, ci} w[co][tap][ci] in[n][y+dy][x+dx][ci]),
// a block a run of 64 pixels of one row, a wave 16 of them against all 16 * kCoTiles channels.
// kMask multiplies by (mask[n][y][x][co] > 0), kRelu clamps at zero, kOutF32 stores FP32 and
// kAccumulate adds the result into the FP32 output.
template <int kCin, int kCoTiles>
__global__ __launch_bounds__(kBlock) void conv3x3_kernel(const __bf16* __restrict__ in, int H, int W,
                                                         const __bf16* __restrict__ w,
                                                         const float* __restrict__ bias, void* out, int out_ch,
                                                         const __bf16* __restrict__ mask, int flags) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (W + kTileX - 1) / kTileX;
    const int y = blockIdx.x / tiles_x;
    const int x0 = (blockIdx.x % tiles_x) * kTileX + wave * 16;
    const int64_t img = (int64_t)blockIdx.y * H * W;
    const int i = lane & 15, kg = lane >> 4;
    const int px = x0 + i;
    constexpr int kK = 9 * kCin;
    constexpr int kSteps = (kK + 31) / 32;
    f32x4 acc[kCoTiles];
#pragma unroll
    for (int t = 0; t < kCoTiles; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int s = 0; s < kSteps; ++s) {
        const int k = 32 * s + 8 * kg;
        const int tap = k / kCin, ci = k % kCin;
        bf16x8 a = zero8();
        if (tap < 9) {
            const int yy = y + tap / 3 - 1, xx = px + tap % 3 - 1;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                a = *(const bf16x8*)(in + (img + (int64_t)yy * W + xx) * kCin + ci);
        }
#pragma unroll
        for (int t = 0; t < kCoTiles; ++t) {
            const bf16x8 b = tap < 9 ? *(const bf16x8*)(w + ((int64_t)(16 * t + i) * 9 + tap) * kCin + ci) : zero8();
            acc[t] = mfma(a, b, acc[t]);
        }
    }
    // lane holds C[4 kg + r][i]: pixel x0 + 4 kg + r, channel 16 t + i
#pragma unroll
    for (int t = 0; t < kCoTiles; ++t) {
        const int co = 16 * t + i;
        if (co >= out_ch) continue;
        const float b = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int p = x0 + 4 * kg + r;
            if (p >= W) continue;
            const int64_t at = (img + (int64_t)y * W + p) * out_ch + co;
            float v = acc[t][r] + b;
            if ((flags & kMask) && !((float)mask[at] > 0.0f)) v = 0.0f;
            if (flags & kRelu) v = v > 0.0f ? v : 0.0f;
            if (flags & kOutF32) {
                float* o = (float*)out;
                o[at] = (flags & kAccumulate) ? o[at] + v : v;
            } else {
                ((__bf16*)out)[at] = (__bf16)v;
            }
        }
    }
}

constexpr int kRows = 2;
constexpr int kCols = 64;

// 4 KB of zeros in global memory: the source of the halo pixels outside the image. A staging load
// takes its lane's own 16 bytes of it, so that the zeros of one load do not all queue on one
// address (one L2 channel).
constexpr int kZeroChunks = 256;
__device__ __align__(16) const unsigned kZeros[4 * kZeroChunks] = {};

// Where the lanes of a store with nothing to store (a pixel past the image edge) write, so that
// every store instruction of the plain epilogue is issued.
__device__ __align__(16) __bf16 kSink[64 * 4];

// s_waitcnt with vmcnt = n and the other counters not waited on (gfx9 encoding: vmcnt in bits
// 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int n>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(n >= 0 && n < 64, "vmcnt holds 6 bits");
    __builtin_amdgcn_s_waitcnt((n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8));
}

// s_waitcnt with lgkmcnt = n and the other counters not waited on.
template <int n>
__device__ __forceinline__ void wait_lgkmcnt() {
    static_assert(n >= 0 && n < 16, "lgkmcnt holds 4 bits");
    __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4) | (n << 8));
}

// The compiler does not see the LDS stores of a global-to-LDS load: LDS reads after the wait
// must not be moved above it or served from values it assumed.
__device__ __forceinline__ void lds_fence() { asm volatile("" ::: "memory"); }

// Stages tile `tt` of an NHWC image into `buf` with global-to-LDS loads (global_load_lds_dwordx4):
// the tile is kTileH x kTileW pixels -- the kRows x kCols pixels of the tile and a ring of kHalo
// pixels about them -- of kCin / 8 16-byte chunks and kPad pad chunks, chunk e of the tile at
// buf + 16 e; each load instruction fills 64 consecutive chunks, lane l the l-th, and wave w
// takes instructions w, w + 4, ... The lanes of pad chunks load nothing (no reader looks at a
// pad: pads only space pixels over the LDS banks), and pixels outside the image are zeros. Nothing passes through registers, so every load
// of the tile is in flight at once; the caller waits (vmcnt) and synchronises before reading. A lane's chunk
// (pixel row ty, column tx, slot cg) moves by 256 chunks from one of its instructions to the
// next, kept up to date by additions rather than divided out each time.
template <int kCin, int kHalo = 1, int kPad = 1>
__device__ __forceinline__ void stage_tile(const __bf16* __restrict__ in, int H, int W, int64_t tt, __bf16* buf) {
    constexpr int kSlots = kCin / 8 + kPad;
    constexpr int kTileW = kCols + 2 * kHalo, kTileH = kRows + 2 * kHalo;
    constexpr int kChunks = kTileH * kTileW * kSlots;
    constexpr int kLoads = (kChunks + 63) / 64;
    constexpr int kPerWave = (kLoads + 3) / 4;
    constexpr int kStepPix = 256 / kSlots, kStepSlot = 256 % kSlots;  // 256 chunks
    static_assert(kPerWave == 1 || kStepPix + 1 < 2 * kTileW, "two wraps of the column at most");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (W + kCols - 1) / kCols, tiles_y = (H + kRows - 1) / kRows;
    const int n = (int)(tt / ((int64_t)tiles_y * tiles_x));
    const int rem = (int)(tt % ((int64_t)tiles_y * tiles_x));
    const int y0 = (rem / tiles_x) * kRows, x0 = (rem % tiles_x) * kCols;
    const __bf16* base = in + ((int64_t)n * H * W + (int64_t)y0 * W + x0) * kCin;
    const bool inside = y0 >= kHalo && y0 + kRows + kHalo <= H && x0 >= kHalo && x0 + kCols + kHalo <= W;
    const int e0 = 64 * wave + lane;
    int pix = e0 / kSlots, cg = e0 % kSlots;
    int ty = pix / kTileW, tx = pix % kTileW;
#pragma unroll 1
    for (int jj = 0; jj < kPerWave; ++jj) {
        const int j = wave + 4 * jj;
        if (j < kLoads && 64 * j + lane < kChunks && cg < kCin / 8) {
            const void* src = kZeros + 4 * ((64 * j + lane) % kZeroChunks);
            const int yy = y0 + ty - kHalo, xx = x0 + tx - kHalo;
            if (inside || (yy >= 0 && yy < H && xx >= 0 && xx < W))
                src = base + ((ty - kHalo) * W + (tx - kHalo)) * kCin + 8 * cg;
            __builtin_amdgcn_global_load_lds(src, buf + 64 * 8 * j, 16, 0, 0);
        }
        // the next instruction of this wave: 256 chunks on
        cg += kStepSlot;
        const int carry = cg >= kSlots ? 1 : 0;
        cg -= carry * kSlots;
        tx += kStepPix + carry;
#pragma unroll
        for (int w = 0; w < 2; ++w)
            if (tx >= kTileW) {
                tx -= kTileW;
                ++ty;
            }
    }
}

// The staged implicit-GEMM 3x3 convolution. Resident blocks walk 2 x 64-pixel tiles; each tile's
// input with its one-pixel halo (4 x 66 pixels, every channel) is staged into LDS by
// global-to-LDS loads, double-buffered so the next tile's loads are in flight while this one
// computes, and serves all 9 taps and all waves, while each wave keeps the weights of its output
// channels in registers for the whole launch. Waves split a tile 2 x 2 (pixels x channels) for
// 64 output channels, 4 x 1 for 16. Pixel stride in LDS is cin + 8 channels, so the 16 lanes
// reading 16 consecutive pixels hit distinct banks.
template <int kCin, int kCoTiles>
__global__ __launch_bounds__(kBlock, 2) void conv3x3_lds_kernel(const __bf16* __restrict__ in, int images, int H,
                                                                int W, const __bf16* __restrict__ w,
                                                                const float* __restrict__ bias, void* out,
                                                                int out_ch, const __bf16* __restrict__ mask,
                                                                int flags) {
    constexpr int kStride = kCin + 8;
    constexpr int kTileW = kCols + 2, kTileH = kRows + 2;
    constexpr int kNT = kCoTiles >= 2 ? 2 : 1;         // channel tiles a wave holds
    constexpr int kWavesN = kCoTiles / kNT;
    constexpr int kWavesM = 4 / kWavesN;
    constexpr int kMT = (kRows * kCols / 16) / kWavesM;  // 16-pixel tiles a wave computes
    constexpr int kK = 9 * kCin;
    constexpr int kSteps = (kK + 31) / 32;
    constexpr int kTileElems = kTileH * kTileW * kStride;
    __shared__ __align__(16) __bf16 staged[2 * kTileElems];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wn = wave % kWavesN, wm = wave / kWavesN;
    const int i = lane & 15, kg = lane >> 4;

    const int tiles_x = (W + kCols - 1) / kCols, tiles_y = (H + kRows - 1) / kRows;
    const int64_t tiles = (int64_t)images * tiles_y * tiles_x;
    if ((int64_t)blockIdx.x < tiles) stage_tile<kCin>(in, H, W, blockIdx.x, staged);

    // this wave's weights: B[k][n] for its kNT channel tiles, every K step
    bf16x8 b[kNT][kSteps];
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
        const int co = 16 * (wn * kNT + t) + i;
#pragma unroll
        for (int s = 0; s < kSteps; ++s) {
            const int k = 32 * s + 8 * kg;
            const int tap = k / kCin, ci = k % kCin;
            b[t][s] = tap < 9 ? *(const bf16x8*)(w + ((int64_t)co * 9 + tap) * kCin + ci) : zero8();
        }
    }
    // The plain epilogue (bf16 out, no mask, no accumulation, every channel) issues exactly
    // kNT * kMT stores a tile and nothing else touches memory after a tile's staging loads, so
    // waiting until at most that many vector memory operations are outstanding waits for the
    // loads (they complete in issue order) without waiting for the previous tile's stores.
    const bool plain = !(flags & (kMask | kOutF32 | kAccumulate)) && out_ch == 16 * kCoTiles;
    // the biases of this lane's output channels, read once (a load in the tile loop would make
    // the wait for it wait for the next tile's staging loads too)
    float bias_r[kNT][4];
#pragma unroll
    for (int t = 0; t < kNT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int co = 16 * (wn * kNT + t) + 4 * kg + q;
            bias_r[t][q] = bias && co < out_ch ? bias[co] : 0.0f;
        }
    // the weights and biases are in registers before the tile loop: no wait inside it is for them
    // (a wait there for a load issued before the loop would also wait for the next tile's staging)
    __builtin_amdgcn_s_waitcnt(0);
    int cur = 0;
    for (int64_t tt = blockIdx.x; tt < tiles; tt += gridDim.x, cur ^= 1) {
        const int n = (int)(tt / ((int64_t)tiles_y * tiles_x));
        const int rem = (int)(tt % ((int64_t)tiles_y * tiles_x));
        const int y0 = (rem / tiles_x) * kRows, x0 = (rem % tiles_x) * kCols;
        const int64_t img = (int64_t)n * H * W;
        // this tile's loads have landed in every wave, and every wave is done with the other buffer.
        // A bare barrier: only LDS is shared here (__syncthreads' release fence would also wait
        // for the previous tile's global stores, which no wave reads).
        if (plain && tt != blockIdx.x) wait_vmcnt<kNT * kMT>();
        else __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS reads are done
        __builtin_amdgcn_s_barrier();
        lds_fence();
        if (tt + gridDim.x < tiles) stage_tile<kCin>(in, H, W, tt + gridDim.x, staged + (cur ^ 1) * kTileElems);
        const __bf16* tile = staged + cur * kTileElems;
        f32x4 acc[kMT][kNT];
#pragma unroll
        for (int m = 0; m < kMT; ++m)
#pragma unroll
            for (int t = 0; t < kNT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < kSteps; ++s) {
            const int k = 32 * s + 8 * kg;
            const int tap = k / kCin, ci = k % kCin;
            const int dy = tap / 3, dx = tap % 3;  // tile offsets of the tap (halo included)
#pragma unroll
            for (int m = 0; m < kMT; ++m) {
                const int mt = wm * kMT + m;
                const int r = mt / (kCols / 16), c = 16 * (mt % (kCols / 16)) + i;
                const bf16x8 a = tap < 9 ? *(const bf16x8*)&tile[((r + dy) * kTileW + c + dx) * kStride + ci] : zero8();
#pragma unroll
                for (int t = 0; t < kNT; ++t) acc[m][t] = mfma(b[t][s], a, acc[m][t]);
            }
        }
        // weights as the row operand: lane holds C[4 kg + q][i] of each (m, t) = channel
        // 16 (wn kNT + t) + 4 kg + q of pixel (row of mt, col 16 (mt % 4) + i), so its 4 values
        // are 4 consecutive channels of one pixel: one 8-byte (bf16) or 16-byte (FP32) access
        if (plain) {
#pragma unroll
            for (int t = 0; t < kNT; ++t) {
                const int co = 16 * (wn * kNT + t) + 4 * kg;
#pragma unroll
                for (int m = 0; m < kMT; ++m) {
                    const int mt = wm * kMT + m;
                    const int yy = y0 + mt / (kCols / 16), xx = x0 + 16 * (mt % (kCols / 16)) + i;
                    bf16x4 r;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float v = acc[m][t][q] + bias_r[t][q];
                        if (flags & kRelu) v = v > 0.0f ? v : 0.0f;
                        r[q] = (__bf16)v;
                    }
                    __bf16* o = yy < H && xx < W ? (__bf16*)out + (img + (int64_t)yy * W + xx) * out_ch + co
                                                 : kSink + 4 * lane;
                    *(bf16x4*)o = r;
                }
            }
            continue;
        }
#pragma unroll
        for (int t = 0; t < kNT; ++t) {
            const int co = 16 * (wn * kNT + t) + 4 * kg;
            if (co >= out_ch) continue;
            const float* bv = bias_r[t];
#pragma unroll
            for (int m = 0; m < kMT; ++m) {
                const int mt = wm * kMT + m;
                const int yy = y0 + mt / (kCols / 16), xx = x0 + 16 * (mt % (kCols / 16)) + i;
                if (yy >= H || xx >= W) continue;
                const int64_t at = (img + (int64_t)yy * W + xx) * out_ch + co;
                float v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] = acc[m][t][q] + bv[q];
                const bool whole = (out_ch & 3) == 0;  // 4 channels in range and the access aligned
                if (flags & kMask) {
                    if (whole) {
                        const bf16x4 mk = *(const bf16x4*)(mask + at);
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (!((float)mk[q] > 0.0f)) v[q] = 0.0f;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (co + q < out_ch && !((float)mask[at + q] > 0.0f)) v[q] = 0.0f;
                    }
                }
                if (flags & kRelu) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) v[q] = v[q] > 0.0f ? v[q] : 0.0f;
                }
                if (flags & kOutF32) {
                    float* o = (float*)out + at;
                    if (whole) {
                        float4 r = make_float4(v[0], v[1], v[2], v[3]);
                        if (flags & kAccumulate) {
                            const float4 p = *(const float4*)o;
                            r.x += p.x;
                            r.y += p.y;
                            r.z += p.z;
                            r.w += p.w;
                        }
                        *(float4*)o = r;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (co + q < out_ch) o[q] = (flags & kAccumulate) ? o[q] + v[q] : v[q];
                    }
                } else {
                    __bf16* o = (__bf16*)out + at;
                    if (whole) {
                        bf16x4 r;
#pragma unroll
                        for (int q = 0; q < 4; ++q) r[q] = (__bf16)v[q];
                        *(bf16x4*)o = r;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (co + q < out_ch) o[q] = (__bf16)v[q];
                    }
                }
            }
        }
    }
}

// How the weight gradient splits its work. One side of the product (the output gradient or the
// input) is read once a K step of 32 pixels, all its channels (kTF 16-channel tiles); the other,
// the "shifted" side, once a tap, 16 channels a wave: the side with fewer channels, so that the
// nine-fold reads fall on the narrower one. A block walks a strip of kCols columns down a segment
// of rows, kRows output rows a step; its waves split the shifted side's channel tiles (kGs) and a
// step's K steps (kGk). kDepth steps' rows are in flight ahead of the one computing.
template <int kCin, int kCout>
struct WgradShape {
    static constexpr bool kShiftX = kCin <= kCout;      // the input is the shifted side
    static constexpr int kCF = kShiftX ? kCout : kCin;  // channels of the side read once a K step
    static constexpr int kCS = kShiftX ? kCin : kCout;  // and of the shifted side
    static constexpr int kTF = kCF / 16;
    static constexpr int kGs = kCS / 16;
    static constexpr bool kWide = kGs == 4;            // 64 channels each side
    static constexpr int kGk = kWide ? 2 : 4;
    static constexpr int kWaves = kGs * kGk;           // 8 (two a SIMD) or 4 (two blocks a CU)
    static constexpr int kRows = kWide ? 1 : 2;        // output rows a step: 2 K steps each
    static constexpr int kDepth = kWide ? 4 : 2;
    static constexpr int kBlocksPerCu = kWide ? 1 : 2;  // as the rings' LDS allows
    static constexpr int kWords = 9 * kCin * kCout + kCout;  // a partial: dw[co][tap][ci], then db[co]
};

// Pad chunks after a pixel's channels in the weight gradient's rows: pixel strides of 40 banks
// (64 channels) and 8 banks (16), so the 8 consecutive pixels of a transposing read's half fall on
// disjoint banks.
template <int kC>
constexpr int wgrad_pad() { return kC == 64 ? 2 : 0; }

// One row of an LDS ring: kCols pixels and kHalo more each side, kC / 8 16-byte chunks and kPad
// pad chunks a pixel, chunk e at 16 e bytes; load instruction j fills chunks 64 j .. 64 j + 63
// and is wave j % kWaves's. kMinLoads: the instructions a row takes from the wave that issues
// fewest.
template <int kC, int kHalo, int kPad, int kWaves = 4>
struct RowStage {
    static constexpr int kSlots = kC / 8 + kPad;
    static constexpr int kPix = kCols + 2 * kHalo;
    static constexpr int kChunks = kPix * kSlots;
    static constexpr int kLoads = (kChunks + 63) / 64;
    static constexpr int kElems = kChunks * 8;
    static constexpr int kMinLoads = kLoads / kWaves;
};

// Stages row y of the image at pixel offset img (columns x0 - kHalo on) into `buf` with
// global-to-LDS loads: pixels outside the image are zeros, pad chunks are not loaded.
template <int kC, int kHalo, int kPad, int kWaves = 4>
__device__ __forceinline__ void stage_row(const __bf16* __restrict__ in, int64_t img, int y, int x0, int H, int W,
                                          __bf16* buf) {
    typedef RowStage<kC, kHalo, kPad, kWaves> R;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool row_in = y >= 0 && y < H;
    const __bf16* base = in + (img + (int64_t)(row_in ? y : 0) * W) * kC;
#pragma unroll
    for (int jj = 0; jj < (R::kLoads + kWaves - 1) / kWaves; ++jj) {
        const int j = wave + kWaves * jj;
        const int e = 64 * j + lane;
        const int px = e / R::kSlots, cg = e % R::kSlots;
        if (j < R::kLoads && e < R::kChunks && cg < kC / 8) {
            const int xx = x0 + px - kHalo;
            const void* src = kZeros + 4 * (e % kZeroChunks);
            if (row_in && xx >= 0 && xx < W) src = base + (int64_t)xx * kC + 8 * cg;
            __builtin_amdgcn_global_load_lds(src, buf + 64 * 8 * j, 16, 0, 0);
        }
    }
}

// s_waitcnt vmcnt(min(n, 63)) for a wave-uniform n.
template <int N = 63>
__device__ __forceinline__ void wait_vmcnt_at_most(int n) {
    if constexpr (N == 0) {
        wait_vmcnt<0>();
    } else {
        if (n >= N) wait_vmcnt<N>();
        else wait_vmcnt_at_most<N - 1>(n);
    }
}

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

// An MFMA operand of 8 pixels of one channel from a pixel-major LDS row, by gfx950's transposing
// read (ds_read_b64_tr_b16): within each 16-lane group, lane i receives element i % 4 of the
// 4-element rows the lanes 4 m + i / 4 address, m = 0..3. With lane (i, kg) addressing pixel
// 4 kg + i / 4, channels c + 4 (i % 4) .. + 3 (`p`), lane i receives channel c + i of pixels
// 4 kg .. 4 kg + 3; a second read 16 pixels on gives 4 kg + 16 .. + 3. So k slot 4 h + m of lane
// group kg is pixel 4 kg + 16 h + m, for every operand read this way.
__device__ __forceinline__ bf16x8 read_pixels(const __bf16* p, int stride) {
    const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((s16x4 __attribute__((address_space(3)))*)p);
    const s16x4 hi =
        __builtin_amdgcn_ds_read_tr16_b64_v4i16((s16x4 __attribute__((address_space(3)))*)(p + 16 * stride));
    const s16x8 v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_bit_cast(bf16x8, v);
}

// The weight and bias gradients of a 3x3 convolution as a GEMM over pixels on the matrix cores:
//   part[p][co][tap][ci] = sum over the block's pixels of dy[pixel][co] x[pixel + off(tap)][ci]
//   part[p][9 cout cin + co] = sum over the block's pixels of dy[pixel][co]
// with p = block kGk + (the wave's K group). The pixels are segments of seg_rows rows of a
// kCols-column strip; a block walks its segments a step (kRows rows) at a time, with the rows of
// the next kDepth steps loading into two LDS rings by global-to-LDS loads: the shifted side's rows
// with one more pixel each end and the rows above and below (zeros outside the image), the other
// side's rows as they are. Every staged row is loaded once. The product's K index is the pixel:
// operands come out of the pixel-major rows by transposing reads. When the output gradient is the
// shifted side the sum runs as sum over q of dy[q - off(tap)][co] x[q][ci] (the same sum: x is
// zero outside the image, dy's ring zero). The bias gradient is one more product, of the output
// gradient against ones.
template <int kCin, int kCout>
__global__ __launch_bounds__((64 * WgradShape<kCin, kCout>::kWaves),
                             (WgradShape<kCin, kCout>::kBlocksPerCu)) void conv3x3_wgrad_kernel(
    const __bf16* __restrict__ x, const __bf16* __restrict__ dy, int images, int H, int W, int seg_rows,
    float* __restrict__ part) {
    typedef WgradShape<kCin, kCout> S;
    constexpr int kTF = S::kTF, P = S::kRows, D = S::kDepth;
    constexpr int kPF = wgrad_pad<S::kCF>(), kPS = wgrad_pad<S::kCS>();
    constexpr int kW = S::kWaves;
    typedef RowStage<S::kCS, 1, kPS, kW> RS;
    typedef RowStage<S::kCF, 0, kPF, kW> RF;
    constexpr int kSS = RS::kSlots * 8, kSF = RF::kSlots * 8;  // LDS pixel strides
    // ring rows: a step reads P + 2 shifted rows and P others, and D steps' rows are in flight
    constexpr int kNs = P * (D + 1) + 2, kNf = P * (D + 1);
    // load instructions a step's rows take from the wave issuing fewest: waiting until this many
    // times the steps issued since are outstanding waits for a step's rows in every wave
    constexpr int kUnitLoads = P * (RS::kMinLoads + RF::kMinLoads);
    __shared__ __align__(16) __bf16 ring_s[kNs * RS::kElems];
    __shared__ __align__(16) __bf16 ring_f[kNf * RF::kElems];
    const __bf16* src_s = S::kShiftX ? x : dy;
    const __bf16* src_f = S::kShiftX ? dy : x;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kg = lane >> 4;
    const int gs = wave % S::kGs, gk = wave / S::kGs;
    const int lp = 4 * kg + (i >> 2), lc = 4 * (i & 3);  // this lane's pixel and channels in a read

    f32x4 acc[9][kTF];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int t = 0; t < kTF; ++t) acc[tap][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int kTB = S::kShiftX ? kTF : 1;  // output-channel tiles of the bias
    f32x4 accb[kTB];
#pragma unroll
    for (int t = 0; t < kTB; ++t) accb[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.0f;

    const int strips = (W + kCols - 1) / kCols, segs_y = (H + seg_rows - 1) / seg_rows;
    const int64_t segs = (int64_t)images * strips * segs_y;
    for (int64_t sg = blockIdx.x; sg < segs; sg += gridDim.x) {
        const int n = (int)(sg / ((int64_t)strips * segs_y));
        const int rem = (int)(sg % ((int64_t)strips * segs_y));
        const int ya = (rem / strips) * seg_rows, x0 = (rem % strips) * kCols;
        const int ye = ya + seg_rows < H ? ya + seg_rows : H;
        const int steps = (ye - ya + P - 1) / P;
        const int64_t img = (int64_t)n * H * W;
        // every wave is done with the rings
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_s_barrier();
        lds_fence();
        // shifted row ya - 1 + j in slot j % kNs; row ya + j of the other side in slot j % kNf.
        // Step k's rows: shifted ya + P k + 1 .. + P, the other side ya + P k .. + P - 1.
        auto issue = [&](int k) {
#pragma unroll
            for (int r = 0; r < P; ++r) {
                const int js = P * k + r + 2, jf = P * k + r;
                stage_row<S::kCS, 1, kPS, kW>(src_s, img, ya - 1 + js, x0, H, W, ring_s + (js % kNs) * RS::kElems);
                stage_row<S::kCF, 0, kPF, kW>(src_f, img, ya + jf, x0, H, W, ring_f + (jf % kNf) * RF::kElems);
            }
        };
        stage_row<S::kCS, 1, kPS, kW>(src_s, img, ya - 1, x0, H, W, ring_s);
        stage_row<S::kCS, 1, kPS, kW>(src_s, img, ya, x0, H, W, ring_s + RS::kElems);
        for (int k = 0; k < D && k < steps; ++k) issue(k);
        for (int k = 0; k < steps; ++k) {
            // this step's rows have landed in every wave, and every wave is done with the slots
            // the step D on takes (the previous step's)
            const int after = steps - 1 - k < D - 1 ? steps - 1 - k : D - 1;
            wait_vmcnt_at_most(after * kUnitLoads);
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
            __builtin_amdgcn_s_barrier();
            lds_fence();
            if (k + D < steps) issue(k + D);
#pragma unroll
            for (int m = 0; m < 2 * P / S::kGk; ++m) {
                const int ks = gk + S::kGk * m;  // K step: row r of the step, columns c0 .. c0 + 31
                const int r = ks >> 1, c0 = 32 * (ks & 1);
                const int j = P * k + r;  // the output row ya + j
                if (ya + j >= ye) continue;
                const __bf16* fr = ring_f + (j % kNf) * RF::kElems;
                bf16x8 fo[kTF];
#pragma unroll
                for (int t = 0; t < kTF; ++t) fo[t] = read_pixels(fr + (c0 + lp) * kSF + 16 * t + lc, kSF);
                if (S::kShiftX && gs == 0) {
#pragma unroll
                    for (int t = 0; t < kTB; ++t) accb[t] = mfma(fo[t], ones, accb[t]);
                }
#pragma unroll
                for (int tr = 0; tr < 3; ++tr) {
                    // the shifted side's row for tap row tr (its ring index: row + 1 - ya)
                    const int js = S::kShiftX ? j + tr : j + 2 - tr;
                    const __bf16* sr = ring_s + (js % kNs) * RS::kElems;
#pragma unroll
                    for (int dxi = 0; dxi < 3; ++dxi) {
                        const int sc = S::kShiftX ? c0 + dxi : c0 + 2 - dxi;  // ring column: image x0 - 1 + sc
                        const bf16x8 so = read_pixels(sr + (sc + lp) * kSS + 16 * gs + lc, kSS);
                        const int tap = 3 * tr + dxi;
#pragma unroll
                        for (int t = 0; t < kTF; ++t)
                            acc[tap][t] = S::kShiftX ? mfma(fo[t], so, acc[tap][t]) : mfma(so, fo[t], acc[tap][t]);
                        if (!S::kShiftX && tap == 4) accb[0] = mfma(so, ones, accb[0]);
                    }
                }
            }
        }
    }
    // lane holds rows 4 kg + q, column i of each tile: with the input shifted, rows are output
    // channels 16 t + 4 kg + q and the column input channel 16 gs + i; otherwise rows are output
    // channels 4 kg + q and the column input channel 16 t + i
    float* out = part + ((int64_t)blockIdx.x * S::kGk + gk) * S::kWords;
#pragma unroll
    for (int t = 0; t < kTF; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int co = S::kShiftX ? 16 * t + 4 * kg + q : 4 * kg + q;
            const int ci = S::kShiftX ? 16 * gs + i : 16 * t + i;
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) out[((int64_t)co * 9 + tap) * kCin + ci] = acc[tap][t][q];
        }
    if (gs == 0 && i == 0) {
#pragma unroll
        for (int t = 0; t < kTB; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) out[9 * kCin * kCout + 16 * t + 4 * kg + q] = accb[t][q];
    }
}

// The streamed 3x3 convolution: resident blocks walk strips of kCols columns down segments of
// seg_rows rows, kR output rows a step, with the input rows of the next kDepth steps loading into
// an LDS ring by global-to-LDS loads (each input row loaded once, with one more pixel each end).
// Each wave keeps the weights of kNT output-channel tiles in registers for the whole launch and
// computes kMT 16-pixel column tiles of the step's rows. With 64 input channels a K step (32
// channels of one tap) lies in one input row, so each pixel operand read from LDS serves every
// output row of the step that reads that input row (up to 3) and every channel tile; with 16 a K
// step spans two taps, and a step is one row. Either way each output accumulates over K = 9 taps x
// cin in the order of conv3x3_lds_kernel (the results are the same bits). With kMasked the mask
// rows stream through a second ring and the epilogue reads them from LDS. Every wave issues the
// same stores every step (pixels and rows past the image edge store to kSink), so the waits can
// count the loads still in flight behind them.
template <int kCin, int kCoTiles, bool kMasked>
struct ConvRows {
    static constexpr int kCo = 16 * kCoTiles;
    static constexpr bool kWide = kCin == 64 && kCoTiles == 4;
    static constexpr int kWaves = kWide ? 8 : 4;
    static constexpr int kNT = kWide ? 2 : kCoTiles;
    static constexpr int kGroupsN = kCoTiles / kNT;           // waves over the channel tiles
    static constexpr int kMT = 4 * kGroupsN / kWaves;         // column tiles a wave computes
    static constexpr int kR = kCin == 64 ? ((kWide && !kMasked) ? 4 : 2) : 1;
    static constexpr int kDepth = kCin == 64 ? 2 : (kMasked ? 5 : 8);
    // two waves a SIMD (so one's epilogue, waits and barriers overlap the other's products): eight
    // waves a block, or two blocks a CU where the rings fit twice in LDS
    static constexpr int kBlocksPerCu = (kWide || (kCin == 64 && kMasked)) ? 1 : 2;
};

template <int kCin, int kCoTiles, bool kMasked>
__global__ __launch_bounds__((64 * ConvRows<kCin, kCoTiles, kMasked>::kWaves),
                             (ConvRows<kCin, kCoTiles, kMasked>::kBlocksPerCu)) void conv3x3_rows_kernel(const __bf16* __restrict__ in, int images, int H,
                                                                 int W, int seg_rows, const __bf16* __restrict__ w,
                                                                 const float* __restrict__ bias, void* out,
                                                                 const __bf16* __restrict__ mask, int flags) {
    typedef ConvRows<kCin, kCoTiles, kMasked> C;
    constexpr int D = C::kDepth, R = C::kR, kNT = C::kNT, kMT = C::kMT;
    constexpr int kW = C::kWaves;
    typedef RowStage<kCin, 1, 1, kW> RX;  // input rows: the pad chunk spaces 16 pixels over the banks
    typedef RowStage<C::kCo, 0, 0, kW> RM;  // mask rows
    constexpr int kStride = RX::kSlots * 8;
    constexpr int kNx = R * (D + 1) + 2, kNm = R * (D + 1);
    constexpr int kK = 9 * kCin;
    constexpr int kSteps = (kK + 31) / 32;
    constexpr int kStepLoads = R * (RX::kMinLoads + (kMasked ? RM::kMinLoads : 0));  // fewest of any wave
    constexpr int kStepStores = R * kMT * kNT;
    __shared__ __align__(16) __bf16 ring_x[kNx * RX::kElems];
    __shared__ __align__(16) __bf16 ring_m[kMasked ? kNm * RM::kElems : 8];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kg = lane >> 4;
    const int wn = wave % C::kGroupsN, wm = wave / C::kGroupsN;
    const bool relu = flags & kRelu, f32 = flags & kOutF32;

    // this wave's weights: B[k][n] for its channel tiles, every K step
    bf16x8 b[kNT][kSteps];
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
        const int co = 16 * (wn * kNT + t) + i;
#pragma unroll
        for (int s = 0; s < kSteps; ++s) {
            const int k = 32 * s + 8 * kg;
            const int tap = k / kCin, ci = k % kCin;
            b[t][s] = tap < 9 ? *(const bf16x8*)(w + ((int64_t)co * 9 + tap) * kCin + ci) : zero8();
        }
    }
    float bias_r[kNT][4];
#pragma unroll
    for (int t = 0; t < kNT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) bias_r[t][q] = bias ? bias[16 * (wn * kNT + t) + 4 * kg + q] : 0.0f;

    const int strips = (W + kCols - 1) / kCols, segs_y = (H + seg_rows - 1) / seg_rows;
    const int64_t segs = (int64_t)images * strips * segs_y;
    for (int64_t sg = blockIdx.x; sg < segs; sg += gridDim.x) {
        const int n = (int)(sg / ((int64_t)strips * segs_y));
        const int rem = (int)(sg % ((int64_t)strips * segs_y));
        const int ya = (rem / strips) * seg_rows, x0 = (rem % strips) * kCols;
        const int ye = ya + seg_rows < H ? ya + seg_rows : H;
        const int steps = (ye - ya + R - 1) / R;
        const int64_t img = (int64_t)n * H * W;
        // every load and store of this wave is done (the weights included), and every wave is done
        // with the rings
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_s_barrier();
        lds_fence();
        // input row ya - 1 + j in slot j % kNx; mask row ya + j in slot j % kNm. Step k's rows:
        // input ya + R k + 1 .. + R, mask ya + R k .. + R - 1.
        auto issue = [&](int k) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int jx = R * k + r + 2, jm = R * k + r;
                stage_row<kCin, 1, 1, kW>(in, img, ya - 1 + jx, x0, H, W, ring_x + (jx % kNx) * RX::kElems);
                if (kMasked)
                    stage_row<C::kCo, 0, 0, kW>(mask, img, ya + jm, x0, H, W, ring_m + (jm % kNm) * RM::kElems);
            }
        };
        stage_row<kCin, 1, 1, kW>(in, img, ya - 1, x0, H, W, ring_x);
        stage_row<kCin, 1, 1, kW>(in, img, ya, x0, H, W, ring_x + RX::kElems);
        for (int k = 0; k < D && k < steps; ++k) issue(k);
        for (int k = 0; k < steps; ++k) {
            // this step's rows have landed in every wave: behind its loads this wave issued the
            // loads of the steps after it and the stores of the steps before
            const int loads_after = steps - 1 - k < D - 1 ? steps - 1 - k : D - 1;
            const int stores_after = k < D - 1 ? k : D - 1;
            wait_vmcnt_at_most(loads_after * kStepLoads + stores_after * kStepStores);
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
            __builtin_amdgcn_s_barrier();
            lds_fence();
            f32x4 acc[R][kMT][kNT];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int m = 0; m < kMT; ++m)
#pragma unroll
                    for (int t = 0; t < kNT; ++t) acc[r][m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
            if constexpr (kCin == 64) {
                // group g = (input row rho of the step, ring index R k + rho; tap column dx): its
                // operands serve output rows r = rho - dy. Each group's LDS reads are issued before
                // the previous group's products, so their latency hides behind them.
                constexpr int kGroups = (R + 2) * 3;
                bf16x8 ab[2][2][kMT];
                auto load_group = [&](int g, bf16x8(&dst)[2][kMT]) {
                    const int rho = g / 3, dx = g % 3;
                    const __bf16* row = ring_x + ((R * k + rho) % kNx) * RX::kElems;
#pragma unroll
                    for (int ch = 0; ch < 2; ++ch)
#pragma unroll
                        for (int m = 0; m < kMT; ++m)
                            dst[ch][m] =
                                *(const bf16x8*)&row[(16 * (wm * kMT + m) + i + dx) * kStride + 32 * ch + 8 * kg];
                };
                load_group(0, ab[0]);
#pragma unroll
                for (int g = 0; g < kGroups; ++g) {
                    // this group's reads (issued before the previous group's products) have landed
                    // before the next group's are issued, so no wait is left for the products
                    wait_lgkmcnt<0>();
                    __builtin_amdgcn_sched_barrier(0);
                    if (g + 1 < kGroups) load_group(g + 1, ab[(g + 1) & 1]);
                    __builtin_amdgcn_sched_barrier(0);
                    const int rho = g / 3, dx = g % 3;
#pragma unroll
                    for (int ch = 0; ch < 2; ++ch)
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const int dy = rho - r;
                            if (dy < 0 || dy > 2) continue;
                            const int s = (3 * dy + dx) * 2 + ch;
#pragma unroll
                            for (int m = 0; m < kMT; ++m)
#pragma unroll
                                for (int t = 0; t < kNT; ++t)
                                    acc[r][m][t] = mfma(b[t][s], ab[g & 1][ch][m], acc[r][m][t]);
                        }
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
#pragma unroll
                for (int s = 0; s < kSteps; ++s) {
                    const int k32 = 32 * s + 8 * kg;
                    const int tap = k32 / kCin, ci = k32 % kCin;
                    const int dy = tap / 3, dx = tap % 3;
#pragma unroll
                    for (int m = 0; m < kMT; ++m) {
                        const bf16x8 a = tap < 9 ? *(const bf16x8*)&ring_x[((R * k + dy) % kNx) * RX::kElems +
                                                                           (16 * (wm * kMT + m) + i + dx) * kStride + ci]
                                                 : zero8();
#pragma unroll
                        for (int t = 0; t < kNT; ++t) acc[0][m][t] = mfma(b[t][s], a, acc[0][m][t]);
                    }
                }
            }
            // lane holds channel 16 (wn kNT + t) + 4 kg + q of pixel (ya + R k + r, x0 + 16 (wm kMT + m) + i)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int yy = ya + R * k + r;
#pragma unroll
                for (int m = 0; m < kMT; ++m) {
                    const int xx = x0 + 16 * (wm * kMT + m) + i;
                    const bool inside = yy < ye && xx < W;
                    const int64_t at = (img + (int64_t)yy * W + xx) * C::kCo;
#pragma unroll
                    for (int t = 0; t < kNT; ++t) {
                        const int co = 16 * (wn * kNT + t) + 4 * kg;
                        float v[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) v[q] = acc[r][m][t][q] + bias_r[t][q];
                        if (kMasked) {
                            const bf16x4 mk = *(const bf16x4*)&ring_m[((R * k + r) % kNm) * RM::kElems +
                                                                       (16 * (wm * kMT + m) + i) * C::kCo + co];
#pragma unroll
                            for (int q = 0; q < 4; ++q)
                                if (!((float)mk[q] > 0.0f)) v[q] = 0.0f;
                        }
                        if (relu) {
#pragma unroll
                            for (int q = 0; q < 4; ++q) v[q] = v[q] > 0.0f ? v[q] : 0.0f;
                        }
                        if (f32) {
                            float* o = inside ? (float*)out + at + co : (float*)kSink + 4 * (lane & 15);
                            *(float4*)o = make_float4(v[0], v[1], v[2], v[3]);
                        } else {
                            bf16x4 rr;
#pragma unroll
                            for (int q = 0; q < 4; ++q) rr[q] = (__bf16)v[q];
                            __bf16* o = inside ? (__bf16*)out + at + co : kSink + 4 * lane;
                            *(bf16x4*)o = rr;
                        }
                    }
                }
            }
            // the rows D steps on, into the slots the previous step read (every wave has passed
            // this step's barrier, so is done with them)
            if (k + D < steps) issue(k + D);
        }
    }
}

// out = sum over the parts of part[p][e] for each of the n words of a partial (the first nw are
// dw, the rest db): four interleaved groups of parts, each summed in order, then the groups in
// order, so the result does not depend on the run.
__global__ __launch_bounds__(kBlock) void wgrad_reduce_kernel(const float* __restrict__ part, int parts, int nw,
                                                              int n, float* __restrict__ dw,
                                                              float* __restrict__ db) {
    __shared__ float s[4][64];
    const int t = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t e = (int64_t)blockIdx.x * 64 + t;
    float sum = 0.0f;
    if (e < n) {
#pragma unroll 8
        for (int p = g; p < parts; p += 4) sum += part[(int64_t)p * n + e];
    }
    s[g][t] = sum;
    __syncthreads();
    if (g == 0 && e < n) {
        const float v = ((s[0][t] + s[1][t]) + s[2][t]) + s[3][t];
        if (e < nw) dw[e] = v;
        else db[e - nw] = v;
    }
}

// wt[ci][8 - tap][co] = w[co][tap][ci]: the kernel that carries output gradients back to inputs.
__global__ void flip_kernel(const __bf16* __restrict__ w, int cout, int cin, __bf16* __restrict__ wt) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= (int64_t)9 * cout * cin) return;
    const int co = (int)(e / (9 * cin)), tap = (int)(e / cin % 9), ci = (int)(e % cin);
    wt[((int64_t)ci * 9 + (8 - tap)) * cout + co] = w[e];
}

// Adam on the FP32 master weights, writing the bf16 copy the convolutions read.
__global__ void adam_kernel(float* __restrict__ p, __bf16* __restrict__ p16, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n) return;
    const float ge = g[e];
    const float me = b1 * m[e] + (1.0f - b1) * ge;
    const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    const float pe = p[e] - lr * (me / c1) / (sqrtf(ve / c2) + eps);
    p[e] = pe;
    if (p16) p16[e] = (__bf16)pe;
}

// The state (planar FP32 h, hu, hv, one image after another) as the network's input: NHWC bf16,
// 16 channels, the first three (q - offset[c]) * scale[c], the rest zero.
__global__ void pack_kernel(const float* __restrict__ h, const float* __restrict__ hu, const float* __restrict__ hv,
                            int64_t pixels, const float* __restrict__ norm, __bf16* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= pixels) return;
    bf16x8 a = zero8(), b = zero8();
    a[0] = (__bf16)((h[p] - norm[0]) * norm[3]);
    a[1] = (__bf16)((hu[p] - norm[1]) * norm[4]);
    a[2] = (__bf16)((hv[p] - norm[2]) * norm[5]);
    *(bf16x8*)(out + p * 16) = a;
    *(bf16x8*)(out + p * 16 + 8) = b;
}

// next[c] = state[c] + delta[p][c] / scale_out[c] for the three state channels (delta FP32, 16
// channels a pixel).
__global__ void unpack_kernel(const float* __restrict__ h, const float* __restrict__ hu,
                              const float* __restrict__ hv, const float* __restrict__ delta, int64_t pixels,
                              const float* __restrict__ norm, float* __restrict__ oh, float* __restrict__ ohu,
                              float* __restrict__ ohv) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= pixels) return;
    oh[p] = h[p] + delta[p * 16] / norm[6];
    ohu[p] = hu[p] + delta[p * 16 + 1] / norm[7];
    ohv[p] = hv[p] + delta[p * 16 + 2] / norm[8];
}

// The loss 1/(3P) sum_c sum_p (delta[p][c] - target_c)^2 with target_c = (next_c - state_c) *
// scale_out[c], in the output's scaled units; writes its gradient (bf16 NHWC, 16 channels, zero
// padding channels) and adds each block's partial loss to loss[0].
__global__ void loss_kernel(const float* __restrict__ delta, const float* __restrict__ h,
                            const float* __restrict__ hu, const float* __restrict__ hv,
                            const float* __restrict__ th, const float* __restrict__ thu,
                            const float* __restrict__ thv, int64_t pixels, const float* __restrict__ norm,
                            __bf16* __restrict__ grad, float* __restrict__ loss) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    float l = 0.0f;
    if (p < pixels) {
        const float r0 = delta[p * 16] - (th[p] - h[p]) * norm[6];
        const float r1 = delta[p * 16 + 1] - (thu[p] - hu[p]) * norm[7];
        const float r2 = delta[p * 16 + 2] - (thv[p] - hv[p]) * norm[8];
        l = r0 * r0 + r1 * r1 + r2 * r2;
        const float k = 2.0f / (3.0f * (float)pixels);
        bf16x8 a = zero8(), b = zero8();
        a[0] = (__bf16)(k * r0);
        a[1] = (__bf16)(k * r1);
        a[2] = (__bf16)(k * r2);
        *(bf16x8*)(grad + p * 16) = a;
        *(bf16x8*)(grad + p * 16 + 8) = b;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l / (3.0f * (float)pixels));
}

// ------------------------------------------------------------------ FP64 HLL

__device__ __forceinline__ double dmn(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmx(double a, double b) { return a < b ? b : a; }

struct S64 {
    double h, hu, hv;
};

__device__ __forceinline__ void flux64(const S64& l, const S64& r, bool along_x, double g, double dry,
                                       double f[3]) {
    const double qnL = along_x ? l.hu : l.hv, qtL = along_x ? l.hv : l.hu;
    const double qnR = along_x ? r.hu : r.hv, qtR = along_x ? r.hv : r.hu;
    const double unL = l.h > dry ? qnL / l.h : 0.0, utL = l.h > dry ? qtL / l.h : 0.0;
    const double unR = r.h > dry ? qnR / r.h : 0.0, utR = r.h > dry ? qtR / r.h : 0.0;
    const double cL = sqrt(g * l.h), cR = sqrt(g * r.h);
    double sL = dmn(unL - cL, unR - cR), sR = dmx(unL + cL, unR + cR);
    if (!(r.h > dry)) {
        sL = unL - cL;
        sR = unL + 2 * cL;
    }
    if (!(l.h > dry)) {
        sL = unR - 2 * cR;
        sR = unR + cR;
    }
    const double lo = dmn(sL, 0.0), hi = dmx(sR, 0.0), span = dmx(hi - lo, 1e-300), lohi = lo * hi;
    const double pL = 0.5 * g * l.h * l.h, pR = 0.5 * g * r.h * r.h;
    f[0] = ((hi * qnL - lo * qnR) + lohi * (r.h - l.h)) / span;
    f[1] = ((hi * (qnL * unL + pL) - lo * (qnR * unR + pR)) + lohi * (qnR - qnL)) / span;
    f[2] = ((hi * (qnL * utL) - lo * (qnR * utR)) + lohi * (qtR - qtL)) / span;
}

// One HLL step in FP64 with a given dt, the operations of the CPU restatement's FP64 twin in its
// order (walls: the ghost's normal momentum negated; open: the cell itself).
__global__ __launch_bounds__(kBlock) void hll64_kernel(int nx, int ny, int4 edges, const double* __restrict__ ih,
                                                       const double* __restrict__ ihu,
                                                       const double* __restrict__ ihv, double* __restrict__ oh,
                                                       double* __restrict__ ohu, double* __restrict__ ohv, double g,
                                                       double rx, double ry, double dry) {
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= (int64_t)nx * ny) return;
    const int x = (int)(cell % nx), y = (int)(cell / nx);
    auto nb = [&](bool edge, int64_t off, int kind, bool along_x) {
        const int64_t at = edge ? cell : cell + off;
        S64 s{ih[at], ihu[at], ihv[at]};
        if (kind == 0 && edge) {
            if (along_x) s.hu = -s.hu;
            else s.hv = -s.hv;
        }
        return s;
    };
    const S64 c{ih[cell], ihu[cell], ihv[cell]};
    const S64 w = nb(x == 0, -1, edges.x, true), e = nb(x == nx - 1, 1, edges.y, true);
    const S64 s = nb(y == 0, -(int64_t)nx, edges.z, false), n = nb(y == ny - 1, nx, edges.w, false);
    double fw[3], fe[3], fs[3], fn[3];
    flux64(w, c, true, g, dry, fw);
    flux64(c, e, true, g, dry, fe);
    flux64(s, c, false, g, dry, fs);
    flux64(c, n, false, g, dry, fn);
    const double h = (c.h - rx * (fe[0] - fw[0])) - ry * (fn[0] - fs[0]);
    oh[cell] = h < 0 ? 0 : h;
    ohu[cell] = (c.hu - rx * (fe[1] - fw[1])) - ry * (fn[2] - fs[2]);
    ohv[cell] = (c.hv - rx * (fe[2] - fw[2])) - ry * (fn[1] - fs[1]);
}

__global__ void f64_to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < n) b[e] = (float)a[e];
}

// The circular dam and Gaussian bumps of the shallow-water scenes on the device: depth 2 inside
// radius r of the centre, 1 outside, plus amp * exp(-d^2 / width^2) per bump; hu = hv = 0.
__global__ void dam_bumps_kernel(int n, int bumps, const double* __restrict__ bp, double radius,
                                 double* __restrict__ h, double* __restrict__ hu, double* __restrict__ hv) {
    const int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cell >= (int64_t)n * n) return;
    const double x = ((double)(cell % n) + 0.5) / n, y = ((double)(cell / n) + 0.5) / n;
    double v = (x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) < radius * radius ? 2.0 : 1.0;
    for (int b = 0; b < bumps; ++b) {
        const double amp = bp[4 * b], width = bp[4 * b + 1], cx = bp[4 * b + 2], cy = bp[4 * b + 3];
        v += amp * exp(-((x - cx) * (x - cx) + (y - cy) * (y - cy)) / (width * width));
    }
    h[cell] = v;
    hu[cell] = 0.0;
    hv[cell] = 0.0;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// The fallback rule, on the device: an application is refused if any depth it gives is negative
// or not finite, or if its volume (sum of depths) moved by more than `tolerance` relative. The
// first pass folds strided shares of the cells per block (sums in double), the second combines
// the blocks in a fixed order, so the verdict is the same on every run.
constexpr int kCheckParts = 256;

struct CheckPart {
    double before, after;
    long long bad;
};

__global__ __launch_bounds__(kBlock) void check_kernel(const float* __restrict__ h0, const float* __restrict__ h1,
                                                       int64_t cells, CheckPart* __restrict__ parts) {
    __shared__ CheckPart lds[kBlock];
    CheckPart p{0.0, 0.0, 0};
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < cells; k += (int64_t)gridDim.x * kBlock) {
        const float a = h0[k], b = h1[k];
        p.before += a;
        p.after += b;
        if (!(b >= 0.0f) || !isfinite(b)) p.bad++;
    }
    lds[threadIdx.x] = p;
    __syncthreads();
    for (int half = kBlock / 2; half > 0; half >>= 1) {
        if (threadIdx.x < half) {
            lds[threadIdx.x].before += lds[threadIdx.x + half].before;
            lds[threadIdx.x].after += lds[threadIdx.x + half].after;
            lds[threadIdx.x].bad += lds[threadIdx.x + half].bad;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = lds[0];
}

// flag[0] = 1 when the application is refused, else 0; flag[1] counts refusals; out (if given)
// = volume before, volume after, bad cells.
__global__ __launch_bounds__(kBlock) void check_final_kernel(const CheckPart* __restrict__ parts, double tolerance,
                                                             int32_t* __restrict__ flag, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    CheckPart t{0.0, 0.0, 0};
    for (int i = 0; i < kCheckParts; ++i) {
        t.before += parts[i].before;
        t.after += parts[i].after;
        t.bad += parts[i].bad;
    }
    const bool refuse = t.bad > 0 || !(fabs(t.after - t.before) <= tolerance * fabs(t.before));
    flag[0] = refuse ? 1 : 0;
    if (refuse) flag[1] += 1;
    if (out) {
        out[0] = t.before;
        out[1] = t.after;
        out[2] = (double)t.bad;
    }
}

}  // namespace
}  // namespace vxr

using namespace vxr;

namespace {
thread_local char g_serr[256] = "no error";
int serr(const char* what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return 0;
    snprintf(g_serr, sizeof g_serr, "%s: %s", what, hipGetErrorString(e));
    return 3;
}
int sfail(const char* msg) {
    snprintf(g_serr, sizeof g_serr, "%s", msg);
    return 1;
}
}  // namespace

// The surrogate's kernels as plain entry points over device pointers on a caller's stream. They
// return 0, 1 (bad argument) or 3 (HIP error); vxs_last_error says which.
// The staged convolution over resident blocks only (as many as the device holds at once), so
// each block loads its weights once and walks tiles; the count is read once per instantiation.
template <int kCin, int kCoTiles>
static int launch_conv(hipStream_t s, int64_t tiles, const __bf16* x, int images, int H, int W, const __bf16* k,
                       const float* bias, void* out, int out_ch, const __bf16* mk, int f) {
    static int resident = 0;
    if (!resident) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, conv3x3_lds_kernel<kCin, kCoTiles>, kBlock, 0) !=
                hipSuccess ||
            cus < 1 || per_cu < 1)
            return 1;
        resident = cus * per_cu;
    }
    const unsigned g = (unsigned)(tiles < resident ? tiles : resident);
    hipLaunchKernelGGL((conv3x3_lds_kernel<kCin, kCoTiles>), dim3(g), dim3(kBlock), 0, s, x, images, H, W, k, bias,
                       out, out_ch, mk, f);
    return 0;
}

// Blocks of a weight-gradient launch: the device's CU count times what a CU holds, taken at the
// first call; the partials and their summation order then depend only on the image size.
template <int kCin, int kCout>
static int wgrad_blocks() {
    static int blocks = 0;
    if (!blocks) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            return 0;
        blocks = cus * WgradShape<kCin, kCout>::kBlocksPerCu;
    }
    return blocks;
}

template <int kCin, int kCout>
static int64_t wgrad_words() {
    const int blocks = wgrad_blocks<kCin, kCout>();
    return blocks ? (int64_t)blocks * WgradShape<kCin, kCout>::kGk * WgradShape<kCin, kCout>::kWords : -1;
}

// The weight gradient: its blocks' partials into `part`, then their sum. Segments of rows about
// as many as the launch has blocks (a whole number of steps each).
template <int kCin, int kCout>
static int launch_wgrad(hipStream_t s, const __bf16* x, const __bf16* dy, int images, int H, int W, float* part,
                        float* dw, float* db) {
    typedef WgradShape<kCin, kCout> S;
    const int blocks = wgrad_blocks<kCin, kCout>();
    if (!blocks) return sfail("wgrad: cannot size the launch");
    const int64_t strip_rows = (int64_t)images * ((W + kCols - 1) / kCols) * H;
    int64_t seg = (strip_rows + blocks - 1) / blocks;
    seg = (seg + S::kRows - 1) / S::kRows * S::kRows;
    if (seg > H) seg = (H + S::kRows - 1) / S::kRows * S::kRows;
    const int64_t segs = (int64_t)images * ((W + kCols - 1) / kCols) * ((H + seg - 1) / seg);
    const int grid = (int)(segs < blocks ? segs : blocks);
    hipLaunchKernelGGL((conv3x3_wgrad_kernel<kCin, kCout>), dim3(grid), dim3(64 * S::kWaves), 0, s, x, dy, images, H, W,
                       (int)seg, part);
    if (int rc = serr("wgrad")) return rc;
    const int nw = 9 * kCin * kCout;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((S::kWords + 63) / 64), dim3(kBlock), 0, s, part, grid * S::kGk,
                       nw, S::kWords, dw, db);
    return serr("wgrad reduce");
}

// The streamed convolution over resident blocks (as many as the device holds at once), in
// segments of rows about as many as there are blocks.
template <int kCin, int kCoTiles, bool kMasked>
static int launch_conv_rows(hipStream_t s, const __bf16* x, int images, int H, int W, const __bf16* k,
                            const float* bias, void* out, const __bf16* mk, int f) {
    static int resident = 0;
    if (!resident) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, conv3x3_rows_kernel<kCin, kCoTiles, kMasked>,
                                                         64 * ConvRows<kCin, kCoTiles, kMasked>::kWaves,
                                                         0) != hipSuccess ||
            cus < 1 || per_cu < 1)
            return 1;
        resident = cus * per_cu;
    }
    constexpr int R = ConvRows<kCin, kCoTiles, kMasked>::kR;
    const int64_t strips = (W + kCols - 1) / kCols;
    int64_t seg = (images * strips * H + resident - 1) / resident;
    seg = (seg + R - 1) / R * R;
    if (seg > H) seg = (H + R - 1) / R * R;
    const int64_t segs = images * strips * ((H + seg - 1) / seg);
    const unsigned g = (unsigned)(segs < resident ? segs : resident);
    hipLaunchKernelGGL((conv3x3_rows_kernel<kCin, kCoTiles, kMasked>), dim3(g),
                       dim3(64 * ConvRows<kCin, kCoTiles, kMasked>::kWaves), 0, s, x, images, H, W,
                       (int)seg, k, bias, out, mk, f);
    return 0;
}

template <int kCin, int kCoTiles>
static int launch_conv_rows(hipStream_t s, const __bf16* x, int images, int H, int W, const __bf16* k,
                            const float* bias, void* out, const __bf16* mk, int f) {
    return (f & kMask) ? launch_conv_rows<kCin, kCoTiles, true>(s, x, images, H, W, k, bias, out, mk, f)
                       : launch_conv_rows<kCin, kCoTiles, false>(s, x, images, H, W, k, bias, out, mk, f);
}

extern "C" {

const char* vxs_last_error(void) { return g_serr; }

// flags: 1 ReLU, 2 mask by (mask > 0), 4 FP32 output, 8 add into the FP32 output, 16 the
// unstaged kernel (operands straight from global memory; kept as a cross-check).
int vxs_conv3x3(void* stream, const void* in, int images, int H, int W, int cin, const void* w, const float* bias,
                int cout, void* out, int out_ch, const void* mask, int flags) {
    if (!in || !w || !out || images < 1 || H < 1 || W < 1) return sfail("conv3x3: bad argument");
    if (out_ch < 1 || out_ch > cout) return sfail("conv3x3: out_ch must be 1..cout");
    if ((flags & kMask) && !mask) return sfail("conv3x3: mask flag without a mask");
    hipStream_t s = (hipStream_t)stream;
    const __bf16* x = (const __bf16*)in;
    const __bf16* k = (const __bf16*)w;
    const __bf16* mk = (const __bf16*)mask;
    if (!(flags & (16 | kAccumulate)) && out_ch == cout) {  // the streamed kernel
        const int f = flags & 15;
        int rc = 1;
        if (cin == 16 && cout == 64)
            rc = launch_conv_rows<16, 4>(s, x, images, H, W, k, bias, out, mk, f);
        else if (cin == 64 && cout == 64)
            rc = launch_conv_rows<64, 4>(s, x, images, H, W, k, bias, out, mk, f);
        else if (cin == 64 && cout == 16)
            rc = launch_conv_rows<64, 1>(s, x, images, H, W, k, bias, out, mk, f);
        else if (cin == 16 && cout == 16)
            rc = launch_conv_rows<16, 1>(s, x, images, H, W, k, bias, out, mk, f);
        else
            return sfail("conv3x3: channels (cin, cout) must be 16 or 64 each");
        if (rc) return sfail("conv3x3: cannot size the launch");
        return serr("conv3x3");
    }
    if (!(flags & 16)) {  // the staged kernel, resident blocks over 2 x 64 tiles
        const int64_t tiles = (int64_t)images * ((H + kRows - 1) / kRows) * ((W + kCols - 1) / kCols);
        const int f = flags & 15;
        int rc = 1;
        if (cin == 16 && cout == 64)
            rc = launch_conv<16, 4>(s, tiles, x, images, H, W, k, bias, out, out_ch, mk, f);
        else if (cin == 64 && cout == 64)
            rc = launch_conv<64, 4>(s, tiles, x, images, H, W, k, bias, out, out_ch, mk, f);
        else if (cin == 64 && cout == 16)
            rc = launch_conv<64, 1>(s, tiles, x, images, H, W, k, bias, out, out_ch, mk, f);
        else if (cin == 16 && cout == 16)
            rc = launch_conv<16, 1>(s, tiles, x, images, H, W, k, bias, out, out_ch, mk, f);
        else
            return sfail("conv3x3: channels (cin, cout) must be 16 or 64 each");
        if (rc) return sfail("conv3x3: cannot size the launch");
        return serr("conv3x3");
    }
    const dim3 grid((unsigned)(H * ((W + kTileX - 1) / kTileX)), (unsigned)images);
    flags &= 15;
    if (cin == 16 && cout == 64)
        hipLaunchKernelGGL((conv3x3_kernel<16, 4>), grid, dim3(kBlock), 0, s, x, H, W, k, bias, out, out_ch, mk, flags);
    else if (cin == 64 && cout == 64)
        hipLaunchKernelGGL((conv3x3_kernel<64, 4>), grid, dim3(kBlock), 0, s, x, H, W, k, bias, out, out_ch, mk, flags);
    else if (cin == 64 && cout == 16)
        hipLaunchKernelGGL((conv3x3_kernel<64, 1>), grid, dim3(kBlock), 0, s, x, H, W, k, bias, out, out_ch, mk, flags);
    else if (cin == 16 && cout == 16)
        hipLaunchKernelGGL((conv3x3_kernel<16, 1>), grid, dim3(kBlock), 0, s, x, H, W, k, bias, out, out_ch, mk, flags);
    else
        return sfail("conv3x3: channels (cin, cout) must be 16 or 64 each");
    return serr("conv3x3");
}

// Scratch words the weight gradient needs: a partial per block and wave group.
int64_t vxs_wgrad_scratch(int cin, int cout) {
    if (cin == 16 && cout == 64) return wgrad_words<16, 64>();
    if (cin == 64 && cout == 64) return wgrad_words<64, 64>();
    if (cin == 64 && cout == 16) return wgrad_words<64, 16>();
    if (cin == 16 && cout == 16) return wgrad_words<16, 16>();
    return -1;
}

// dw[co][tap][ci] and db[co] of a 3x3 convolution from its input x and output gradient dy (both
// NHWC bf16 over `images` images of H x W), reduced in a fixed order.
int vxs_conv3x3_wgrad(void* stream, const void* x, const void* dy, int images, int H, int W, int cin, int cout,
                      float* scratch, float* dw, float* db) {
    if (!x || !dy || !scratch || !dw || !db || images < 1 || H < 1 || W < 1) return sfail("wgrad: bad argument");
    hipStream_t s = (hipStream_t)stream;
    const __bf16* a = (const __bf16*)x;
    const __bf16* g = (const __bf16*)dy;
    if (cin == 16 && cout == 64) return launch_wgrad<16, 64>(s, a, g, images, H, W, scratch, dw, db);
    if (cin == 64 && cout == 64) return launch_wgrad<64, 64>(s, a, g, images, H, W, scratch, dw, db);
    if (cin == 64 && cout == 16) return launch_wgrad<64, 16>(s, a, g, images, H, W, scratch, dw, db);
    if (cin == 16 && cout == 16) return launch_wgrad<16, 16>(s, a, g, images, H, W, scratch, dw, db);
    return sfail("wgrad: channels (cin, cout) must be 16 or 64 each");
}

int vxs_flip(void* stream, const void* w, int cout, int cin, void* wt) {
    if (!w || !wt) return sfail("flip: bad argument");
    hipLaunchKernelGGL(flip_kernel, dim3(blocks_for((int64_t)9 * cout * cin)), dim3(kBlock), 0,
                       (hipStream_t)stream, (const __bf16*)w, cout, cin, (__bf16*)wt);
    return serr("flip");
}

int vxs_adam(void* stream, float* p, void* p16, const float* g, float* m, float* v, int64_t n, float lr, float b1,
             float b2, float eps, int t) {
    if (!p || !g || !m || !v || t < 1) return sfail("adam: bad argument");
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, p, (__bf16*)p16, g, m,
                       v, n, lr, b1, b2, eps, c1, c2);
    return serr("adam");
}

// norm: offsets (3), input scales (3), output scales (3) on the device.
int vxs_pack(void* stream, const float* h, const float* hu, const float* hv, int64_t pixels, const float* norm,
             void* out) {
    if (!h || !hu || !hv || !norm || !out) return sfail("pack: bad argument");
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(pixels)), dim3(kBlock), 0, (hipStream_t)stream, h, hu, hv, pixels,
                       norm, (__bf16*)out);
    return serr("pack");
}

int vxs_unpack(void* stream, const float* h, const float* hu, const float* hv, const float* delta, int64_t pixels,
               const float* norm, float* oh, float* ohu, float* ohv) {
    if (!h || !hu || !hv || !delta || !norm || !oh || !ohu || !ohv) return sfail("unpack: bad argument");
    hipLaunchKernelGGL(unpack_kernel, dim3(blocks_for(pixels)), dim3(kBlock), 0, (hipStream_t)stream, h, hu, hv,
                       delta, pixels, norm, oh, ohu, ohv);
    return serr("unpack");
}

int vxs_loss(void* stream, const float* delta, const float* h, const float* hu, const float* hv, const float* th,
             const float* thu, const float* thv, int64_t pixels, const float* norm, void* grad, float* loss) {
    if (!delta || !h || !th || !norm || !grad || !loss) return sfail("loss: bad argument");
    hipLaunchKernelGGL(loss_kernel, dim3(blocks_for(pixels)), dim3(kBlock), 0, (hipStream_t)stream, delta, h, hu, hv,
                       th, thu, thv, pixels, norm, (__bf16*)grad, loss);
    return serr("loss");
}

// One FP64 HLL step with a fixed dt; edges west, east, south, north: 0 wall, 1 open.
int vxs_hll64(void* stream, int nx, int ny, const int* edges, const double* h, const double* hu, const double* hv,
              double* oh, double* ohu, double* ohv, double g, double dx, double dy, double dry, double dt) {
    if (nx < 1 || ny < 1 || !edges || !h || !hu || !hv || !oh || !ohu || !ohv) return sfail("hll64: bad argument");
    const int4 e = make_int4(edges[0], edges[1], edges[2], edges[3]);
    hipLaunchKernelGGL(hll64_kernel, dim3(blocks_for((int64_t)nx * ny)), dim3(kBlock), 0, (hipStream_t)stream, nx, ny,
                       e, h, hu, hv, oh, ohu, ohv, g, dt / dx, dt / dy, dry);
    return serr("hll64");
}

// bump_params: amp, width, cx, cy per bump (device memory).
int vxs_dam_bumps(void* stream, int n, int bumps, const double* bump_params, double radius, double* h, double* hu,
                  double* hv) {
    if (n < 1 || bumps < 0 || (bumps && !bump_params) || !h || !hu || !hv) return sfail("dam_bumps: bad argument");
    hipLaunchKernelGGL(dam_bumps_kernel, dim3(blocks_for((int64_t)n * n)), dim3(kBlock), 0, (hipStream_t)stream, n,
                       bumps, bump_params, radius, h, hu, hv);
    return serr("dam_bumps");
}

// One FP32 HLL step of exactly `ticks` ticks (dt = ticks * 1e-6 s, as the tick rule rounds it):
// the shallow-water step's kernel and arithmetic (bit for bit the CPU restatement's step whose
// budget caps it at `ticks`), walls or open edges as `edges` (west, east, south, north: 0 wall,
// 1 open), params the 5 device words of ShallowWater.params. With `gate` (a device word) the step
// runs only if *gate is nonzero. No speed, budget or clamp words are kept.
int vxs_hll32(void* stream, int nx, int ny, const int* edges, const int32_t* params, int32_t ticks,
              const int32_t* gate, const float* h, const float* hu, const float* hv, float* oh, float* ohu,
              float* ohv) {
    if (nx < 1 || ny < 1 || !edges || !params || ticks < 1 || !h || !hu || !hv || !oh || !ohu || !ohv)
        return sfail("hll32: bad argument");
    vxr::SweArgs a{};
    a.nx = nx;
    a.ny = ny;
    for (int k = 0; k < 4; ++k) a.edges[k] = edges[k] ? 1 : 0;
    a.in[0] = h;
    a.in[1] = hu;
    a.in[2] = hv;
    a.out[0] = oh;
    a.out[1] = ohu;
    a.out[2] = ohv;
    a.params = params;
    a.fixed_ticks = ticks;
    a.gate = gate;
    const vxr::Launch l{(hipStream_t)stream, nullptr};
    if (vxr::swe_step(l, a, 32)) return serr("hll32");
    return serr("hll32");
}

// Words of scratch vxs_fallback_check takes.
int64_t vxs_check_scratch(void) { return (int64_t)kCheckParts * (int64_t)(sizeof(CheckPart) / 4); }

// The surrogate's fallback rule on the device, for an application that took depths h0 to h1
// (cells each): flag[0] = 1 if any h1 is negative or not finite, or the volume moved by more than
// `tolerance` relative, else 0; flag[1] counts the refusals. out (device, optional) = the two
// volumes (sums of depths) and the bad cells. Asynchronous; a fixed summation order.
int vxs_fallback_check(void* stream, const float* h0, const float* h1, int64_t cells, double tolerance,
                       void* scratch, int32_t* flag, double* out) {
    if (!h0 || !h1 || cells < 1 || !scratch || !flag || !(tolerance >= 0)) return sfail("fallback_check: bad argument");
    CheckPart* parts = (CheckPart*)scratch;
    hipLaunchKernelGGL(check_kernel, dim3(kCheckParts), dim3(kBlock), 0, (hipStream_t)stream, h0, h1, cells, parts);
    hipLaunchKernelGGL(check_final_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, parts, tolerance, flag, out);
    return serr("fallback_check");
}

int vxs_f64_to_f32(void* stream, const double* a, float* b, int64_t n) {
    if (!a || !b) return sfail("f64_to_f32: bad argument");
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, a, b, n);
    return serr("f64_to_f32");
}

}  // extern "C"
