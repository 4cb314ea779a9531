/*
 * vxs.h -- the surrogate's entry points in libvexelhip.so: a learned stand-in for 8 HLL steps
 * of the shallow-water patch (a residual 3x3 convolutional net in bf16 on the matrix cores, FP32
 * accumulation), the FP64 HLL that judges it, and the fallback that replaces a refused
 * application by plain HLL steps.
 *
 * None of these has a counterpart in the reference, which steps the patch only with the HLL
 * kernel (core/stencil/ShallowWater.java:166-384, dispatched by
 * gui/PatchSimulation.java:58-76,197-213). Their nearest reference relatives are named per
 * entry point. The host that drives them (training, judging, the fallback rule) is
 * vexelray-sim-fluid_amd/vexhip/surrogate.py.
 *
 * Conventions: `stream` is a hipStream_t (NULL: the null stream); every pointer but `stream`
 * is device memory; every call is asynchronous on `stream`; images are NHWC bf16 (16-bit words
 * passed as void*); the status is 0 on success, 1 for a bad argument or 3 for a HIP error, with
 * a message for vxs_last_error().
 *
 * Reference paths are shortened as in vxr.h: core/ = vexelray-sim-fluid-core/src/main/java/
 * dev/vexelray/sim/fluid/, gui/ = vexelray-sim-fluid-gui/src/main/java/dev/vexelray/sim/fluid/gui/.
 */
#ifndef VEXEL_VXS_H
#define VEXEL_VXS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The message of the calling thread's last failed vxs_* call (thread-local storage). */
const char* vxs_last_error(void);

/* ---- the network's layers (no reference counterpart) ---- */

/* flags for vxs_conv3x3 */
#define VXS_CONV_RELU 1       /* max(0, .) on the output */
#define VXS_CONV_MASK 2       /* multiply the output by (mask > 0), mask NHWC bf16 with out_ch channels */
#define VXS_CONV_F32_OUT 4    /* FP32 output instead of bf16 */
#define VXS_CONV_ACCUMULATE 8 /* add into the FP32 output (with VXS_CONV_F32_OUT) */
#define VXS_CONV_UNSTAGED 16  /* the unstaged kernel: operands straight from global memory (a cross-check) */

/* A 3x3 same-padded convolution: out[.., :out_ch] = conv(in, w) + bias. in: images x H x W x cin;
 * w: cout x 9 taps x cin bf16; bias: cout floats; (cin, cout) in {(16,64), (64,64), (64,16), (16,16)}. */
int vxs_conv3x3(void* stream, const void* in, int images, int H, int W, int cin, const void* w, const float* bias,
                int cout, void* out, int out_ch, const void* mask, int flags);

/* Words of FP32 scratch vxs_conv3x3_wgrad needs for channels (cin, cout), at any image size;
 * -1 for channels it does not take. */
int64_t vxs_wgrad_scratch(int cin, int cout);

/* The weight and bias gradients of a 3x3 convolution from its input x and output gradient dy
 * (NHWC bf16, images x H x W): dw[co][tap][ci] and db[co] in FP32, summed in an order fixed by
 * the image size (per block of the launch, then over blocks). (cin, cout) as vxs_conv3x3. */
int vxs_conv3x3_wgrad(void* stream, const void* x, const void* dy, int images, int H, int W, int cin, int cout,
                      float* scratch, float* dw, float* db);

/* The weights of the transposed convolution (input gradient): wt[ci][8 - tap][co] = w[co][tap][ci]. */
int vxs_flip(void* stream, const void* w, int cout, int cin, void* wt);

/* One Adam step over n FP32 parameters p (with a bf16 copy p16 when non-NULL), gradient g,
 * moments m and v, step number t >= 1. */
int vxs_adam(void* stream, float* p, void* p16, const float* g, float* m, float* v, int64_t n, float lr, float b1,
             float b2, float eps, int t);

/* The patch state (h, hu, hv FP32 planes) into the network's 16-channel bf16 input; norm: 3
 * offsets, 3 input scales, 3 output scales (device). */
int vxs_pack(void* stream, const float* h, const float* hu, const float* hv, int64_t pixels, const float* norm,
             void* out);

/* The state after one application: o = state + the network's output delta (FP32, 16 channels),
 * de-normalised by norm. */
int vxs_unpack(void* stream, const float* h, const float* hu, const float* hv, const float* delta, int64_t pixels,
               const float* norm, float* oh, float* ohu, float* ohv);

/* The training loss: the mean over pixels and the 3 fields of the squared difference between
 * the network's delta and the target change (th - h, thu - hu, thv - hv) times the output scales,
 * added into *loss (FP32), and its gradient with respect to delta into grad (bf16, 16 channels,
 * zero past the third). */
int vxs_loss(void* stream, const float* delta, const float* h, const float* hu, const float* hv, const float* th,
             const float* thu, const float* thv, int64_t pixels, const float* norm, void* grad, float* loss);

/* ---- the judge and the training data ---- */

/* One FP64 HLL step with a fixed dt: ShallowWater.kernel's arithmetic in binary64
 * (core/stencil/ShallowWater.java:166-384). edges: west, east, south, north; 0 wall, 1 open
 * (core/stencil/Edges.java). */
int vxs_hll64(void* stream, int nx, int ny, const int* edges, const double* h, const double* hu, const double* hv,
              double* oh, double* ohu, double* ohv, double g, double dx, double dy, double dry, double dt);

/* A dam-and-bumps scene in FP64 on an n x n patch of the unit square: a circular dam of depth 2
 * and radius `radius` about the centre in water of depth 1, plus `bumps` Gaussian humps
 * (bump_params: amp, width, cx, cy each, drawn by the host), at rest. */
int vxs_dam_bumps(void* stream, int n, int bumps, const double* bump_params, double radius, double* h, double* hu,
                  double* hv);

/* FP64 planes to FP32 (round to nearest). */
int vxs_f64_to_f32(void* stream, const double* a, float* b, int64_t n);

/* ---- the fallback ---- */

/* One FP32 HLL step of exactly `ticks` ticks (dt = ticks * 1e-6 s): the kernel and arithmetic of
 * vxr_swe_dispatch's step (core/stencil/ShallowWater.java:166-384), bit for bit the step whose
 * budget caps it at `ticks`; params: the 5 device words of ShallowWater.params
 * (core/stencil/ShallowWater.java:248-271). With `gate` (a device word) the step runs only if
 * *gate is nonzero, so a refused application's replacement needs no host round trip. */
int vxs_hll32(void* stream, int nx, int ny, const int* edges, const int32_t* params, int32_t ticks,
              const int32_t* gate, const float* h, const float* hu, const float* hv, float* oh, float* ohu,
              float* ohv);

/* Words of scratch vxs_fallback_check takes. */
int64_t vxs_check_scratch(void);

/* The fallback rule on the device for an application that took depths h0 to h1: flag[0] = 1 if
 * any h1 is negative or not finite, or the volume (sum of depths) moved by more than `tolerance`
 * relative, else 0; flag[1] counts refusals. out (optional, 3 doubles): the two volumes and the
 * bad cells. A fixed summation order. The reference's analogue is Diagnostics.broken()
 * (core/stencil/Diagnostics.java:43-91: non-finite or negative cells). */
int vxs_fallback_check(void* stream, const float* h0, const float* h1, int64_t cells, double tolerance,
                       void* scratch, int32_t* flag, double* out);

#ifdef __cplusplus
}
#endif

#endif /* VEXEL_VXS_H */
