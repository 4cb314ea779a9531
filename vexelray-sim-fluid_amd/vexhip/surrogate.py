"""A convolutional surrogate of the HLL shallow-water step, its training on FP64 HLL
trajectories, and its fallback to plain HLL.

No reference counterpart: the reference has only the HLL solver (core/stencil/ShallowWater.java).
The network maps a state (h, hu, hv) on an n x n grid to the state ``STEPS`` HLL steps of a fixed
``dt`` later: a residual stack of six 3x3 convolutions (16 -> 64, four 64 -> 64, 64 -> 16; the 3
state channels padded to 16), ReLU between, bf16 weights and activations, FP32 accumulation on
the matrix cores (csrc/surrogate.hip). The last layer's output, scaled back, is added to the
state. Its targets and its judge are the FP64 HLL step (vxs_hll64), the CPU restatement's FP64
twin on the device. ``Surrogate.advance`` applies it under the fallback rule: a device check
(vxs_fallback_check) refuses an application that gives a negative or non-finite depth or moves the
volume by more than 1e-3, and then STEPS plain FP32 HLL steps of the same dt (vxs_hll32, the
shallow-water step's own kernel) replace it -- queued behind the check and gated on its flag word,
so no host round trip decides it.

Every kernel is libvexelhip.so's; torch only allocates device memory and names the stream.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _native, scenes

LAYERS = ((16, 64), (64, 64), (64, 64), (64, 64), (64, 64), (64, 16))
STEPS = 8            # HLL steps one application stands for
G = 9.81
DRY = 1e-6
COURANT = 0.3        # of the fixed dt against the fastest wave the scenes reach
FASTEST = 6.0        # m/s: sqrt(g * 2.6) plus a current of about 1 m/s
TOLERANCE = 1e-3     # the fallback rule's relative volume drift
BLOCKS = 16          # blocks of STEPS steps a training trajectory runs

_torch = None


def _t():
    """torch, for device memory and the stream. It must be imported before libvexelhip.so is
    loaded, so that both use torch's HIP runtime."""
    global _torch
    if _torch is None:
        if _native._lib is not None and not _native._torch_first:
            raise RuntimeError("import torch before libvexelhip.so is loaded: the surrogate's buffers are torch's, "
                               "and both must use one HIP runtime")
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("the surrogate runs on a HIP device; none is visible")
        _torch = torch
    return _torch


def _p(t) -> int:
    return t.data_ptr()


def _stream() -> int:
    return _t().cuda.current_stream().cuda_stream


def fixed_ticks(n: int) -> int:
    """The fixed step on an n x n unit square in whole ticks of 1e-6 s (the shallow-water clock's
    unit), so that the FP32 fallback's step is the tick rule's own."""
    return max(1, int(COURANT * (1.0 / n) / FASTEST * 1e6))


def fixed_dt(n: int) -> float:
    """The fixed step of the training and judging HLL runs on an n x n unit square."""
    return fixed_ticks(n) * 1e-6


def swe_words(n: int) -> np.ndarray:
    """ShallowWater.params' five words for the unit square at n x n (the Courant number is not
    read by a fixed-tick step)."""
    f = np.array([G, 0.45, 1.0 / n, 1.0 / n, DRY], np.float32)
    return f.view(np.int32)


def bump_params(seed: int, bumps: int = 32) -> np.ndarray:
    """swe_dam_bumps's bump draws for `seed`: amp, width, cx, cy per bump."""
    rng = scenes.JavaRandom(seed)
    out = np.empty((bumps, 4), np.float64)
    for b in range(bumps):
        out[b, 0] = 0.05 + 0.25 * rng.next_double()
        out[b, 1] = 0.01 + 0.04 * rng.next_double()
        out[b, 2] = rng.next_double()
        out[b, 3] = rng.next_double()
    return out


class Hll64:
    """FP64 HLL on the device with walls all round and a fixed dt (vxs_hll64)."""

    def __init__(self, n: int):
        torch = _t()
        self.n = n
        self.dt = fixed_dt(n)
        self.a = [torch.empty(n * n, dtype=torch.float64, device="cuda") for _ in range(3)]
        self.b = [torch.empty_like(self.a[0]) for _ in range(3)]
        self.edges = (ctypes.c_int * 4)(0, 0, 0, 0)

    def scene(self, seed: int) -> None:
        """The dam-and-bumps initial state of `seed` (scenes.swe_dam_bumps on the device)."""
        torch = _t()
        bp = torch.from_numpy(bump_params(seed).ravel()).to("cuda")
        _native.vxs_check(_native.lib().vxs_dam_bumps(_stream(), self.n, 32, _p(bp), 0.2, *map(_p, self.a)))

    def load(self, h, hu, hv) -> None:
        for dst, src in zip(self.a, (h, hu, hv)):
            dst.copy_(src.reshape(-1).to(dst.dtype))

    def step(self, count: int) -> None:
        lib = _native.lib()
        d = 1.0 / self.n
        for _ in range(count):
            _native.vxs_check(lib.vxs_hll64(_stream(), self.n, self.n, self.edges, *map(_p, self.a),
                                            *map(_p, self.b), G, d, d, DRY, self.dt))
            self.a, self.b = self.b, self.a

    def state32(self):
        return [x.to(_t().float32) for x in self.a]


class Surrogate:
    """The network, its FP32 master weights and Adam state, and buffers for `images` images of
    n x n."""

    def __init__(self, n: int, images: int = 1, seed: int = 0):
        torch = _t()
        self.n, self.images = n, images
        gen = np.random.default_rng(seed)
        self.w32, self.b32, self.w16 = [], [], []
        for li, (cin, cout) in enumerate(LAYERS):
            fan_in = 9 * (3 if li == 0 else cin)
            w = gen.standard_normal((cout, 9, cin)) * math.sqrt(2.0 / fan_in)
            if li == 0:
                w[:, :, 3:] = 0.0          # padding input channels
            if li == len(LAYERS) - 1:
                w *= 1e-2
                w[3:] = 0.0                # padding output channels stay zero
            self.w32.append(torch.tensor(w.astype(np.float32).ravel(), device="cuda"))
            self.b32.append(torch.zeros(cout, dtype=torch.float32, device="cuda"))
            self.w16.append(self.w32[-1].to(torch.bfloat16))
        self.m = [torch.zeros_like(p) for p in self.w32 + self.b32]
        self.v = [torch.zeros_like(p) for p in self.w32 + self.b32]
        self.t = 0
        # offsets (3), input scales (3), output scales (3)
        self.norm = torch.tensor([1.0, 0.0, 0.0, 2.0, 4.0, 4.0, 1.0, 1.0, 1.0], dtype=torch.float32, device="cuda")
        self._alloc(images)

    def _alloc(self, images: int) -> None:
        torch = _t()
        px = images * self.n * self.n
        self.images = images
        self.x0 = torch.empty(px * 16, dtype=torch.bfloat16, device="cuda")
        self.acts = [torch.empty(px * 64, dtype=torch.bfloat16, device="cuda") for _ in range(5)]
        self.delta = torch.empty(px * 16, dtype=torch.float32, device="cuda")
        self.grads = None

    def _alloc_grads(self) -> None:
        torch = _t()
        px = self.images * self.n * self.n
        self.g16 = torch.empty(px * 16, dtype=torch.bfloat16, device="cuda")
        self.g64 = [torch.empty(px * 64, dtype=torch.bfloat16, device="cuda") for _ in range(2)]
        self.dw = [torch.empty_like(w) for w in self.w32]
        self.db = [torch.empty_like(b) for b in self.b32]
        self.wt = [torch.empty_like(w) for w in self.w16]
        words = max(_native.lib().vxs_wgrad_scratch(cin, cout) for cin, cout in LAYERS)
        self.scratch = torch.empty(words, dtype=torch.float32, device="cuda")
        self.loss_acc = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.grads = True

    def _alloc_fallback(self) -> None:
        torch = _t()
        px = self.images * self.n * self.n
        self.flag = torch.zeros(2, dtype=torch.int32, device="cuda")
        self.check_out = torch.zeros(3, dtype=torch.float64, device="cuda")
        self.check_scratch = torch.empty(_native.lib().vxs_check_scratch(), dtype=torch.int32, device="cuda")
        self.params32 = torch.from_numpy(swe_words(self.n)).to("cuda")
        self.hll_tmp = [[torch.empty(px, dtype=torch.float32, device="cuda") for _ in range(3)] for _ in range(2)]
        self.edges = (ctypes.c_int * 4)(0, 0, 0, 0)

    def advance(self, h, hu, hv, out=None):
        """One application under the fallback rule (an image at a time, walls all round): the
        network's state STEPS steps on, unless the device check refuses it -- then STEPS plain
        FP32 HLL steps of fixed_ticks(n) each from the same state, bit for bit. Asynchronous."""
        if self.images != 1:
            raise ValueError("advance takes one image at a time")
        if getattr(self, "flag", None) is None:
            self._alloc_fallback()
        lib = _native.lib()
        st = _stream()
        out = self.forward(h, hu, hv, out)
        _native.vxs_check(lib.vxs_fallback_check(st, _p(h), _p(out[0]), self.n * self.n, TOLERANCE,
                                                 _p(self.check_scratch), _p(self.flag), _p(self.check_out)))
        ticks = fixed_ticks(self.n)
        src = (h, hu, hv)
        for k in range(STEPS):
            dst = out if k == STEPS - 1 else self.hll_tmp[k % 2]
            _native.vxs_check(lib.vxs_hll32(st, self.n, self.n, self.edges, _p(self.params32), ticks,
                                            _p(self.flag), *map(_p, src), *map(_p, dst)))
            src = dst
        return out

    def fallbacks(self) -> int:
        """Applications the fallback rule has refused so far (a readback)."""
        if getattr(self, "flag", None) is None:
            return 0
        return int(self.flag[1].item())

    def set_output_scale(self, scales) -> None:
        self.norm[6:9] = _t().tensor([float(s) for s in scales], dtype=_t().float32)

    # ---------------------------------------------------------------- forward

    def _conv(self, li, src, dst, relu, out_f32=False):
        cin, cout = LAYERS[li]
        flags = (1 if relu else 0) | (4 if out_f32 else 0)
        _native.vxs_check(_native.lib().vxs_conv3x3(_stream(), _p(src), self.images, self.n, self.n, cin,
                                                    _p(self.w16[li]), _p(self.b32[li]), cout, _p(dst), cout, None,
                                                    flags))

    def forward(self, h, hu, hv, out=None):
        """The state STEPS steps on, as three flat FP32 tensors (written into `out` if given)."""
        torch = _t()
        px = self.images * self.n * self.n
        lib = _native.lib()
        _native.vxs_check(lib.vxs_pack(_stream(), _p(h), _p(hu), _p(hv), px, _p(self.norm), _p(self.x0)))
        src = self.x0
        for li in range(5):
            self._conv(li, src, self.acts[li], relu=True)
            src = self.acts[li]
        self._conv(5, src, self.delta, relu=False, out_f32=True)
        if out is None:
            out = [torch.empty_like(h), torch.empty_like(hu), torch.empty_like(hv)]
        _native.vxs_check(lib.vxs_unpack(_stream(), _p(h), _p(hu), _p(hv), _p(self.delta), px, _p(self.norm),
                                         *map(_p, out)))
        return out

    # ---------------------------------------------------------------- training

    def train_step(self, state, target, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8) -> float:
        """One Adam step on the mean squared error of the scaled change; returns the loss."""
        if self.grads is None:
            self._alloc_grads()
        lib = _native.lib()
        st = _stream()
        px = self.images * self.n * self.n
        self.forward(*state)
        self.loss_acc.zero_()
        _native.vxs_check(lib.vxs_loss(st, _p(self.delta), *map(_p, state), *map(_p, target), px, _p(self.norm),
                                       _p(self.g16), _p(self.loss_acc)))
        g = self.g16
        for li in range(5, -1, -1):
            cin, cout = LAYERS[li]
            x = self.x0 if li == 0 else self.acts[li - 1]
            _native.vxs_check(lib.vxs_conv3x3_wgrad(st, _p(x), _p(g), self.images, self.n, self.n, cin, cout,
                                                    _p(self.scratch), _p(self.dw[li]), _p(self.db[li])))
            if li > 0:
                _native.vxs_check(lib.vxs_flip(st, _p(self.w16[li]), cout, cin, _p(self.wt[li])))
                dst = self.g64[li % 2]
                # the input gradient, kept where the ReLU that made this layer's input passed
                _native.vxs_check(lib.vxs_conv3x3(st, _p(g), self.images, self.n, self.n, cout, _p(self.wt[li]),
                                                  None, cin, _p(dst), cin, _p(x), 2))
                g = dst
        self.t += 1
        params = self.w32 + self.b32
        grads = self.dw + self.db
        copies = self.w16 + [None] * len(self.b32)
        for k, (p, gr, c) in enumerate(zip(params, grads, copies)):
            _native.vxs_check(lib.vxs_adam(st, _p(p), _p(c) if c is not None else None, _p(gr), _p(self.m[k]),
                                           _p(self.v[k]), p.numel(), lr, b1, b2, eps, self.t))
        return float(self.loss_acc.item())

    # ---------------------------------------------------------------- weights

    def state_dict(self) -> dict:
        d = {f"w{k}": w.cpu().numpy() for k, w in enumerate(self.w32)}
        d.update({f"b{k}": b.cpu().numpy() for k, b in enumerate(self.b32)})
        d["norm"] = self.norm.cpu().numpy()
        return d

    def load_state_dict(self, d: dict) -> None:
        torch = _t()
        for k in range(len(LAYERS)):
            self.w32[k].copy_(torch.from_numpy(np.asarray(d[f"w{k}"], np.float32)))
            self.b32[k].copy_(torch.from_numpy(np.asarray(d[f"b{k}"], np.float32)))
            self.w16[k].copy_(self.w32[k].to(torch.bfloat16))
        self.norm.copy_(torch.from_numpy(np.asarray(d["norm"], np.float32)))


# ------------------------------------------------------------------ data, training, judging

def trajectory_pair(hll: Hll64, seed: int, block: int):
    """(state, target) in FP32: the scene of `seed` after `block` blocks of STEPS FP64 HLL steps,
    and STEPS steps later."""
    hll.scene(seed)
    hll.step(STEPS * block)
    state = hll.state32()
    hll.step(STEPS)
    return state, hll.state32()


def calibrate(model: Surrogate, hll: Hll64, seeds, rng) -> list:
    """Output scales: 1 / (3 std) of each channel's change over a block, on a few samples."""
    torch = _t()
    sums = np.zeros(3)
    for s in seeds:
        a, b = trajectory_pair(hll, int(s), int(rng.integers(BLOCKS)))
        for c in range(3):
            sums[c] += float(torch.std(b[c] - a[c]).item())
    std = sums / len(seeds)
    scales = [1.0 / max(3.0 * x, 1e-6) for x in std]
    model.set_output_scale(scales)
    return scales


def train(n=1024, seeds=range(2000), iters=None, lr=1e-3, seed=0, log=None, model=None):
    """Trains on one pair per trajectory: seed k's scene after a random number of blocks. Returns
    the model and the loss history."""
    rng = np.random.default_rng(seed)
    hll = Hll64(n)
    model = model or Surrogate(n, 1, seed)
    seeds = list(seeds)
    calibrate(model, hll, seeds[:8], rng)
    iters = iters or len(seeds)
    losses = []
    t0 = time.perf_counter()
    for it in range(iters):
        s = seeds[it % len(seeds)]
        state, target = trajectory_pair(hll, s, int(rng.integers(BLOCKS)))
        rate = lr * 0.5 * (1 + math.cos(math.pi * it / iters))
        losses.append(model.train_step(state, target, lr=rate))
        if log and (it % 100 == 0 or it == iters - 1):
            log(f"[surrogate] iter {it} loss {losses[-1]:.3e} ({time.perf_counter() - t0:.1f} s)")
    return model, losses


def evaluate(model: Surrogate, n=1024, seeds=range(10000, 10200), seed=1):
    """Relative L2 error in h against the FP64 HLL, one random block per scene, one application
    under the fallback rule (what Surrogate.advance returns); with the error of keeping the state
    (persistence) for scale, and the applications the rule refused."""
    torch = _t()
    rng = np.random.default_rng(seed)
    hll = Hll64(n)
    errs, keep = [], []
    before = model.fallbacks()
    for s in seeds:
        hll.scene(int(s))
        hll.step(STEPS * int(rng.integers(BLOCKS)))
        state = hll.state32()
        hll.step(STEPS)
        truth = hll.a[0]
        pred = model.advance(*state)
        ref = torch.linalg.vector_norm(truth).item()
        errs.append(torch.linalg.vector_norm(pred[0].double() - truth).item() / ref)
        keep.append(torch.linalg.vector_norm(state[0].double() - truth).item() / ref)
    return {"rel_l2_h_mean": float(np.mean(errs)), "rel_l2_h_max": float(np.max(errs)),
            "persistence_rel_l2_h_mean": float(np.mean(keep)), "fallbacks": model.fallbacks() - before,
            "scenes": len(errs)}
