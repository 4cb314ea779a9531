"""The surrogate's kernels against a plain PyTorch FP32 reference of the same operation (on the
CPU), the FP64 HLL step against the CPU restatement bit for bit, and a short training run."""
import ctypes

import numpy as np
import pytest

import oracle
from vexhip import _native, scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _bf(shape, gen, scale=1.0):
    return (torch.from_numpy(gen.standard_normal(shape).astype(np.float32)) * scale).to(torch.bfloat16)


def _conv_ref(x_nhwc, w_ctc, bias, relu):
    """FP32 reference: x [N,H,W,Ci], w [Co, 9, Ci] -> [N,H,W,Co] with zero padding."""
    co, _, ci = w_ctc.shape
    w = w_ctc.float().reshape(co, 3, 3, ci).permute(0, 3, 1, 2).contiguous()
    y = F.conv2d(x_nhwc.float().permute(0, 3, 1, 2).contiguous(), w, bias, padding=1)
    if relu:
        y = torch.relu(y)
    return y.permute(0, 2, 3, 1).contiguous()


@pytest.mark.parametrize("shape", [(2, 37, 70), (1, 21, 200)], ids=["edges", "interior"])
@pytest.mark.parametrize("cin,cout,relu,f32", [(16, 64, True, False), (64, 64, True, False), (64, 16, False, True),
                                               (16, 64, False, True)])
def test_conv_matches_torch(gpu, cin, cout, relu, f32, shape):
    """Both staged paths: 37 x 70 has every 2 x 64 tile on an image edge (rows not a multiple
    of the tile either); 21 x 200 has tiles clear of every edge, staged without bounds checks."""
    gen = np.random.default_rng(cin + cout)
    n, h, w = shape
    x = _bf((n, h, w, cin), gen)
    k = _bf((cout, 9, cin), gen, 0.1)
    b = torch.from_numpy(gen.standard_normal(cout).astype(np.float32))
    print(f"[surrogate] conv {cin}->{cout}", flush=True)
    dx, dk, db = x.cuda(), k.cuda(), b.cuda()
    outs = []
    for staged in (0, 16):  # the LDS-staged kernel (default) and the unstaged one
        out = torch.empty((n, h, w, cout), dtype=torch.float32 if f32 else torch.bfloat16, device="cuda")
        flags = (1 if relu else 0) | (4 if f32 else 0) | staged
        _native.vxs_check(_native.lib().vxs_conv3x3(torch.cuda.current_stream().cuda_stream, dx.data_ptr(), n, h, w,
                                                    cin, dk.data_ptr(), db.data_ptr(), cout, out.data_ptr(), cout,
                                                    None, flags))
        outs.append(out.float().cpu())
    ref = _conv_ref(x, k, b, relu)
    tol = 1e-4 if f32 else 1e-2  # FP32 sums in another order; the bf16 store rounds to 2^-8
    for got in outs:
        err = (got - ref).abs() / (ref.abs() + 1.0)
        assert float(err.max()) <= tol, float(err.max())
    # both walk K in the same order through the same matrix-core sequence
    assert torch.equal(outs[0], outs[1])


def test_conv_masked_input_gradient_matches_autograd(gpu):
    """The input gradient as the forward kernel over flipped weights, masked by the ReLU that made
    the input: equal to autograd's through conv2d and the ReLU."""
    gen = np.random.default_rng(5)
    n, h, w, cin, cout = 1, 20, 33, 64, 16
    x = _bf((n, h, w, cin), gen)
    k = _bf((cout, 9, cin), gen, 0.1)
    dy = _bf((n, h, w, cout), gen)
    xr = x.float().requires_grad_(True)
    y = _conv_ref(xr, k, None, False)
    (y * dy.float()).sum().backward()
    ref = xr.grad * (x.float() > 0)
    s = torch.cuda.current_stream().cuda_stream
    lib = _native.lib()
    dk, ddy = k.cuda(), dy.cuda()
    wt = torch.empty_like(dk)
    _native.vxs_check(lib.vxs_flip(s, dk.data_ptr(), cout, cin, wt.data_ptr()))
    out = torch.empty((n, h, w, cin), dtype=torch.float32, device="cuda")
    mask = x.cuda()
    _native.vxs_check(lib.vxs_conv3x3(s, ddy.data_ptr(), n, h, w, cout, wt.data_ptr(), None, cin, out.data_ptr(),
                                      cin, mask.data_ptr(), 2 | 4))
    got = out.cpu()
    assert float(((got - ref).abs() / (ref.abs() + 1.0)).max()) <= 1e-4


@pytest.mark.parametrize("shape", [(2, 23, 41), (1, 21, 200), (1, 130, 1024), (600, 4, 9)],
                         ids=["edges", "interior", "long-segments", "many-segments"])
@pytest.mark.parametrize("cin,cout", [(16, 64), (64, 64), (64, 16), (16, 16)])
def test_weight_gradient_matches_autograd(gpu, cin, cout, shape):
    """23 x 41: odd rows, a partial strip, every row on an image edge; 21 x 200: strips clear of
    the side edges; 130 x 1024: segments of several steps with rows in flight ahead; 600 images
    of 4 x 9: more segments than the launch has blocks, so blocks walk several. Twice, equal."""
    gen = np.random.default_rng(cin * 3 + cout)
    n, h, w = shape
    x = _bf((n, h, w, cin), gen)
    dy = _bf((n, h, w, cout), gen)
    k = torch.zeros((cout, 9, cin), requires_grad=True)
    b = torch.zeros(cout, requires_grad=True)
    y = _conv_ref(x, k, b, False)
    (y * dy.float()).sum().backward()
    print(f"[surrogate] wgrad {cin}->{cout} {shape}", flush=True)
    s = torch.cuda.current_stream().cuda_stream
    lib = _native.lib()
    words = lib.vxs_wgrad_scratch(cin, cout)
    assert words > 0
    scratch = torch.empty(words, dtype=torch.float32, device="cuda")
    dx, ddy = x.cuda(), dy.cuda()
    runs = []
    for _ in range(2):
        dw = torch.empty(cout * 9 * cin, dtype=torch.float32, device="cuda")
        db = torch.empty(cout, dtype=torch.float32, device="cuda")
        _native.vxs_check(lib.vxs_conv3x3_wgrad(s, dx.data_ptr(), ddy.data_ptr(), n, h, w, cin, cout,
                                                scratch.data_ptr(), dw.data_ptr(), db.data_ptr()))
        runs.append((dw.cpu(), db.cpu()))
    got, gotb = runs[0]
    ref = k.grad
    # FP32 sums of bf16 products in another order than torch's
    assert float(((got.reshape(cout, 9, cin) - ref).abs() / (ref.abs().max() + 1e-6)).max()) <= 1e-4
    assert float(((gotb - b.grad).abs() / (b.grad.abs().max() + 1e-6)).max()) <= 1e-4
    assert torch.equal(runs[0][0], runs[1][0]) and torch.equal(runs[0][1], runs[1][1])


def test_weight_gradient_refuses_other_channels(gpu):
    lib = _native.lib()
    assert lib.vxs_wgrad_scratch(32, 64) == -1
    rc = lib.vxs_conv3x3_wgrad(None, 1, 1, 1, 8, 8, 32, 64, 1, 1, 1)
    assert rc == 1 and b"16 or 64" in lib.vxs_last_error()


def test_hll64_bit_exact_with_oracle(gpu):
    from vexhip import surrogate

    nx = ny = 40
    h, hu, hv = scenes.swe_dam_bumps(nx, seed=3)
    h, hu, hv = (a.astype(np.float64) for a in (h, hu, hv))
    hu = hu + 0.1 * np.sin(np.arange(nx * ny))  # currents both ways
    hll = surrogate.Hll64(nx)
    hll.load(*(torch.from_numpy(a).cuda() for a in (h, hu, hv)))
    dt = hll.dt
    hll.step(12)
    ref = (h, hu, hv)
    for _ in range(12):
        ref = oracle.swe_step_f64(nx, ny, (0, 0, 0, 0), *ref, 9.81, 1.0 / nx, 1.0 / ny, 1e-6, dt)
    for got, exp in zip(hll.a, ref):
        np.testing.assert_array_equal(got.cpu().numpy().view(np.int64), np.asarray(exp).view(np.int64))


def test_device_scene_matches_host(gpu):
    from vexhip import surrogate

    n = 96
    hll = surrogate.Hll64(n)
    hll.scene(7)
    h, _, _ = scenes.swe_dam_bumps(n, seed=7)
    np.testing.assert_allclose(hll.a[0].float().cpu().numpy(), h, rtol=0, atol=1e-6)


def test_training_reduces_the_loss(gpu):
    from vexhip import surrogate

    print("[surrogate] train 60 iterations at 64^2", flush=True)
    model, losses = surrogate.train(n=64, seeds=range(60), iters=60, lr=2e-3)
    assert np.isfinite(losses).all()
    assert np.mean(losses[-10:]) < 0.5 * np.mean(losses[:5])
    result = surrogate.evaluate(model, n=64, seeds=range(10000, 10004))
    assert result["scenes"] == 4 and np.isfinite(result["rel_l2_h_mean"])


def _hll32(n_x, n_y, params, ticks, src, dst, gate=None):
    edges = (ctypes.c_int * 4)(0, 0, 0, 0)
    _native.vxs_check(_native.lib().vxs_hll32(torch.cuda.current_stream().cuda_stream, n_x, n_y, edges,
                                              params.data_ptr(), ticks, None if gate is None else gate.data_ptr(),
                                              *(t.data_ptr() for t in src), *(t.data_ptr() for t in dst)))


def test_hll32_fixed_ticks_matches_oracle(gpu):
    """The fallback's FP32 step (vxs_hll32): eight steps of a fixed number of ticks are the CPU
    restatement's shallow-water steps with the budget capping each at that many ticks, bit for bit;
    with its gate word at zero the step writes nothing."""
    nx, ny, ticks = 48, 40, 7
    h, hu, hv = scenes.swe_dam_bumps(48, seed=5)
    h = h.reshape(48, 48)[:ny].ravel().copy()
    hu = (0.2 * np.sin(np.arange(nx * ny) / 7.0)).astype(np.float32)
    hv = (0.1 * np.cos(np.arange(nx * ny) / 5.0)).astype(np.float32)
    params = oracle.swe_params(9.81, 0.45, 1.0 / nx, 1.0 / ny)
    ref = oracle.SweStepper(nx, ny)
    ref.load(h, hu, hv, params, oracle.swe_fastest_wave(h, hu, hv, 9.81, 1e-6))
    cur = [torch.from_numpy(a.astype(np.float32)).cuda() for a in (h, hu, hv)]
    nxt = [torch.empty_like(t) for t in cur]
    dparams = torch.from_numpy(np.asarray(params, np.int32)).cuda()
    for _ in range(8):
        ref.write_budget(ticks)
        ref.step(1)
        _hll32(nx, ny, dparams, ticks, cur, nxt)
        cur, nxt = nxt, cur
    for got, exp in zip(cur, ref.read()):
        np.testing.assert_array_equal(got.cpu().numpy().view(np.int32), exp.view(np.int32))
    gate = torch.zeros(1, dtype=torch.int32, device="cuda")
    for t in nxt:
        t.fill_(123.0)
    _hll32(nx, ny, dparams, ticks, cur, nxt, gate)
    assert all(bool((t == 123.0).all()) for t in nxt)


def test_fallback_replaces_a_refused_application(gpu):
    """Surrogate.advance under the fallback rule. A network whose depth output is driven negative is
    refused on the device, and the result is eight plain FP32 HLL steps of the surrogate's fixed dt
    from the same state, bit for bit, with the refusal counted. A network that returns the state
    unchanged (no volume drift, no negative depth) is accepted: the result is its own output."""
    from vexhip import surrogate

    n = 64
    model = surrogate.Surrogate(n)
    hll = surrogate.Hll64(n)
    hll.scene(11)
    hll.step(surrogate.STEPS * 2)
    state = hll.state32()
    model.b32[5][0] = -50.0  # the depth channel's change: far below zero everywhere
    got = model.advance(*state)
    assert model.fallbacks() == 1
    params = torch.from_numpy(surrogate.swe_words(n)).cuda()
    cur = [t.clone() for t in state]
    nxt = [torch.empty_like(t) for t in cur]
    for _ in range(surrogate.STEPS):
        _hll32(n, n, params, surrogate.fixed_ticks(n), cur, nxt)
        cur, nxt = nxt, cur
    for g, e in zip(got, cur):
        assert torch.equal(g, e)
    model.b32[5].zero_()
    model.w32[5].zero_()
    model.w16[5].zero_()
    kept = model.advance(*state)
    assert model.fallbacks() == 1
    for g, e in zip(kept, model.forward(*state)):
        assert torch.equal(g, e)
    for g, e in zip(kept, state):
        assert torch.equal(g, e)


def test_trained_surrogate_beats_persistence(gpu):
    """What the network gains: trained at 128^2 for 400 iterations, on held-out scenes its mean
    relative L2 error in depth against the FP64 HLL, one application under the fallback rule, is
    under half that of keeping the state unchanged."""
    from vexhip import surrogate

    print("[surrogate] train 400 iterations at 128^2", flush=True)
    model, losses = surrogate.train(n=128, seeds=range(200), iters=400, lr=2e-3)
    result = surrogate.evaluate(model, n=128, seeds=range(10000, 10020))
    print(f"[surrogate] 128^2: {result}", flush=True)
    assert result["scenes"] == 20
    assert result["rel_l2_h_mean"] < 0.5 * result["persistence_rel_l2_h_mean"], result
