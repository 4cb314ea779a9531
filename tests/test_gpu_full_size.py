"""Parity at the configurations' own sizes.

- 3D MLS-MPM, 8,388,608 particles on 256^3: one substep per kernel against the oracle -- every float
  P2G schedule's grid within FP32 reassociation, the fixed-point one bit for bit, the grid update bit
  for bit from the oracle's sums,
  G2P bit for bit (g2p=1) and within 1e-5 (g2p=2) from the oracle's velocities.
- HLL on 4096^2: 100 steps bit for bit (the oracle's step on host threads), and with hll_fast
  within 1e-5.
- 256 batched 2D scenes: 20 steps with graphs and the autotuner; every scene's mass, and the
  largest, the largest-grid and the smallest scene bit for bit against single-scene runs and the
  oracle.
- 2D dam block at 8,192 and 65,536 particles: 1,000 fixed-point substeps bit for bit; the float
  schedules within 1e-5 over 100.
- The surrogate's whole six-layer forward at 256^2 against a torch FP32 reference of the same
  bf16 weights and activations.
Patterns: test/particle/Flip3Test.java:57-118, test/stencil/ShallowWaterTest.java:166-195."""
import os

import numpy as np
import pytest

import oracle
from vexhip import _native, particles, patch, scenes

try:  # before libvexelhip.so loads: the surrogate's buffers are torch's, on one HIP runtime
    import torch
except ImportError:  # the surrogate case skips
    torch = None

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _close(got, exp, what, rel=1e-5):
    err = np.abs(got.astype(np.float64) - exp.astype(np.float64))
    tol = rel * np.maximum(1.0, np.abs(exp.astype(np.float64)))
    bad = err > tol
    assert not bad.any(), f"{what}: {int(bad.sum())} of {bad.size} off, worst {float(err.max()):.3e}"


def _bits_equal(got, exp, what):
    np.testing.assert_array_equal(np.asarray(got).view(np.int32), np.asarray(exp).view(np.int32), err_msg=what)


def test_mpm3_two_cubes_substep_per_kernel(gpu):
    s = scenes.mpm3_two_cubes()
    count = len(s["x"])
    g, bulk, rho0 = scenes.mpm3_units(256)
    words = particles.params3(5e-5, 0, g, 0, bulk, rho0)
    rng = np.random.default_rng(2)
    # a moving, compressed state (the ranges of Flip3Test's scatter case): velocities in [-2, 2],
    # C in [-1, 1], J in [0.8, 1.2]
    u = rng.uniform(-2.0, 2.0, (3, count)).astype(np.float32)
    sim = particles.ParticleSimulation3(256, 256, 256, count)
    sim.load((s["x"], s["y"], s["z"]), (u[0], u[1], u[2]), s["m"], words)
    sim.sim.write("j", rng.uniform(0.8, 1.2, count).astype(np.float32))
    for name in oracle.FIELDS3[8:]:
        sim.sim.write(name, rng.uniform(-1.0, 1.0, count).astype(np.float32))
    print("[full] 3D: sort, then read the sorted state", flush=True)
    sim.sort()
    f = {name: sim.sim.read(name).copy() for name in oracle.FIELDS3}
    ref = oracle.Mpm3(256, 256, 256, f, words)
    ref.clear()
    ref.scatter()
    mass = float(np.sum(f["m"], dtype=np.float64))
    for p2g in (0, 1):
        print(f"[full] 3D: P2G schedule {p2g}", flush=True)
        sim.sim.set_option("p2g", p2g)
        sim.sim.run_pass("clear")
        sim.sim.run_pass("scatter")
        for name in ("gm", "gmu", "gmv", "gmw"):
            _close(sim.sim.read(name), ref.g[name], f"p2g={p2g} {name}")
        assert abs(float(np.sum(sim.sim.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass
    print("[full] 3D: grid update from the oracle's sums", flush=True)
    for name in ("gm", "gmu", "gmv", "gmw"):
        sim.sim.write(name, ref.g[name])
    sim.sim.run_pass("grid")
    ref.grid()
    for name in ("gu", "gv", "gw"):
        _bits_equal(sim.sim.read(name), ref.g[name], f"grid {name}")
    after = oracle.Mpm3(256, 256, 256, f, words)
    after.g = {k: v.copy() for k, v in ref.g.items()}
    after.advect()
    for g2p in (1, 2):
        print(f"[full] 3D: G2P {g2p}", flush=True)
        sim.sim.set_option("g2p", g2p)
        for name in oracle.FIELDS3:
            sim.sim.write(name, f[name])
        for name in ("gu", "gv", "gw"):
            sim.sim.write(name, ref.g[name])
        sim.sim.run_pass("advect")
        for name in oracle.FIELDS3:
            if g2p == 1:
                _bits_equal(sim.sim.read(name), after.p[name], f"g2p=1 {name}")
            else:
                _close(sim.sim.read(name), after.p[name], f"g2p=2 {name}")
    print("[full] 3D: P2G schedule 3 after its own sort", flush=True)
    for name in oracle.FIELDS3:
        sim.sim.write(name, f[name])
    sim.sim.set_option("p2g", 3)
    sim.sort()
    sim.sim.run_pass("clear")
    sim.sim.run_pass("scatter")
    for name in ("gm", "gmu", "gmv", "gmw"):
        _close(sim.sim.read(name), ref.g[name], f"p2g=3 {name}")
    print("[full] 3D: P2G schedule 2 (fixed point) against the oracle's fixed-point scatter", flush=True)
    f3 = {name: sim.sim.read(name).copy() for name in oracle.FIELDS3}
    del ref, after
    fixed = oracle.Mpm3(256, 256, 256, f3, words, fixed=True)
    fixed.scatter()
    sim.sim.set_option("p2g", 2)
    sim.sim.run_pass("clear")
    sim.sim.run_pass("scatter")
    for name in ("gm", "gmu", "gmv", "gmw"):
        _bits_equal(sim.sim.read(name), fixed.g[name], f"p2g=2 {name}")
    sim.close()


def _worst(got, exp):
    err = np.abs(got.astype(np.float64) - exp.astype(np.float64))
    return float(np.max(err / np.maximum(1.0, np.abs(exp.astype(np.float64)))))


def test_hll_4096_hundred_steps_bit_exact(gpu):
    """100 steps: the default step bit for bit, and the step with hll_fast (refined hardware
    reciprocals and square roots) within 1e-5 * max(1, |e|) of the same oracle run, on the same
    clock (its budget word equal to the oracle's)."""
    n = 4096
    h, hu, hv = scenes.swe_dam_bumps(n, seed=0)
    words = patch.params(9.81, 0.45, 1.0 / n, 1.0 / n)
    first = patch.fastest_wave(h, hu, hv, 9.81, 1e-6)
    runs = {}
    for fast in (0, 1):
        print(f"[full] HLL 4096^2 hll_fast={fast}", flush=True)
        sim = patch.PatchSimulation(n, n)
        sim.sim.set_option("hll_fast", fast)
        sim.load(h, hu, hv, words)
        sim.write_budget(2**31 - 1)
        sim.dispatch(100)
        runs[fast] = (sim.read(), int(sim.sim.read_words("speed")[0]), int(sim.sim.read_words("clamped")[0]),
                      sim.read_budget())
        sim.close()
    print(f"[full] HLL 4096^2: the oracle on {THREADS} threads", flush=True)
    ref = oracle.SweStepper(n, n, threads=THREADS)
    ref.load(h, hu, hv, words, first)
    ref.write_budget(2**31 - 1)
    ref.step(100)
    exp = ref.read()
    got, got_speed, got_clamped, _ = runs[0]
    for a, b, name in zip(got, exp, ("h", "hu", "hv")):
        _bits_equal(a, b, name)
    assert got_speed == ref.fastest_wave_bits()
    assert got_clamped == ref.clamped()
    fast, _, fast_clamped, fast_budget = runs[1]
    assert fast_budget == ref.read_budget()
    for a, b, name in zip(fast, exp, ("h", "hu", "hv")):
        print(f"[full] hll_fast {name}: worst {_worst(a, b):.2e} of max(1, |e|)", flush=True)
        _close(a, b, f"hll_fast {name}")
    assert fast_clamped == ref.clamped()


def _single2(n, scene, words, steps, sort_every, p2g=2):
    sim = _native.Sim(_native.KIND_MPM2, (n, n, 1), len(scene["x"]))
    sim.set_option("p2g", p2g)
    count = len(scene["x"])
    z = np.zeros(count, np.float32)
    for name, v in (("x", scene["x"]), ("y", scene["y"]), ("u", z), ("v", z), ("m", scene["m"]),
                    ("j", np.ones(count, np.float32))):
        sim.write(name, v)
    sim.write("params", words)
    sim.advance(steps, sort_every, 0)
    return sim


def test_batch_256_scenes(gpu):
    layouts = scenes.mpm2_batch_layouts(256)
    loaded = [scenes.mpm2_batch_scene(l, 1000 + k) for k, l in enumerate(layouts)]
    sizes = [(l["n"], l["n"]) for l in layouts]
    counts = [l["count"] for l in layouts]
    batch = particles.ParticleBatch(sizes, counts)
    batch.load([x[0] for x in loaded], [x[1] for x in loaded])
    batch.sim.autotune(2)
    batch.sim.set_option("graph", 1)
    print(f"[full] batch: {sum(counts)} particles in 256 scenes, 20 steps", flush=True)
    batch.advance(20, 10)
    for k in range(256):
        grid = batch.scene_grid(k, names=("gm",))["gm"]
        mass = 0.25 * counts[k]
        assert abs(float(np.sum(grid, dtype=np.float64)) - mass) <= 1e-6 * mass, k
    picks = {int(np.argmax(counts)), int(np.argmin(counts)),
             max(range(256), key=lambda k: (sizes[k][0], counts[k]))}
    for k in sorted(picks):
        n = sizes[k][0]
        scene, words = loaded[k]
        print(f"[full] batch scene {k}: {counts[k]} particles on {n}^2", flush=True)
        single = _single2(n, scene, words, 20, 10)
        got = batch.scene(k)
        for name in oracle.FIELDS2:
            _bits_equal(got[name], single.read(name), f"scene {k} vs single: {name}")
        single.close()
        f = dict(scene)
        f["u"] = f["v"] = np.zeros(counts[k], np.float32)
        ref = oracle.Mpm2(n, n, f, words, fixed=True)
        ref.advance(20, 10)
        for name in oracle.FIELDS2:
            _bits_equal(got[name], ref.p[name], f"scene {k} vs oracle: {name}")
    batch.close()


@pytest.mark.parametrize("ppc", [4, 32], ids=["8192", "65536"])
def test_dam_block_2d(gpu, ppc):
    n = 128
    s = scenes.mpm2_dam_block(ppc=ppc)
    count = len(s["x"])
    h = 1.0 / (n - 1)
    words = particles.params2(1e-4, 0, -9.8 / h, 10.0 / h ** 2, 1.0)
    f = dict(s)
    f["u"] = f["v"] = np.zeros(count, np.float32)
    print(f"[full] 2D dam block, {count} particles: 1,000 fixed-point substeps", flush=True)
    dev = _single2(n, s, words, 1000, 10, p2g=2)
    ref = oracle.Mpm2(n, n, f, words, fixed=True)
    ref.advance(1000, 10)
    for name in oracle.FIELDS2:
        _bits_equal(dev.read(name), ref.p[name], f"{name} after 1,000 steps")
    dev.close()
    mass = float(np.sum(s["m"], dtype=np.float64))
    ref = oracle.Mpm2(n, n, f, words)
    ref.advance(100, 10)
    for p2g in (0, 1):
        dev = _single2(n, s, words, 100, 10, p2g=p2g)
        for name in ("x", "y"):
            np.testing.assert_allclose(dev.read(name), ref.p[name], rtol=1e-5, atol=1e-5, err_msg=f"p2g={p2g} {name}")
        assert abs(float(np.sum(dev.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass
        dev.close()


def test_surrogate_forward_256(gpu):
    if torch is None:
        pytest.skip("torch is not importable")
    F = torch.nn.functional
    from vexhip import surrogate

    n = 256
    model = surrogate.Surrogate(n, seed=3)
    gen = torch.Generator().manual_seed(5)
    for b in model.b32:
        b.copy_((torch.rand(b.shape, generator=gen) - 0.5) * 0.2)
    hll = surrogate.Hll64(n)
    hll.scene(21)
    hll.step(40)
    state = hll.state32()
    out = model.forward(*state)
    got = model.delta.float().cpu().reshape(n, n, 16)
    norm = model.norm.cpu()
    st = [t.cpu().reshape(n, n) for t in state]
    x = torch.zeros(1, 16, n, n)
    for c in range(3):
        x[0, c] = ((st[c] - norm[c]) * norm[3 + c]).to(torch.bfloat16).float()
    for li, (cin, cout) in enumerate(surrogate.LAYERS):
        w = model.w16[li].float().cpu().reshape(cout, 3, 3, cin).permute(0, 3, 1, 2).contiguous()
        y = F.conv2d(x, w, model.b32[li].cpu(), padding=1)
        if li < len(surrogate.LAYERS) - 1:
            x = torch.relu(y).to(torch.bfloat16).float()
    ref = y[0].permute(1, 2, 0)
    scale = float(ref.abs().max())
    err = (got - ref).abs()
    print(f"[full] surrogate 256^2: max error {float(err.max()) / scale:.2e}, mean {float(err.mean()) / scale:.2e} "
          f"of the largest output", flush=True)
    assert float(err.max()) <= 2e-2 * scale
    assert float(err.mean()) <= 1e-3 * scale
    for c in range(3):
        expect = st[c] + got[:, :, c] / norm[6 + c]
        assert torch.equal(out[c].cpu().reshape(n, n), expect)
