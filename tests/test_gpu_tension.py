"""3D surface tension on the device (core/particle/Tension3.java, Flip3Step's tension step in
core/particle/Flip3Step.java:88-109): every tension pass bit for bit with the oracle from the same
grid, whole tension steps within FP32 reassociation of P2G, and the reference's own physics tests
(test/particle/Tension3Test.java): a round drop's Laplace pressure 2 sigma / R, and a drop on the
floor that comes to rest."""
import math

import numpy as np
import pytest

import oracle
from gpu_rigs import maker3
from vexhip import particles, scenes

pytestmark = pytest.mark.gpu

PASSES = [f"blur{k}" for k in range(oracle.TENSION_BLURS)] + ["stress", "grid"]


def _blob(n, lo, hi, ppc, seed):
    x, y, z = scenes.jittered_block3(lo, tuple(h - l for l, h in zip(lo, hi)), ppc, scenes.JavaRandomFloats(seed))
    count = len(x)
    rng = np.random.default_rng(seed)
    f = {"x": x, "y": y, "z": z, "m": np.full(count, np.float32(1.0 / ppc)),
         "u": rng.standard_normal(count).astype(np.float32), "v": rng.standard_normal(count).astype(np.float32),
         "w": rng.standard_normal(count).astype(np.float32),
         "j": rng.uniform(0.9, 1.1, count).astype(np.float32)}
    return f


def _load(rig, f, params):
    for name in oracle.FIELDS3:
        if name in f:
            rig.write(name, f[name])
    rig.params(params)


@pytest.mark.parametrize("shape", [(24, 20, 18), (33, 17, 29)], ids=["even", "odd"])
def test_tension_passes_bit_exact(gpu, shape):
    """From the oracle's own scatter: the six blurs, the stress and the grid with tension, then
    G2P, every buffer the oracle's bit for bit (walls, ring and a drop touching neither)."""
    n = shape
    f = _blob(n, (3, 2, 4), (n[0] - 6, n[1] - 2, n[2] - 9), 8, 31)
    params = oracle.mpm3_params(1e-3, 0.25, -40.0, 0.5, 400.0, 1.0, 35.0)
    ref = oracle.Mpm3(*n, f, params, tension=True)
    ref.clear()
    ref.scatter()
    rig = maker3(p2g=1, g2p=1, tension=True)(*n, len(f["x"]))
    _load(rig, f, params)
    for g in ("gm", "gmu", "gmv", "gmw"):
        rig.write(g, ref.g[g])
    for p in PASSES:
        rig.run([p])
        getattr(ref, "blur" if p.startswith("blur") else p)(*([int(p[4])] if p.startswith("blur") else []))
        if p == "stress":
            names = oracle.STRESS3
        elif p == "grid":
            names = ("gu", "gv", "gw")
        else:
            names = ("c0" if int(p[-1]) % 2 == 0 else "c1",)
        for name in names:
            np.testing.assert_array_equal(rig.read(name).view(np.int32), ref.g[name].view(np.int32),
                                          err_msg=f"{p}: {name}")
    assert np.abs(ref.g["txy"]).max() > 0 and ref.g["c1"].max() > 0.3
    rig.run(["advect"])
    ref.advect()
    for name in oracle.FIELDS3:
        np.testing.assert_array_equal(rig.read(name).view(np.int32), ref.p[name].view(np.int32), err_msg=name)


def _tension_passes_match(rig, n, params, passes):
    """Hands the device's grid sums to a fresh oracle, runs `passes` on both, and compares every
    tension buffer and velocity over the whole grid bit for bit (so a node the device skipped
    must hold what the oracle's full pass wrote there)."""
    ref = oracle.Mpm3(*n, {"x": np.zeros(1, np.float32), "y": np.zeros(1, np.float32),
                           "z": np.zeros(1, np.float32), "m": np.zeros(1, np.float32)}, params, tension=True)
    for g in ("gm", "gmu", "gmv", "gmw"):
        ref.g[g] = rig.read(g)
    for p in passes:
        rig.run([p])
        if p.startswith("blur"):
            ref.blur(int(p[4]))
        else:
            getattr(ref, p)()
    names = ["c0", "c1"] + list(oracle.STRESS3) + (["gu", "gv", "gw"] if "grid" in passes else [])
    for name in names:
        got, exp = rig.read(name).view(np.int32), ref.g[name].view(np.int32)
        bad = np.flatnonzero(got != exp)
        assert bad.size == 0, f"{name}: {bad.size} nodes differ, first {bad[:5]}"
    return ref


def test_sparse_region_equals_the_full_passes(gpu):
    """A drop thrown across the box, stepped with sorts: every few steps the six blurs, the stress
    and the grid run on the device's own scatter, and every node of every tension buffer and
    velocity equals the oracle's full-grid passes, bit for bit -- including the bricks the region
    skipped and the ones the drop left behind."""
    n = (48, 40, 44)
    f = _blob(n, (6, 14, 8), (18, 26, 20), 8, 77)
    f["u"] = np.full_like(f["u"], 120.0)
    f["v"] = np.full_like(f["v"], 40.0)
    f["w"] = np.full_like(f["w"], 90.0)
    f["j"] = np.ones_like(f["j"])
    params = oracle.mpm3_params(2e-3, 0, -60.0, 0, 400.0, 1.0, 30.0)
    rig = maker3(tension=True)(*n, len(f["x"]))
    _load(rig, f, params)
    for k in range(60):
        if k % 10 == 0:
            rig.run(["sort"])
        if k % 12 == 5:
            rig.run(["clear", "scatter"])
            _tension_passes_match(rig, n, params, PASSES)
            rig.run(["advect"])
        else:
            rig.run("step")
    x = rig.read("x")
    assert x.max() - f["x"].max() > 8  # it crossed bricks


def test_passes_out_of_order_fall_back_to_dense(gpu):
    """Blurs run past the reach of the region (the colour spreading further than a brick), a
    stress of that colour, then ordinary steps again: each pass equals the oracle's over the whole
    grid, the ones past the reach run dense and the region is rebuilt after."""
    n = (40, 36, 40)
    f = _blob(n, (12, 10, 14), (22, 20, 24), 8, 5)
    params = oracle.mpm3_params(1e-3, 0, -30.0, 0, 300.0, 1.0, 20.0)
    rig = maker3(tension=True)(*n, len(f["x"]))
    _load(rig, f, params)
    rig.run("step")
    rig.run(["clear", "scatter"])
    _tension_passes_match(rig, n, params, ["blur0"] + ["blur1", "blur2"] * 6 + ["blur1", "stress", "grid"])
    rig.run(["advect"])
    for _ in range(3):
        rig.run(["clear", "scatter"])
        _tension_passes_match(rig, n, params, PASSES)
        rig.run(["advect"])


def test_tension_steps_track_the_oracle(gpu):
    """Ten whole tension steps of a falling drop: positions within 1e-5 of the oracle's (P2G sums
    in another order), grid mass equal to particle mass."""
    n = (28, 28, 28)
    f = _blob(n, (8, 9, 7), (18, 19, 17), 8, 12)
    for k in "uvw":
        f[k] = np.zeros_like(f[k])
    f["j"] = np.ones_like(f["j"])
    params = oracle.mpm3_params(5e-4, 0, -100.0, 0, 2000.0, 1.0, 200.0)
    ref = oracle.Mpm3(*n, f, params, tension=True)
    rig = maker3(tension=True)(*n, len(f["x"]))
    _load(rig, f, params)
    for k in range(10):
        if k % 5 == 0:
            ref.sort()
            rig.run(["sort"])
        ref.step()
        rig.run("step")
    for name in ("x", "y", "z"):
        np.testing.assert_allclose(rig.read(name), ref.p[name], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(rig.read("j"), ref.p["j"], rtol=1e-4, atol=1e-5)
    for name in ("u", "v", "w"):
        np.testing.assert_allclose(rig.read(name), ref.p[name], rtol=1e-3, atol=1e-3)
    mass = float(np.sum(f["m"], dtype=np.float64))
    assert abs(float(np.sum(rig.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass


def test_tension_without_the_flag_is_refused(gpu):
    sim = particles.ParticleSimulation3(16, 16, 16, 64)
    with pytest.raises(Exception):
        sim.sim.run_pass("blur0")
    with pytest.raises(Exception):
        particles.ParticleSimulation(16, 16, 64, heat=True)
    plain = particles.ParticleSimulation(16, 16, 64)
    with pytest.raises(Exception):
        plain.sim.run_pass("stress")


# ------------------------------------------------------------------ Tension3Test

N, PPC, SOUND = 40, 8, 200.0


def _inside(col, row, layer, centre, radius):
    return math.sqrt((col + 0.5 - centre) ** 2 + (row + 0.5 - centre) ** 2 + (layer + 0.5 - centre) ** 2) <= radius


def _pressure(radius, sigma):
    """Tension3Test.pressure (:31-115): a drop at rest, no gravity; the mean B(1/J - 1) of its
    inner half over the later steps, 2 sigma / R, R, the centre of mass's speed, the fastest
    particle."""
    bulk = SOUND * SOUND
    dt = particles.stable_step(bulk, 1, 20, 0.3)
    steps = math.ceil(0.4 / dt)
    centre = N / 2.0
    cells = [(c, r, l) for l in range(N) for r in range(N) for c in range(N) if _inside(c, r, l, centre, radius)]
    count = len(cells) * PPC
    draws = scenes.JavaRandomFloats(5).next_floats(3 * count).reshape(count, 3)
    base = np.repeat(np.array(cells, np.float32), PPC, axis=0)
    at = base + draws * np.float32(0.999)
    m = np.full(count, np.float32(1.0 / PPC))
    effective = (3 * len(cells) / (4 * math.pi)) ** (1.0 / 3.0)
    sim = particles.ParticleSimulation3(N, N, N, count, tension=True)
    sim.load([at[:, 0], at[:, 1], at[:, 2]], [np.zeros(count, np.float32)] * 3, m,
             particles.params3(dt, 0, 0, 0, bulk, 1, sigma))
    total, samples, fastest, drift = 0.0, 0, 0.0, 0.0
    s = 0
    while s < steps:
        sim.advance(1, sort_every=25)
        if s >= steps // 3 and s % 5 == 0:
            pos = np.stack(sim.positions()).astype(np.float64)
            vel = np.stack(sim.velocities()).astype(np.float64)
            j = sim.compression().astype(np.float64)
            fastest = max(fastest, float(np.sqrt((vel ** 2).sum(0)).max()))
            inner = np.sqrt(((pos - centre) ** 2).sum(0)) < 0.5 * effective
            total += float(np.mean(bulk * (1 / j[inner] - 1)))
            mean = vel.sum(1)
            drift = max(drift, float(np.sqrt(mean @ mean)) / count)
            samples += 1
        s += 1
    return total / samples, 2 * sigma / effective, effective, drift, fastest


def test_a_drops_pressure_is_twice_sigma_over_its_radius(gpu):
    """Tension3Test.aDropsPressureIsTwiceSigmaOverItsRadius (:127-148), its bounds unchanged."""
    radius = 8.0
    sigma = 0.01 * SOUND * SOUND * radius
    none = _pressure(radius, 0)
    one = _pressure(radius, sigma)
    two = _pressure(radius, 2 * sigma)
    wide = _pressure(1.5 * radius, sigma)
    print(f"[tension3] GPU: no tension p {none[0]:.0f}; 2 sigma/R = {one[1]:.0f} gives p {one[0]:.0f} "
          f"({one[0] / one[1]:.2f} of it, centre of mass at {one[3]:.3f}, fastest {one[4]:.1f}); twice gives "
          f"{two[0]:.0f}; R {wide[2]:.1f} gives {wide[0]:.0f} against {wide[1]:.0f}", flush=True)
    assert abs(none[0]) < 0.15 * one[1]
    ratio = (one[0] - none[0]) / one[1]
    assert 0.3 < ratio < 1.1, ratio
    doubled = (two[0] - none[0]) / (one[0] - none[0])
    assert 1.6 < doubled < 2.4, doubled
    wider = (wide[0] - none[0]) / (one[0] - none[0])
    assert 0.45 < wider < 0.9, wider
    assert one[3] < 0.5, one[3]


def test_a_drop_on_the_floor_comes_to_rest(gpu):
    """Tension3Test.aDropOnTheFloorComesToRest (:159-233): a column falls and gathers into one
    drop on the floor; after four seconds it neither turns over nor keeps its energy."""
    n = 40
    g = 9.81 * (n - 1)
    fall = math.sqrt(2 * g * 30)
    bulk = 25 * fall * fall
    capillary = 10.0 * (n - 3) / 45
    dt = particles.stable_step(bulk, 1, 2.5 * fall, 0.2)
    count = 16 * 30 * 20 * PPC
    draws = scenes.JavaRandomFloats(1).next_floats(3 * count).reshape(count, 3)
    cells = np.array([(c, r, l) for l in range(12, 32) for r in range(30) for c in range(16)], np.float32)
    cell = np.repeat(cells, PPC, axis=0)
    s = np.tile(np.arange(PPC), len(cells))
    bits = np.stack([(s >> a) & 1 for a in range(3)], 1).astype(np.float32)
    at = np.float32(1.0) + cell + ((bits + np.float32(0.25)) + np.float32(0.5) * draws) / np.float32(2)
    sim = particles.ParticleSimulation3(n, n, n, count, tension=True)
    sim.load([at[:, 0], at[:, 1], at[:, 2]], [np.zeros(count, np.float32)] * 3,
             np.full(count, np.float32(1.0 / PPC)), particles.params3(dt, 0, -g, 0, bulk, 1, g * capillary ** 2))
    settle = math.ceil(4 / dt)
    steps = math.ceil(6 / dt)
    rise, energy, samples = 0.0, 0.0, 0
    done = 0
    # the reference samples after step s (from 0) when s >= settle and s % 500 == 0; the steps
    # after the last sample change nothing it measures
    for s in (s for s in range(steps) if s >= settle and s % 500 == 0):
        sim.advance(s + 1 - done, sort_every=20)
        done = s + 1
        pos = np.stack(sim.positions()).astype(np.float64)
        vel = np.stack(sim.velocities()).astype(np.float64)
        cx, cz = pos[0].mean(), pos[2].mean()
        energy += float(0.5 * (vel ** 2).sum() / count)
        middle = np.hypot(pos[0] - cx, pos[2] - cz) < 3
        rise += float(vel[1][middle].sum()) / max(int(middle.sum()), 1)
        samples += 1
    rise /= samples
    energy /= samples
    print(f"[tension3] GPU: a settled drop rises up its middle at {rise:.2f}, kinetic energy {energy:.1f}, "
          f"{steps} steps", flush=True)
    assert abs(rise) < 5, rise
    assert energy < 120, energy


def test_tension_steps_replayed_as_graphs(gpu):
    """vxr_advance with graph replay (option graph = 1) on a tension simulation: the same steps as
    eager advancing, within P2G's float-atomic reassociation; the first segment after the load
    runs eagerly (the region is rebuilt), the rest replay captured segments."""
    n = 32
    f = _blob((n, n, n), (10, 11, 9), (20, 21, 19), 8, 44)
    count = len(f["x"])
    words = particles.params3(5e-4, 0, -100.0, 0, 2000.0, 1.0, 200.0)
    runs = []
    for graph in (0, 1):
        sim = particles.ParticleSimulation3(n, n, n, count, tension=True)
        zero = np.zeros(count, np.float32)
        sim.load((f["x"], f["y"], f["z"]), (zero, zero, zero), f["m"], words)
        sim.sim.set_option("graph", graph)
        sim.advance(25, sort_every=10)
        sim.advance(15, sort_every=10)
        slots = sim.slots()
        pos = []
        for v in sim.positions():
            out = np.empty_like(v)
            out[slots] = v
            pos.append(out)
        runs.append(pos)
        sim.close()
    for a, b in zip(*runs):
        np.testing.assert_allclose(b, a, rtol=1e-5, atol=2e-5)
