"""2D surface tension on the device (core/particle/Tension.java, FlipStep(..., tension)'s passes in
core/particle/FlipStep.java:111-143): every tension pass bit for bit with the oracle from the same
grid, whole tension steps bit for bit with the fixed-point P2G (eager and replayed as graphs), and
the reference's own physics test (test/particle/TensionTest.java:193-210): a round drop's
Laplace pressure sigma / R."""
import numpy as np
import pytest

import oracle
import tension2_cases as cases
from vexhip import _native, particles

pytestmark = pytest.mark.gpu

PASSES = [f"blur{k}" for k in range(oracle.TENSION_BLURS)] + ["stress", "grid"]
TENSION_BUFFERS = ("c0", "c1") + oracle.STRESS2


def _blob(n, radius, seed):
    """A disc of particles, 4 a cell, with random velocities and J near one."""
    rng = np.random.default_rng(seed)
    c = np.array(n, np.float64) / 2
    cells = [(x, y) for y in range(n[1]) for x in range(n[0]) if np.hypot(x + 0.5 - c[0], y + 0.5 - c[1]) <= radius]
    base = np.repeat(np.array(cells, np.float32), 4, axis=0)
    at = base + rng.uniform(0, 0.999, base.shape).astype(np.float32)
    count = len(at)
    return {"x": at[:, 0].copy(), "y": at[:, 1].copy(), "m": np.full(count, np.float32(0.25)),
            "u": rng.standard_normal(count).astype(np.float32), "v": rng.standard_normal(count).astype(np.float32),
            "j": rng.uniform(0.9, 1.1, count).astype(np.float32)}


def _sim(n, count, p2g):
    sim = _native.Sim(_native.KIND_MPM2, (n[0], n[1], 1), count, _native.FLAG_TENSION)
    sim.set_option("p2g", p2g)
    return sim


def _load(sim, f, params):
    for name in oracle.FIELDS2:
        if name in f:
            sim.write(name, f[name])
        elif name == "j":
            sim.write(name, np.ones(len(f["x"]), np.float32))
        else:
            sim.write(name, np.zeros(len(f["x"]), np.float32))
    sim.write("params", params)


@pytest.mark.parametrize("n", [(24, 20), (33, 17)], ids=["even", "ragged"])
def test_tension_passes_bit_exact(gpu, n):
    """From the oracle's own scatter, written into the device's grid: the six blurs, the stress
    and the grid with tension, each compared with the oracle's pass bit for bit."""
    print(f"[tension2 passes] {n}", flush=True)
    f = _blob(n, min(n) / 3, 7)
    params = oracle.mpm2_params(1e-3, 0.0, -9.8, 400.0, 1.0, 30.0)
    ref = oracle.Mpm2(*n, f, params, tension=True)
    ref.clear()
    ref.scatter()
    sim = _sim(n, len(f["x"]), 1)
    _load(sim, f, params)
    for g in ("gm", "gmu", "gmv"):
        sim.write(g, ref.g[g])
    for p in PASSES:
        if p.startswith("blur"):
            ref.blur(int(p[4:]))
        elif p == "stress":
            ref.stress()
        else:
            ref.grid()
        sim.run_pass(p)
        for name in TENSION_BUFFERS + ("gu", "gv"):
            np.testing.assert_array_equal(sim.read(name).view(np.int32), ref.g[name].view(np.int32),
                                          err_msg=f"{p}: {name}")
    assert np.abs(ref.g["txx"]).max() > 0 and np.abs(ref.g["gu"]).max() > 0
    sim.close()


def _compare_state(sim, ref, what):
    for name in oracle.FIELDS2:
        np.testing.assert_array_equal(sim.read(name).view(np.int32), ref.p[name].view(np.int32),
                                      err_msg=f"{what}: {name}")
    for name in ("gm", "gu", "gv") + TENSION_BUFFERS:
        np.testing.assert_array_equal(sim.read(name).view(np.int32), ref.g[name].view(np.int32),
                                      err_msg=f"{what}: {name}")


@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
def test_fixed_point_tension_steps_bit_exact(gpu, graph):
    """Whole tension steps with the fixed-point P2G (its sums decoded before the first blur),
    sorted every 10, against the oracle's: every particle field, the grid and the tension
    buffers bit for bit after 25 steps; the same when the steps replay as captured graphs."""
    n = (40, 36)
    print(f"[tension2 steps] {n} graph={graph}", flush=True)
    f = _blob(n, 11, 9)
    params = oracle.mpm2_params(5e-4, 0.0, -9.8, 400.0, 1.0, 40.0)
    ref = oracle.Mpm2(*n, f, params, fixed=True, tension=True)
    sim = _sim(n, len(f["x"]), 2)
    sim.set_option("graph", graph)
    _load(sim, f, params)
    done = 0
    for chunk in (1, 9, 15):
        ref.advance(chunk, 10)
        sim.advance(chunk, 10, done)
        done += chunk
        _compare_state(sim, ref, f"after {done} steps")
    assert np.abs(ref.g["txx"]).max() > 0
    sim.close()


def test_float_tension_steps_track_the_oracle(gpu):
    """With the float P2G (sums in another order than the oracle's) 20 steps stay within 1e-4 of
    the oracle's positions and velocities, and the mass on the grid is the particles' to 1e-6."""
    n = (40, 36)
    f = _blob(n, 11, 4)
    params = oracle.mpm2_params(5e-4, 0.0, -9.8, 400.0, 1.0, 40.0)
    ref = oracle.Mpm2(*n, f, params, tension=True)
    sim = _sim(n, len(f["x"]), 1)
    _load(sim, f, params)
    ref.advance(20, 10)
    sim.advance(20, 10, 0)
    for name in ("x", "y", "u", "v"):
        np.testing.assert_allclose(sim.read(name), ref.p[name], rtol=1e-4, atol=1e-4, err_msg=name)
    mass = float(np.sum(f["m"], dtype=np.float64))
    assert abs(float(np.sum(sim.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass
    sim.close()


class _Gpu:
    p2g = 1

    def __init__(self, f, params):
        count = len(f["x"])
        self.sim = particles.ParticleSimulation(cases.N, cases.N, count, tension=True)
        self.sim.sim.set_option("p2g", self.p2g)
        self.sim.load(f["x"], f["y"], f["u"], f["v"], f["m"], params)

    def advance(self, count, sort_every):
        self.sim.advance(count, sort_every)

    def state(self):
        return self.sim.positions() + self.sim.velocities() + [self.sim.compression()]


class _GpuFixed(_Gpu):
    p2g = 2


@pytest.mark.parametrize("backend", [_Gpu, _GpuFixed], ids=["float", "fixed"])
def test_a_drops_pressure_is_sigma_over_its_radius(gpu, backend):
    """TensionTest.aDropsPressureIsSigmaOverItsRadius on the device, its bounds unchanged."""
    cases.judge(backend, f"GPU p2g={backend.p2g}")
