"""The HIP HLL step: the reference's physics tests on the device, and bit-for-bit parity with the
oracle over many steps, grids, edges and tile shapes."""
import sys

import numpy as np
import pytest

import oracle
import swe_cases as cases
from gpu_rigs import GpuSwe
from vexhip import patch, scenes

pytestmark = pytest.mark.gpu


def make(nx, ny, edges=(0, 0, 0, 0)):
    return GpuSwe(nx, ny, edges)


ARGS = (make, oracle.swe_params, oracle.swe_fastest_wave)


def test_ritter_dam_break_converges(gpu):
    coarse, fine = cases.case_ritter_converges(*ARGS)
    print(f"[ritter] GPU: L1 {coarse:.4f} at 200 cells, {fine:.4f} at 400", flush=True)


def test_step_follows_fastest_wave(gpu):
    cases.case_step_follows_fastest_wave(*ARGS)


def test_no_step_passes_the_end(gpu):
    cases.case_no_step_passes_the_end(*ARGS)


def test_closed_box_holds_its_water(gpu):
    drift = cases.case_closed_box_holds_its_water(*ARGS)
    print(f"[box] GPU drift {drift:.2e}", flush=True)


def test_lake_at_rest(gpu):
    cases.case_lake_at_rest(*ARGS)


def _parity(nx, ny, edges, h, hu, hv, steps, courant=0.45, dx=1.0, rows=None, budget=2**31 - 1):
    print(f"[swe parity] {nx}x{ny} edges={edges} steps={steps} rows={rows}", flush=True)
    params = oracle.swe_params(9.81, courant, dx, dx)
    first = oracle.swe_fastest_wave(h, hu, hv, 9.81, 1e-6)
    cpu = oracle.SweStepper(nx, ny, edges)
    cpu.load(h, hu, hv, params, first)
    cpu.write_budget(budget)
    dev = GpuSwe(nx, ny, edges, rows=rows)
    dev.load(h, hu, hv, params, first)
    dev.write_budget(budget)
    for chunk in (1, steps // 2 - 1, steps - steps // 2):
        cpu.step(chunk)
        dev.step(chunk)
        a, b = cpu.read(), dev.read()
        for f in range(3):
            np.testing.assert_array_equal(b[f].view(np.int32), a[f].view(np.int32), err_msg=f"field {f}")
        assert dev.read_budget() == cpu.read_budget()
        assert dev.fastest_wave_bits() == cpu.fastest_wave_bits()
        assert dev.clamped() == cpu.clamped()


def test_bit_exact_hump_walls(gpu):
    h = cases.hump(48)
    z = np.zeros_like(h)
    _parity(48, 48, (0, 0, 0, 0), h, z, z, 60, courant=0.4)


def test_bit_exact_dry_front_1d(gpu):
    h = cases.dam_break(200)
    z = np.zeros_like(h)
    _parity(200, 1, (0, 0, 0, 0), h, z, z, 100, dx=0.5)


@pytest.mark.parametrize("rows", [1, 2, 4, 7, 32, 256])
def test_bit_exact_bumps_mixed_edges(gpu, rows):
    """A ragged grid (not a multiple of the 62-column strip or of the rows a wave marches),
    open and wall edges mixed, currents in both directions, a dry corner."""
    nx, ny = 131, 77
    r = np.random.default_rng(3)
    h = (1 + 0.3 * r.random(nx * ny)).astype(np.float32)
    h[: nx * 5] = 0  # a dry band along the south edge
    hu = (0.2 * r.standard_normal(nx * ny)).astype(np.float32) * (h > 0)
    hv = (0.2 * r.standard_normal(nx * ny)).astype(np.float32) * (h > 0)
    _parity(nx, ny, (1, 0, 0, 1), h, hu, hv, 40, rows=rows)


def test_bit_exact_budget_runs_dry(gpu):
    """The step capped by a small budget, then steps that move nothing (:188-189, :237-240)."""
    h = cases.hump(32)
    z = np.zeros_like(h)
    _parity(32, 32, (0, 0, 0, 0), h, z, z, 30, budget=1500)


def test_bit_exact_dam_and_bumps_512(gpu):
    """The 4096^2 benchmark's shallow-water scene (circular dam plus 32 seeded bumps) at 512^2,
    100 steps, against the oracle bit for bit."""
    h, hu, hv = scenes.swe_dam_bumps(512, seed=0)
    _parity(512, 512, (0, 0, 0, 0), h, hu, hv, 100, courant=0.45, dx=1.0 / 512)


def _fast_within(nx, ny, edges, h, hu, hv, steps, dx, rows=None, rel=1e-5):
    """hll_fast (refined hardware reciprocals and square roots in place of the correctly rounded
    divisions and roots) against the oracle's exact step: every field within rel * max(1, |e|)
    after `steps` steps. The clock may slip: a step's ticks are the floor of a quotient of the
    fastest wave speed, which the fast step gives to within a few ulps, so where that quotient
    lies within a few ulps of a whole tick the two steps round apart by one tick (1e-6 s)."""
    params = oracle.swe_params(9.81, 0.45, dx, dx)
    first = oracle.swe_fastest_wave(h, hu, hv, 9.81, 1e-6)
    cpu = oracle.SweStepper(nx, ny, edges)
    cpu.load(h, hu, hv, params, first)
    cpu.write_budget(2**31 - 1)
    cpu.step(steps)
    dev = GpuSwe(nx, ny, edges, rows=rows)
    dev.sim.set_option("hll_fast", 1)
    dev.load(h, hu, hv, params, first)
    dev.write_budget(2**31 - 1)
    dev.step(steps)
    slip = dev.read_budget() - cpu.read_budget()
    assert abs(slip) <= steps
    worst = 0.0
    for f, (a, b) in enumerate(zip(dev.read(), cpu.read())):
        err = np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(1.0, np.abs(b.astype(np.float64)))
        worst = max(worst, float(err.max()))
        assert float(err.max()) <= rel, f"field {f}: worst {float(err.max()):.3e}"
    print(f"[swe fast] {nx}x{ny} {steps} steps: worst {worst:.2e} of max(1, |e|), clock slip {slip} ticks",
          flush=True)
    return worst


def test_fast_dam_and_bumps_512(gpu):
    """The benchmark's scene at 512^2, 100 steps with hll_fast: within 1e-5 of the reference's bits."""
    h, hu, hv = scenes.swe_dam_bumps(512, seed=0)
    worst = _fast_within(512, 512, (0, 0, 0, 0), h, hu, hv, 100, 1.0 / 512)
    assert worst > 0.0  # the fast step does differ: the option reaches the kernel


@pytest.mark.parametrize("rows", [2, 32])
def test_fast_mixed_edges_and_dry_band(gpu, rows):
    """hll_fast on the ragged grid with open and wall edges and a dry band (reciprocals of dry
    depths discarded), 40 steps."""
    nx, ny = 131, 77
    r = np.random.default_rng(3)
    h = (1 + 0.3 * r.random(nx * ny)).astype(np.float32)
    h[: nx * 5] = 0
    hu = (0.2 * r.standard_normal(nx * ny)).astype(np.float32) * (h > 0)
    hv = (0.2 * r.standard_normal(nx * ny)).astype(np.float32) * (h > 0)
    _fast_within(nx, ny, (1, 0, 0, 1), h, hu, hv, 40, 1.0, rows=rows)


def test_full_size_4096_conserves_and_repeats(gpu):
    """4096^2, 100 steps: closed walls hold the water to 1e-6 relative, nothing is clamped, and a
    second run from the same state is the same bits (no order-dependent reduction)."""
    n = 4096
    h, hu, hv = scenes.swe_dam_bumps(n, seed=0)
    runs = []
    for _ in range(2):
        sim = patch.PatchSimulation(n, n)
        sim.load(h, hu, hv, patch.params(9.81, 0.45, 1.0 / n, 1.0 / n))
        sim.write_budget(2**31 - 1)
        sim.dispatch(100)
        out = sim.read()
        runs.append((out, sim.clamped(), sim.read_budget()))
        sim.close()
    (a, ca, ba), (b, cb, bb) = runs
    for f in range(3):
        np.testing.assert_array_equal(a[f].view(np.int32), b[f].view(np.int32))
    assert ca == 0 and cb == 0 and ba == bb
    before = float(np.sum(h, dtype=np.float64))
    after = float(np.sum(a[0], dtype=np.float64))
    print(f"[swe 4096] mass drift {(after - before) / before:.2e}", flush=True)
    assert abs(after - before) <= 1e-6 * before
    assert np.all(np.isfinite(a[0])) and np.all(a[0] >= 0)


def test_patch_simulation_advance(gpu):
    """PatchSimulation.advance's budget loop: a slice of time is spent exactly, in steps sized by
    the kernel, and time past the step allowance is reported behind (gui/PatchSimulation.java:120-145)."""
    n = 64
    h = cases.hump(n)
    z = np.zeros_like(h)
    sim = patch.PatchSimulation(n, n)
    sim.load(h, z, z, patch.params(9.81, 0.45, 1.0, 1.0))
    a = sim.advance(2_000_000, 10_000)   # two seconds
    assert a.behind_ticks == 0 and a.spent_ticks == 2_000_000 and sim.now() == 2_000_000
    assert a.steps > 10
    b = sim.advance(2_000_000, 3)
    assert b.steps == 3 and b.behind_ticks > 0 and b.spent_ticks + b.behind_ticks == 2_000_000
    assert sim.clamped() == 0
    assert 0 < sim.allowed_step(0.45, 1, 1) < 1
    sim.close()
