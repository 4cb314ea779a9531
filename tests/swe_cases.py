"""The reference's shallow-water physics tests, written once over a backend.

Restates test/stencil/Stepper.java (the harness: set, step, runUntil, ticks) and the cases of
test/stencil/ShallowWaterTest.java. A backend supplies load / dispatch / write_budget /
read_budget / read / clamped; the oracle's ``SweStepper`` is one, the HIP runner another.
Each backend is judged against the physics on its own, as the reference judges its CPU and GPU.
"""
import math

import numpy as np

from vexhip import clock as vclock

G = 9.81
COURANT = 0.45
FOREVER = 1e9
LENGTH = 100.0
DAM = 50.0
H0 = 1.0
T = 5.0
DRY = 1e-6


class Harness:
    """Stepper.java's host side: a clock, budgets granted from it, steps dispatched."""

    def __init__(self, backend):
        self.b = backend
        self.clock = vclock.Clock()
        self.end = 0

    def set(self, h, hu, hv, params, end, fastest_bits):
        self.b.load(h, hu, hv, params, fastest_bits)
        self.clock = vclock.Clock()
        self.end = vclock.ticks(end)
        self.b.write_budget(self.clock.grant(self.end))

    def step(self, n):
        self.b.step(n)

    def ticks(self):
        return self.clock.now() + self.clock.outstanding() - self.b.read_budget()

    def time(self):
        return vclock.seconds(self.ticks())

    def run_until(self, batch, limit):
        while self.ticks() < self.end:
            if self.b.read_budget() == 0:
                self.clock.settle(0)
                self.b.write_budget(self.clock.grant(self.end))
            if self.b.dispatched >= limit:
                raise AssertionError(f"the clock is at {self.time()} after {limit} steps")
            self.step(batch)


def ritter(x, t):
    c0 = math.sqrt(G * H0)
    xi = (x - DAM) / t
    if xi <= -c0:
        return H0
    if xi >= 2 * c0:
        return 0.0
    root = 2 * c0 - xi
    return root * root / (9 * G)


def dam_break(cells):
    dx = LENGTH / cells
    centres = (np.arange(cells) + 0.5) * dx
    return np.where(centres < DAM, np.float32(H0), np.float32(0)).astype(np.float32)


def ritter_error(make, params_fn, fastest_fn, cells):
    """ShallowWaterTest.ritterError (:380-402)."""
    dx = LENGTH / cells
    hz = np.zeros(cells, np.float32)
    h = dam_break(cells)
    params = params_fn(G, COURANT, dx, dx, DRY)
    s = Harness(make(cells, 1))
    s.set(h, hz, hz, params, T, fastest_fn(h, hz, hz, G, DRY))
    s.run_until(32, 100_000)
    assert s.ticks() == vclock.ticks(T), "the clock is an integer, so it lands on the end exactly"
    assert s.b.clamped() == 0, "depths were clamped, so the front created water"
    depth = s.b.read()[0]
    assert np.all(np.isfinite(depth)) and np.all(depth >= 0)
    exact = np.array([ritter((k + 0.5) * dx, T) for k in range(cells)])
    return float(np.sum(np.abs(depth.astype(np.float64) - exact)) * dx / (np.sum(exact) * dx))


def case_ritter_converges(make, params_fn, fastest_fn):
    coarse = ritter_error(make, params_fn, fastest_fn, 200)
    fine = ritter_error(make, params_fn, fastest_fn, 400)
    assert fine < 0.03, fine
    assert fine < coarse * 0.8, (coarse, fine)
    return coarse, fine


def case_step_follows_fastest_wave(make, params_fn, fastest_fn):
    """ShallowWaterTest.theStepFollowsTheFastestWave (:424-446)."""
    cells = 200
    dx = LENGTH / cells
    c0 = math.sqrt(G * H0)
    h = dam_break(cells)
    hz = np.zeros(cells, np.float32)
    s = Harness(make(cells, 1))
    s.set(h, hz, hz, params_fn(G, COURANT, dx, dx, DRY), FOREVER, fastest_fn(h, hz, hz, G, DRY))
    s.step(1)
    first = s.time()
    assert abs(first - COURANT * dx / c0) <= 1.5e-6, first
    s.step(99)
    before = s.time()
    s.step(10)
    later = (s.time() - before) / 10
    assert first * 0.4 < later < first * 0.7, (first, later)


def case_no_step_passes_the_end(make, params_fn, fastest_fn):
    """ShallowWaterTest.noStepPassesTheEnd (:449-469)."""
    cells = 200
    dx = LENGTH / cells
    h = dam_break(cells)
    hz = np.zeros(cells, np.float32)
    s = Harness(make(cells, 1))
    s.set(h, hz, hz, params_fn(G, COURANT, dx, dx, DRY), 1.0, fastest_fn(h, hz, hz, G, DRY))
    s.run_until(16, 10_000)
    at_end = s.b.read()
    t = s.ticks()
    assert t == vclock.ticks(1.0)
    s.step(50)
    assert s.ticks() == t, "the clock moved past the end"
    after = s.b.read()
    for f in range(3):
        np.testing.assert_array_equal(at_end[f], after[f])


def hump(n):
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    rx = x - n * 0.35
    ry = y - n * 0.6
    return (1 + 0.5 * np.exp(-(rx * rx + ry * ry) / 20)).astype(np.float32).ravel()


def case_closed_box_holds_its_water(make, params_fn, fastest_fn):
    """ShallowWaterTest.aClosedBoxHoldsItsWater (:478-507)."""
    n = 48
    h = hump(n)
    z = np.zeros(n * n, np.float32)
    s = Harness(make(n, n))
    s.set(h, z, z, params_fn(G, 0.4, 1, 1, DRY), FOREVER, fastest_fn(h, z, z, G, DRY))
    s.step(300)
    state = s.b.read()
    before = float(np.sum(h.astype(np.float64)))
    after = float(np.sum(state[0].astype(np.float64)))
    for f in state:
        assert np.all(np.isfinite(f))
    assert abs(after - before) <= before * 1e-5, (before, after)
    assert s.b.clamped() == 0
    return (after - before) / before


def case_lake_at_rest(make, params_fn, fastest_fn):
    """ShallowWaterTest.aLakeAtRestStaysAtRest (:518-540)."""
    n = 32
    h = np.ones(n * n, np.float32)
    z = np.zeros(n * n, np.float32)
    s = Harness(make(n, n))
    s.set(h, z, z, params_fn(G, 0.4, 1, 1, DRY), FOREVER, fastest_fn(h, z, z, G, DRY))
    for steps in (100, 900):
        s.step(steps)
        state = s.b.read()
        assert np.max(np.abs(state[0].astype(np.float64) - 1)) < 1e-6
        assert max(np.max(np.abs(state[1])), np.max(np.abs(state[2]))) < 1e-6
