"""test/particle/TensionTest.java restated for any backend: a round drop of liquid at rest in a
vacuum, no gravity, whose pressure must exceed the vacuum's by sigma / R (Laplace, in two
dimensions)."""
import math

import numpy as np

from vexhip import scenes

N, PPC, SOUND = 64, 4, 200.0  # TensionTest.java:29-31


def _inside(x, y, cx, cy, radius):
    return math.hypot(x - cx, y - cy) <= radius


def drop(radius, seed=1):
    """TensionTest.pressure's drop (:43-62): PPC particles jittered in every cell of the N x N
    grid whose centre lies within `radius` of the middle, x then y drawn per particle with
    nextFloat * 0.999f, mass 1 / PPC. Returns the fields and the number of cells."""
    c = N / 2.0
    cells = [(col, row) for row in range(N) for col in range(N) if _inside(col + 0.5, row + 0.5, c, c, radius)]
    count = len(cells) * PPC
    draws = scenes.JavaRandomFloats(seed).next_floats(2 * count).reshape(count, 2)
    base = np.repeat(np.array(cells, np.float32), PPC, axis=0)
    at = base + draws * np.float32(0.999)
    f = {"x": at[:, 0].copy(), "y": at[:, 1].copy(), "m": np.full(count, np.float32(1.0 / PPC)),
         "u": np.zeros(count, np.float32), "v": np.zeros(count, np.float32), "j": np.ones(count, np.float32)}
    return f, len(cells)


def pressure(make, radius, sigma, seed=1):
    """TensionTest.pressure (:33-99): the drop stepped for 0.5 s at dt = Flip.stableStep(B, 1, 20,
    0.3), sorted every 10 steps; from a third of the way on, every 5 steps, the mean B (1/J - 1)
    of the particles within half the drop's effective radius of the centre. `make(fields,
    params)` returns a backend with advance(count, sort_every) and state() -> x, y, u, v, j.
    Returns (mean pressure, sigma / R, R, the fastest particle)."""
    bulk = SOUND * SOUND
    dt = 0.3 / (math.sqrt(bulk / 1.0) + 20)  # Flip.stableStep (Flip.java:214-216)
    steps = math.ceil(0.5 / dt)
    f, cells = drop(radius, seed)
    effective = math.sqrt(cells / math.pi)
    params = np.zeros(28, np.int32)
    params[:6] = np.array([dt, 0, 0, bulk, 1, sigma], np.float64).astype(np.float32).view(np.int32)
    sim = make(f, params)
    c = N / 2.0
    total, samples, fastest = 0.0, 0, 0.0
    for s in range(steps):
        sim.advance(1, 10)
        if s >= steps // 3 and s % 5 == 0:
            x, y, u, v, j = (np.asarray(a, np.float64) for a in sim.state())
            fastest = max(fastest, float(np.hypot(u, v).max()))
            inner = np.hypot(x - c, y - c) < 0.5 * effective
            total += float(np.mean(bulk * (1 / j[inner] - 1)))
            samples += 1
    return total / samples, sigma / effective, effective, fastest


def judge(make, label):
    """TensionTest.aDropsPressureIsSigmaOverItsRadius (:193-210), its bounds unchanged."""
    sigma = 0.02 * SOUND * SOUND * 12
    none = pressure(make, 12, 0)
    one = pressure(make, 12, sigma)
    two = pressure(make, 12, 2 * sigma)
    wide = pressure(make, 18, sigma)
    print(f"[tension2] {label}: no tension p {none[0]:.0f}; sigma/R = {one[1]:.0f} gives p {one[0]:.0f} "
          f"({one[0] / one[1]:.2f} of it, fastest {one[3]:.1f}); twice gives {two[0]:.0f}; R {wide[2]:.1f} gives "
          f"{wide[0]:.0f} against {wide[1]:.0f}", flush=True)
    assert abs(none[0]) < 0.15 * one[1], f"a drop with no tension has pressure {none[0]}"
    ratio = (one[0] - none[0]) / one[1]
    assert 0.4 < ratio < 1.1, f"the pressure is {ratio} of sigma/R"
    doubled = (two[0] - none[0]) / (one[0] - none[0])
    assert 1.6 < doubled < 2.4, f"twice the tension gives {doubled} times the pressure"
    wider = (wide[0] - none[0]) / (one[0] - none[0])
    assert 0.45 < wider < 0.85, f"a drop of 1.5 times the radius has {wider} of the pressure, against 2/3"
    return none, one, two, wide
