"""The oracle's 2D surface tension (core/particle/Tension.java:98-185, Flip.gridWithTension in
core/particle/Flip.java:342-412) pinned by what the mathematics fixes: the blur keeps a constant
and spreads an impulse as [1 2 1]/4 along each axis; a node of rest mass reads as colour one and
more is held to one; the stress of a linear colour is the closed form (c_y^2, c_x^2, -c_x c_y)/|g|
with each component in its place; a uniform stress gives no force anywhere; the face fluxes
cancel pairwise, so the force of a drop sums to zero over the grid; with sigma = 0 the tension
grid is the plain grid; and the reference's own physics test, a drop's Laplace pressure sigma/R
(test/particle/TensionTest.java:193-210), on the oracle's whole tension steps."""
import numpy as np
import pytest

import oracle
import tension2_cases as cases

N = (12, 9)


def _rig(fields=None, sigma=0.0, rho0=1.0, dt=1e-3, n=N, fixed=False):
    count = 4
    f = fields or {"x": np.full(count, 5.5, np.float32), "y": np.full(count, 4.5, np.float32),
                   "m": np.full(count, 0.25, np.float32)}
    return oracle.Mpm2(*n, f, oracle.mpm2_params(dt, 0.0, -9.0, 100.0, rho0, sigma), fixed=fixed, tension=True)


def _field(n=N):
    y, x = np.meshgrid(np.arange(n[1]), np.arange(n[0]), indexing="ij")
    return x.ravel(), y.ravel()


@pytest.mark.parametrize("value", [1.0, 0.5, 0.375])
def test_blur_keeps_a_constant(value):
    r = _rig()
    r.g["c0"][:] = np.float32(value)
    r.blur(1)  # unscaled: c0 -> c1
    np.testing.assert_array_equal(r.g["c1"], np.float32(value))


def test_blur_weights_are_one_two_one():
    """A unit impulse away from the walls spreads as [1 2 1]/4 along each axis, in place."""
    r = _rig()
    at = 4 * N[0] + 6
    r.g["c0"][:] = 0
    r.g["c0"][at] = 1.0
    r.blur(1)
    w = np.array([0.25, 0.5, 0.25])
    exp = np.zeros(N[::-1])
    exp[3:6, 5:8] = w[:, None] * w[None, :]
    np.testing.assert_array_equal(r.g["c1"].reshape(N[::-1]), exp.astype(np.float32))
    assert float(r.g["c1"].sum(dtype=np.float64)) == 1.0


@pytest.mark.parametrize("rho0", [1.0, 2.5])
def test_rest_mass_reads_as_colour_one(rho0):
    """gm = rho0 on every node blurs to exactly one (the 2D blur has no wall correction: the held
    reads repeat the wall ring); three times that is held to one; half of it reads a half."""
    r = _rig(rho0=rho0)
    r.g["gm"][:] = np.float32(rho0)
    r.blur(0)
    np.testing.assert_array_equal(r.g["c0"], np.float32(1.0))
    r.g["gm"][:] = np.float32(3.0 * rho0)
    r.blur(0)
    np.testing.assert_array_equal(r.g["c0"], np.float32(1.0))
    r.g["gm"][:] = np.float32(0.5 * rho0)
    r.blur(0)
    np.testing.assert_array_equal(r.g["c0"], np.float32(0.5))


def test_stress_of_a_linear_colour():
    """colour = a x + b y: inside, (c_x, c_y) = (a, b) and (txx, tyy, txy) = (b^2, a^2, -a b)/|g|
    with |g| = sqrt(a^2 + b^2 + 1e-8); on the outer column x = 0 the held reads see x = 1 twice, so
    c_x = 0 there."""
    a, b = 0.0625, -0.03125
    r = _rig()
    x, y = _field()
    r.g["c1"][:] = (a * x + b * y).astype(np.float32)
    r.stress()
    inner = (x > 1) & (x < N[0] - 2) & (y > 1) & (y < N[1] - 2)
    length = np.sqrt(a * a + b * b + 1e-8)
    np.testing.assert_allclose(r.g["txx"][inner], b * b / length, rtol=2e-7)
    np.testing.assert_allclose(r.g["tyy"][inner], a * a / length, rtol=2e-7)
    np.testing.assert_allclose(r.g["txy"][inner], -a * b / length, rtol=2e-7)
    edge = (x == 0) & (y > 1) & (y < N[1] - 2)
    l0 = np.sqrt(b * b + 1e-8)
    np.testing.assert_allclose(r.g["txx"][edge], b * b / l0, rtol=2e-7)
    assert np.all(r.g["tyy"][edge] == 0) and np.all(r.g["txy"][edge] == 0)


def test_a_uniform_stress_gives_no_force():
    """Every node's east and west (north and south) fluxes of a uniform stress cancel, on the wall
    ring too, where the face into the wall joins the node to itself; and a thin colour (at most
    MIN_COLOUR) opens no face."""
    t = (0.5, 0.25, 0.375)
    r = _rig()
    r.g["c1"][:] = np.float32(1.0)
    for name, v in zip(oracle.STRESS2, t):
        r.g[name][:] = np.float32(v)
    for node in range(N[0] * N[1]):
        np.testing.assert_array_equal(r.force(node), np.zeros(2, np.float32))
    # one node's stress raised: its four neighbours feel half the difference across their face
    at = 4 * N[0] + 6
    r.g["txx"][at] = np.float32(1.5)
    np.testing.assert_array_equal(r.force(at + 1), np.array([-0.5, 0.0], np.float32))
    np.testing.assert_array_equal(r.force(at - 1), np.array([0.5, 0.0], np.float32))
    np.testing.assert_array_equal(r.force(at + N[0]), np.zeros(2, np.float32))
    r.g["c1"][at + 1] = np.float32(0.3)
    np.testing.assert_array_equal(r.force(at + 1), np.zeros(2, np.float32))


def test_forces_sum_to_zero_over_the_grid():
    """Every open face pairs two nodes with equal and opposite flux, so the forces of a random
    blob sum to zero (to float32 rounding of the per-node sums)."""
    rng = np.random.default_rng(3)
    r = _rig(n=(16, 14))
    x, y = _field((16, 14))
    blob = (x > 2) & (x < 13) & (y > 2) & (y < 11)
    r.g["c1"][:] = np.where(blob, rng.uniform(0.2, 1.0, x.size), 0.0).astype(np.float32)
    r.stress()
    total = np.zeros(2)
    scale = 0.0
    for node in range(x.size):
        f = r.force(node).astype(np.float64)
        total += f
        scale = max(scale, float(np.abs(f).max()))
    assert scale > 0.05
    np.testing.assert_allclose(total, 0.0, atol=1e-5 * scale)


def test_sigma_zero_is_the_plain_grid():
    """With sigma = 0 the tension grid is Flip.grid bit for bit (the colour and stress of a real
    scatter on hand); with tension only nodes with mass move."""
    rng = np.random.default_rng(5)
    count = 400
    f = {"x": rng.uniform(2, 9, count).astype(np.float32), "y": rng.uniform(2, 6, count).astype(np.float32),
         "m": np.full(count, 0.25, np.float32), "u": rng.standard_normal(count).astype(np.float32),
         "v": rng.standard_normal(count).astype(np.float32)}
    t = _rig(f, sigma=0.0)
    p = oracle.Mpm2(*N, f, t.params)
    for r in (t, p):
        r.clear()
        r.scatter()
    for k in range(oracle.TENSION_BLURS):
        t.blur(k)
    t.stress()
    assert np.abs(t.g["tyy"]).max() > 0
    t.grid()
    p.grid()
    for g in ("gu", "gv"):
        np.testing.assert_array_equal(t.g[g], p.g[g], err_msg=g)
    s = _rig(f, sigma=50.0)
    s.g = {k: v.copy() for k, v in t.g.items()}
    s.grid()
    moved = s.g["gu"] != t.g["gu"]
    assert moved.any()
    assert np.all(t.g["gm"][moved] > 1e-9)


class _Oracle:
    def __init__(self, f, params):
        self.r = oracle.Mpm2(cases.N, cases.N, f, params, tension=True)

    def advance(self, count, sort_every):
        self.r.advance(count, sort_every)

    def state(self):
        return [self.r.p[k] for k in ("x", "y", "u", "v", "j")]


def test_a_drops_pressure_is_sigma_over_its_radius():
    """TensionTest.aDropsPressureIsSigmaOverItsRadius on the oracle's whole tension steps."""
    cases.judge(_Oracle, "oracle")
