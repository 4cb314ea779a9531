"""The oracle's 3D surface tension (core/particle/Tension3.java:69-196, Flip3.gridWithTension in
core/particle/Flip3.java:263-311) pinned by what the mathematics fixes: the blur keeps a constant
and its weights sum to one; a full node -- a rest mass inside, half of one for each wall on the
wall ring -- reads as colour one; the stress of a linear colour is the closed form
(|g|^2 - g_a g_b)/|g| with every component in its place; the force of a uniform stress is zero
inside and only the shear past a wall; the fluxes cancel pairwise so the force sums to zero over
the grid; and with sigma = 0 the tension grid is the plain grid."""
import numpy as np
import pytest

import oracle

N = (10, 9, 8)


def _rig(fields=None, sigma=0.0, rho0=1.0, dt=1e-3, n=N):
    count = 8
    f = fields or {"x": np.full(count, 4.5, np.float32), "y": np.full(count, 4.5, np.float32),
                   "z": np.full(count, 4.5, np.float32), "m": np.full(count, 0.125, np.float32)}
    return oracle.Mpm3(*n, f, oracle.mpm3_params(dt, 0.0, -9.0, 0.0, 100.0, rho0, sigma), tension=True)


def _field(n=N):
    nx, ny, nz = n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


@pytest.mark.parametrize("value", [1.0, 0.5, 0.375])
def test_blur_keeps_a_constant(value):
    r = _rig()
    r.g["gm"][:] = np.float32(value)
    r.g["c0"][:] = np.float32(value)
    r.blur(1)  # unscaled: c0 -> c1
    np.testing.assert_array_equal(r.g["c1"], np.float32(value))


def test_blur_weights_are_one_two_one():
    """A unit impulse in the middle spreads as [1 2 1]/4 along each axis."""
    r = _rig()
    at = (4 * N[1] + 4) * N[0] + 5
    r.g["c0"][:] = 0
    r.g["c0"][at] = 1.0
    r.blur(1)
    w = np.array([0.25, 0.5, 0.25])
    exp = np.zeros(N[::-1])
    exp[3:6, 3:6, 4:7] = w[:, None, None] * w[None, :, None] * w[None, None, :]
    np.testing.assert_array_equal(r.g["c1"].reshape(N[::-1]), exp.astype(np.float32))
    assert float(r.g["c1"].sum(dtype=np.float64)) == 1.0


@pytest.mark.parametrize("rho0", [1.0, 2.5])
def test_full_nodes_read_as_colour_one(rho0):
    """gm = rho0 inside, halved once for each axis on which the node sits on the wall ring (the
    outer layer is never read): every blurred colour is exactly one. Leaving out one halving
    reads the ring as half full and the colour there drops below one."""
    r = _rig(rho0=rho0)
    x, y, z = _field()
    halves = np.zeros(x.size)
    for c, size in ((x, N[0]), (y, N[1]), (z, N[2])):
        halves += (c == 1) | (c == size - 2)
    r.g["gm"][:] = (rho0 * 0.5 ** halves).astype(np.float32)
    r.blur(0)
    np.testing.assert_array_equal(r.g["c0"], np.float32(1.0))
    # an over-full node is held to one
    r.g["gm"][:] *= np.float32(3.0)
    r.blur(0)
    np.testing.assert_array_equal(r.g["c0"], np.float32(1.0))
    # a ring node filled only as an inner one is (one halving left out) reads full too (held to one),
    # while an inner node at half a rest mass does not
    r.g["gm"][:] = np.float32(rho0)
    r.blur(0)
    inner = (x > 2) & (x < N[0] - 3) & (y > 2) & (y < N[1] - 3) & (z > 2) & (z < N[2] - 3)
    np.testing.assert_array_equal(r.g["c0"][inner], np.float32(1.0))
    r.g["gm"][:] = np.float32(0.5 * rho0)
    r.blur(0)
    assert np.all(r.g["c0"][inner] == np.float32(0.5))


def test_stress_of_a_linear_colour():
    """colour = a x + b y + c z: inside, g = (a, b, c) and T_ab = (|g|^2 delta_ab - g_a g_b)/|g|
    with |g| = sqrt(|g|^2 + 1e-8); each of the six components is checked, so a swapped one fails."""
    a, b, c = 0.0625, -0.03125, 0.015625
    r = _rig()
    x, y, z = _field()
    r.g["c1"][:] = (a * x + b * y + c * z).astype(np.float32)
    r.stress()
    inner = (x > 1) & (x < N[0] - 2) & (y > 1) & (y < N[1] - 2) & (z > 1) & (z < N[2] - 2)
    g = np.array([a, b, c])
    length = np.sqrt(g @ g + 1e-8)
    exp = {"txx": (b * b + c * c) / length, "tyy": (a * a + c * c) / length, "tzz": (a * a + b * b) / length,
           "txy": -a * b / length, "txz": -a * c / length, "tyz": -b * c / length}
    for name, e in exp.items():
        np.testing.assert_allclose(r.g[name][inner], e, rtol=2e-7, err_msg=name)
    # on the outer layer the held reads see one side only: at x = 0 both x reads are x = 1
    edge = (x == 0) & (y > 1) & (y < N[1] - 2) & (z > 1) & (z < N[2] - 2)
    g0 = np.array([0.0, b, c])
    l0 = np.sqrt(g0 @ g0 + 1e-8)
    np.testing.assert_allclose(r.g["txy"][edge], 0.0, atol=0)
    np.testing.assert_allclose(r.g["txx"][edge], (b * b + c * c) / l0, rtol=2e-7)
    np.testing.assert_allclose(r.g["tyz"][edge], -b * c / l0, rtol=2e-7)


def test_stress_vanishes_where_the_colour_is_flat():
    r = _rig()
    r.g["c1"][:] = np.float32(0.7)
    r.stress()
    for name in oracle.STRESS3:
        assert np.all(r.g[name] == 0), name


def _uniform(r, t):
    r.g["c1"][:] = np.float32(1.0)
    for name, v in zip(oracle.STRESS3, t):
        r.g[name][:] = np.float32(v)


def test_force_of_a_uniform_stress():
    """Inside, the six face fluxes of a uniform stress cancel. On a node of the wall ring at
    x = 1, the -x face is in the wall: its normal flux -T_xx still crosses, its shear does not, so
    the force is (0, T_xy, T_xz) (what the +x face carries)."""
    t = (0.5, 0.25, 0.125, 0.375, -0.0625, 0.1875)
    r = _rig()
    _uniform(r, t)
    inner = (4 * N[1] + 4) * N[0] + 5
    np.testing.assert_array_equal(r.force(inner), np.zeros(3, np.float32))
    ring = (4 * N[1] + 4) * N[0] + 1
    np.testing.assert_array_equal(r.force(ring), np.array([0.0, t[3], t[4]], np.float32))
    ring_y = (4 * N[1] + N[1] - 2) * N[0] + 5  # y = ny - 2: the +y face is in the wall
    np.testing.assert_array_equal(r.force(ring_y), np.array([-t[3], 0.0, -t[5]], np.float32))
    # outside the ring (x = 0) and where the colour is thin there is no force
    np.testing.assert_array_equal(r.force((4 * N[1] + 4) * N[0]), np.zeros(3, np.float32))
    r.g["c1"][inner] = np.float32(0.3)
    np.testing.assert_array_equal(r.force(inner), np.zeros(3, np.float32))


def test_forces_sum_to_zero_over_the_grid():
    """Away from the walls every open face pairs two nodes with equal and opposite flux, so the
    forces of a random blob sum to zero (to float32 rounding of the per-node sums)."""
    rng = np.random.default_rng(3)
    r = _rig()
    x, y, z = _field()
    blob = (x > 2) & (x < N[0] - 3) & (y > 2) & (y < N[1] - 3) & (z > 2) & (z < N[2] - 3)
    colour = np.where(blob, rng.uniform(0.2, 1.0, x.size), 0.0).astype(np.float32)
    r.g["c1"][:] = colour
    r.stress()
    total = np.zeros(3)
    scale = 0.0
    for node in range(x.size):
        f = r.force(node).astype(np.float64)
        total += f
        scale = max(scale, float(np.abs(f).max()))
    assert scale > 0.05
    np.testing.assert_allclose(total, 0.0, atol=1e-5 * scale)


def test_sigma_zero_is_the_plain_grid():
    """With sigma = 0 the tension grid is Flip3.grid: the same velocities wherever there is mass
    (the colour and stress of a real scatter on hand)."""
    rng = np.random.default_rng(5)
    count = 480
    f = {"x": rng.uniform(2, 8, count).astype(np.float32), "y": rng.uniform(2, 7, count).astype(np.float32),
         "z": rng.uniform(2, 6, count).astype(np.float32), "m": np.full(count, 0.125, np.float32),
         "u": rng.standard_normal(count).astype(np.float32), "v": rng.standard_normal(count).astype(np.float32),
         "w": rng.standard_normal(count).astype(np.float32)}
    t = _rig(f, sigma=0.0)
    p = oracle.Mpm3(*N, f, t.params)
    for r in (t, p):
        r.clear()
        r.scatter()
    for k in range(oracle.TENSION_BLURS):
        t.blur(k)
    t.stress()
    assert np.abs(t.g["txx"]).max() > 0
    t.grid()
    p.grid()
    for g in ("gu", "gv", "gw"):
        np.testing.assert_array_equal(t.g[g], p.g[g], err_msg=g)
    # and with tension the velocities move, only where there is mass
    s = _rig(f, sigma=50.0)
    s.g = {k: v.copy() for k, v in t.g.items()}
    s.grid()
    moved = s.g["gu"] != t.g["gu"]
    assert moved.any()
    assert np.all(t.g["gm"][moved] > 1e-9)


def test_step_conserves_momentum_without_gravity():
    """Whole tension steps of a resting cube of water (no gravity): the tension moves the
    particles, but it is internal, so the mean velocity stays near zero."""
    rng = np.random.default_rng(8)
    n = (16, 16, 16)
    count = 8 * 6 ** 3
    f = {a: rng.uniform(5, 11, count).astype(np.float32) for a in "xyz"}
    f["m"] = np.full(count, 0.125, np.float32)
    r = oracle.Mpm3(*n, f, oracle.mpm3_params(1e-3, 0, 0, 0, 100.0, 1.0, 40.0), tension=True)
    r.step(3)
    speed = np.sqrt(r.p["u"] ** 2 + r.p["v"] ** 2 + r.p["w"] ** 2)
    assert speed.max() > 1e-3
    mean = np.array([r.p[k].mean(dtype=np.float64) for k in "uvw"])
    assert np.abs(mean).max() < 0.02 * speed.max()
