"""The oracle's fixed-point 2D P2G (orc_mpm2_scatter_fixed): what integer sums must give whatever
the order, and agreement with the float scatter within FP32 rounding."""
import numpy as np

import oracle
from vexhip import scenes


def _state(n=48, seed=4):
    x, y = scenes.jittered_block2((2, 3), (20, 25), 4, scenes.JavaRandomFloats(seed))
    r = np.random.default_rng(seed)
    c = len(x)
    f = {"x": x, "y": y, "m": np.full(c, 0.25, np.float32),
         "u": (30 * r.standard_normal(c)).astype(np.float32), "v": (30 * r.standard_normal(c)).astype(np.float32),
         "j": (1 + 0.05 * r.standard_normal(c)).astype(np.float32)}
    for k in ("c00", "c01", "c10", "c11"):
        f[k] = r.standard_normal(c).astype(np.float32)
    return n, f, oracle.mpm2_params(1e-3, 0, -50, 400, 1)


def _scatter(n, f, params):
    s = oracle.Mpm2(n, n, f, params, fixed=True)
    s.scatter()
    return s


def test_any_order_gives_the_same_bits():
    """Integer addition associates: the particles shuffled give every node's sums bit for bit."""
    n, f, params = _state()
    a = _scatter(n, f, params)
    order = oracle.JavaRandom(9).shuffle_order(len(f["x"]))
    b = _scatter(n, {k: v[order] for k, v in f.items()}, params)
    for qa, qb in zip(a.q, b.q):
        np.testing.assert_array_equal(qa, qb)
    for g in ("gm", "gmu", "gmv"):
        np.testing.assert_array_equal(a.g[g].view(np.int32), b.g[g].view(np.int32))


def test_split_runs_sum_to_the_whole():
    """Scattering the two halves of the particles separately and adding the integer grids gives
    the whole run's grid exactly (grouped_eq_grid)."""
    n, f, params = _state()
    half = len(f["x"]) // 2
    whole = _scatter(n, f, params)
    a = _scatter(n, {k: v[:half] for k, v in f.items()}, params)
    b = _scatter(n, {k: v[half:] for k, v in f.items()}, params)
    for qw, qa, qb in zip(whole.q, a.q, b.q):
        np.testing.assert_array_equal(qw, qa + qb)


def test_matches_float_scatter_and_conserves():
    n, f, params = _state()
    fx = _scatter(n, f, params)
    fl = oracle.Mpm2(n, n, f, params)
    fl.scatter()
    for g in ("gm", "gmu", "gmv"):
        ref = fl.g[g].astype(np.float64)
        assert np.all(np.abs(fx.g[g] - ref) <= 1e-5 * np.maximum(1.0, np.abs(ref))), g
    mass = float(np.sum(f["m"], dtype=np.float64))
    assert abs(float(np.sum(fx.q[0])) * 2.0 ** -40 - mass) <= 1e-6 * mass
    # momentum: sum_k w_k m (v + C d_k) = m v, since sum_k w_k d_k = 0 (the stress term too)
    for q, vel in ((fx.q[1], "u"), (fx.q[2], "v")):
        p = float(np.sum(f["m"].astype(np.float64) * f[vel]))
        assert abs(float(np.sum(q)) * 2.0 ** -32 - p) <= 1e-5 * max(1.0, abs(p)) + 1e-3


def test_quantum_is_the_rounded_amount():
    """A lone particle at a node-aligned offset: its mass shares are the quadratic weights times m,
    each a binary32 product rounded to the nearest 2^-40, so the integer total is the exact sum of
    those products within half a quantum each."""
    params = oracle.mpm2_params(1e-3, 0, 0, 0, 1)
    f = {"x": np.array([5.0], np.float32), "y": np.array([7.0], np.float32), "m": np.array([0.3], np.float32)}
    s = _scatter(12, f, params)
    w = np.array([0.125, 0.75, 0.125])
    shares = np.outer(w, w) * np.float32(0.3)
    nz = s.q[0][s.q[0] != 0]
    expect = np.sort(np.rint(shares.astype(np.float32).astype(np.float64) * 2.0 ** 40).ravel())
    np.testing.assert_array_equal(np.sort(nz), expect.astype(np.int64))
    exact = float(np.sum(shares.astype(np.float32).astype(np.float64))) * 2.0 ** 40
    assert abs(int(s.q[0].sum()) - exact) <= 9 * 0.5


# ------------------------------------------------------------------ 3D (orc_mpm3_scatter_fixed)

def _state3(n=(20, 18, 16), seed=6):
    x, y, z = scenes.jittered_block3((2, 3, 2), (10, 9, 8), 8, scenes.JavaRandomFloats(seed))
    r = np.random.default_rng(seed)
    c = len(x)
    f = {"x": x, "y": y, "z": z, "m": np.full(c, 0.125, np.float32),
         "j": (1 + 0.05 * r.standard_normal(c)).astype(np.float32)}
    for k in ("u", "v", "w"):
        f[k] = (30 * r.standard_normal(c)).astype(np.float32)
    for k in oracle.FIELDS3[8:]:
        f[k] = r.standard_normal(c).astype(np.float32)
    return n, f, oracle.mpm3_params(1e-3, 0, -50, 0, 400, 1)


def _scatter3(n, f, params):
    s = oracle.Mpm3(*n, f, params, fixed=True)
    s.scatter()
    return s


def test_3d_any_order_gives_the_same_bits():
    """Integer addition associates: the 3D particles shuffled give every node's four sums bit for
    bit (a float scatter in another order would not)."""
    n, f, params = _state3()
    a = _scatter3(n, f, params)
    order = oracle.JavaRandom(11).shuffle_order(len(f["x"]))
    b = _scatter3(n, {k: v[order] for k, v in f.items()}, params)
    for qa, qb in zip(a.q, b.q):
        np.testing.assert_array_equal(qa, qb)
    for g in ("gm", "gmu", "gmv", "gmw"):
        np.testing.assert_array_equal(a.g[g].view(np.int32), b.g[g].view(np.int32))
    fl = [oracle.Mpm3(*n, ff, params) for ff in (f, {k: v[order] for k, v in f.items()})]
    for r in fl:
        r.clear()
        r.scatter()
    assert any(not np.array_equal(fl[0].g[g], fl[1].g[g]) for g in ("gmu", "gmv", "gmw"))


def test_3d_split_runs_sum_to_the_whole():
    n, f, params = _state3()
    half = len(f["x"]) // 3
    whole = _scatter3(n, f, params)
    a = _scatter3(n, {k: v[:half] for k, v in f.items()}, params)
    b = _scatter3(n, {k: v[half:] for k, v in f.items()}, params)
    for qw, qa, qb in zip(whole.q, a.q, b.q):
        np.testing.assert_array_equal(qw, qa + qb)


def test_3d_matches_float_scatter_and_conserves():
    """The decoded sums are the float scatter's within its own rounding; the mass is the
    particles' and the momentum sum_p m v (the stencil's first moments vanish)."""
    n, f, params = _state3()
    fx = _scatter3(n, f, params)
    fl = oracle.Mpm3(*n, f, params)
    fl.clear()
    fl.scatter()
    for g in ("gm", "gmu", "gmv", "gmw"):
        ref = fl.g[g].astype(np.float64)
        assert np.all(np.abs(fx.g[g] - ref) <= 1e-5 * np.maximum(1.0, np.abs(ref))), g
    mass = float(np.sum(f["m"], dtype=np.float64))
    assert abs(float(np.sum(fx.q[0])) * 2.0 ** -40 - mass) <= 1e-6 * mass
    for q, vel in ((fx.q[1], "u"), (fx.q[2], "v"), (fx.q[3], "w")):
        p = float(np.sum(f["m"].astype(np.float64) * f[vel]))
        assert abs(float(np.sum(q)) * 2.0 ** -32 - p) <= 1e-5 * max(1.0, abs(p)) + 1e-3


def test_3d_quantum_is_the_rounded_amount():
    """A lone particle at rest on a node-aligned offset: its 27 mass shares are the quadratic
    weights' products times m, each rounded to the nearest 2^-40; its momentum shares are zero."""
    params = oracle.mpm3_params(1e-3, 0, 0, 0, 0, 1)
    f = {"x": np.array([5.0], np.float32), "y": np.array([7.0], np.float32), "z": np.array([4.0], np.float32),
         "m": np.array([0.3], np.float32)}
    s = _scatter3((12, 12, 12), f, params)
    w = np.array([0.125, 0.75, 0.125], np.float32)
    shares = (w[None, None, :] * (w[None, :, None] * w[:, None, None])) * np.float32(0.3)
    nz = s.q[0][s.q[0] != 0]
    expect = np.sort(np.rint(shares.astype(np.float32).astype(np.float64) * 2.0 ** 40).ravel())
    np.testing.assert_array_equal(np.sort(nz), expect.astype(np.int64))
    for q in s.q[1:]:
        assert not q.any()


def test_3d_fixed_steps_are_bit_exact_repeats():
    """Whole fixed-point substeps from a shuffled start equal those from the sorted one bit for
    bit once the particle order is undone: P2G's sums do not depend on the order, and the grid
    update and G2P are per node and per particle."""
    n, f, params = _state3()
    order = oracle.JavaRandom(5).shuffle_order(len(f["x"]))
    a = oracle.Mpm3(*n, f, params, fixed=True)
    b = oracle.Mpm3(*n, {k: v[order] for k, v in f.items()}, params, fixed=True)
    a.step(3)
    b.step(3)
    for k in oracle.FIELDS3:
        np.testing.assert_array_equal(a.p[k][order].view(np.int32), b.p[k].view(np.int32), err_msg=k)
