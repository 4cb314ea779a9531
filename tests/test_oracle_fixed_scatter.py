"""The oracle's FixedScatter (orc_fixed_scatter) held to the reference's own tests of it
(test/particle/FixedScatterTest.java), run on the CPU restatement: exact conservation, no
momentum without mass, wrapping partial sums, the exact floor of the split share, closeness to
the FP64 deposit within the quanta, and the finest scale rule (vexhip.fixed_scatter.FixedPoint)."""
import numpy as np

import oracle
from vexhip.fixed_scatter import FixedPoint

SCALE = FixedPoint.for_range(16, 1.5, 8)


def test_for_range_is_the_finest_scale_that_fits():
    """FixedScatterTest.forRangeIsTheFinestScaleThatFits (:245-251)."""
    assert SCALE.fits(16, 1.5, 8)
    assert not (FixedPoint(SCALE.mass_bits + 1, SCALE.velocity_bits).fits(16, 1.5, 8)
                and FixedPoint(SCALE.mass_bits, SCALE.velocity_bits + 1).fits(16, 1.5, 8))


def test_the_split_share_is_the_exact_floor():
    """theSplitShareIsTheExactFloor (:232-243): over edges and random weights and masses."""
    one = 1 << oracle.FIXED_OFFSET_BITS
    r = oracle.JavaRandom(71)
    for t in range(100_000):
        q = (0 if t % 2 == 0 else one * one) if t < 4 else r.next_int(one * one + 1)
        mass = (0 if t < 2 else oracle.FIXED_MAX_MASS - 1) if t < 4 else r.next_int(oracle.FIXED_MAX_MASS)
        assert oracle.fixed_split(q, mass) == (q * mass) >> (2 * oracle.FIXED_OFFSET_BITS), (q, mass)


def test_mass_and_momentum_are_conserved_exactly():
    """massAndMomentumAreConservedExactly (:100-122)."""
    n = 64
    p = oracle.shuffled(oracle.jittered_particles(n - 1, 4, oracle.JavaRandom(21)), oracle.JavaRandom(22))
    g = oracle.fixed_scatter(n, p, SCALE.mass_bits, SCALE.velocity_bits)
    mq = (p["m"] * np.float32(SCALE.mass_scale())).astype(np.int64)
    uq = (p["u"] * np.float32(SCALE.velocity_scale())).astype(np.int64)
    vq = (p["v"] * np.float32(SCALE.velocity_scale())).astype(np.int64)
    assert int(g[0].astype(np.int64).sum()) == int(mq.sum())
    assert int(g[1].astype(np.int64).sum()) == int((mq * uq).sum())
    assert int(g[2].astype(np.int64).sum()) == int((mq * vq).sum())


def test_a_massless_node_holds_no_momentum():
    """aMasslessNodeHoldsNoMomentum (:130-160): particles of a few quanta, one in every other cell."""
    n = 32
    allp = oracle.jittered_particles(n - 1, 1, oracle.JavaRandom(31))
    keep = ((allp["x"].astype(np.int32) % 2) == 0) & ((allp["y"].astype(np.int32) % 2) == 0)
    p = {k: v[keep].copy() for k, v in allp.items()}
    r = oracle.JavaRandom(32)
    for i in range(len(p["x"])):
        p["m"][i] = np.float32(np.float32(1 + r.next_int(3)) / np.float32(SCALE.mass_scale()))
        p["u"][i] = np.float32(4 * r.next_gaussian())
        p["v"][i] = np.float32(4 * r.next_gaussian())
    g = oracle.fixed_scatter(n, p, SCALE.mass_bits, SCALE.velocity_bits)
    assert (g[0] >= 0).all()
    empty = g[0] == 0
    assert (g[1][empty] == 0).all() and (g[2][empty] == 0).all()
    assert int(empty.sum()) > n * n // 2
    mq = (p["m"] * np.float32(SCALE.mass_scale())).astype(np.int64)
    assert int(g[0].astype(np.int64).sum()) == int(mq.sum())


def _pairs(n=16):
    half = oracle.jittered_particles(n - 1, 2, oracle.JavaRandom(41))
    count = 2 * len(half["x"])
    p = {k: np.empty(count, np.float32) for k in half}
    for k in ("x", "y", "m"):
        p[k][0::2] = half[k]
        p[k][1::2] = half[k]
    p["u"][0::2], p["u"][1::2] = 1000, -1000
    p["v"][0::2], p["v"][1::2] = -700, 700
    return p


def test_partial_sums_may_overflow_when_the_total_fits():
    """partialSumsMayOverflowWhenTheTotalFits (:168-195): opposite velocities in pairs, momentum reads 0."""
    coarse = FixedPoint(20, 16)
    assert not coarse.fits(16, 1.5, 1000)
    p = _pairs()
    for q in (p, oracle.shuffled(p, oracle.JavaRandom(42))):
        g = oracle.fixed_scatter(16, q, coarse.mass_bits, coarse.velocity_bits)
        assert (g[1] == 0).all() and (g[2] == 0).all()


def test_the_grid_is_the_exact_deposit_to_within_its_quanta():
    """theGridIsTheExactDepositToWithinItsQuanta (:198-214)."""
    n = 64
    p = oracle.jittered_particles(n - 1, 4, oracle.JavaRandom(61))
    g = oracle.fixed_scatter(n, p, SCALE.mass_bits, SCALE.velocity_bits)
    exact = oracle.scatter_f64(n, p)
    mq = SCALE.mass(1)
    vq = 2.0 ** -SCALE.velocity_bits
    mass_tol = 16 * (4 * mq + 1.5 * 2.0 / (1 << oracle.FIXED_OFFSET_BITS))
    mass = g[0].astype(np.float64) * 2.0 ** -SCALE.mass_bits
    mom = [g[f].astype(np.float64) * 2.0 ** -(SCALE.mass_bits + SCALE.velocity_bits) for f in (1, 2)]
    assert np.all(np.abs(exact[0] - mass) <= mass_tol)
    mom_tol = mass_tol * 6 + exact[0] * vq
    for f in (0, 1):
        assert np.all(np.abs(exact[1 + f] - mom[f]) <= mom_tol)


def test_any_order_gives_the_same_grid():
    n = 17
    p = oracle.jittered_particles(n - 1, 64, oracle.JavaRandom(7))
    a = oracle.fixed_scatter(n, p, 20, 10)
    b = oracle.fixed_scatter(n, oracle.shuffled(p, oracle.JavaRandom(8)), 20, 10)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
