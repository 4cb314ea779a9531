"""The oracle's sort keys and stable order, against numpy and the reference's cellStarts."""
import numpy as np
import pytest

import mpm_cases
import oracle


def test_stable_order_is_numpy_stable_argsort():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 300, 20000).astype(np.int32)
    np.testing.assert_array_equal(oracle.stable_order(keys), np.argsort(keys, kind="stable"))


def test_cell_is_the_kernels_cell():
    """Scatter.cell (core/particle/Scatter.java:473-477): row-major over (n-1)^2 cells, the far
    edge clamped into the last cell, an escaped particle clamped to the nearest."""
    n = 8
    assert oracle.mpm2_cell(2.5, 3.5, n, n) == 3 * 7 + 2
    assert oracle.mpm2_cell(7.0, 7.0, n, n) == 6 * 7 + 6
    assert oracle.mpm2_cell(-3.0, 100.0, n, n) == 6 * 7 + 0
    p = mpm_cases.jittered2(n - 1, 3, 9)
    keys = oracle.mpm2_keys(p["x"], p["y"], n, n)
    np.testing.assert_array_equal(keys, np.repeat(np.arange(49), 3))


def test_cell_starts_refuses_particles_out_of_order():
    """ScatterTest.cellStartsRefusesParticlesOutOfOrder (test/particle/ScatterTest.java:111-123)."""
    p = mpm_cases.jittered2(7, 2, 3)
    order = oracle.JavaRandom(4).shuffle_order(len(p["x"]))
    x, y = p["x"][order], p["y"][order]
    with pytest.raises(ValueError, match="sorted by cell"):
        oracle.mpm2_cell_starts(x, y, 8, 8)
    keys = oracle.mpm2_keys(x, y, 8, 8)
    s = oracle.stable_order(keys)
    starts = oracle.mpm2_cell_starts(x[s], y[s], 8, 8)
    assert len(starts) == 50 and starts[-1] == len(x)
    np.testing.assert_array_equal(np.diff(starts), 2)
    np.testing.assert_array_equal(starts, oracle.bucket_starts(keys, 49))


def test_sort_by_cell_keeps_each_cells_particles():
    """SortTest.theDeviceSortsByCell's contract (test/particle/SortTest.java:118-140) on the oracle's
    stable sort: starts are the host's, the buffer is in cell order, each cell keeps its particles."""
    n = 40
    p = mpm_cases.jittered2(n - 1, 3, 31)
    order = oracle.JavaRandom(32).shuffle_order(len(p["x"]))
    p = {k: v[order] for k, v in p.items()}
    count = len(p["x"])
    s = oracle.Mpm2(n, n, {**p, "j": np.ones(count)}, oracle.mpm2_params(1e-3, 0, 0, 0, 1))
    s.sort()
    keys = oracle.mpm2_keys(s.p["x"], s.p["y"], n, n)
    assert np.all(np.diff(keys) >= 0)
    expected = oracle.mpm2_cell_starts(np.sort(p["x"]), np.sort(p["y"]), n, n) if False else None
    starts = oracle.mpm2_cell_starts(s.p["x"], s.p["y"], n, n)
    np.testing.assert_array_equal(starts, oracle.bucket_starts(oracle.mpm2_keys(p["x"], p["y"], n, n), (n - 1) ** 2))
    a = sorted(zip(keys.tolist(), s.p["x"].view(np.int32).tolist(), s.p["m"].view(np.int32).tolist()))
    k0 = oracle.mpm2_keys(p["x"], p["y"], n, n)
    b = sorted(zip(k0.tolist(), p["x"].view(np.int32).tolist(), p["m"].view(np.int32).tolist()))
    assert a == b
    del expected


def test_brick_keys_are_a_bijection_of_the_base_cell():
    """The 3D key renumbers Stencil3's base cell brick-major; distinct cells get distinct keys,
    and every key fits the reported bit count."""
    n = 20
    cells = np.arange(1, n - 2)
    cz, cy, cx = np.meshgrid(cells, cells, cells, indexing="ij")
    x = (cx.ravel() + 0.75).astype(np.float32)
    y = (cy.ravel() + 0.75).astype(np.float32)
    z = (cz.ravel() + 0.75).astype(np.float32)
    m = np.ones_like(x)
    s = oracle.Mpm3(n, n, n, {"x": x, "y": y, "z": z, "m": m}, oracle.mpm3_params(1e-3, 0, 0, 0, 0, 1))
    keys = s.keys()
    assert len(np.unique(keys)) == len(keys)
    bits = oracle.mpm3_key_bits(n, n, n)
    assert keys.max() < (1 << bits)
    # the inactive key is one past the largest cell's (the base cell n - 3 on every axis)
    assert oracle.mpm3_inactive_key(n, n, n) == keys.max() + 1 < (1 << bits)
    # cells within one 8x8x8 brick share its 512-key block
    base = ((z - 0.5).astype(np.int64), (y - 0.5).astype(np.int64), (x - 0.5).astype(np.int64))
    brick = (base[0] // 8) * 9 + (base[1] // 8) * 3 + base[2] // 8
    np.testing.assert_array_equal(keys // 512, brick)
    assert oracle.mpm3_key_bits(256, 256, 256) == 24


@pytest.mark.parametrize("n", [20, 256, 258])
def test_inactive_slots_sort_last_in_slot_order(n):
    """A slot without mass (m <= 0, or NaN) keys past every cell; the stable sort then puts the
    active slots first, by cell, and the inactive ones after them in their slot order. The key
    bits grow only when the last brick is full (n - 2 a multiple of 8)."""
    rng = np.random.default_rng(n)
    count = 5000
    x, y, z = (rng.uniform(0, n, count).astype(np.float32) for _ in range(3))
    m = rng.uniform(0.1, 1, count).astype(np.float32)
    dead = rng.random(count) < 0.3
    m[dead] = 0.0
    m[np.flatnonzero(dead)[:5]] = -1.0
    m[np.flatnonzero(dead)[5:10]] = np.nan
    s = oracle.Mpm3(n, n, n, {"x": x, "y": y, "z": z, "m": m}, oracle.mpm3_params(1e-3, 0, 0, 0, 0, 1))
    keys = s.keys()
    inactive = oracle.mpm3_inactive_key(n, n, n)
    np.testing.assert_array_equal(keys == inactive, dead)
    assert keys[~dead].max() < inactive < (1 << oracle.mpm3_key_bits(n, n, n))
    order = oracle.stable_order(keys)
    active = int((~dead).sum())
    assert not dead[order[:active]].any()
    np.testing.assert_array_equal(order[active:], np.flatnonzero(dead))
    full = (n - 2) % 8 == 0
    cells = ((n - 2 + 7) // 8) ** 3 * 512
    assert oracle.mpm3_key_bits(n, n, n) == int(np.ceil(np.log2(cells + (1 if full else 0))))


def test_layer_order_by_brute_force():
    """The group-by-group order (P2G schedule 3's layout) against a direct construction: per
    group (key >> 8) in key order, its cells ranked by count (most first, counts above 255 equal,
    ties by cell), 64 to a generation; generation by generation, round r takes the r-th particle of
    each cell in rank order; the inactive keys keep their places at the end."""
    rng = np.random.default_rng(11)
    inactive = 256 * 5 + 17
    heavy = np.full(300, 256 + 5)  # a cell past the count clamp, and one just at it
    edge = np.full(255, 256 + 9)
    keys = np.sort(np.concatenate([rng.integers(0, inactive, 6000), heavy, edge, np.full(40, inactive),
                                   rng.integers(inactive + 1, inactive + 50, 10)])).astype(np.int32)
    got = oracle.mpm3_layer_order(keys, inactive)
    expect = []
    pos = {}
    for p, k in enumerate(keys):
        if k < inactive:
            pos.setdefault(int(k) >> 8, {}).setdefault(int(k) & 255, []).append(p)
    for group in sorted(pos):
        cells = pos[group]
        ranked = sorted(cells, key=lambda c: (-min(len(cells[c]), 255), c))
        for g in range(0, len(ranked), 64):
            gen = ranked[g:g + 64]
            for r in range(max(len(cells[c]) for c in gen)):
                for c in gen:
                    if len(cells[c]) > r:
                        expect.append(cells[c][r])
    live = len(expect)
    np.testing.assert_array_equal(got[:live], expect)
    np.testing.assert_array_equal(got[live:], np.arange(live, len(keys)))
    assert sorted(got) == list(range(len(keys)))
