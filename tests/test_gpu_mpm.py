"""The HIP MLS-MPM passes: the reference's particle tests on the device (both P2G schedules), and
parity with the oracle per pass -- grid update and G2P bit for bit, P2G within FP32
reassociation (1e-5 max(1, |e|) per node, mass to 1e-6) -- and the sort bit for bit."""
import numpy as np
import pytest

import mpm_cases as cases
import oracle
from gpu_rigs import GpuRig, maker2, maker3
from vexhip import _native, particles, scenes

pytestmark = pytest.mark.gpu

# The segmented schedule with float wave tiles (read-add-write rounds, option p2g_ints 0) ...
SCHEDULES = [dict(p2g=0, g2p=0), dict(p2g=1, g2p=1, run=4, p2g_ints=0), dict(p2g=1, g2p=2, run=4, p2g_ints=0),
             dict(p2g=1, g2p=2, run=1, p2g_ints=0), dict(p2g=1, g2p=2, run=2, p2g_ints=0)]
IDS = ["atomic", "seg-exact", "seg-fast", "seg-run1", "seg-run2"]
# ... and with integer wave tiles (p2g_ints 1, the default): runs rounded to each wave's
# power-of-two units and added by LDS integer atomics.
INTS = [dict(p2g=1, g2p=2, run=4, p2g_ints=1), dict(p2g=1, g2p=2, run=2, p2g_ints=1),
        dict(p2g=1, g2p=1, run=1, p2g_ints=1)]
INT_IDS = ["ints-run4", "ints-run2", "ints-run1"]
SCHEDULES += INTS
IDS += INT_IDS


@pytest.mark.parametrize("opts", SCHEDULES, ids=IDS)
def test_lone_particle_falls_3d(gpu, opts):
    cases.case_lone_particle_falls_3d(maker3(**opts))


@pytest.mark.parametrize("opts", SCHEDULES, ids=IDS)
def test_scatter_conserves_3d(gpu, opts):
    cases.case_scatter_conserves_3d(maker3(**opts))


@pytest.mark.parametrize("opts", SCHEDULES, ids=IDS)
def test_affine_field_is_reproduced_3d(gpu, opts):
    cases.case_affine_field_is_reproduced_3d(maker3(**opts))


def test_compressed_particle_pushes_out_3d(gpu):
    cases.case_compressed_particle_pushes_out_3d(maker3())


def test_walls_hold_3d(gpu):
    cases.case_walls_hold_3d(maker3())


@pytest.mark.parametrize("opts", SCHEDULES[:3] + INTS[:1], ids=IDS[:3] + INT_IDS[:1])
def test_dam_break_slab_3d(gpu, opts):
    r = cases.case_dam_break_slab_3d(maker3(**opts))
    print(f"[flip3] GPU {opts}: moved {r[0]:.2f} of {r[1]:.2f}, J {r[2]:.3f}..{r[3]:.3f}, "
          f"height {r[4]:.1f}, w/u {r[5]:.3f}", flush=True)


@pytest.mark.parametrize("opts", [dict(p2g=0), dict(p2g=1, run=4)], ids=["atomic", "runs"])
def test_lone_particle_falls_2d(gpu, opts):
    cases.case_lone_particle_falls_2d(maker2(**opts))


@pytest.mark.parametrize("opts", [dict(p2g=0), dict(p2g=1, run=4), dict(p2g=1, run=2)], ids=["atomic", "run4", "run2"])
def test_scatter_matches_fp64_splat_2d(gpu, opts):
    cases.case_scatter_matches_fp64_splat_2d(maker2(**opts))
    cases.case_scatter_matches_fp64_splat_2d(maker2(**opts), order_seed=42)


@pytest.mark.parametrize("opts", [dict(p2g=0), dict(p2g=1, run=4), dict(p2g=1, run=2)], ids=["atomic", "run4", "run2"])
def test_broken_runs_2d(gpu, opts):
    cases.case_broken_runs_2d(maker2(**opts))


def test_affine_field_is_reproduced_2d(gpu):
    cases.case_affine_field_is_reproduced_2d(maker2())


@pytest.mark.parametrize("opts", [dict(p2g=0), dict(p2g=1, run=4)], ids=["atomic", "runs"])
def test_dam_break_2d(gpu, opts):
    r = cases.case_dam_break_2d(maker2(**opts))
    print(f"[flip] GPU {opts}: moved {r[0]:.2f} of {r[1]:.2f}, J {r[2]:.3f}..{r[3]:.3f}, height {r[4]:.1f}",
          flush=True)


# ------------------------------------------------------------------ per-pass parity, 3D

def _state3(n, count, seed):
    """A seeded state with motion, C and J away from rest: positions jittered over the interior."""
    f = cases.scatter_inputs_3d(n, count, seed)
    return f


def _close(got, exp, what):
    err = np.abs(got.astype(np.float64) - exp.astype(np.float64))
    tol = 1e-5 * np.maximum(1.0, np.abs(exp.astype(np.float64)))
    bad = err > tol
    assert not bad.any(), f"{what}: {int(bad.sum())} nodes off, worst {float(err.max()):.3e}"


@pytest.mark.parametrize("opts", SCHEDULES, ids=IDS)
def test_scatter_parity_3d(gpu, opts):
    n, count = 40, 60_000
    f = _state3(n, count, 7)
    ref = oracle.Mpm3(n, n, n, f, oracle.mpm3_params(1e-3, 0, -50, 0, 400.0, 1))
    ref.clear()
    ref.scatter()
    rig = maker3(**opts)(n, n, n, count)
    cases.load3(rig, f)
    rig.params(ref.params)
    rig.run(["clear", "scatter"])
    for g in ("gm", "gmu", "gmv", "gmw"):
        _close(rig.read(g), ref.g[g], g)
    mass = float(np.sum(f["m"], dtype=np.float64))
    assert abs(float(np.sum(rig.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass


@pytest.mark.parametrize("opts", [SCHEDULES[2]] + INTS, ids=[IDS[2]] + INT_IDS)
def test_sorted_scatter_parity_3d(gpu, opts):
    """P2G of a dense block (8 particles a cell, the headline's density) after a sort, so most
    waves take one brick's cells: every node within 1e-5 max(1, |e|) of the oracle's scatter of
    the same sorted particles, mass to 1e-6."""
    n = 40
    x, y, z = scenes.jittered_block3((3, 3, 3), (21, 18, 17), 8, scenes.JavaRandomFloats(31))
    count = len(x)
    order = oracle.JavaRandom(32).shuffle_order(count)
    rng = np.random.default_rng(5)
    f = {"x": x[order], "y": y[order], "z": z[order], "m": np.full(count, 0.125, np.float32),
         "j": (1.0 + 0.2 * rng.standard_normal(count)).astype(np.float32)}
    for name in ("u", "v", "w"):
        f[name] = (3.0 * rng.standard_normal(count)).astype(np.float32)
    for name in oracle.FIELDS3[8:]:
        f[name] = (0.5 * rng.standard_normal(count)).astype(np.float32)
    ref = oracle.Mpm3(n, n, n, f, oracle.mpm3_params(1e-3, 0, -50, 0, 400.0, 1))
    ref.sort()
    ref.clear()
    ref.scatter()
    rig = maker3(**opts)(n, n, n, count)
    cases.load3(rig, f)
    rig.params(ref.params)
    rig.run(["sort", "clear", "scatter"])
    for name in ("x", "y", "z"):
        np.testing.assert_array_equal(rig.read(name).view(np.int32), ref.p[name].view(np.int32), err_msg=name)
    for g in ("gm", "gmu", "gmv", "gmw"):
        _close(rig.read(g), ref.g[g], g)
    mass = float(np.sum(f["m"], dtype=np.float64))
    assert abs(float(np.sum(rig.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass


@pytest.mark.parametrize("opts", [SCHEDULES[2], INTS[0]], ids=[IDS[2], INT_IDS[0]])
def test_scatter_carries_a_nan_3d(gpu, opts):
    """A particle with a NaN velocity: the scatter leaves NaN in exactly the momentum nodes the
    oracle's scatter does (its 27 nodes' u sums) and nothing else non-finite -- with integer tiles
    its wave falls back to the direct atomics, which carry the NaN as the float schedules do."""
    n, count = 32, 30_000
    f = _state3(n, count, 11)
    f["u"] = f["u"].copy()
    f["u"][12_345] = np.float32("nan")
    ref = oracle.Mpm3(n, n, n, f, oracle.mpm3_params(1e-3, 0, -50, 0, 400.0, 1))
    ref.clear()
    ref.scatter()
    rig = maker3(**opts)(n, n, n, count)
    cases.load3(rig, f)
    rig.params(ref.params)
    rig.run(["clear", "scatter"])
    for g in ("gm", "gmu", "gmv", "gmw"):
        got, exp = rig.read(g), ref.g[g]
        np.testing.assert_array_equal(np.isnan(got), np.isnan(exp), err_msg=g)
        ok = ~np.isnan(exp)
        _close(got[ok], exp[ok], g)


def test_grid_and_advect_bit_exact_3d(gpu):
    """From the oracle's own scatter, the device's grid update and both G2P variants are the
    oracle's results bit for bit."""
    n, count = 32, 30_000
    f = _state3(n, count, 9)
    ref = oracle.Mpm3(n, n, n, f, oracle.mpm3_params(2e-3, 0.5, -50, 0.25, 400.0, 1))
    ref.clear()
    ref.scatter()
    for g2p in (0, 1):
        rig = maker3(p2g=1, g2p=g2p)(n, n, n, count)
        cases.load3(rig, f)
        rig.params(ref.params)
        for g in ("gm", "gmu", "gmv", "gmw"):
            rig.write(g, ref.g[g])
        rig.run(["grid"])
        exp = oracle.Mpm3(n, n, n, f, ref.params)
        exp.g = {k: v.copy() for k, v in ref.g.items()}
        exp.grid()
        for g in ("gu", "gv", "gw"):
            np.testing.assert_array_equal(rig.read(g).view(np.int32), exp.g[g].view(np.int32), err_msg=g)
        rig.run(["advect"])
        exp.advect()
        for name in oracle.FIELDS3:
            np.testing.assert_array_equal(rig.read(name).view(np.int32), exp.p[name].view(np.int32),
                                          err_msg=f"g2p={g2p} field {name}")


LAYERS = dict(p2g=3, g2p=2, run=4)


@pytest.mark.parametrize("sort", [False, True], ids=["unsorted", "sorted"])
@pytest.mark.parametrize("opts", SCHEDULES[:3] + INTS[:2] + [LAYERS], ids=IDS[:3] + INT_IDS[:2] + ["layers"])
def test_steps_track_the_oracle_3d(gpu, opts, sort):
    """Twenty whole steps of a falling, compressing block: positions within 1e-5 relative of the
    oracle's, grid mass equal to particle mass to 1e-6. Sorted, the segmented P2G and the tiled
    G2P take one brick of particles a block (the ranges the sort leaves)."""
    n = 32
    x, y, z = scenes.jittered_block3((3, 3, 3), (10, 14, 9), 8, scenes.JavaRandomFloats(5))
    count = len(x)
    m = np.full(count, 0.125, np.float32)
    params = oracle.mpm3_params(5e-4, 0, -200.0, 0, 2000.0, 1)
    ref = oracle.Mpm3(n, n, n, {"x": x, "y": y, "z": z, "m": m}, params)
    rig = maker3(**opts)(n, n, n, count)
    for name, v in (("x", x), ("y", y), ("z", z), ("m", m), ("j", np.ones(count, np.float32))):
        rig.write(name, v)
    rig.params(params)
    for k in range(20):
        if sort and k % 10 == 0:
            ref.sort(layers=opts.get("p2g") == 3)
            rig.run(["sort"])
        ref.step()
        rig.run("step")
    for name in ("x", "y", "z"):
        np.testing.assert_allclose(rig.read(name), ref.p[name], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(rig.read("j"), ref.p["j"], rtol=1e-4, atol=1e-5)
    mass = float(np.sum(m, dtype=np.float64))
    assert abs(float(np.sum(rig.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass


def test_fast_advect_close_3d(gpu):
    """The contracted G2P (g2p=2) from the oracle's own grid: every particle field within
    1e-5 max(1, |e|) of the reference-order result."""
    n, count = 32, 30_000
    f = _state3(n, count, 9)
    ref = oracle.Mpm3(n, n, n, f, oracle.mpm3_params(2e-3, 0.5, -50, 0.25, 400.0, 1))
    ref.clear()
    ref.scatter()
    ref.grid()
    rig = maker3(p2g=1, g2p=2)(n, n, n, count)
    cases.load3(rig, f)
    rig.params(ref.params)
    for g in oracle.GRID3:
        rig.write(g, ref.g[g])
    rig.run(["advect"])
    ref.advect()
    for name in oracle.FIELDS3:
        _close(rig.read(name), ref.p[name], name)


# ------------------------------------------------------------------ sort

def test_sort_3d_bit_exact(gpu):
    """The device sort of shuffled particles, a fifth of them without mass: keys ascending, and
    the permutation the host's stable sort by the same brick key gives (the massless slots last,
    in slot order), every field bit for bit; the slot map is that permutation and the active
    count is the number of slots with mass."""
    n = 48
    x, y, z = scenes.jittered_block3((2, 2, 2), (40, 30, 20), 2, scenes.JavaRandomFloats(21))
    count = len(x)
    order = oracle.JavaRandom(22).shuffle_order(count)
    f = {"x": x[order], "y": y[order], "z": z[order]}
    rng = np.random.default_rng(1)
    for name in oracle.FIELDS3[3:]:
        f[name] = rng.standard_normal(count).astype(np.float32)
    f["m"] = np.abs(f["m"]) + np.float32(0.1)
    dead = rng.random(count) < 0.2
    f["m"][dead] = 0.0
    f["m"][np.flatnonzero(dead)[:7]] = -0.5
    ref = oracle.Mpm3(n, n, n, f, oracle.mpm3_params(1e-3, 0, 0, 0, 0, 1))
    rig = maker3()(n, n, n, count)
    cases.load3(rig, f)
    np.testing.assert_array_equal(rig.sim.read_words("slot"), np.arange(count))
    rig.run(["sort"])
    keys = rig.sim.read_words("keys")
    perm = ref.sort()
    np.testing.assert_array_equal(keys, ref.keys())
    assert np.all(np.diff(keys) >= 0)
    for name in oracle.FIELDS3:
        np.testing.assert_array_equal(rig.read(name).view(np.int32), ref.p[name].view(np.int32), err_msg=name)
    assert perm.shape == (count,)
    np.testing.assert_array_equal(rig.sim.read_words("slot"), perm)
    active = int(np.count_nonzero(~dead))
    assert int(rig.sim.read_words("active")[0]) == active
    np.testing.assert_array_equal(perm[active:], np.flatnonzero(dead))


@pytest.mark.parametrize("p2g", [1, 3], ids=["seg", "layers"])
def test_inactive_slots_skipped_after_sort_3d(gpu, p2g):
    """A block with 40% of its slots emptied, sorted every 5 steps: the sort moves the empty slots
    behind the rest and P2G and G2P stop at the last active one. Twenty steps track the oracle
    (which visits every slot and skips the massless ones) slot for slot within 1e-5, and the
    empty slots come through untouched, bit for bit."""
    n = 32
    x, y, z = scenes.jittered_block3((3, 3, 3), (12, 14, 10), 8, scenes.JavaRandomFloats(9))
    count = len(x)
    rng = np.random.default_rng(4)
    m = np.full(count, 0.125, np.float32)
    dead = rng.random(count) < 0.4
    m[dead] = 0.0
    u = rng.standard_normal(count).astype(np.float32)
    params = oracle.mpm3_params(5e-4, 0, -200.0, 0, 2000.0, 1)
    f = {"x": x, "y": y, "z": z, "m": m, "u": u}
    ref = oracle.Mpm3(n, n, n, f, params)
    rig = maker3(p2g=p2g)(n, n, n, count)
    for name, v in (("x", x), ("y", y), ("z", z), ("m", m), ("u", u), ("j", np.ones(count, np.float32))):
        rig.write(name, v)
    rig.params(params)
    for k in range(20):
        if k % 5 == 0:
            rig.run(["sort"])
            assert int(rig.sim.read_words("active")[0]) == count - int(dead.sum())
        ref.step()
        rig.run("step")
    slots = rig.sim.read_words("slot")

    def by_slot(name):
        v = rig.read(name)
        out = np.empty_like(v)
        out[slots] = v
        return out

    for name in ("x", "y", "z"):
        np.testing.assert_allclose(by_slot(name), ref.p[name], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(by_slot("j"), ref.p["j"], rtol=1e-4, atol=1e-5)
    # the oracle never touches a massless slot either
    for name in oracle.FIELDS3:
        np.testing.assert_array_equal(by_slot(name)[dead].view(np.int32), ref.p[name][dead].view(np.int32),
                                      err_msg=name)
    mass = float(np.sum(m, dtype=np.float64))
    assert abs(float(np.sum(rig.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass


def _device_fields(rig):
    return {name: rig.read(name) for name in oracle.FIELDS3}


def _launches(rig, kernel):
    return rig.sim.kernel_time(kernel)[1]


def test_sort_3d_layers_bit_exact(gpu):
    """With P2G by layers (p2g=3) the sort lays each layer of 64 cells out rank-major: the keys,
    every field, the slot map and the active count are those of the oracle's restatement of that
    order, bit for bit, the massless slots last in slot order."""
    n = 48
    x, y, z = scenes.jittered_block3((2, 2, 2), (40, 30, 20), 2, scenes.JavaRandomFloats(21))
    count = len(x)
    order = oracle.JavaRandom(22).shuffle_order(count)
    f = {"x": x[order], "y": y[order], "z": z[order]}
    rng = np.random.default_rng(1)
    for name in oracle.FIELDS3[3:]:
        f[name] = rng.standard_normal(count).astype(np.float32)
    f["m"] = np.abs(f["m"]) + np.float32(0.1)
    dead = rng.random(count) < 0.2
    f["m"][dead] = 0.0
    ref = oracle.Mpm3(n, n, n, f, oracle.mpm3_params(1e-3, 0, 0, 0, 0, 1))
    rig = maker3(**LAYERS)(n, n, n, count)
    cases.load3(rig, f)
    rig.run(["sort"])
    perm = ref.sort(layers=True)
    np.testing.assert_array_equal(rig.sim.read_words("keys"), ref.keys())
    for name in oracle.FIELDS3:
        np.testing.assert_array_equal(rig.read(name).view(np.int32), ref.p[name].view(np.int32), err_msg=name)
    np.testing.assert_array_equal(rig.sim.read_words("slot"), perm)
    active = int(np.count_nonzero(~dead))
    assert int(rig.sim.read_words("active")[0]) == active
    np.testing.assert_array_equal(perm[active:], np.flatnonzero(dead))


def test_scatter_parity_3d_layers(gpu):
    """P2G by layers after a sort: each grid sum within FP32 reassociation (1e-5 max(1, |e|)) of
    the oracle's from the same particles, the mass to 1e-6. Then an advect through a strong grid
    velocity carries a fifth of the particles out of the cells the sort filed them under, and P2G
    (the moved ones by direct atomics) again matches the oracle from the device's particles."""
    n, count = 40, 60_000
    f = _state3(n, count, 7)
    params = oracle.mpm3_params(1e-3, 0, -50, 0, 400.0, 1)
    rig = maker3(**LAYERS)(n, n, n, count)
    cases.load3(rig, f)
    rig.params(params)
    rig.sim.timing(True)
    rig.run(["sort", "clear", "scatter"])
    assert _launches(rig, "p2g3_layer") == 1
    mass = float(np.sum(f["m"][f["m"] > 0], dtype=np.float64))
    for stage in ("sorted", "moved"):
        ref = oracle.Mpm3(n, n, n, _device_fields(rig), params)
        ref.clear()
        ref.scatter()
        for g in ("gm", "gmu", "gmv", "gmw"):
            _close(rig.read(g), ref.g[g], f"{stage} {g}")
        assert abs(float(np.sum(rig.read("gm"), dtype=np.float64)) - mass) <= 1e-6 * mass
        if stage == "sorted":
            keys_before = ref.keys()
            rng = np.random.default_rng(3)
            for g in ("gu", "gv", "gw"):
                rig.write(g, rng.uniform(-400, 400, n ** 3).astype(np.float32))
            rig.run(["advect", "clear", "scatter"])
            moved = float(np.mean(oracle.Mpm3(n, n, n, _device_fields(rig), params).keys() != keys_before))
            assert moved > 0.1, moved
    assert _launches(rig, "p2g3_layer") == 2
    rig.sim.timing(False)


def test_layers_fall_back_after_a_write_3d(gpu):
    """A host write of positions or masses (or a lod pass) leaves P2G by layers without a layout
    that covers every particle: until the next sort it takes the sorted-segmented schedule, and
    the result still matches the oracle."""
    n, count = 32, 20_000
    f = _state3(n, count, 5)
    params = oracle.mpm3_params(1e-3, 0, -50, 0, 400.0, 1)
    rig = maker3(**LAYERS)(n, n, n, count)
    cases.load3(rig, f)
    rig.params(params)
    rig.run(["sort"])
    m = rig.read("m")
    m[::3] = 0.25
    rig.write("m", m)
    rig.sim.timing(True)
    rig.run(["clear", "scatter"])
    assert _launches(rig, "p2g3_layer") == 0 and _launches(rig, "p2g3_seg") == 1
    ref = oracle.Mpm3(n, n, n, _device_fields(rig), params)
    ref.clear()
    ref.scatter()
    for g in ("gm", "gmu", "gmv", "gmw"):
        _close(rig.read(g), ref.g[g], g)
    rig.run(["sort", "clear", "scatter"])
    assert _launches(rig, "p2g3_layer") == 1
    rig.sim.timing(False)


def test_slot_writes_are_checked(gpu):
    """The slot map takes only a permutation of the slots; the active count is read-only."""
    rig = maker3()(16, 16, 16, 64)
    with pytest.raises(_native.InvalidArgument):
        rig.sim.write("slot", np.zeros(64, np.int32))
    with pytest.raises(_native.InvalidArgument):
        rig.sim.write("slot", np.arange(1, 65, dtype=np.int32))
    with pytest.raises(_native.InvalidArgument):
        rig.sim.write("active", np.zeros(1, np.int32))
    rig.sim.write("slot", np.arange(64, dtype=np.int32)[::-1].copy())
    np.testing.assert_array_equal(rig.sim.read_words("slot"), np.arange(64)[::-1])


def _sorter2(n, p, count):
    rig = maker2()(n, n, count)
    for k in ("x", "y", "u", "v", "m"):
        rig.write(k, p[k])
    rig.write("j", np.ones(count, np.float32))
    return rig


def test_sort_2d_by_cell_matches_host(gpu):
    """SortTest.theDeviceSortsByCell (test/particle/SortTest.java:51-61): 259^2 cells, one jittered
    particle a cell, shuffled; the starts are the host's cellStarts exactly and the order is the
    host's stable sort by Scatter.cell, bit for bit."""
    n = 260
    p = cases.jittered2(n - 1, 1, 21)
    order = oracle.JavaRandom(22).shuffle_order(len(p["x"]))
    p = {k: v[order] for k, v in p.items()}
    count = len(p["x"])
    rig = _sorter2(n, p, count)
    rig.run(["sort"])
    keys = oracle.mpm2_keys(p["x"], p["y"], n, n)
    perm = oracle.stable_order(keys)
    for k in ("x", "y", "u", "v", "m"):
        np.testing.assert_array_equal(rig.read(k).view(np.int32), p[k][perm].view(np.int32), err_msg=k)
    starts = rig.sim.read_words("starts")
    np.testing.assert_array_equal(starts, oracle.mpm2_cell_starts(p["x"][perm], p["y"][perm], n, n))


def test_second_sort_starts_clean(gpu):
    """SortTest.aSecondSortStartsFromZeroCounts (:68-81): a second sort of other particles."""
    n = 40
    first = cases.jittered2(n - 1, 3, 31)
    o1 = oracle.JavaRandom(32).shuffle_order(len(first["x"]))
    first = {k: v[o1] for k, v in first.items()}
    second = cases.jittered2(n - 1, 3, 33)
    o2 = oracle.JavaRandom(34).shuffle_order(len(second["x"]))
    second = {k: v[o2] for k, v in second.items()}
    count = len(first["x"])
    rig = _sorter2(n, first, count)
    rig.run(["sort"])
    for k in ("x", "y", "u", "v", "m"):
        rig.write(k, second[k])
    rig.run(["sort"])
    perm = oracle.stable_order(oracle.mpm2_keys(second["x"], second["y"], n, n))
    for k in ("x", "y", "u", "v", "m"):
        np.testing.assert_array_equal(rig.read(k).view(np.int32), second[k][perm].view(np.int32))
    np.testing.assert_array_equal(rig.sim.read_words("starts"),
                                  oracle.mpm2_cell_starts(second["x"][perm], second["y"][perm], n, n))


# ------------------------------------------------------------------ the 3D workload at full size

def test_config_two_cubes_full_size(gpu):
    """The 3D workload at its real shape: 8,388,608 particles in two cubes of a 256^3 box. The sort
    is bit-exact against the host's stable sort of the same keys; after a step in each P2G
    schedule the grid holds the particles' mass to 1e-6 relative, and the two schedules' grids
    agree within FP32 reassociation."""
    s = scenes.mpm3_two_cubes()
    count = len(s["x"])
    assert count == 8_388_608
    g, bulk, rho0 = scenes.mpm3_units(256)
    params = particles.params3(5e-5, 0, g, 0, bulk, rho0)
    sim = particles.ParticleSimulation3(256, 256, 256, count)
    z = np.zeros(count, np.float32)
    sim.load((s["x"], s["y"], s["z"]), (z, z, z), s["m"], params)
    keys_host = np.empty(count, np.int32)
    oracle.lib.orc_mpm3_keys(256, 256, 256, count, s["x"], s["y"], s["z"], s["m"], keys_host)
    perm = np.argsort(keys_host, kind="stable")
    sim.sort()
    np.testing.assert_array_equal(sim.sim.read_words("keys"), keys_host[perm])
    np.testing.assert_array_equal(sim.positions()[0].view(np.int32), s["x"][perm].view(np.int32))
    mass = float(np.sum(s["m"], dtype=np.float64))
    grids = []
    for p2g in (0, 1):
        sim.sim.set_option("p2g", p2g)
        sim.sim.run_pass("clear")
        sim.sim.run_pass("scatter")
        gm = sim.grid()
        grids.append(gm)
        assert abs(float(np.sum(gm[0], dtype=np.float64)) - mass) <= 1e-6 * mass
    for a, b, name in zip(grids[0], grids[1], ("gm", "gmu", "gmv", "gmw")):
        _close(b, a, name)
    # P2G by layers: the sort's layer-by-layer order bit for bit, and the same grid
    sim.sim.set_option("p2g", 3)
    sim.sort()
    sorted_keys = keys_host[perm]
    lperm = perm[oracle.mpm3_layer_order(sorted_keys, oracle.mpm3_inactive_key(256, 256, 256))]
    np.testing.assert_array_equal(sim.sim.read_words("keys"), keys_host[lperm])
    np.testing.assert_array_equal(sim.positions()[1].view(np.int32), s["y"][lperm].view(np.int32))
    sim.sim.run_pass("clear")
    sim.sim.run_pass("scatter")
    gm = sim.grid()
    assert abs(float(np.sum(gm[0], dtype=np.float64)) - mass) <= 1e-6 * mass
    for a, b, name in zip(grids[0], gm, ("gm", "gmu", "gmv", "gmw")):
        _close(b, a, name)
    sim.close()


def test_timing_one_kernel(gpu):
    """vxr_timing_only: event pairs around the named kernel's launches alone; vxr_timing times
    every kernel again."""
    x, y, z = scenes.jittered_block3((2, 2, 2), (6, 8, 6), 8, scenes.JavaRandomFloats(3))
    count = len(x)
    sim = particles.ParticleSimulation3(16, 16, 16, count)
    zero = np.zeros(count, np.float32)
    sim.load((x, y, z), (zero, zero, zero), np.full(count, 0.125, np.float32),
             particles.params3(1e-3, 0, -100, 0, 500, 1))
    sim.sim.timing_only("p2g3_seg")
    sim.advance(6, sort_every=5)
    ms, n = sim.sim.kernel_time("p2g3_seg")
    assert n == 6 and ms > 0
    assert sim.sim.kernel_time("grid3") == (0.0, 0)
    sim.sim.timing(True)
    sim.advance(2, sort_every=5)
    assert sim.sim.kernel_time("p2g3_seg")[1] == 2 and sim.sim.kernel_time("grid3")[1] == 2
    sim.sim.timing(False)
    with pytest.raises(_native.InvalidArgument):
        sim.sim.timing_only("")
    sim.close()


def test_runner_api_3d(gpu):
    """ParticleSimulation3's surface: load, advance with a sort cadence, read-backs, active count."""
    x, y, z = scenes.jittered_block3((2, 2, 2), (6, 8, 6), 8, scenes.JavaRandomFloats(3))
    count = len(x)
    sim = particles.ParticleSimulation3(16, 16, 16, count)
    zero = np.zeros(count, np.float32)
    sim.load((x, y, z), (zero, zero, zero), np.full(count, 0.125, np.float32),
             particles.params3(1e-3, 0, -100, 0, 500, 1))
    sim.advance(12, sort_every=5)
    assert sim.steps() == 12 and sim.active() == count
    j = sim.compression()
    assert j.shape == (count,) and np.all(np.isfinite(j))
    pos = sim.positions()
    assert all(np.all((p >= 1) & (p <= 14)) for p in pos)
    with pytest.raises(_native.Unsupported):
        particles.ParticleSimulation3(16, 16, 16, 10, plastic=True)
    with pytest.raises(_native.InvalidArgument):
        sim.sim.write("x", np.zeros(count + 1, np.float32))
    sim.close()
