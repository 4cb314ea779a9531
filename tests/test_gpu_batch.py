"""Many 2D scenes in one launch, the fixed-point P2G, hipGraph replay and the autotuner.

The fixed-point P2G makes a 2D step deterministic, so here parity is bit for bit throughout: a
single scene against the oracle's fixed-point restatement (orc_mpm2_scatter_fixed) over whole
steps and sorts, every scene of a batch against its own single-scene run, graph replays against
eager runs, and tuned runs against untuned ones."""
import numpy as np
import pytest

import oracle
from gpu_rigs import GpuSwe
from vexhip import _native, particles, scenes

pytestmark = pytest.mark.gpu

FIELDS = oracle.FIELDS2


def _scene(n, block, count, seed, speed=0.0):
    x, y = scenes.jittered_block2((1, 1), block, 4, scenes.JavaRandomFloats(seed))
    x, y = x[:count].copy(), y[:count].copy()
    r = np.random.default_rng(seed)
    u = (speed * r.standard_normal(count)).astype(np.float32)
    v = (speed * r.standard_normal(count)).astype(np.float32)
    return {"x": x, "y": y, "u": u, "v": v, "m": np.full(count, 0.25, np.float32)}


def _params(n):
    h = 1.0 / (n - 1)
    return oracle.mpm2_params(1e-4 * 127.0 / (n - 1), 0, -9.8 / h, 10.0 / h ** 2, 1)


def _single(n, sc, params, **opts):
    sim = _native.Sim(_native.KIND_MPM2, (n, n, 1), len(sc["x"]))
    sim.set_option("p2g", 2)
    for k, v in opts.items():
        sim.set_option(k, v)
    count = len(sc["x"])
    for name in ("x", "y", "u", "v", "m"):
        sim.write(name, sc[name])
    sim.write("j", np.ones(count, np.float32))
    sim.write("params", params)
    return sim


def _assert_fields(read, expect, what):
    for name in FIELDS:
        np.testing.assert_array_equal(read(name).view(np.int32), expect[name].view(np.int32),
                                      err_msg=f"{what}: {name}")


@pytest.mark.parametrize("run", [1, 2, 4])
def test_fixed_p2g_bit_exact_with_oracle(gpu, run):
    """One pass at a time, then 30 steps with a sort every 10: every particle field, the decoded
    grid and the starts are the oracle's bits."""
    n = 64
    sc = _scene(n, (20, 30), 2400, 5, speed=20.0)
    params = _params(n)
    print(f"[batch] fixed p2g single scene run={run}", flush=True)
    dev = _single(n, sc, params, run=run)
    ref = oracle.Mpm2(n, n, sc, params, fixed=True)
    for p in ("clear", "scatter", "grid"):
        dev.run_pass(p)
        getattr(ref, p)()
    for g in ("gm", "gmu", "gmv", "gu", "gv"):
        np.testing.assert_array_equal(dev.read(g).view(np.int32), ref.g[g].view(np.int32), err_msg=g)
    dev.run_pass("advect")
    ref.advect()
    _assert_fields(dev.read, ref.p, "first step")
    dev.advance(30, 10, 1)
    ref.steps = 1
    ref.advance(30, 10)
    _assert_fields(dev.read, ref.p, "30 steps")
    # the whole grid of the last step: the chunks clear and grid skipped hold what the oracle's
    # full passes give them
    for g in ("gm", "gmu", "gmv", "gu", "gv"):
        np.testing.assert_array_equal(dev.read(g).view(np.int32), ref.g[g].view(np.int32), err_msg=f"30 steps: {g}")
    dev.run(_native.GROUP_SORT)
    ref.sort()
    _assert_fields(dev.read, ref.p, "sorted")
    np.testing.assert_array_equal(dev.read_words("starts"), oracle.mpm2_cell_starts(ref.p["x"], ref.p["y"], n, n))


@pytest.mark.parametrize("case", ["heavy", "fast"])
def test_fixed_p2g_wide_amounts(gpu, case):
    """Rounded amounts of 2^40 and more: every particle at rho0 = 1000 (mass 250 at 4 a cell, the
    same motion as rho0 = 1 with mass 0.25), or one particle in 97 at 8000 nodes/s. The blocks that
    hold such amounts sum in whole 64-bit integers, the others in split 32-bit pairs; the oracle's
    bits either way, over one pass and 20 steps with a sort every 10."""
    n = 64
    sc = _scene(n, (20, 30), 2400, 5, speed=20.0)
    h = 1.0 / (n - 1)
    if case == "heavy":
        sc["m"] = np.full(len(sc["x"]), 250.0, np.float32)
        params = oracle.mpm2_params(1e-4 * 127.0 / (n - 1), 0, -9.8 / h, 10.0 / h ** 2, 1000)
    else:
        sc["u"][::97] = np.float32(8000.0)
        params = _params(n)
    print(f"[batch] fixed p2g wide amounts: {case}", flush=True)
    dev = _single(n, sc, params)
    ref = oracle.Mpm2(n, n, sc, params, fixed=True)
    for p in ("clear", "scatter", "grid"):
        dev.run_pass(p)
        getattr(ref, p)()
    for g in ("gm", "gmu", "gmv", "gu", "gv"):
        np.testing.assert_array_equal(dev.read(g).view(np.int32), ref.g[g].view(np.int32), err_msg=g)
    # the largest amounts pass 2^40 (mass in units of 2^-40, momentum of 2^-32; a particle's largest
    # weight is at least 1/4)
    assert max(float(sc["m"].max()) * 2.0 ** 40, float((sc["m"] * np.abs(sc["u"])).max()) * 2.0 ** 32) / 4 >= 2.0 ** 40
    dev.run_pass("advect")
    ref.advect()
    dev.advance(20, 10, 1)
    ref.steps = 1
    ref.advance(20, 10)
    _assert_fields(dev.read, ref.p, f"{case}: 20 steps")
    for g in ("gm", "gmu", "gmv", "gu", "gv"):
        np.testing.assert_array_equal(dev.read(g).view(np.int32), ref.g[g].view(np.int32), err_msg=f"{case}: {g}")
    dev.close()


@pytest.mark.parametrize("heavy", [False, True], ids=["pairs", "wide"])
def test_fixed_p2g_whole_chunks(gpu, heavy):
    """A scene of 270,112 particles (264 chunks of 1,024), so that P2G takes whole chunks a block
    (four particles a thread) rather than the quarter chunks of small scenes: with mass 0.25 every
    block adds split pairs, with rho0 = 1000 and mass 250 every block keeps the double-precision
    fold. One pass and 10 steps with a sort, bit-exact with the oracle."""
    n = 512
    sc = _scene(n, (184, 367), 270112, 21, speed=10.0)
    assert len(sc["x"]) == 270112
    h = 1.0 / (n - 1)
    params = _params(n)
    if heavy:
        sc["m"] = np.full(len(sc["x"]), 250.0, np.float32)
        params = oracle.mpm2_params(1e-4 * 127.0 / (n - 1), 0, -9.8 / h, 10.0 / h ** 2, 1000)
    print(f"[batch] fixed p2g whole chunks, heavy={heavy}", flush=True)
    dev = _single(n, sc, params)
    ref = oracle.Mpm2(n, n, sc, params, fixed=True)
    for p in ("clear", "scatter", "grid"):
        dev.run_pass(p)
        getattr(ref, p)()
    for g in ("gm", "gmu", "gmv", "gu", "gv"):
        np.testing.assert_array_equal(dev.read(g).view(np.int32), ref.g[g].view(np.int32), err_msg=g)
    dev.run_pass("advect")
    ref.advect()
    dev.advance(10, 10, 1)
    ref.steps = 1
    ref.advance(10, 10)
    _assert_fields(dev.read, ref.p, f"heavy={heavy}: 10 steps")
    dev.close()


def _batch_case():
    # grids of several sizes, counts off the 1024 padding, one scene of a few particles
    spec = [(64, (20, 30), 2400, 11), (16, (4, 6), 5, 12), (128, (40, 70), 11000, 13), (40, (12, 20), 960, 14),
            (96, (30, 51), 6001, 15), (64, (10, 10), 400, 16)]
    sizes = [(n, n) for n, _, _, _ in spec]
    counts = [c for _, _, c, _ in spec]
    sc = [_scene(n, b, c, seed, speed=5.0) for n, b, c, seed in spec]
    params = [_params(n) for n, _, _, _ in spec]
    return sizes, counts, sc, params


def test_batch_scenes_equal_single_runs(gpu):
    sizes, counts, sc, params = _batch_case()
    print(f"[batch] {len(sizes)} scenes vs single runs", flush=True)
    batch = particles.ParticleBatch(sizes, counts)
    batch.load(sc, params)
    batch.advance(25, 10)
    for k, (size, s, p) in enumerate(zip(sizes, sc, params)):
        single = _single(size[0], s, p)
        single.advance(25, 10, 0)
        got = batch.scene(k)
        _assert_fields(lambda name: got[name], {name: single.read(name) for name in FIELDS}, f"scene {k}")
        grid = batch.scene_grid(k)
        for g in ("gm", "gmu", "gmv", "gu", "gv"):
            np.testing.assert_array_equal(grid[g].view(np.int32), single.read(g).view(np.int32),
                                          err_msg=f"scene {k}: {g}")
        np.testing.assert_array_equal(batch.scene_starts(k), single.read_words("starts"), err_msg=f"scene {k}")
        single.close()
    # and the two smallest scenes against the oracle itself
    for k in (1, 3):
        n = sizes[k][0]
        ref = oracle.Mpm2(n, n, sc[k], params[k], fixed=True)
        ref.advance(25, 10)
        got = batch.scene(k)
        _assert_fields(lambda name: got[name], ref.p, f"scene {k} vs oracle")
    batch.close()


def _clusters(n, count, seed, size=5.0):
    """count particles in two small clusters at opposite corners of an n x n grid, so that the
    sorted keys jump over many chunks of 256 cells and the last particle sits far from the end."""
    r = np.random.default_rng(seed)
    half = count // 2
    lo = r.uniform(1.0, 1.0 + size, (half, 2))
    hi = r.uniform(0.6 * (n - 1), 0.7 * (n - 1), (count - half, 2))
    xy = np.concatenate([lo, hi]).astype(np.float32)
    return {"x": xy[:, 0].copy(), "y": xy[:, 1].copy(), "u": np.zeros(count, np.float32),
            "v": np.zeros(count, np.float32), "m": np.full(count, 0.25, np.float32)}


def test_batch_starts_sparse_scenes(gpu):
    """Every scene's cellStarts after a sort against the oracle's, on scenes whose particles leave
    long runs of empty chunks between them and after them: one particle, exactly kBatchAlign
    particles, a scene of one chunk, and one whose first chunk holds more sorted keys than the
    LDS window of 2048 (its cells' bounds are searched in global memory)."""
    spec = [(256, 1024, 31, 5.0), (64, 1, 32, 5.0), (512, 3000, 33, 5.0), (8, 40, 34, 5.0), (128, 2048, 35, 5.0),
            (64, 6000, 36, 1.5)]
    sizes = [(n, n) for n, _, _, _ in spec]
    sc = [_clusters(n, c, seed, size) for n, c, seed, size in spec]
    print(f"[batch] sparse starts over {len(spec)} scenes", flush=True)
    batch = particles.ParticleBatch(sizes, [c for _, c, _, _ in spec])
    batch.load(sc, [_params(n) for n, _, _, _ in spec])
    batch.sim.run(_native.GROUP_SORT)
    for k, (n, _, _, _) in enumerate(spec):
        got = batch.scene(k, names=("x", "y"))
        np.testing.assert_array_equal(batch.scene_starts(k), oracle.mpm2_cell_starts(got["x"], got["y"], n, n),
                                      err_msg=f"scene {k}")
    batch.close()


def test_graph_replays_match_eager(gpu):
    """The same batch stepped eagerly and by replayed graphs, several calls of different lengths
    (so graphs with and without a leading sort, of several step counts, each replayed more than
    once), with reads and synchronisations between: the same bits after every call."""
    sizes, counts, sc, params = _batch_case()
    print("[batch] graph vs eager", flush=True)
    eager = particles.ParticleBatch(sizes, counts)
    graph = particles.ParticleBatch(sizes, counts)
    graph.sim.set_option("graph", 1)
    for b in (eager, graph):
        b.load(sc, params)
    for count in (7, 10, 13, 10, 25, 3, 20):
        eager.advance(count, 10)
        graph.advance(count, 10)
        eager.finish()
        graph.finish()
        for name in FIELDS:
            np.testing.assert_array_equal(graph.sim.read(name).view(np.int32), eager.sim.read(name).view(np.int32),
                                          err_msg=f"after {graph.steps()} steps: {name}")
    eager.close()
    graph.close()


def test_autotune_keeps_results(gpu):
    sizes, counts, sc, params = _batch_case()
    print("[batch] autotune", flush=True)
    plain = particles.ParticleBatch(sizes, counts)
    tuned = particles.ParticleBatch(sizes, counts)
    for b in (plain, tuned):
        b.load(sc, params)
    chosen = tuned.sim.autotune(2)
    assert chosen["run"] in (1, 2, 4) and 4 <= chosen["radix_bits"] <= 8
    assert chosen["blocks"] in (0, 1024, 2048, 4096, 8192, 16384)
    for b in (plain, tuned):
        b.advance(22, 10)
    for name in FIELDS:
        np.testing.assert_array_equal(tuned.sim.read(name).view(np.int32), plain.sim.read(name).view(np.int32),
                                      err_msg=name)
    plain.close()
    tuned.close()


def test_batch_refuses_bad_scenes(gpu):
    with pytest.raises(_native.InvalidArgument):
        particles.ParticleBatch([(3, 8)], [10])
    with pytest.raises(_native.InvalidArgument):
        particles.ParticleBatch([(8, 8)], [0])
    b = particles.ParticleBatch([(8, 8), (16, 16)], [10, 20])
    with pytest.raises(_native.InvalidArgument):
        b.sim.set_option("p2g", 1)
    with pytest.raises(ValueError):
        b.load([{"x": np.zeros(3)}], [oracle.mpm2_params(1e-3, 0, 0, 0, 1)])
    b.close()


def test_mpm3_graph_and_autotune(gpu):
    """3D: a tuned handle replaying graphs tracks the oracle over 20 steps with two sorts, as the
    eager default does (float atomics: within 1e-5)."""
    n = 32
    x, y, z = scenes.jittered_block3((3, 3, 3), (10, 14, 9), 8, scenes.JavaRandomFloats(5))
    count = len(x)
    m = np.full(count, 0.125, np.float32)
    params = oracle.mpm3_params(5e-4, 0, -200.0, 0, 2000.0, 1)
    ref = oracle.Mpm3(n, n, n, {"x": x, "y": y, "z": z, "m": m}, params)
    print("[batch] 3D graph + autotune", flush=True)
    sim = particles.ParticleSimulation3(n, n, n, count)
    z0 = np.zeros(count, np.float32)
    sim.load((x, y, z), (z0, z0, z0), m, params)
    sim.sim.set_option("graph", 1)
    chosen = sim.sim.autotune(2)
    assert chosen["run"] in (1, 2, 4)
    sim.advance(20, sort_every=10)
    for k in range(20):
        if k % 10 == 0:
            ref.sort()
        ref.step()
    got = sim.positions()
    for a, name in zip(got, "xyz"):
        np.testing.assert_allclose(a, ref.p[name], rtol=1e-5, atol=1e-5)
    sim.close()


def test_swe_autotune_keeps_parity(gpu):
    nx, ny = 131, 77
    r = np.random.default_rng(3)
    h = (1 + 0.3 * r.random(nx * ny)).astype(np.float32)
    z = np.zeros_like(h)
    params = oracle.swe_params(9.81, 0.45, 1.0, 1.0)
    first = oracle.swe_fastest_wave(h, z, z, 9.81, 1e-6)
    cpu = oracle.SweStepper(nx, ny, (0, 0, 0, 0))
    dev = GpuSwe(nx, ny)
    for b in (cpu, dev):
        b.load(h, z, z, params, first)
        b.write_budget(2**31 - 1)
        b.step(5)
    print("[batch] swe autotune", flush=True)
    chosen = dev.sim.autotune(2)
    assert chosen["rows"] in (2, 4, 8, 16, 32, 48, 64, 128)
    for b in (cpu, dev):
        b.step(10)
    a, g = cpu.read(), dev.read()
    for f in range(3):
        np.testing.assert_array_equal(g[f].view(np.int32), a[f].view(np.int32))
    assert dev.clamped() == cpu.clamped() and dev.fastest_wave_bits() == cpu.fastest_wave_bits()
