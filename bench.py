"""Benchmark of the hot path on one MI355X.

Default workload: the 3D MLS-MPM/APIC substep at 8,388,608 particles in a 256^3 box (two seeded
cubes of water, 8 particles a cell, dt = 5e-5), sorted-segmented P2G with the particles
re-sorted by cell every 10 substeps (the sort's cost is inside the timed region). A "step" is
one substep: clear, P2G, grid update, G2P. Prints one JSON line; beside the headline it carries
the dominant kernel's roofline from HIP events on the library's own stream, the HLL step on the
4096^2 shallow-water scene, and the CPU oracle timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "vexelray-sim-fluid_amd"))

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
SORT_EVERY = 10

# Algorithmic bytes per unit (particle, node, element) of each kernel: what the pass must move
# through HBM at least, as derived in the design notes.
BYTES = {
    "p2g3_seg": ("particle", 68), "p2g3_atomic": ("particle", 68), "p2g3_layer": ("particle", 68),
    "p2g3_fixed": ("particle", 68),
    "g2p3_fast": ("particle", 84), "g2p3_tiled": ("particle", 84), "g2p3": ("particle", 84),
    "grid3": ("node", 28), "mpm3_clear": ("node", 16),
    "keys3": ("particle", 20), "radix_hist": ("element", 4), "radix_scatter": ("element", 16),
    "radix_scan": ("none", 0), "permute": ("particle", 148),
    "hll_step": ("cell", 24),
    # 3D tension, dense over the grid: a blur reads one colour (or gm) and writes one; the stress
    # reads the colour and writes six components; the grid with tension reads four sums, the
    # colour and six components and writes three velocities
    "tension3_blur": ("node", 8), "tension3_stress": ("node", 28), "grid3_tension": ("node", 56),
    # batched 2D: 10 particle fields in for P2G; x y j in and 9 fields out for G2P; three 64-bit
    # sums in and five floats out per node for the grid update; three 64-bit sums cleared per node
    "batch2_p2g": ("particle", 40), "batch2_g2p": ("particle", 48), "batch2_grid": ("node", 44),
    "batch2_clear": ("node", 24), "batch2_keys": ("particle", 16), "batch2_starts": ("none", 0),
}


# Passes that visit only the flagged bricks (3D) or the chunks P2G reached (batched 2D): their
# BYTES are per node of the whole grid, an upper bound on what they move, so no rate is derived.
SPARSE = {"mpm3_clear", "fixed3_decode", "grid3", "grid3_tension", "tension3_stress", "tension3_blur", "batch2_grid", "batch2_clear"}


def kernel_table(sim, names, units, bytes_override=None):
    """Per-kernel launches and mean duration from the library's HIP-event timers (on its own
    stream), with the rate algorithmic bytes / duration for dense passes. A sparse pass reports
    its dense-grid bytes as a bound and no rate; a rate above the HBM peak (a working set served
    from cache) is reported as none, with the reason."""
    out = {}
    for k in names:
        ms, n = sim.kernel_time(k)
        if n == 0:
            continue
        unit, per = (bytes_override or {}).get(k, BYTES.get(k, ("none", 0)))
        count = units.get(unit, 0)
        avg_ms = ms / n
        entry = {"launches": n, "avg_us": round(avg_ms * 1e3, 3), "total_ms": round(ms, 3)}
        if k in SPARSE:
            entry["dense_bytes_bound"] = per * count
            entry["achieved_gbs"] = None
            entry["note"] = "visits flagged bricks or chunks only: no rate from dense-grid bytes"
        else:
            gbs = per * count / (avg_ms * 1e-3) / 1e9 if per and avg_ms > 0 else None
            entry["algorithmic_bytes"] = per * count
            if gbs is not None and gbs > PEAK_HBM_GBS:
                entry["note"] = f"{gbs:.0f} GB/s of algorithmic bytes exceeds the HBM peak: served from cache"
                gbs = None
            entry["achieved_gbs"] = None if gbs is None else round(gbs, 1)
        out[k] = entry
    return out


P2G_KERNEL = {0: "p2g3_atomic", 1: "p2g3_seg", 2: "p2g3_fixed", 3: "p2g3_layer"}


def bench_mpm3(steps, warmup, p2g, g2p, run, tension=False, lod_level=None, p2g_ints=1):
    """The 3D substep on the two-cubes scene; with ``tension`` Flip3Step's tension step (six
    blurs, the stress and the grid with the capillary force), sigma = 0.01 B 8, the scale
    Tension3Test gives a drop of radius 8 nodes. With ``lod_level`` every cell's group of 8 is
    first thinned to that many slots by 60 lod calls (Flip3.lod), so the sorts move the emptied
    slots behind the rest and P2G and G2P stop at the last active one."""
    from vexhip import particles, scenes

    s = scenes.mpm3_two_cubes()
    count = len(s["x"])
    g, bulk, rho0 = scenes.mpm3_units(256)
    words = particles.params3(5e-5, 0, g, 0, bulk, rho0, 0.01 * bulk * 8 if tension else 0.0)
    sim = particles.ParticleSimulation3(256, 256, 256, count, tension=tension)
    sim.sim.set_option("p2g", p2g)
    sim.sim.set_option("g2p", g2p)
    sim.sim.set_option("run", run)
    sim.sim.set_option("p2g_ints", p2g_ints)
    z = np.zeros(count, np.float32)
    sim.load((s["x"], s["y"], s["z"]), (z, z, z), s["m"], words)
    if lod_level is not None:
        sim.sim.write("target", np.full(count // 8, np.float32(lod_level)))
        sim.sim.run_pass("lod", 60)
    sim.advance(warmup, sort_every=SORT_EVERY)
    sim.finish()
    # HIP events between kernels keep the next launch from queueing behind the running one, so the
    # timed region carries an event pair around the dominant kernel (P2G) alone
    dominant = P2G_KERNEL[p2g]
    sim.sim.timing_only(dominant)
    t0 = time.perf_counter()
    sim.advance(steps, sort_every=SORT_EVERY)
    sim.finish()
    t1 = time.perf_counter()
    names = ["mpm3_clear", "p2g3_seg", "p2g3_atomic", "p2g3_layer", "p2g3_fixed", "fixed3_decode", "grid3",
             "g2p3_fast", "g2p3_tiled", "g2p3",
             "keys3", "radix_hist", "radix_scan", "radix_scatter", "layers3", "permute", "tension3_blur",
             "tension3_stress", "grid3_tension"]
    units = {"particle": count, "node": 256 ** 3, "element": count}
    live = kernel_table(sim.sim, [dominant], units)
    # every other kernel's figures from two more sort periods with all of them timed
    sim.sim.timing(True)
    sim.advance(2 * SORT_EVERY, sort_every=SORT_EVERY)
    sim.finish()
    table = kernel_table(sim.sim, names, units)
    for k, v in live.items():
        table[k] = dict(v, timed_region=True)
    sim.sim.timing(False)
    mass = float(np.sum(sim.grid()[0], dtype=np.float64))
    if lod_level is not None:
        active = sim.active()
        sim.close()
        return active, t1 - t0, table, mass / float(np.sum(s["m"], dtype=np.float64)) - 1.0
    sim.close()
    return count, t1 - t0, table, mass / float(np.sum(s["m"], dtype=np.float64)) - 1.0


def bench_hll(steps, n=4096, fast=False, rows=None):
    """The HLL step on the seeded dam-and-bumps scene; with `fast` the hll_fast option (refined
    hardware reciprocals and square roots, within 1e-5 of the reference's bits after 100 steps)."""
    from vexhip import patch, scenes

    h, hu, hv = scenes.swe_dam_bumps(n, seed=0)
    sim = patch.PatchSimulation(n, n)
    sim.sim.set_option("hll_fast", 1 if fast else 0)
    params = patch.params(9.81, 0.45, 1.0 / n, 1.0 / n)
    sim.load(h, hu, hv, params)
    # rows a wave marches: the default, 32, unless given (untuned here so that a profile of this
    # command times only the measured launches)
    if rows is not None:
        sim.sim.set_option("rows", rows)
    rows = sim.sim.option("rows")
    sim.write_budget(2**31 - 1)
    sim.dispatch(10)
    sim.sim.finish()
    t0 = time.perf_counter()
    sim.dispatch(steps)
    sim.sim.finish()
    t1 = time.perf_counter()
    # the kernel's own time from as many steps again with an event pair around each launch (the
    # step is one kernel, and events between launches keep each from queueing behind the last)
    sim.sim.timing(True)
    sim.dispatch(steps)
    sim.sim.finish()
    table = kernel_table(sim.sim, ["hll_step"], {"cell": n * n})
    sim.close()
    return n * n * steps / (t1 - t0), table, rows


def bench_batch2(steps, scenes=256, sort_every=SORT_EVERY, graph=True, tune=True):
    """The batched 2D workload: 256 seeded scenes (counts log-uniform in [4096, 262144], grids of
    64^2 .. 512^2), every scene stepped by one launch per pass; particle-substeps/s over all."""
    from vexhip import particles, scenes as sc_mod

    layouts = sc_mod.mpm2_batch_layouts(scenes)
    loaded = [sc_mod.mpm2_batch_scene(l, 1000 + k) for k, l in enumerate(layouts)]
    sizes = [(l["n"], l["n"]) for l in layouts]
    counts = [l["count"] for l in layouts]
    batch = particles.ParticleBatch(sizes, counts)
    batch.load([x[0] for x in loaded], [x[1] for x in loaded])
    chosen = batch.sim.autotune(2) if tune else None
    batch.sim.set_option("graph", 1 if graph else 0)
    batch.advance(sort_every, sort_every)  # warm: builds the graphs
    batch.finish()
    t0 = time.perf_counter()
    batch.advance(steps, sort_every)
    batch.finish()
    t1 = time.perf_counter()
    total = sum(counts)
    # per-kernel times from an eager stretch (graphs are replayed only with timing off)
    batch.sim.timing(True)
    batch.advance(2 * sort_every, sort_every)
    names = ["batch2_clear", "batch2_p2g", "batch2_grid", "batch2_g2p", "batch2_keys", "radix_hist", "radix_scan",
             "radix_scatter", "permute", "batch2_starts"]
    nodes = sum(n * n for n, _ in sizes)
    # the 2D permutation moves 10 fields: index + 10 gathers + 10 stores
    table = kernel_table(batch.sim, names, {"particle": total, "node": nodes, "element": int(batch.poff[-1])},
                         {"permute": ("element", 84)})
    batch.sim.timing(False)
    grid = batch.sim.read("gm")
    mass = float(np.sum(grid, dtype=np.float64)) / (0.25 * total) - 1.0
    batch.close()
    return {"value": round(total * steps / (t1 - t0), 1), "unit": "particle-substeps/s", "steps": steps,
            "ms_per_step": round((t1 - t0) / steps * 1e3, 4), "scenes": scenes, "particles": total,
            "sort_every": sort_every, "graph": graph, "autotuned": chosen, "grid_mass_drift": mass,
            "kernels": table,
            "workload": "batched 2D MLS-MPM: 256 seeded scenes, counts log-uniform in [4096, 262144], grids "
                        "64^2..512^2, water columns at 4 per cell, fixed-point P2G, sort every 10"}


def bench_small2(steps, graph, ppc=4):
    """The 2D reference-sized scene alone: a 32x64-cell column on 128^2 at `ppc` a cell -- 8,192
    particles at 4, the launch-bound case that graphs address, and 65,536 at 32."""
    from vexhip import particles, scenes as sc_mod

    s = sc_mod.mpm2_dam_block(ppc=ppc)
    count = len(s["x"])
    h = 1.0 / 127
    words = particles.params2(1e-4, 0, -9.8 / h, 10.0 / h ** 2, 1.0)
    sim = particles.ParticleSimulation(128, 128, count)
    z = np.zeros(count, np.float32)
    sim.load(s["x"], s["y"], z, z, s["m"], words)
    sim.sim.set_option("p2g", 2)
    sim.sim.set_option("graph", 1 if graph else 0)
    sim.advance(SORT_EVERY, SORT_EVERY)
    sim.finish()
    t0 = time.perf_counter()
    sim.advance(steps, SORT_EVERY)
    sim.finish()
    t1 = time.perf_counter()
    sim.close()
    return {"value": round(count * steps / (t1 - t0), 1), "us_per_step": round((t1 - t0) / steps * 1e6, 2),
            "graph": graph}


def bench_surrogate(iters, n=1024):
    """The convolutional surrogate of 8 HLL steps at 1024^2: trained here (deterministically) on one
    pair from each of 2,000 seeded dam-and-bumps trajectories, judged against the FP64 HLL on 200
    held-out seeds; its time per application against the FP32 HLL's for 8 steps."""
    import torch

    from vexhip import patch, scenes as sc_mod, surrogate

    t0 = time.perf_counter()
    model, losses = surrogate.train(n=n, seeds=range(2000), iters=iters)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    quality = surrogate.evaluate(model, n=n, seeds=range(10000, 10200))
    hll = surrogate.Hll64(n)
    hll.scene(0)
    hll.step(64)
    state = hll.state32()
    out = model.forward(*state)
    model.advance(*state, out=out)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    reps = 20
    # an application as the fallback rule runs it: the network, the device check, and the eight
    # HLL steps queued behind it, gated on its flag (each exits at once when the check passes)
    fallbacks_before = model.fallbacks()
    ev[0].record()
    for _ in range(reps):
        model.advance(*state, out=out)
    ev[1].record()
    torch.cuda.synchronize()
    app_ms = ev[0].elapsed_time(ev[1]) / reps
    timed_fallbacks = model.fallbacks() - fallbacks_before
    ev[0].record()
    for _ in range(reps):
        model.forward(*state, out=out)
    ev[1].record()
    layer = (64, 64)
    ev[2].record()
    for _ in range(reps):
        model._conv(1, model.acts[0], model.acts[1], True)
    ev[3].record()
    torch.cuda.synchronize()
    fwd_ms = ev[0].elapsed_time(ev[1]) / reps
    conv_ms = ev[2].elapsed_time(ev[3]) / reps
    ev[0].record()
    for _ in range(reps):
        hll.step(surrogate.STEPS)
    ev[1].record()
    torch.cuda.synchronize()
    hll64_ms = ev[0].elapsed_time(ev[1]) / reps
    h, hu, hv = sc_mod.swe_dam_bumps(n, seed=0)
    sim = patch.PatchSimulation(n, n)
    params = patch.params(9.81, 0.45, 1.0 / n, 1.0 / n)
    sim.load(h, hu, hv, params)
    hll_rows = sim.sim.autotune(2)["rows"]  # the HLL at its best launch shape for this size
    sim.load(h, hu, hv, params)
    sim.write_budget(2**31 - 1)
    sim.dispatch(surrogate.STEPS)
    sim.sim.finish()
    sim.sim.timing(True)
    sim.dispatch(surrogate.STEPS * reps)
    ms, launches = sim.sim.kernel_time("hll_step")
    sim.close()
    hll32_ms = ms / launches * surrogate.STEPS
    flop = 2.0 * n * n * 9 * sum(ci * co for ci, co in ((3, 64), (64, 64), (64, 64), (64, 64), (64, 64), (64, 3)))
    conv_flop = 2.0 * n * n * 9 * layer[0] * layer[1]
    cells = n * n * surrogate.STEPS
    return {"workload": f"residual 3x3 conv net (16->64, 4x 64->64, 64->16), bf16 MFMA, FP32 accumulate, one "
                        f"application = {surrogate.STEPS} HLL steps of fixed dt at {n}^2; trained on 2,000 "
                        f"dam-and-bumps trajectories (seeds 0-1999), judged on seeds 10000-10199 vs FP64 HLL",
            "train_iters": iters, "train_seconds": round(train_s, 1), "final_loss": losses[-1],
            "quality": quality,
            "surrogate_ms_per_application": round(app_ms, 4),
            "surrogate_application": "network + device check + the 8 gated fallback HLL launches; "
                                     f"{timed_fallbacks} of the {reps} timed applications fell back",
            "surrogate_forward_ms": round(fwd_ms, 4),
            "surrogate_cell_updates_per_s": round(cells / (app_ms * 1e-3), 1),
            "hll_fp32_ms_per_8_steps": round(hll32_ms, 4), "hll_fp32_rows": hll_rows,
            "hll_fp32_cell_updates_per_s": round(cells / (hll32_ms * 1e-3), 1),
            "hll_fp64_ms_per_8_steps": round(hll64_ms, 4),
            "surrogate_speed_vs_hll_fp32": round(hll32_ms / app_ms, 4),
            "surrogate_forward_speed_vs_hll_fp32": round(hll32_ms / fwd_ms, 4),
            "roofline": {"kernel": "conv3x3_rows 64->64", "bound": "mfma",
                         "achieved": round(conv_flop / (conv_ms * 1e-3) / 1e12, 1), "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": round(conv_flop / (conv_ms * 1e-3) / 1e12 / 2500.0, 4),
                         "avg_us": round(conv_ms * 1e3, 2)},
            "forward_tflops": round(flop / (fwd_ms * 1e-3) / 1e12, 1)}


def oracle_native():
    """Builds the oracle for the host it runs on (g++ -O3 -march=native -ffp-contract=off
    -fopenmp, oracle/Makefile's `native` target) into a temporary directory and points the
    oracle's loader at it, before the oracle is first imported. Returns the build's description;
    when the build fails the prebuilt x86-64-v3 library is timed instead, and the description
    says so."""
    import subprocess
    import tempfile

    out = tempfile.mkdtemp(prefix="vexel_oracle_")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"DIR={out}"],
                       capture_output=True, text=True)
    if r.returncode == 0:
        os.environ["VEXEL_ORACLE_LIB"] = os.path.join(out, "liboracle_native.so")
        return "g++ -O3 -march=native -ffp-contract=off -fopenmp (oracle/Makefile native)"
    return "prebuilt g++ -O3 -march=x86-64-v3 -ffp-contract=off (native build failed: " + r.stderr[-200:] + ")"


def _median_run(run, reps=5):
    """One warm-up call of `run`, then the median wall time of `reps` more."""
    import statistics

    run()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def _omp_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "16") or 16)


def cpu_baseline_hll(n, steps_per_run=1):
    """The oracle's HLL step (ShallowWater.kernel restated, one call per cell as the reference's
    CPU Stepper makes them) on the same seeded dam-and-bumps scene: `steps_per_run` steps per
    timed run, one warm-up run then the median of five, on one thread and on the host's CPU share
    (orc_swe_step_omp: the same bits, rows split over threads)."""
    import oracle
    from vexhip import scenes

    h, hu, hv = scenes.swe_dam_bumps(n, seed=0)
    params = oracle.swe_params(9.81, 0.45, 1.0 / n, 1.0 / n)
    first = oracle.swe_fastest_wave(h, hu, hv, 9.81, 1e-6)
    out = []
    for threads in (1, _omp_threads()):
        st = oracle.SweStepper(n, n, threads=threads)
        st.load(h, hu, hv, params, first)
        st.write_budget(2**31 - 1)
        t = _median_run(lambda: st.step(steps_per_run))
        out.append({"value": round(n * n * steps_per_run / t, 1), "unit": "cell-updates/s", "cores": threads,
                    "kind": "port", "seconds_per_run": round(t, 4),
                    "sample": f"{steps_per_run} HLL step(s) of the {n}^2 scene per run, one warm-up run, median "
                              f"of 5, on {threads} thread(s)"})
    return out


def cpu_baseline_small2(ppc=4, substeps=200):
    """The 2D reference-sized scene (the 32x64-cell column on 128^2) through the oracle's
    fixed-point substep, sorted every 10 as ParticleSimulation.advance sorts: runs of `substeps`,
    one warm-up run then the median of five. On one thread; then the host's CPU share stepping
    one independent copy each (the throughput a CPU gets from this scene size)."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    from vexhip import particles, scenes

    s = scenes.mpm2_dam_block(ppc=ppc)
    count = len(s["x"])
    h = 1.0 / 127
    words = particles.params2(1e-4, 0, -9.8 / h, 10.0 / h ** 2, 1.0)
    ref = oracle.Mpm2(128, 128, s, words, fixed=True)
    t = _median_run(lambda: ref.advance(substeps, SORT_EVERY))
    one = {"value": round(count * substeps / t, 1), "unit": "particle-substeps/s", "cores": 1, "kind": "port",
           "sample": f"{substeps} fixed-point substeps of the {count}-particle scene per run (sort every 10), one "
                     f"warm-up run, median of 5: {t:.3f} s"}
    threads = _omp_threads()
    copies = [oracle.Mpm2(128, 128, s, words, fixed=True) for _ in range(threads)]
    with ThreadPoolExecutor(threads) as pool:
        t = _median_run(lambda: list(pool.map(lambda r: r.advance(substeps, SORT_EVERY), copies)))
    many = {"value": round(threads * count * substeps / t, 1), "unit": "particle-substeps/s", "cores": threads,
            "kind": "port", "sample": f"{threads} independent copies of the scene, one a thread, {substeps} "
                                      f"substeps each per run, one warm-up run, median of 5: {t:.3f} s"}
    return one, many


def cpu_baseline_batch(scenes=256, sample_every=4, substeps=2):
    """The batched 2D workload through the oracle's fixed-point substep, every scene its own
    state, sorted every 10: runs of `substeps` substeps of every scene, one warm-up run then the
    median of five. On one thread over every `sample_every`-th scene (a bounded sample of the
    same mix of sizes); on the host's CPU share over all scenes, a scene a task."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    from vexhip import scenes as sc_mod

    layouts = sc_mod.mpm2_batch_layouts(scenes)
    states = []
    for k, l in enumerate(layouts):
        fields, words = sc_mod.mpm2_batch_scene(l, 1000 + k)
        states.append(oracle.Mpm2(l["n"], l["n"], fields, words, fixed=True))
    sample = states[::sample_every]
    part = sum(r.count for r in sample)
    t = _median_run(lambda: [r.advance(substeps, SORT_EVERY) for r in sample])
    one = {"value": round(part * substeps / t, 1), "unit": "particle-substeps/s", "cores": 1, "kind": "port",
           "sample": f"every {sample_every}th scene ({len(sample)} scenes, {part} particles), {substeps} "
                     f"fixed-point substeps a run (sort every 10), one warm-up run, median of 5: {t:.3f} s"}
    threads = _omp_threads()
    total = sum(r.count for r in states)
    big_first = sorted(states, key=lambda r: -r.count)
    with ThreadPoolExecutor(threads) as pool:
        t = _median_run(lambda: list(pool.map(lambda r: r.advance(substeps, SORT_EVERY), big_first)))
    many = {"value": round(total * substeps / t, 1), "unit": "particle-substeps/s", "cores": threads,
            "kind": "port", "sample": f"all {len(states)} scenes ({total} particles) on {threads} threads, a scene "
                                      f"a task, {substeps} substeps a run, one warm-up run, median of 5: {t:.3f} s"}
    return one, many


def cpu_baseline_mpm3(threads=None):
    """The oracle on whole substeps of the same 8,388,608-particle workload, one warm-up substep
    then the median of five: the scalar restatement on one thread (about 2 s a substep, some
    13 s in all), and its threaded twin on the host's CPU share (OMP_NUM_THREADS, 16 by
    default): the same arithmetic, P2G summed per thread."""
    import statistics

    import oracle
    from vexhip import scenes

    s = scenes.mpm3_two_cubes()
    count = len(s["x"])
    g, bulk, rho0 = scenes.mpm3_units(256)
    words = oracle.mpm3_params(5e-5, 0, g, 0, bulk, rho0)
    fields = {"x": s["x"], "y": s["y"], "z": s["z"], "m": s["m"]}

    def median_substep(step):
        step()  # warm-up: first touch of the grid (and of the threads' private boxes)
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            step()
            times.append(time.perf_counter() - t0)
        return statistics.median(times)

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    ref = oracle.Mpm3(256, 256, 256, fields, words)
    t = median_substep(lambda: ref.step(1))
    one = {"value": round(count / t, 1), "unit": "particle-substeps/s", "cores": 1, "kind": "port",
           "sample": f"substeps (clear, scatter, grid, advect) of the full {count}-particle 256^3 workload by the "
                     f"scalar oracle: one warm-up, median of 5, {t:.2f} s", "seconds": round(t, 3),
           "host_cores": os.cpu_count(), "host_cpu": model}
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    par = oracle.Mpm3(256, 256, 256, fields, words)
    t = median_substep(lambda: par.step_omp(1, threads))
    many = {"value": round(count / t, 1), "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"substeps of the same workload by the oracle's threaded step on {threads} threads: one "
                      f"warm-up, median of 5, {t:.3f} s", "seconds": round(t, 4)}
    return one, many


def copy_calibration(mib):
    """A plain device copy of a known size, timed under the same HBM counters (FETCH_SIZE,
    WRITE_SIZE) as the kernels, to scale those counters to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}), flush=True)


TRAFFIC_SOURCE = ("profiles/traffic.json: committed rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over this "
                  "command, each scaled by a plain copy's reading; not measured in this run")


def profile_entry(name, kernel):
    """A kernel's entry in a counter summary committed under profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("kernels", {}).get(kernel)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic.json (two rocprofv3 --pmc passes
    and the copy calibration); None when the kernel has no entry."""
    entry = profile_entry("traffic.json", kernel)
    return None if entry is None else entry.get("hbm_bytes_per_launch")


def roofline_of(kernel, entry):
    gbs = entry.get("achieved_gbs")
    return {"kernel": kernel, "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": None if gbs is None else round(gbs / PEAK_HBM_GBS, 4), "traffic": measured_traffic(kernel),
            "traffic_source": TRAFFIC_SOURCE, "bytes_per_launch": entry.get("algorithmic_bytes"),
            "avg_us": entry["avg_us"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--p2g", type=int, default=1)
    ap.add_argument("--g2p", type=int, default=2)
    ap.add_argument("--run", type=int, default=4)
    ap.add_argument("--p2g-ints", type=int, default=1)
    ap.add_argument("--hll-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hll", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--surrogate-iters", type=int, default=2000)
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--tension-steps", type=int, default=50)
    ap.add_argument("--no-lod", action="store_true")
    ap.add_argument("--no-tension", action="store_true")
    ap.add_argument("--atomic-steps", type=int, default=20)
    ap.add_argument("--no-atomic", action="store_true")
    ap.add_argument("--no-fixed", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the path is single-GPU; "
              "the design notes say how it would shard)", file=sys.stderr)
        sys.exit(2)

    if not args.no_surrogate:
        import torch  # noqa: F401  (first: the surrogate's buffers are torch's, on torch's HIP runtime)
    from vexhip import _native
    if _native.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    print(f"[bench] 3D MPM: 8,388,608 particles, 256^3, {args.warmup} warmup + {args.steps} steps", file=sys.stderr,
          flush=True)
    count, secs, table, drift = bench_mpm3(args.steps, args.warmup, args.p2g, args.g2p, args.run,
                                           p2g_ints=args.p2g_ints)
    value = count * args.steps / secs
    dominant = max(table.items(), key=lambda kv: kv[1]["total_ms"])
    dk, dv = dominant
    roofline = roofline_of(dk, dv)
    line = {
        "metric": "particle-substeps/s (3D MPM) & cell-updates/s (HLL); surrogate vs HLL at ≤1% err",
        "value": round(value, 1), "unit": "particle-substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(secs / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3D MLS-MPM/APIC substep, 8,388,608 particles (two seeded cubes, 8 per cell), "
                               "256^3 nodes, dt=5e-5, sorted-segmented P2G, sort every 10 substeps",
                   "particles": count, "grid": [256, 256, 256], "sort_every": SORT_EVERY,
                   "p2g": ["atomic", "sorted-segmented", "fixed-point", "layers"][args.p2g], "g2p": ["gather", "lds-tile", "lds-tile-contracted"][args.g2p],
                   "run": args.run, "p2g_tile": ["float read-add-write", "int32 LDS atomics"][args.p2g_ints]},
        "roofline": roofline, "kernels": table, "grid_mass_drift": drift,
        "kernel_timing": "roofline.kernel by HIP events around each of its launches in the timed region, no events "
                         "around the other kernels; the other kernels over 20 further substeps with every launch "
                         "timed (events between kernels stall the launch queue: about 7% of a substep)",
    }
    if not args.no_atomic:
        print(f"[bench] 3D MPM with the direct atomic P2G: same scene, {args.atomic_steps} steps", file=sys.stderr,
              flush=True)
        acount, asecs, atable, adrift = bench_mpm3(args.atomic_steps, 3, 0, args.g2p, args.run)
        ak = atable["p2g3_atomic"]
        atomics = 108 * acount
        counted = profile_entry("atomics.json", "p2g3_atomic")
        line["p2g_atomic"] = {
            "workload": "the same 3D substep with P2G as the reference schedules it (Flip3.java:193-246): a "
                        "thread a particle, its 27 nodes x 4 sums straight into the grid by float atomics; "
                        "sorted every 10 substeps like the headline",
            "value": round(acount * args.atomic_steps / asecs, 1), "unit": "particle-substeps/s",
            "steps": args.atomic_steps, "ms_per_step": round(asecs / args.atomic_steps * 1e3, 4),
            "roofline": roofline_of("p2g3_atomic", ak),
            "atomics_per_launch": atomics, "atomics_per_s": round(atomics / (ak["avg_us"] * 1e-6), 1),
            "atomics_counted": counted, "kernels": atable, "grid_mass_drift": adrift}
    if not args.no_fixed:
        print(f"[bench] 3D MPM with the fixed-point P2G: same scene, {args.tension_steps} steps", file=sys.stderr,
              flush=True)
        fcount, fsecs, ftable, fdrift = bench_mpm3(args.tension_steps, 5, 2, 1, args.run)
        line["fixed3"] = {"value": round(fcount * args.tension_steps / fsecs, 1), "unit": "particle-substeps/s",
                          "steps": args.tension_steps, "ms_per_step": round(fsecs / args.tension_steps * 1e3, 4),
                          "workload": "the same 3D substep, deterministic: P2G in 64-bit fixed point (option p2g 2) "
                                      "and the reference-order G2P (g2p 1), so every substep is the oracle's bit "
                                      "for bit in any particle order",
                          "kernels": ftable, "grid_mass_drift": fdrift}
    if not args.no_tension:
        print(f"[bench] 3D MPM with tension: same scene, {args.tension_steps} steps", file=sys.stderr, flush=True)
        tcount, tsecs, ttable, tdrift = bench_mpm3(args.tension_steps, 5, args.p2g, args.g2p, args.run, tension=True,
                                                   p2g_ints=args.p2g_ints)
        line["tension3"] = {"value": round(tcount * args.tension_steps / tsecs, 1), "unit": "particle-substeps/s",
                            "steps": args.tension_steps, "ms_per_step": round(tsecs / args.tension_steps * 1e3, 4),
                            "workload": "the same 3D scene stepped as Flip3Step's tension step (clear, P2G, six "
                                        "blurs, stress, grid with tension, G2P), sigma = 0.01*B*8, sort every 10",
                            "kernels": ttable, "grid_mass_drift": tdrift}
    if not args.no_lod:
        line["lod3"] = {"workload": "the same 3D scene after 60 Flip3.lod calls thin every cell's 8 slots to "
                                    "`level`; the sort every 10 substeps moves the emptied slots last and P2G/G2P "
                                    "stop at the active count", "levels": {}}
        for level in (4, 1):
            print(f"[bench] 3D MPM thinned to {level} a cell, {args.tension_steps} steps", file=sys.stderr, flush=True)
            lcount, lsecs, ltable, _ = bench_mpm3(args.tension_steps, 10, args.p2g, args.g2p, args.run, lod_level=level,
                                                 p2g_ints=args.p2g_ints)
            line["lod3"]["levels"][str(level)] = {
                "active_particles": lcount, "ms_per_step": round(lsecs / args.tension_steps * 1e3, 4),
                "active_particle_substeps_per_s": round(lcount * args.tension_steps / lsecs, 1),
                "p2g_us": ltable.get("p2g3_seg", {}).get("avg_us"), "g2p_us": ltable.get("g2p3_fast", {}).get("avg_us")}
    if not args.no_hll:
        print(f"[bench] HLL 4096^2, {args.hll_steps} steps", file=sys.stderr, flush=True)
        rate, htable, hrows = bench_hll(args.hll_steps)
        hk = htable["hll_step"]
        line["hll"] = {"value": round(rate, 1), "unit": "cell-updates/s", "steps": args.hll_steps, "rows": hrows,
                       "workload": "HLL 4096^2 circular dam + 32 seeded bumps, walls, CFL 0.45",
                       "roofline": roofline_of("hll_step", hk),
                       "kernel_timing": "value over launches with no events between them; the kernel's time "
                                        "from as many launches again, each between an event pair"}
        frate, ftable, frows = bench_hll(args.hll_steps, fast=True, rows=48)
        line["hll"]["fast"] = {"value": round(frate, 1), "unit": "cell-updates/s", "rows": frows,
                               "avg_us": ftable["hll_step"]["avg_us"],
                               "achieved_gbs": ftable["hll_step"]["achieved_gbs"],
                               "note": "option hll_fast: a refined hardware reciprocal per face and per cell and the "
                                       "hardware square root; within 1e-5 of the reference's bits after 100 steps "
                                       "(tests), not bit-exact"}
    if not args.no_batch:
        print(f"[bench] batched 2D: 256 scenes, {args.batch_steps} steps", file=sys.stderr, flush=True)
        line["batch2d"] = bench_batch2(args.batch_steps)
        print("[bench] 2D 8,192 particles on 128^2, eager and graph", file=sys.stderr, flush=True)
        line["small2d"] = {"eager": bench_small2(500, False), "graph": bench_small2(500, True),
                           "workload": "2D MLS-MPM, 8,192 particles, 128^2, fixed-point P2G, sort every 10"}
        print("[bench] 2D 65,536 particles on 128^2 (32 a cell), graph", file=sys.stderr, flush=True)
        line["small2d_65536"] = {"graph": bench_small2(500, True, ppc=32),
                                 "workload": "2D MLS-MPM, 65,536 particles (the 32x64-cell column at 32 a cell), "
                                             "128^2, fixed-point P2G, sort every 10"}
    if not args.no_surrogate:
        print(f"[bench] surrogate: train {args.surrogate_iters} iterations at 1024^2, judge, time",
              file=sys.stderr, flush=True)
        line["surrogate"] = bench_surrogate(args.surrogate_iters)
    if not args.no_cpu:
        build = oracle_native()
        print(f"[bench] CPU oracle baselines ({build}): 3D substeps, median of 5, one thread then the host's "
              "CPU share", file=sys.stderr, flush=True)
        line["cpu_baseline"], line["cpu_baseline_threaded"] = cpu_baseline_mpm3()
        line["cpu_baseline"]["build"] = line["cpu_baseline_threaded"]["build"] = build
        if "hll" in line:
            print("[bench] CPU oracle baseline: HLL 4096^2", file=sys.stderr, flush=True)
            line["hll"]["cpu_baseline"], line["hll"]["cpu_baseline_threaded"] = cpu_baseline_hll(4096)
        if "small2d" in line:
            print("[bench] CPU oracle baseline: 2D 8,192 particles", file=sys.stderr, flush=True)
            line["small2d"]["cpu_baseline"], line["small2d"]["cpu_baseline_threaded"] = cpu_baseline_small2()
        if "batch2d" in line:
            print("[bench] CPU oracle baseline: batched 2D", file=sys.stderr, flush=True)
            line["batch2d"]["cpu_baseline"], line["batch2d"]["cpu_baseline_threaded"] = cpu_baseline_batch()
        if "surrogate" in line:
            print("[bench] CPU oracle baseline: 8 HLL steps at 1024^2", file=sys.stderr, flush=True)
            one, many = cpu_baseline_hll(1024, steps_per_run=8)
            line["surrogate"]["cpu_hll_8_steps"] = {
                "one_thread_ms": round(one["seconds_per_run"] * 1e3, 3),
                "threaded_ms": round(many["seconds_per_run"] * 1e3, 3), "threads": many["cores"],
                "sample": "8 FP32 HLL steps of the 1024^2 dam-and-bumps scene (seed 0) by the oracle, one warm-up "
                          "run, median of 5"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
