"""The device FixedScatter (vxr_fixed_scatter) against the oracle's restatement of the host grid,
bit for bit: FixedScatterTest.everyScheduleAndOrderGivesTheHostsGrid and
theSegmentedSumKeepsBrokenRunsApart (test/particle/FixedScatterTest.java:44-98)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _device(p):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in p.items()}


def _run(n, p, scale, schedule):
    from vexhip import fixed_scatter as fs

    d = _device(p)
    grid = fs.scatter(n, d["x"], d["y"], d["u"], d["v"], d["m"], scale, schedule)
    torch.cuda.synchronize()
    return [g.cpu().numpy() for g in grid]


@pytest.mark.parametrize("ppc", [4, 64])
def test_every_schedule_and_order_gives_the_hosts_grid(gpu, ppc):
    from vexhip.fixed_scatter import DIRECT, SEGMENTED, FixedPoint

    n = 64 if ppc == 4 else 17
    srt = oracle.jittered_particles(n - 1, ppc, oracle.JavaRandom(7))
    rnd = oracle.shuffled(srt, oracle.JavaRandom(8))
    for scale in (FixedPoint.for_range(4 * ppc, 1.5, 8), FixedPoint(20, 10)):
        expected = oracle.fixed_scatter(n, srt, scale.mass_bits, scale.velocity_bits)
        for schedule in (DIRECT, SEGMENTED):
            for name, p in (("sorted", srt), ("random", rnd)):
                print(f"[fixed scatter] ppc {ppc} {scale} schedule {schedule} {name}", flush=True)
                got = _run(n, p, scale, schedule)
                for f in range(3):
                    np.testing.assert_array_equal(got[f], expected[f], err_msg=f"{scale} {schedule} {name} {f}")


def test_the_segmented_sum_keeps_broken_runs_apart(gpu):
    from vexhip.fixed_scatter import SEGMENTED, FixedPoint

    n = 8
    cells = ((1, 1), (2, 1), (1, 2))
    pattern = (0, 0, 1, 0, 2, 2, 2, 0, 1, 1, 0)
    count = 301
    r = oracle.JavaRandom(51)
    p = {k: np.empty(count, np.float32) for k in ("x", "y", "u", "v", "m")}
    for k in range(count):
        c = cells[pattern[k % len(pattern)]]
        p["x"][k] = np.float32(np.float32(c[0]) + np.float32(r.next_float()) * np.float32(0.999))
        p["y"][k] = np.float32(np.float32(c[1]) + np.float32(r.next_float()) * np.float32(0.999))
        p["u"][k] = np.float32(r.next_gaussian())
        p["v"][k] = np.float32(r.next_gaussian())
        p["m"][k] = np.float32(np.float32(0.5) + np.float32(r.next_float()))
    scale = FixedPoint.for_range(512, 1.5, 8)
    expected = oracle.fixed_scatter(n, p, scale.mass_bits, scale.velocity_bits)
    got = _run(n, p, scale, SEGMENTED)
    for f in range(3):
        np.testing.assert_array_equal(got[f], expected[f])


def test_refuses_bad_scales(gpu):
    from vexhip import _native

    lib = _native.lib()
    assert lib.vxr_fixed_scatter(None, 8, 0, None, None, None, None, None, 31, 10, 0, 1, 1, 1) == _native.EINVAL
    assert lib.vxr_fixed_scatter(None, 8, 0, None, None, None, None, None, 10, 10, 2, 1, 1, 1) == _native.EINVAL
