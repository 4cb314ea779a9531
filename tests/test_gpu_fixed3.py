"""3D P2G in fixed point (option p2g = 2) on the device: every amount the reference computes,
rounded to a 64-bit integer (mass 2^-40, momentum 2^-32) and summed as integers, so the grid is
the oracle's (orc_mpm3_scatter_fixed) bit for bit in any particle order; with the reference-order
G2P (g2p = 1) whole 3D substeps, sorts and surface tension included, are the oracle's bit for
bit, eager and replayed as graphs."""
import numpy as np
import pytest

import mpm_cases as cases
import oracle
from gpu_rigs import maker3
from vexhip import scenes

pytestmark = pytest.mark.gpu

FIXED = dict(p2g=2, g2p=1, run=4)


def _bits(got, exp, what):
    np.testing.assert_array_equal(np.asarray(got).view(np.int32), np.asarray(exp).view(np.int32), err_msg=what)


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_fixed_scatter_bit_exact(gpu, order):
    """Sorted, each wave's particles fit its LDS tile; shuffled, a wave's stencils spread over
    the box and go straight to the grid's 64-bit atomics. Either way the decoded sums equal the
    oracle's bit for bit, some particles without mass included."""
    n, count = 40, 60_000
    f = cases.scatter_inputs_3d(n, count, 7)
    f["m"] = f["m"].copy()
    f["m"][::97] = 0.0
    params = oracle.mpm3_params(1e-3, 0, -50, 0, 400.0, 1)
    rig = maker3(**FIXED)(n, n, n, count)
    cases.load3(rig, f)
    rig.params(params)
    if order == "sorted":
        rig.run(["sort"])
        f = {name: rig.read(name) for name in oracle.FIELDS3}
    else:
        perm = oracle.JavaRandom(3).shuffle_order(count)
        f = {k: v[perm] for k, v in f.items()}
        cases.load3(rig, f)
    print(f"[fixed3] scatter {order}", flush=True)
    ref = oracle.Mpm3(n, n, n, f, params, fixed=True)
    ref.scatter()
    rig.sim.timing(True)
    rig.run(["clear", "scatter"])
    assert rig.sim.kernel_time("p2g3_fixed")[1] == 1
    rig.sim.timing(False)
    for g in ("gm", "gmu", "gmv", "gmw"):
        _bits(rig.read(g), ref.g[g], f"{order} {g}")
    # a second scatter over the same particles adds nothing stale: the clear zeroed the sums
    rig.run(["clear", "scatter"])
    for g in ("gm", "gmu", "gmv", "gmw"):
        _bits(rig.read(g), ref.g[g], f"{order} again {g}")


def _steps(tension, graph):
    n = 32
    x, y, z = scenes.jittered_block3((3, 3, 3), (10, 14, 9), 8, scenes.JavaRandomFloats(5))
    count = len(x)
    m = np.full(count, 0.125, np.float32)
    sigma = 0.01 * 2000.0 * 4 if tension else 0.0
    params = oracle.mpm3_params(5e-4, 0, -200.0, 0, 2000.0, 1, sigma)
    fields = {"x": x, "y": y, "z": z, "m": m}
    ref = oracle.Mpm3(n, n, n, fields, params, tension=tension, fixed=True)
    rig = maker3(tension=tension, **FIXED)(n, n, n, count)
    for name, v in (("x", x), ("y", y), ("z", z), ("m", m), ("j", np.ones(count, np.float32))):
        rig.write(name, v)
    rig.params(params)
    rig.sim.set_option("graph", graph)
    done = 0
    for chunk in (1, 9, 15):
        for k in range(done, done + chunk):
            if k % 10 == 0:
                ref.sort()
            ref.step()
        rig.sim.advance(chunk, 10, done)
        done += chunk
        for name in oracle.FIELDS3:
            _bits(rig.read(name), ref.p[name], f"after {done} steps: {name}")
        for g in ("gm", "gu", "gv", "gw"):
            _bits(rig.read(g), ref.g[g], f"after {done} steps: {g}")
    return ref


@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
def test_fixed_steps_bit_exact(gpu, graph):
    """25 whole substeps of a falling block, sorted every 10: every particle field and the grid
    bit for bit with the oracle's fixed-point substeps, after 1, 10 and 25 steps."""
    print(f"[fixed3] steps graph={graph}", flush=True)
    ref = _steps(False, graph)
    assert float(np.abs(ref.p["v"]).max()) > 0.1


def test_fixed_tension_steps_bit_exact(gpu):
    """The same with Flip3Step's tension passes between the scatter and the grid: the fixed-point
    sums make the whole tension step the oracle's bit for bit."""
    print("[fixed3] tension steps", flush=True)
    ref = _steps(True, 0)
    assert float(np.abs(ref.g["txx"]).max()) > 0
