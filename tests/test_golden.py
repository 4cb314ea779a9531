"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle reproduces them bit for bit on the CPU, and the device path reproduces them -- the HLL
run and the sort bit for bit, the MLS-MPM runs within FP32 reassociation of the P2G sums."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as golden  # noqa: E402
import oracle  # noqa: E402

NAMES = list(golden.CASES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    inputs, expected = golden.load(name)
    case, run = golden.CASES[name]
    fresh = case()
    for k, v in fresh.items():
        np.testing.assert_array_equal(np.asarray(v), inputs[k], err_msg=f"input {k}")
    out = run(inputs)
    assert set(out) == set(expected)
    for k, v in out.items():
        np.testing.assert_array_equal(np.asarray(v).view(np.int32) if np.asarray(v).dtype == np.float32 else v,
                                      expected[k].view(np.int32) if expected[k].dtype == np.float32 else expected[k],
                                      err_msg=k)


def test_fixtures_hold_physics():
    """What the fixtures must satisfy whatever produced them: the closed box keeps its water, the
    particle runs keep their mass on the grid and stay inside the walls, the keys are sorted by
    the stored order."""
    i, e = golden.load("swe_dam_bumps_32")
    n = int(i["n"])
    v0 = float(np.sum(i["h0"], dtype=np.float64))
    assert abs(float(np.sum(e["h"], dtype=np.float64)) - v0) <= 1e-5 * v0
    assert int(e["clamped"]) == 0
    for name, axes in (("mpm3_block_12", "xyz"), ("mpm2_block_16", "xy")):
        i, e = golden.load(name)
        n = int(i["n"])
        mass = float(np.sum(i["m0"], dtype=np.float64))
        assert abs(float(np.sum(e["gm"], dtype=np.float64)) - mass) <= 1e-6 * mass
        for a in axes:
            assert e[a].min() >= 1.0 and e[a].max() <= n - 2
    i, e = golden.load("sort3_keys_24")
    assert np.all(np.diff(e["keys"][e["order"]]) >= 0)
    assert sorted(e["order"].tolist()) == list(range(len(e["order"])))


# ------------------------------------------------------------------ the device path

@pytest.mark.gpu
def test_device_hll_fixture(gpu):
    from gpu_rigs import GpuSwe

    i, e = golden.load("swe_dam_bumps_32")
    n = int(i["n"])
    print(f"[golden] hll {n}x{n} {int(i['steps'])} steps", flush=True)
    dev = GpuSwe(n, n)
    dev.load(i["h0"], i["hu0"], i["hv0"], i["params"], int(i["first"]))
    dev.write_budget(2**31 - 1)
    dev.step(int(i["steps"]))
    for k, got in zip(("h", "hu", "hv"), dev.read()):
        np.testing.assert_array_equal(got.view(np.int32), e[k].view(np.int32), err_msg=k)
    assert dev.read_budget() == int(e["budget"])
    assert dev.fastest_wave_bits() == int(e["speed"])
    assert dev.clamped() == int(e["clamped"])


@pytest.mark.gpu
def test_device_sort_fixture(gpu):
    from gpu_rigs import maker3

    i, e = golden.load("sort3_keys_24")
    n = int(i["n"])
    count = len(i["x0"])
    print(f"[golden] sort3 {n}^3 {count} particles", flush=True)
    rig = maker3()(n, n, n, count)
    rig.write("x", i["x0"])
    rig.write("y", i["y0"])
    rig.write("z", i["z0"])
    rig.write("m", np.arange(1, count + 1, dtype=np.float32))  # index + 1 rides along (all active)
    rig.run(["sort"])
    np.testing.assert_array_equal(rig.sim.read_words("keys"), e["keys"][e["order"]])
    np.testing.assert_array_equal(rig.read("m").astype(np.int64) - 1, e["order"])


SCHEDULES = [dict(p2g=0, g2p=0), dict(p2g=1, g2p=1, run=4), dict(p2g=1, g2p=2, run=4)]


@pytest.mark.gpu
@pytest.mark.parametrize("opts", SCHEDULES, ids=["atomic", "seg-exact", "seg-fast"])
def test_device_mpm3_fixture(gpu, opts):
    from gpu_rigs import maker3

    i, e = golden.load("mpm3_block_12")
    n, count = int(i["n"]), len(i["x0"])
    print(f"[golden] mpm3 {n}^3 {count} particles {opts}", flush=True)
    rig = maker3(**opts)(n, n, n, count)
    for k in ("x", "y", "z", "m"):
        rig.write(k, i[k + "0"])
    rig.write("j", np.ones(count, np.float32))
    rig.params(i["params"])
    for _ in range(int(i["steps"])):
        rig.run("step")
    for k in oracle.FIELDS3:
        np.testing.assert_allclose(rig.read(k), e[k], rtol=1e-5, atol=1e-5, err_msg=k)
    np.testing.assert_allclose(rig.read("gm"), e["gm"], rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("opts", SCHEDULES[:2], ids=["atomic", "seg"])
def test_device_mpm2_fixture(gpu, opts):
    from gpu_rigs import maker2

    i, e = golden.load("mpm2_block_16")
    n, count = int(i["n"]), len(i["x0"])
    print(f"[golden] mpm2 {n}^2 {count} particles {opts}", flush=True)
    rig = maker2(**opts)(n, n, count)
    for k in ("x", "y", "m"):
        rig.write(k, i[k + "0"])
    rig.write("j", np.ones(count, np.float32))
    rig.params(i["params"])
    for _ in range(int(i["steps"])):
        rig.run("step")
    for k in oracle.FIELDS2:
        np.testing.assert_allclose(rig.read(k), e[k], rtol=1e-5, atol=1e-5, err_msg=k)
    np.testing.assert_allclose(rig.read("gm"), e["gm"], rtol=1e-5, atol=1e-6)
