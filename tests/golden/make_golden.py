"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

The reference holds no fixture files (its tests draw their inputs from seeds and check physics
and closed forms), so these vectors come from the oracle once its functions pass those checks
(tests/test_oracle_*.py). They freeze small whole runs so that a change to either the oracle or
the device path shows up against stored numbers:

  swe_dam_bumps_32.npz   HLL, 32^2 circular dam + 32 seeded bumps, walls, C = 0.45, 100 steps
  mpm3_block_12.npz      3D MLS-MPM, 12^3 box, a 4x5x3-cell jittered block at 8 a cell, 5 substeps
  mpm2_block_16.npz      2D MLS-MPM, 16^2 box, a 5x7-cell jittered block at 4 a cell, 5 substeps
  sort3_keys_24.npz      3D brick keys and the stable order of 3,000 scattered particles in 24^3

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "vexelray-sim-fluid_amd"))

import oracle  # noqa: E402
from vexhip import scenes  # noqa: E402


def swe_case():
    n, steps = 32, 100
    h, hu, hv = scenes.swe_dam_bumps(n, seed=0)
    params = oracle.swe_params(9.81, 0.45, 1.0 / n, 1.0 / n)
    first = oracle.swe_fastest_wave(h, hu, hv, 9.81, 1e-6)
    return dict(n=n, steps=steps, h0=h, hu0=hu, hv0=hv, params=params, first=np.int32(first))


def swe_run(c):
    s = oracle.SweStepper(int(c["n"]), int(c["n"]), (0, 0, 0, 0))
    s.load(c["h0"], c["hu0"], c["hv0"], c["params"], int(c["first"]))
    s.write_budget(2**31 - 1)
    s.step(int(c["steps"]))
    h, hu, hv = s.read()
    return dict(h=h, hu=hu, hv=hv, budget=np.int32(s.read_budget()), speed=np.int32(s.fastest_wave_bits()),
                clamped=np.int32(s.clamped()))


def mpm3_case():
    n, steps = 12, 5
    x, y, z = scenes.jittered_block3((2, 2, 3), (4, 5, 3), 8, scenes.JavaRandomFloats(12))
    m = np.full(len(x), 0.125, np.float32)
    params = oracle.mpm3_params(2e-3, 0, -40.0, 0, 300.0, 1)
    return dict(n=n, steps=steps, x0=x, y0=y, z0=z, m0=m, params=params)


def mpm3_run(c):
    n = int(c["n"])
    s = oracle.Mpm3(n, n, n, {"x": c["x0"], "y": c["y0"], "z": c["z0"], "m": c["m0"]}, c["params"])
    s.step(int(c["steps"]))
    out = {f: s.p[f] for f in oracle.FIELDS3}
    out["gm"] = s.g["gm"]
    return out


def mpm2_case():
    n, steps = 16, 5
    x, y = scenes.jittered_block2((1, 1), (5, 7), 4, scenes.JavaRandomFloats(16))
    m = np.full(len(x), 0.25, np.float32)
    params = oracle.mpm2_params(2e-3, 0, -40.0, 300.0, 1)
    return dict(n=n, steps=steps, x0=x, y0=y, m0=m, params=params)


def mpm2_run(c):
    n = int(c["n"])
    s = oracle.Mpm2(n, n, {"x": c["x0"], "y": c["y0"], "m": c["m0"]}, c["params"])
    s.step(int(c["steps"]))
    out = {f: s.p[f] for f in oracle.FIELDS2}
    out["gm"] = s.g["gm"]
    return out


def sort3_case():
    n, count = 24, 3000
    r = np.random.default_rng(24)
    pos = r.uniform(0.0, n, size=(3, count)).astype(np.float32)
    return dict(n=n, x0=pos[0], y0=pos[1], z0=pos[2])


def sort3_run(c):
    n = int(c["n"])
    m = np.ones_like(c["x0"])  # every particle active
    s = oracle.Mpm3(n, n, n, {"x": c["x0"], "y": c["y0"], "z": c["z0"], "m": m},
                    oracle.mpm3_params(1e-3, 0, 0, 0, 0, 1))
    keys = s.keys()
    return dict(keys=keys, order=oracle.stable_order(keys))


CASES = {
    "swe_dam_bumps_32": (swe_case, swe_run),
    "mpm3_block_12": (mpm3_case, mpm3_run),
    "mpm2_block_16": (mpm2_case, mpm2_run),
    "sort3_keys_24": (sort3_case, sort3_run),
}


def load(name):
    """Inputs and expected outputs of a fixture: (inputs, expected) as dicts of arrays."""
    with np.load(os.path.join(HERE, name + ".npz")) as f:
        data = {k: f[k] for k in f.files}
    inputs = {k[3:]: v for k, v in data.items() if k.startswith("in_")}
    expected = {k[4:]: v for k, v in data.items() if k.startswith("out_")}
    return inputs, expected


def main():
    for name, (case, run) in CASES.items():
        c = case()
        out = run(c)
        arrays = {"in_" + k: np.asarray(v) for k, v in c.items()}
        arrays.update({"out_" + k: np.asarray(v) for k, v in out.items()})
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, sum(a.nbytes for a in arrays.values()), "bytes")


if __name__ == "__main__":
    main()
