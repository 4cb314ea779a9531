"""Device diagnostics against the oracle's Diagnostics.of and ParticleDiagnostics.of restatements:
counts, minima and maxima exactly, the volume to 1e-12 relative (another summation order)."""
import numpy as np
import pytest

import oracle
from gpu_rigs import GpuSwe
from vexhip import _native

pytestmark = pytest.mark.gpu


def test_swe_diagnostics_match_the_oracle(gpu):
    nx, ny = 301, 203
    r = np.random.default_rng(11)
    h = (1 + 0.5 * r.random(nx * ny)).astype(np.float32)
    hu = (0.8 * r.standard_normal(nx * ny)).astype(np.float32)
    hv = (0.8 * r.standard_normal(nx * ny)).astype(np.float32)
    h[:500] = 0.0                      # dry
    h[500:520] = 5e-4                  # wet but shallower than the Froude floor
    h[600:603] = -0.25                 # negative
    h[700], hu[701], hv[702] = np.nan, np.inf, -np.inf
    dev = GpuSwe(nx, ny)
    dev.load(h, hu, hv, oracle.swe_params(9.81, 0.45, 1.0, 1.0), 0)
    print("[diagnostics] swe", flush=True)
    for args in ((9.81, 1e-6, 0.01, 1.0, 1.0), (9.81, 1e-6, 2e-3, 0.5, 0.25)):
        got = dev.sim.swe_diagnostics(*args)
        exp = oracle.swe_diagnostics(h, hu, hv, *args)
        for k, v in exp.items():
            if k == "volume":
                assert abs(got[k] - v) <= 1e-12 * abs(v), (k, got[k], v)
            else:
                assert got[k] == v, (k, got[k], v)


def _j_values(count, seed):
    r = np.random.default_rng(seed)
    j = r.uniform(0.7, 1.3, count).astype(np.float32)
    j[3], j[17], j[29], j[31], j[40] = np.nan, 0.0, -1.0, np.inf, 0.5
    return j


def test_particle_diagnostics_match_the_oracle(gpu):
    count = 100_003
    sim = _native.Sim(_native.KIND_MPM3, (32, 32, 32), count)
    j = _j_values(count, 5)
    sim.write("j", j)
    print("[diagnostics] particles 3D", flush=True)
    got = sim.particle_diagnostics(50.0, 300.0, 1e-3)
    exp = oracle.particle_diagnostics(j, 50.0, 300.0, 1e-3)
    assert got == exp
    sim.close()


def test_particle_diagnostics_of_a_batch(gpu):
    sizes, counts = [(16, 16), (32, 32), (8, 8)], [1500, 2000, 7]
    sim = _native.Sim.batch([s[0] for s in sizes], [s[1] for s in sizes], counts)
    po, _, _ = sim.batch_layout()
    full = np.zeros(int(po[-1]), np.float32)  # pads hold J = 0
    parts = []
    for k, c in enumerate(counts):
        jk = _j_values(c, 20 + k) if c > 100 else np.ones(c, np.float32)
        full[po[k]:po[k] + c] = jk
        parts.append(jk)
    sim.write("j", full)
    print("[diagnostics] particles batch", flush=True)
    got = sim.particle_diagnostics(1.0, 10.0, 1e-2)
    exp = oracle.particle_diagnostics(np.concatenate(parts), 1.0, 10.0, 1e-2)
    assert got == exp
    sim.close()
