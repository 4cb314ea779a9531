"""vxr_particle_diagnostics_noalloc, the entry point that allocates nothing (its sort lives in the
library's fixed device array), against vxr_particle_diagnostics on the same particles: two calls
in a row, the second smaller, bit for bit; and past its limit a clear refusal."""
import ctypes

import numpy as np
import pytest

from vexhip import _native

pytestmark = pytest.mark.gpu


def _handle_with_j(j):
    sim = _native.Sim(_native.KIND_MPM3, (16, 16, 16), len(j), 0)
    sim.write("j", np.asarray(j, np.float32))
    return sim


def _noalloc(sim, count, speed=2.0, sound=7.0, dt=1e-3):
    out = (ctypes.c_double * 7)()
    stream = ctypes.c_void_p()
    _native.check(_native.lib().vxr_stream(sim.h, ctypes.byref(stream)))
    _native.check(_native.lib().vxr_particle_diagnostics_noalloc(stream, sim.device_ptr("j"), count, speed, sound,
                                                                 dt, out))
    return np.array(list(out))


def _reference(j, speed=2.0, sound=7.0, dt=1e-3):
    sim = _handle_with_j(j)
    d = sim.particle_diagnostics(speed, sound, dt)
    keys = ("min_density", "max_density", "low_density", "high_density", "mach", "courant", "non_finite")
    out = np.array([d[k] for k in keys], np.float64)
    sim.close()
    return out


def test_noalloc_matches_the_allocating_entry_point(gpu):
    rng = np.random.default_rng(17)
    n1, n2 = 600_000, 70_001
    j = rng.uniform(0.6, 1.4, n1).astype(np.float32)
    j[rng.integers(0, n1, 300)] = np.nan
    j[rng.integers(0, n1, 200)] = 0.0
    j[rng.integers(0, n1, 100)] = -2.0
    j[:n2][rng.integers(0, n2, 50)] = np.inf
    print(f"[no_alloc] {n1} then {n2} particles", flush=True)
    sim = _handle_with_j(j)
    first = _noalloc(sim, n1)
    np.testing.assert_array_equal(first.view(np.int64), _reference(j).view(np.int64))
    second = _noalloc(sim, n2)  # a prefix of the same J, after the larger call
    np.testing.assert_array_equal(second.view(np.int64), _reference(j[:n2]).view(np.int64))
    again = _noalloc(sim, n1)
    np.testing.assert_array_equal(again.view(np.int64), first.view(np.int64))
    sim.close()


def test_noalloc_refuses_past_its_limit(gpu):
    sim = _handle_with_j(np.ones(64, np.float32))
    out = (ctypes.c_double * 7)()
    rc = _native.lib().vxr_particle_diagnostics_noalloc(None, sim.device_ptr("j"), (1 << 20) + 1, 1.0, 1.0, 1.0, out)
    assert rc == 1
    assert b"1048576" in _native.lib().vxr_last_error()
    rc = _native.lib().vxr_particle_diagnostics_noalloc(None, sim.device_ptr("j"), 0, 1.0, 1.0, 1.0, out)
    assert rc == 1
    sim.close()
