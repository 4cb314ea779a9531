"""The C ABI library without a GPU: it loads, exports exactly what include/vexel/vxr.h and vxs.h declare,
and refuses bad arguments before touching a device. The runners' host logic is checked here too."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from vexhip import _native, clock, patch, particles

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"vxr": os.path.join(ROOT, "include", "vexel", "vxr.h"),
           "vxs": os.path.join(ROOT, "include", "vexel", "vxs.h")}


def declared(prefix):
    text = open(HEADERS[prefix]).read()
    return sorted(set(re.findall(rf"\b({prefix}_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("prefix,must", [("vxr", ("vxr_create", "vxr_swe_dispatch")),
                                         ("vxs", ("vxs_conv3x3", "vxs_hll32", "vxs_fallback_check"))])
def test_library_exports_every_declared_symbol(prefix, must):
    """Every entry point a header declares is exported, and the library exports no other of
    that prefix."""
    names = declared(prefix)
    assert all(m in names for m in must)
    lib = _native.lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(rf"\b({prefix}_[a-z0-9_]+)\b", out)))
    assert exported == names


@pytest.mark.parametrize("name", ["vxr.h", "vxs.h"])
def test_header_compiles_as_c(name):
    src = f'#include "vexel/{name}"\nint main(void){{ return 0; }}\n'
    r = subprocess.run(["gcc", "-x", "c", "-std=c99", "-Wall", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_arguments_refused_before_the_device():
    lib = _native.lib()
    h = ctypes.c_void_p()
    dims = (ctypes.c_int32 * 3)(8, 8, 8)
    assert lib.vxr_create(99, dims, 10, 0, ctypes.byref(h)) == _native.EINVAL
    assert b"unknown kind" in lib.vxr_last_error()
    # surface tension is built in 2D and 3D; heat and 3D plasticity are not
    assert lib.vxr_create(_native.KIND_MPM2, dims, 10, _native.FLAG_HEAT, ctypes.byref(h)) == _native.EUNSUPPORTED
    assert lib.vxr_create(_native.KIND_MPM2, dims, 10, _native.FLAG_TENSION | _native.FLAG_HEAT,
                          ctypes.byref(h)) == _native.EUNSUPPORTED
    assert lib.vxr_create(_native.KIND_MPM3, dims, 10, _native.FLAG_PLASTIC, ctypes.byref(h)) == _native.EUNSUPPORTED
    assert lib.vxr_create(_native.KIND_MPM3, dims, 10, _native.FLAG_PLASTIC | _native.FLAG_TENSION,
                          ctypes.byref(h)) == _native.EUNSUPPORTED
    assert lib.vxr_create(_native.KIND_MPM3, dims, 0, 0, ctypes.byref(h)) == _native.EINVAL
    assert b"needs particles" in lib.vxr_last_error()
    small = (ctypes.c_int32 * 3)(3, 8, 8)
    assert lib.vxr_create(_native.KIND_MPM2, small, 10, 0, ctypes.byref(h)) == _native.EINVAL
    zero = (ctypes.c_int32 * 3)(0, 8, 1)
    assert lib.vxr_create(_native.KIND_SWE, zero, 0, 0, ctypes.byref(h)) == _native.EINVAL
    assert lib.vxr_write(None, b"x", None, 0) == _native.EINVAL
    assert lib.vxr_run(None, 0, 1) == _native.EINVAL
    n = ctypes.c_int(-1)
    assert lib.vxr_device_count(ctypes.byref(n)) == _native.OK and n.value >= 0
    # a batch of scenes is refused on its shapes before any device call
    nx = (ctypes.c_int32 * 2)(8, 3)
    ny = (ctypes.c_int32 * 2)(8, 8)
    cnt = (ctypes.c_int64 * 2)(10, 10)
    assert lib.vxr_create_batch(0, nx, ny, cnt, 0, ctypes.byref(h)) == _native.EINVAL
    assert lib.vxr_create_batch(2, nx, ny, cnt, 0, ctypes.byref(h)) == _native.EINVAL
    assert b"scene 1" in lib.vxr_last_error()
    assert lib.vxr_create_batch(1, nx, ny, cnt, _native.FLAG_HEAT, ctypes.byref(h)) == _native.EUNSUPPORTED
    assert lib.vxr_create(_native.KIND_MPM2_BATCH, dims, 10, 0, ctypes.byref(h)) == _native.EINVAL
    assert lib.vxr_advance(None, 1, 10, 0) == _native.EINVAL
    assert lib.vxr_autotune(None, 1) == _native.EINVAL
    # the fixed-point scatter checks its scale and schedule before queueing anything
    assert lib.vxr_fixed_scatter(None, 8, 0, None, None, None, None, None, 31, 10, 0, 1, 1, 1) == _native.EINVAL
    assert lib.vxr_fixed_scatter(None, 8, 0, None, None, None, None, None, 10, 10, 2, 1, 1, 1) == _native.EINVAL
    assert lib.vxr_fixed_scatter(None, 1, 0, None, None, None, None, None, 10, 10, 0, 1, 1, 1) == _native.EINVAL


def test_params_words():
    w = patch.params(9.81, 0.45, 0.25, 0.5)
    assert w.dtype == np.int32 and w.view(np.float32).tolist() == [np.float32(v) for v in (9.81, 0.45, 0.25, 0.5, 1e-6)]
    with pytest.raises(ValueError):
        patch.params(9.81, 0.6, 1, 1)
    assert len(particles.params2(1e-3, 0, -9.8, 10, 1)) == 28
    assert particles.params3(1e-3, 0, -9.8, 0, 10, 1).view(np.float32)[4] == np.float32(10)
    with pytest.raises(ValueError):
        particles.params3(0, 0, 0, 0, 1, 1)


def test_fastest_wave_matches_the_oracle():
    import oracle
    r = np.random.default_rng(0)
    h = r.random(5000).astype(np.float32)
    h[::7] = 0
    hu = r.standard_normal(5000).astype(np.float32)
    hv = r.standard_normal(5000).astype(np.float32)
    assert patch.fastest_wave(h, hu, hv, 9.81, 1e-6) == oracle.swe_fastest_wave(h, hu, hv, 9.81, 1e-6)


def test_clock_grant_settle():
    c = clock.Clock()
    assert c.grant(clock.ticks(1e9)) == 2**31 - 1      # a grant fits an i32
    c.settle(2**31 - 1 - 500)
    assert c.now() == 500 and c.outstanding() == 0
    assert c.grant(1500) == 1000
    c.settle(0)
    assert c.now() == 1500
    assert clock.ticks(0.0000015) == 2 and clock.seconds(2_500_000) == 2.5
