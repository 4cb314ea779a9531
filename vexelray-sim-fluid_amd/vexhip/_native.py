"""ctypes binding of libvexelhip.so's C ABI (include/vexel/vxr.h).

Loads the in-tree library and raises when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvexelhip.so")

OK, EINVAL, ESTATE, EDEVICE, ENOMEM, EUNSUPPORTED = range(6)
KIND_MPM2, KIND_MPM3, KIND_SWE, KIND_MPM2_BATCH = 1, 2, 3, 4
GROUP_STEP, GROUP_SORT, GROUP_LOD = 0, 1, 2
FLAG_TENSION, FLAG_HEAT, FLAG_CONVECTION, FLAG_RELAX, FLAG_PUMP, FLAG_PLASTIC = (1 << k for k in range(6))
FLAG_OPEN_WEST, FLAG_OPEN_EAST, FLAG_OPEN_SOUTH, FLAG_OPEN_NORTH = (1 << k for k in range(16, 20))


class VxrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"vxr error {code}: {message}")
        self.code = code


class InvalidArgument(VxrError, ValueError):
    """VXR_EINVAL: the Java shim's IllegalArgumentException."""


class InvalidState(VxrError):
    """VXR_ESTATE: the Java shim's IllegalStateException."""


class Unsupported(VxrError):
    """VXR_EUNSUPPORTED: a flag outside the built path."""


_lib = None
_torch_first = False  # whether torch (and its HIP runtime) was loaded before this library


def lib() -> ctypes.CDLL:
    global _lib, _torch_first
    if _lib is None:
        _torch_first = "torch" in sys.modules
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run vexelray-sim-fluid_amd/build.py")
        l = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, cs = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_char_p
        sig = {
            "vxr_last_error": (cs, []),
            "vxr_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
            "vxr_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(i32), i64, ctypes.c_uint32,
                                          ctypes.POINTER(vp)]),
            "vxr_destroy": (ctypes.c_int, [vp]),
            "vxr_write": (ctypes.c_int, [vp, cs, vp, i64]),
            "vxr_read": (ctypes.c_int, [vp, cs, vp, i64]),
            "vxr_length": (ctypes.c_int, [vp, cs, ctypes.POINTER(i64)]),
            "vxr_run": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
            "vxr_run_pass": (ctypes.c_int, [vp, cs, ctypes.c_int]),
            "vxr_swe_dispatch": (ctypes.c_int, [vp, ctypes.c_int]),
            "vxr_clear": (ctypes.c_int, [vp]),
            "vxr_finish": (ctypes.c_int, [vp]),
            "vxr_device_ptr": (ctypes.c_int, [vp, cs, ctypes.POINTER(vp)]),
            "vxr_stream": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
            "vxr_set_option": (ctypes.c_int, [vp, cs, i64]),
            "vxr_get_option": (ctypes.c_int, [vp, cs, ctypes.POINTER(i64)]),
            "vxr_timing": (ctypes.c_int, [vp, ctypes.c_int]),
            "vxr_timing_only": (ctypes.c_int, [vp, ctypes.c_char_p]),
            "vxr_kernel_time": (ctypes.c_int, [vp, cs, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]),
            "vxr_create_batch": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                                ctypes.POINTER(i64), ctypes.c_uint32, ctypes.POINTER(vp)]),
            "vxr_batch_layout": (ctypes.c_int, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)]),
            "vxr_scene_count": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
            "vxr_advance": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, i64]),
            "vxr_autotune": (ctypes.c_int, [vp, ctypes.c_int]),
            "vxr_swe_diagnostics": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
            "vxr_particle_diagnostics": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                        ctypes.POINTER(ctypes.c_double)]),
            "vxr_particle_diagnostics_noalloc": (ctypes.c_int, [vp, vp, i64, ctypes.c_double, ctypes.c_double,
                                                                ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
            "vxr_fixed_scatter": (ctypes.c_int, [vp, ctypes.c_int, i64, vp, vp, vp, vp, vp, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, vp, vp, vp]),
            # the shallow-water surrogate's kernels (csrc/surrogate.hip)
            "vxs_last_error": (cs, []),
            "vxs_conv3x3": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp,
                                           ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int]),
            "vxs_wgrad_scratch": (i64, [ctypes.c_int, ctypes.c_int]),
            "vxs_conv3x3_wgrad": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, vp, vp, vp]),
            "vxs_flip": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, vp]),
            "vxs_adam": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_int]),
            "vxs_pack": (ctypes.c_int, [vp, vp, vp, vp, i64, vp, vp]),
            "vxs_unpack": (ctypes.c_int, [vp, vp, vp, vp, vp, i64, vp, vp, vp, vp]),
            "vxs_loss": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp]),
            "vxs_hll64": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), vp, vp, vp, vp,
                                         vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double]),
            "vxs_dam_bumps": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_double, vp, vp, vp]),
            "vxs_f64_to_f32": (ctypes.c_int, [vp, vp, vp, i64]),
            "vxs_hll32": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), vp, i32, vp,
                                         vp, vp, vp, vp, vp, vp]),
            "vxs_check_scratch": (i64, []),
            "vxs_fallback_check": (ctypes.c_int, [vp, vp, vp, i64, ctypes.c_double, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc == OK:
        return
    msg = lib().vxr_last_error().decode()
    cls = {EINVAL: InvalidArgument, ESTATE: InvalidState, EUNSUPPORTED: Unsupported}.get(rc, VxrError)
    raise cls(rc, msg)


def vxs_check(rc: int) -> None:
    """Status of a surrogate kernel call: 0, 1 (bad argument) or 3 (HIP error)."""
    if rc == 0:
        return
    msg = lib().vxs_last_error().decode()
    raise (InvalidArgument if rc == 1 else VxrError)(rc, msg)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().vxr_device_count(ctypes.byref(n)))
    return n.value


class Sim:
    """One vxr_sim handle: named resident buffers and the passes over them."""

    def __init__(self, kind: int, dims, particles: int = 0, flags: int = 0):
        d = (ctypes.c_int32 * 3)(*(list(dims) + [1, 1, 1])[:3])
        h = ctypes.c_void_p()
        check(lib().vxr_create(kind, d, int(particles), int(flags), ctypes.byref(h)))
        self.h = h

    @classmethod
    def batch(cls, nx, ny, particles, flags: int = 0) -> "Sim":
        """A batch of 2D scenes (vxr_create_batch): scene k is nx[k] x ny[k] with particles[k]."""
        n = len(nx)
        if not (len(ny) == n and len(particles) == n):
            raise ValueError("nx, ny and particles need one entry per scene")
        h = ctypes.c_void_p()
        check(lib().vxr_create_batch(n, (ctypes.c_int32 * n)(*map(int, nx)), (ctypes.c_int32 * n)(*map(int, ny)),
                                     (ctypes.c_int64 * n)(*map(int, particles)), int(flags), ctypes.byref(h)))
        self = cls.__new__(cls)
        self.h = h
        return self

    def scene_count(self) -> int:
        n = ctypes.c_int()
        check(lib().vxr_scene_count(self.h, ctypes.byref(n)))
        return n.value

    def batch_layout(self):
        """(particle, node, start) offsets of each scene, scenes + 1 entries each."""
        n = self.scene_count() + 1
        po, go, so = (ctypes.c_int64 * n)(), (ctypes.c_int64 * n)(), (ctypes.c_int64 * n)()
        check(lib().vxr_batch_layout(self.h, po, go, so))
        return np.array(po[:], np.int64), np.array(go[:], np.int64), np.array(so[:], np.int64)

    def advance(self, count: int, sort_every: int, first_step: int) -> None:
        check(lib().vxr_advance(self.h, int(count), int(sort_every), int(first_step)))

    def swe_diagnostics(self, g, dry, dt, dx, dy) -> dict:
        out = (ctypes.c_double * 9)()
        check(lib().vxr_swe_diagnostics(self.h, g, dry, dt, dx, dy, out))
        keys = ("volume", "min_depth", "max_depth", "max_speed", "max_froude", "max_courant", "non_finite",
                "negative", "wet")
        return dict(zip(keys, list(out)))

    def particle_diagnostics(self, speed, sound, dt) -> dict:
        out = (ctypes.c_double * 7)()
        check(lib().vxr_particle_diagnostics(self.h, speed, sound, dt, out))
        keys = ("min_density", "max_density", "low_density", "high_density", "mach", "courant", "non_finite")
        d = dict(zip(keys, list(out)))
        d["non_finite"] = int(d["non_finite"])
        return d

    def autotune(self, reps: int = 3) -> dict:
        check(lib().vxr_autotune(self.h, int(reps)))
        return {k: self.option(k) for k in ("p2g", "run", "radix_bits", "rows", "blocks")}

    def close(self) -> None:
        if self.h:
            lib().vxr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def length(self, name: str) -> int:
        n = ctypes.c_int64()
        check(lib().vxr_length(self.h, name.encode(), ctypes.byref(n)))
        return n.value

    def write(self, name: str, values) -> None:
        a = np.ascontiguousarray(values)
        if a.dtype not in (np.float32, np.int32, np.uint32):
            a = np.ascontiguousarray(a, dtype=np.float32 if a.dtype.kind == "f" else np.int32)
        check(lib().vxr_write(self.h, name.encode(), a.ctypes.data, a.size))

    def read_words(self, name: str, n: int | None = None) -> np.ndarray:
        n = self.length(name) if n is None else n
        out = np.empty(n, np.int32)
        check(lib().vxr_read(self.h, name.encode(), out.ctypes.data, n))
        return out

    def read(self, name: str) -> np.ndarray:
        return self.read_words(name).view(np.float32)

    def run(self, group: int, repeat: int = 1) -> None:
        check(lib().vxr_run(self.h, group, repeat))

    def run_pass(self, name: str, repeat: int = 1) -> None:
        check(lib().vxr_run_pass(self.h, name.encode(), repeat))

    def swe_dispatch(self, count: int) -> None:
        check(lib().vxr_swe_dispatch(self.h, count))

    def clear(self) -> None:
        check(lib().vxr_clear(self.h))

    def finish(self) -> None:
        check(lib().vxr_finish(self.h))

    def device_ptr(self, name: str) -> int:
        p = ctypes.c_void_p()
        check(lib().vxr_device_ptr(self.h, name.encode(), ctypes.byref(p)))
        return p.value or 0

    def set_option(self, key: str, value: int) -> None:
        check(lib().vxr_set_option(self.h, key.encode(), int(value)))

    def option(self, key: str) -> int:
        v = ctypes.c_int64()
        check(lib().vxr_get_option(self.h, key.encode(), ctypes.byref(v)))
        return v.value

    def timing(self, enable: bool) -> None:
        check(lib().vxr_timing(self.h, 1 if enable else 0))

    def timing_only(self, kernel: str) -> None:
        """Times the named kernel's launches alone (no events around the others)."""
        check(lib().vxr_timing_only(self.h, kernel.encode()))

    def kernel_time(self, kernel: str):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(lib().vxr_kernel_time(self.h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
