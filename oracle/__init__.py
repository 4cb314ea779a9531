"""CPU oracle of the vexel-hip hot path -- test infrastructure only.

A ctypes face over ``oracle/liboracle.so``, the C++ restatement of the reference's kernels
(see ``oracle.h`` for the file:line each one follows and what pins it). Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this module, and
only as the checker: the product library never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VEXEL_ORACLE_LIB names another build of the same sources (the CPU baseline's -march=native one)
LIB_PATH = os.environ.get("VEXEL_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_i = ctypes.c_int
_c_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def build() -> None:
    """Compiles liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C {HERE}`")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_rng_new": (_vp, [_c_i64]),
        "orc_rng_free": (None, [_vp]),
        "orc_rng_next_int": (ctypes.c_int32, [_vp]),
        "orc_rng_next_int_bound": (ctypes.c_int32, [_vp, ctypes.c_int32]),
        "orc_rng_next_float": (ctypes.c_float, [_vp]),
        "orc_rng_next_double": (ctypes.c_double, [_vp]),
        "orc_rng_next_gaussian": (ctypes.c_double, [_vp]),
        "orc_rng_pattern": (None, [_vp, ctypes.c_char_p, _c_i64, _f32p]),
        "orc_rng_shuffle_order": (None, [_vp, ctypes.c_int32, _i32p]),
        "orc_swe_step": (None, [_c_i, _c_i, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                                _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]),
        "orc_swe_step_omp": (None, [_c_i, _c_i, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                                _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _c_i]),
        "orc_swe_fastest_wave": (ctypes.c_int32, [_c_i64, _f32p, _f32p, _f32p, ctypes.c_double,
                                                  ctypes.c_double]),
        "orc_swe_diagnostics": (None, [_c_i64, _f32p, _f32p, _f32p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, _f64p]),
        "orc_swe_step_f64": (None, [_c_i, _c_i, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double]),
        "orc_mpm3_clear": (None, [_c_i, _c_i, _c_i, _f32p, _f32p, _f32p, _f32p]),
        "orc_mpm3_scatter": (None, [_c_i, _c_i, _c_i, _c_i64, _vp, _i32p, _f32p, _f32p, _f32p, _f32p]),
        "orc_mpm3_scatter_fixed": (None, [_c_i, _c_i, _c_i, _c_i64, _vp, _i32p, _i64p, _i64p, _i64p, _i64p]),
        "orc_mpm3_decode_fixed": (None, [_c_i64, _i64p, _i64p, _i64p, _i64p, _f32p, _f32p, _f32p, _f32p]),
        "orc_mpm3_grid": (None, [_c_i, _c_i, _c_i, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
        "orc_mpm3_advect": (None, [_c_i, _c_i, _c_i, _c_i64, _vp, _i32p, _f32p, _f32p, _f32p]),
        "orc_mpm3_step": (None, [_c_i, _c_i, _c_i, _c_i64, _vp, _i32p, _vp]),
        "orc_mpm3_key_bits": (_c_i, [_c_i, _c_i, _c_i]),
        "orc_mpm3_step_omp": (None, [_c_i, _c_i, _c_i, _c_i64, _vp, _i32p, _vp, _c_i]),
        "orc_mpm3_keys": (None, [_c_i, _c_i, _c_i, _c_i64, _f32p, _f32p, _f32p, _f32p, _i32p]),
        "orc_mpm3_inactive_key": (ctypes.c_int32, [_c_i, _c_i, _c_i]),
        "orc_mpm2_clear": (None, [_c_i, _c_i, _f32p, _f32p, _f32p]),
        "orc_mpm2_scatter": (None, [_c_i, _c_i, _c_i64, _vp, _i32p, _f32p, _f32p, _f32p]),
        "orc_mpm2_grid": (None, [_c_i, _c_i, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
        "orc_mpm2_advect": (None, [_c_i, _c_i, _c_i64, _vp, _i32p, _f32p, _f32p]),
        "orc_mpm2_step": (None, [_c_i, _c_i, _c_i64, _vp, _i32p, _vp]),
        "orc_mpm2_scatter_fixed": (None, [_c_i, _c_i, _c_i64, _vp, _i32p, _i64p, _i64p, _i64p]),
        "orc_mpm2_decode_fixed": (None, [_c_i64, _i64p, _i64p, _i64p, _f32p, _f32p, _f32p]),
        "orc_mpm2_cell": (ctypes.c_int32, [ctypes.c_float, ctypes.c_float, _c_i, _c_i]),
        "orc_mpm2_keys": (None, [_c_i, _c_i, _c_i64, _f32p, _f32p, _i32p]),
        "orc_mpm2_cell_starts": (_c_i64, [_c_i, _c_i, _c_i64, _f32p, _f32p, _i32p]),
        "orc_mpm2_sort": (None, [_c_i, _c_i, _c_i64, _vp]),
        "orc_stable_order": (None, [_c_i64, _i32p, _i32p]),
        "orc_tension3_blur": (None, [_c_i, _c_i, _c_i, _f32p, _f32p, _i32p, _c_i]),
        "orc_tension3_stress": (None, [_c_i, _c_i, _c_i, _f32p, _vp]),
        "orc_tension3_force": (None, [_c_i, _c_i, _c_i, _f32p, _vp, _c_i64, _f32p]),
        "orc_mpm3_grid_tension": (None, [_c_i, _c_i, _c_i, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                         _vp, _f32p]),
        "orc_tension2_blur": (None, [_c_i, _c_i, _f32p, _f32p, _i32p, _c_i]),
        "orc_tension2_stress": (None, [_c_i, _c_i, _f32p, _f32p, _f32p, _f32p]),
        "orc_tension2_force": (None, [_c_i, _c_i, _f32p, _f32p, _f32p, _f32p, _c_i64, _f32p]),
        "orc_mpm2_grid_tension": (None, [_c_i, _c_i, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                         _f32p]),
        "orc_mpm3_lod": (None, [_c_i, _c_i, _c_i, _c_i64, _vp, _f32p, _f32p, _f32p]),
        "orc_mpm3_lod_target": (None, [_c_i64, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
        "orc_particle_diagnostics": (None, [_c_i64, _f32p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            _f64p]),
        "orc_fixed_split": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
        "orc_fixed_scatter": (None, [_c_i, _c_i64, _f32p, _f32p, _f32p, _f32p, _f32p, _c_i, _c_i, _i32p, _i32p,
                                     _i32p]),
        "orc_scatter_f64": (None, [_c_i, _c_i64, _f32p, _f32p, _f32p, _f32p, _f32p, _f64p, _f64p, _f64p]),
        "orc_bucket_starts": (None, [_c_i64, _i32p, ctypes.c_int32, _i32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Particle field orders: core/particle/Flip3.java:112-115 and core/particle/FlipStep.java:95-97.
FIELDS3 = ("x", "y", "z", "u", "v", "w", "m", "j",
           "c00", "c01", "c02", "c10", "c11", "c12", "c20", "c21", "c22")
FIELDS2 = ("x", "y", "u", "v", "m", "j", "c00", "c01", "c10", "c11")
GRID3 = ("gm", "gmu", "gmv", "gmw", "gu", "gv", "gw")
GRID2 = ("gm", "gmu", "gmv", "gu", "gv")


def f32_bits(values) -> np.ndarray:
    """Float.floatToRawIntBits of each value, after a (float) cast."""
    return np.asarray(values, dtype=np.float64).astype(np.float32).view(np.int32).copy()


def bits_f32(words) -> np.ndarray:
    return np.asarray(words, dtype=np.int32).view(np.float32).copy()


# ------------------------------------------------------------------ java.util.Random

class JavaRandom:
    """java.util.Random, through the oracle's C port."""

    def __init__(self, seed: int):
        self._h = lib.orc_rng_new(int(seed))

    def __del__(self):
        if getattr(self, "_h", None):
            lib.orc_rng_free(self._h)
            self._h = None

    def next_int(self, bound: int | None = None) -> int:
        if bound is None:
            return lib.orc_rng_next_int(self._h)
        return lib.orc_rng_next_int_bound(self._h, int(bound))

    def next_float(self) -> float:
        return lib.orc_rng_next_float(self._h)

    def next_double(self) -> float:
        return lib.orc_rng_next_double(self._h)

    def next_gaussian(self) -> float:
        return lib.orc_rng_next_gaussian(self._h)

    def pattern(self, pattern: str, count: int) -> np.ndarray:
        """``count`` rows of draws in the order ``pattern`` names them ('f', 'g', 'd')."""
        out = np.empty(count * len(pattern), dtype=np.float32)
        lib.orc_rng_pattern(self._h, pattern.encode(), count, out)
        return out.reshape(count, len(pattern))

    def shuffle_order(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.int32)
        lib.orc_rng_shuffle_order(self._h, n, out)
        return out


# ------------------------------------------------------------------ shallow water

WALL, OPEN = 0, 1


def swe_params(g, courant, dx, dy, dry=1e-6, allow_unstable=False) -> np.ndarray:
    """ShallowWater.params (core/stencil/ShallowWater.java:248-271)."""
    if not courant > 0:
        raise ValueError(f"a Courant number must be positive, got {courant}")
    if not allow_unstable and not courant <= 0.5:
        raise ValueError(f"a first-order 2D step needs a Courant number in (0, 1/2], got {courant}")
    return f32_bits([g, courant, dx, dy, dry])


def swe_fastest_wave(h, hu, hv, g, dry) -> int:
    return lib.orc_swe_fastest_wave(h.size, _c(h), _c(hu), _c(hv), float(g), float(dry))


def swe_diagnostics(h, hu, hv, g, dry, dt, dx, dy) -> dict:
    out = np.zeros(9)
    lib.orc_swe_diagnostics(h.size, _c(h), _c(hu), _c(hv), g, dry, dt, dx, dy, out)
    keys = ("volume", "min_depth", "max_depth", "max_speed", "max_froude", "max_courant",
            "non_finite", "negative", "wet")
    return dict(zip(keys, out.tolist()))


def _c(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


class SweStepper:
    """The reference test harness's CPU stepper (test/stencil/Stepper.java:26-221) over the
    oracle: two states ping-pong, three speed words rotate, two budget words ping-pong."""

    def __init__(self, nx: int, ny: int, edges=(WALL, WALL, WALL, WALL), threads: int = 1):
        if nx < 1 or ny < 1:
            raise ValueError(f"a patch needs at least one cell each way, got {nx} x {ny}")
        self.nx, self.ny = nx, ny
        self.threads = threads  # > 1: orc_swe_step_omp, the same bits on that many host threads
        self.edges = np.asarray(edges, dtype=np.int32)
        self.dispatched = 0
        self.state = [[np.zeros(nx * ny, np.float32) for _ in range(3)] for _ in range(2)]
        self.speed = [np.zeros(1, np.int32) for _ in range(3)]
        self.budget = [np.zeros(1, np.int32) for _ in range(2)]
        self._clamped = np.zeros(1, np.int32)
        self.params = np.zeros(5, np.int32)

    def load(self, h, hu, hv, params, first_speed):
        self.state[0] = [_c(h).ravel().copy(), _c(hu).ravel().copy(), _c(hv).ravel().copy()]
        self.state[1] = [np.zeros(self.nx * self.ny, np.float32) for _ in range(3)]
        self.speed = [np.array([first_speed], np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32)]
        self.params = np.asarray(params, dtype=np.int32).copy()
        self._clamped[0] = 0
        self.dispatched = 0

    def dispatch(self):
        k = self.dispatched
        out, inn = self.state[(k + 1) % 2], self.state[k % 2]
        args = (self.nx, self.ny, self.edges, out[0], out[1], out[2], inn[0], inn[1], inn[2], self.params,
                self.speed[k % 3], self.speed[(k + 1) % 3], self.speed[(k + 2) % 3], self.budget[k % 2],
                self.budget[(k + 1) % 2], self._clamped)
        if self.threads > 1:
            lib.orc_swe_step_omp(*args, int(self.threads))
        else:
            lib.orc_swe_step(*args)
        self.dispatched += 1

    def step(self, n: int):
        for _ in range(n):
            self.dispatch()

    def write_budget(self, ticks: int):
        self.budget[self.dispatched % 2][0] = ticks

    def read_budget(self) -> int:
        return int(self.budget[self.dispatched % 2][0])

    def read(self):
        return [a.copy() for a in self.state[self.dispatched % 2]]

    def clamped(self) -> int:
        return int(self._clamped[0])

    def fastest_wave_bits(self) -> int:
        return int(self.speed[self.dispatched % 3][0])


def swe_step_f64(nx, ny, edges, h, hu, hv, g, dx, dy, dry, dt):
    out = [np.empty(nx * ny) for _ in range(3)]
    lib.orc_swe_step_f64(nx, ny, np.asarray(edges, np.int32), out[0], out[1], out[2],
                         _c(h, np.float64), _c(hu, np.float64), _c(hv, np.float64), g, dx, dy, dry, dt)
    return out


# ------------------------------------------------------------------ MPM

def mpm3_params(dt, gx, gy, gz, bulk, rho0, sigma=0.0) -> np.ndarray:
    """Flip3.params (core/particle/Flip3.java:89-103)."""
    if not (dt > 0) or not (bulk >= 0) or not (rho0 > 0):
        raise ValueError(f"dt > 0, bulk >= 0 and rho0 > 0, got dt {dt}, bulk {bulk}, rho0 {rho0}")
    return f32_bits([dt, gx, gy, gz, bulk, rho0, sigma])


def mpm2_params(dt, gx, gy, bulk, rho0, sigma=0.0) -> np.ndarray:
    """Flip.params (core/particle/Flip.java:145-211): 16 values, the slice bounds, the pump's ten."""
    if not (dt > 0) or not (bulk >= 0) or not (rho0 > 0):
        raise ValueError(f"dt > 0, bulk >= 0 and rho0 > 0, got dt {dt}, bulk {bulk}, rho0 {rho0}")
    words = np.zeros(28, np.int32)
    words[:6] = f32_bits([dt, gx, gy, bulk, rho0, sigma])
    return words


def _ptrs(arrays):
    for a in arrays:
        assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])


TENSION_BLURS = 6                      # Flip.TENSION_BLURS
STRESS3 = ("txx", "tyy", "tzz", "txy", "txz", "tyz")


class Mpm3:
    """A 3D particle state on the CPU, stepped pass by pass as Flip3Step lists them. With
    ``tension`` the step is Flip3Step's tension one (core/particle/Flip3Step.java:88-109): clear,
    scatter, six blurs (gm -> c0 -> c1 -> ...), stress, grid with tension, advect. With ``fixed``
    the scatter is the fixed-point one (orc_mpm3_scatter_fixed), its sums (kept in self.q) decoded
    into gm gmu gmv gmw."""

    def __init__(self, nx, ny, nz, fields: dict, params, tension=False, fixed=False):
        self.tension = tension
        self.fixed = fixed
        self.n = (nx, ny, nz)
        count = len(fields["x"])
        self.p = {f: _c(fields.get(f, np.zeros(count)), np.float32).copy() for f in FIELDS3}
        if "j" not in fields:
            self.p["j"][:] = 1
        self.count = count
        nodes = nx * ny * nz
        self.g = {f: np.zeros(nodes, np.float32) for f in GRID3}
        if tension:
            for f in ("c0", "c1") + STRESS3:
                self.g[f] = np.zeros(nodes, np.float32)
        self.params = np.asarray(params, np.int32)

    def _pp(self):
        return _ptrs([self.p[f] for f in FIELDS3])

    def blur(self, k):
        """Pass blur k: gm (scaled to a colour) into c0 for k = 0, then c0 and c1 in turn."""
        src = "gm" if k == 0 else ("c0" if k % 2 == 1 else "c1")
        dst = "c0" if k % 2 == 0 else "c1"
        lib.orc_tension3_blur(*self.n, self.g[src], self.g[dst], self.params, 1 if k == 0 else 0)

    def colour(self):
        return "c0" if (TENSION_BLURS - 1) % 2 == 0 else "c1"

    def stress(self):
        lib.orc_tension3_stress(*self.n, self.g[self.colour()], _ptrs([self.g[f] for f in STRESS3]))

    def force(self, node):
        out = np.zeros(3, np.float32)
        lib.orc_tension3_force(*self.n, self.g[self.colour()], _ptrs([self.g[f] for f in STRESS3]), int(node), out)
        return out

    def clear(self):
        lib.orc_mpm3_clear(*self.n, self.g["gm"], self.g["gmu"], self.g["gmv"], self.g["gmw"])

    def scatter(self):
        if self.fixed:
            nodes = self.n[0] * self.n[1] * self.n[2]
            self.q = [np.zeros(nodes, np.int64) for _ in range(4)]
            lib.orc_mpm3_scatter_fixed(*self.n, self.count, self._pp(), self.params, *self.q)
            lib.orc_mpm3_decode_fixed(nodes, *self.q, self.g["gm"], self.g["gmu"], self.g["gmv"], self.g["gmw"])
            return
        lib.orc_mpm3_scatter(*self.n, self.count, self._pp(), self.params,
                             self.g["gm"], self.g["gmu"], self.g["gmv"], self.g["gmw"])

    def grid(self):
        if self.tension:
            lib.orc_mpm3_grid_tension(*self.n, self.params, *[self.g[f] for f in GRID3],
                                      _ptrs([self.g[f] for f in STRESS3]), self.g[self.colour()])
            return
        lib.orc_mpm3_grid(*self.n, self.params, *[self.g[f] for f in GRID3])

    def advect(self):
        lib.orc_mpm3_advect(*self.n, self.count, self._pp(), self.params,
                            self.g["gu"], self.g["gv"], self.g["gw"])

    def step(self, n=1):
        if self.tension or self.fixed:
            for _ in range(n):
                self.clear()
                self.scatter()
                if self.tension:
                    for k in range(TENSION_BLURS):
                        self.blur(k)
                    self.stress()
                self.grid()
                self.advect()
            return
        grid = _ptrs([self.g[f] for f in GRID3])
        for _ in range(n):
            lib.orc_mpm3_step(*self.n, self.count, self._pp(), self.params, grid)

    def step_omp(self, n=1, threads=4):
        """Plain steps on `threads` host threads (the CPU baseline's step; its P2G sums in another
        order than step's)."""
        grid = _ptrs([self.g[f] for f in GRID3])
        for _ in range(n):
            lib.orc_mpm3_step_omp(*self.n, self.count, self._pp(), self.params, grid, int(threads))

    def keys(self):
        out = np.empty(self.count, np.int32)
        lib.orc_mpm3_keys(*self.n, self.count, self.p["x"], self.p["y"], self.p["z"], self.p["m"], out)
        return out

    def lod(self, level, target, stay):
        """Flip3.lod over count // 8 groups; level, target and stay are float32 arrays, updated."""
        lib.orc_mpm3_lod(*self.n, self.count // 8, self._pp(), level, target, stay)

    def lod_target(self, level, target, view):
        """Flip3.lodTarget; writes target."""
        lib.orc_mpm3_lod_target(self.count // 8, self.p["x"], self.p["y"], self.p["z"], level, target,
                                _c(np.asarray(view, np.float32)))

    def sort(self, layers=False):
        """The sort by brick key; with ``layers`` the layer-by-layer order of P2G schedule 3.
        Returns the permutation applied."""
        keys = self.keys()
        order = stable_order(keys)
        if layers:
            order = order[mpm3_layer_order(keys[order], mpm3_inactive_key(*self.n))]
        for f in FIELDS3:
            self.p[f] = self.p[f][order].copy()
        return order


MPM3_GROUP_SHIFT = 8  # a group of P2G schedule 3: the 256 cells whose keys share key >> 8


def mpm3_layer_order(sorted_keys, inactive_key) -> np.ndarray:
    """The group-by-group order of the 3D sort for P2G schedule 3, as a permutation of the
    key-sorted particles (no reference counterpart: the reference's 3D step has no sort; this is
    the library's layout). A group is the keys sharing key >> 8 (256 cells, half a sort brick).
    Its cells are placed by particle count, most first, counts compared as min(count, 255), ties by
    cell number; place k is generation k // 64, lane k % 64. The group's particles go generation
    by generation, and within one rank-major: the first particle of each cell in lane order, then
    every second, and so on (each cell's particles keep their sorted order). The inactive keys
    (from ``inactive_key`` on) keep their places at the end."""
    keys = np.asarray(sorted_keys, np.int64)
    n = len(keys)
    live = int(np.searchsorted(keys, inactive_key, side="left"))
    order = np.arange(n, dtype=np.int64)
    if live:
        k = keys[:live]
        rank = np.arange(live) - np.searchsorted(k, k, side="left")  # place within the cell
        cells, inverse, counts = np.unique(k, return_inverse=True, return_counts=True)
        group, cell = cells >> MPM3_GROUP_SHIFT, cells & ((1 << MPM3_GROUP_SHIFT) - 1)
        by_count = np.lexsort((cell, -np.minimum(counts, 255), group))
        place = np.empty(len(cells), np.int64)
        place[by_count] = np.arange(len(cells)) - np.searchsorted(group[by_count], group[by_count], side="left")
        p = place[inverse]
        order[:live] = np.lexsort((p % 64, rank, p // 64, k >> MPM3_GROUP_SHIFT))
    return order


def mpm3_key_bits(nx, ny, nz) -> int:
    return lib.orc_mpm3_key_bits(nx, ny, nz)


def mpm3_inactive_key(nx, ny, nz) -> int:
    return lib.orc_mpm3_inactive_key(nx, ny, nz)


STRESS2 = ("txx", "tyy", "txy")


class Mpm2:
    """A 2D particle state on the CPU, stepped as FlipStep lists its passes. With ``fixed`` the
    scatter is the deterministic fixed-point one (orc_mpm2_scatter_fixed) and its sums are
    decoded into gm, gmu, gmv before the grid update. With ``tension`` the step is
    FlipStep(..., tension)'s (core/particle/FlipStep.java:111-143): clear, scatter, six blurs
    (gm -> c0 -> c1 -> ...), stress, grid with tension, advect."""

    def __init__(self, nx, ny, fields: dict, params, fixed=False, tension=False):
        self.fixed = fixed
        self.tension = tension
        self.n = (nx, ny)
        count = len(fields["x"])
        self.p = {f: _c(fields.get(f, np.zeros(count)), np.float32).copy() for f in FIELDS2}
        if "j" not in fields:
            self.p["j"][:] = 1
        self.count = count
        self.g = {f: np.zeros(nx * ny, np.float32) for f in GRID2}
        if tension:
            for f in ("c0", "c1") + STRESS2:
                self.g[f] = np.zeros(nx * ny, np.float32)
        self.params = np.asarray(params, np.int32)
        self.steps = 0

    def blur(self, k):
        """Pass blur k: gm (scaled to a colour) into c0 for k = 0, then c0 and c1 in turn."""
        src = "gm" if k == 0 else ("c0" if k % 2 == 1 else "c1")
        dst = "c0" if k % 2 == 0 else "c1"
        lib.orc_tension2_blur(*self.n, self.g[src], self.g[dst], self.params, 1 if k == 0 else 0)

    def colour(self):
        return "c0" if (TENSION_BLURS - 1) % 2 == 0 else "c1"

    def stress(self):
        lib.orc_tension2_stress(*self.n, self.g[self.colour()], *[self.g[f] for f in STRESS2])

    def force(self, node):
        out = np.zeros(2, np.float32)
        lib.orc_tension2_force(*self.n, self.g[self.colour()], *[self.g[f] for f in STRESS2], int(node), out)
        return out

    def _pp(self):
        return _ptrs([self.p[f] for f in FIELDS2])

    def clear(self):
        lib.orc_mpm2_clear(*self.n, self.g["gm"], self.g["gmu"], self.g["gmv"])

    def scatter(self):
        if self.fixed:
            nodes = self.n[0] * self.n[1]
            q = [np.zeros(nodes, np.int64) for _ in range(3)]
            lib.orc_mpm2_scatter_fixed(*self.n, self.count, self._pp(), self.params, *q)
            self.q = q
            lib.orc_mpm2_decode_fixed(nodes, *q, self.g["gm"], self.g["gmu"], self.g["gmv"])
            return
        lib.orc_mpm2_scatter(*self.n, self.count, self._pp(), self.params,
                             self.g["gm"], self.g["gmu"], self.g["gmv"])

    def grid(self):
        if self.tension:
            lib.orc_mpm2_grid_tension(*self.n, self.params, *[self.g[f] for f in GRID2],
                                      *[self.g[f] for f in STRESS2], self.g[self.colour()])
            return
        lib.orc_mpm2_grid(*self.n, self.params, *[self.g[f] for f in GRID2])

    def advect(self):
        lib.orc_mpm2_advect(*self.n, self.count, self._pp(), self.params, self.g["gu"], self.g["gv"])

    def step(self, n=1):
        if self.fixed or self.tension:
            for _ in range(n):
                self.clear()
                self.scatter()
                if self.tension:
                    for k in range(TENSION_BLURS):
                        self.blur(k)
                    self.stress()
                self.grid()
                self.advect()
            return
        grid = _ptrs([self.g[f] for f in GRID2])
        for _ in range(n):
            lib.orc_mpm2_step(*self.n, self.count, self._pp(), self.params, grid)

    def sort(self):
        lib.orc_mpm2_sort(*self.n, self.count, self._pp())

    def advance(self, count, sort_every):
        """ParticleSimulation.advance (gui/ParticleSimulation.java:175-183)."""
        for _ in range(count):
            if self.steps % sort_every == 0:
                self.sort()
            self.step()
            self.steps += 1


def mpm2_cell(x, y, nx, ny) -> int:
    return lib.orc_mpm2_cell(float(x), float(y), nx, ny)


def mpm2_keys(x, y, nx, ny) -> np.ndarray:
    out = np.empty(len(x), np.int32)
    lib.orc_mpm2_keys(nx, ny, len(x), _c(x), _c(y), out)
    return out


def mpm2_cell_starts(x, y, nx, ny) -> np.ndarray:
    """Scatter.cellStarts; raises ValueError on particles out of cell order, as the reference does."""
    starts = np.empty((nx - 1) * (ny - 1) + 1, np.int32)
    r = lib.orc_mpm2_cell_starts(nx, ny, len(x), _c(x), _c(y), starts)
    if r != 0:
        raise ValueError(f"particle {-1 - r} is out of order; the particles must be sorted by cell")
    return starts


def stable_order(keys) -> np.ndarray:
    keys = _c(keys, np.int32)
    out = np.empty(len(keys), np.int32)
    lib.orc_stable_order(len(keys), keys, out)
    return out


def bucket_starts(keys, cells) -> np.ndarray:
    keys = _c(keys, np.int32)
    out = np.empty(cells + 1, np.int32)
    lib.orc_bucket_starts(len(keys), keys, cells, out)
    return out


# ------------------------------------------------------------------ the fixed-point bilinear scatter

FIXED_OFFSET_BITS = 10
FIXED_MAX_MASS = 1 << 21


def fixed_split(q: int, mass: int) -> int:
    """Quanta's share floor(q * M / 2^20) in i32 (core/particle/FixedScatter.java:186-193)."""
    return lib.orc_fixed_split(int(q), int(mass))


def fixed_scatter(n, p: dict, mass_bits, velocity_bits, grid=None):
    """FixedScatter's deposit into three i32 grids (added to `grid` if given), the host restatement
    of test/particle/FixedScatterTest.java:402-432."""
    grid = [np.zeros(n * n, np.int32) for _ in range(3)] if grid is None else grid
    lib.orc_fixed_scatter(n, len(p["x"]), *(_c(p[k]) for k in ("x", "y", "u", "v", "m")), int(mass_bits),
                          int(velocity_bits), *grid)
    return grid


def scatter_f64(n, p: dict):
    """The FP64 bilinear deposit of test/particle/ScatterTest.java:429-450."""
    grid = [np.zeros(n * n, np.float64) for _ in range(3)]
    lib.orc_scatter_f64(n, len(p["x"]), *(_c(p[k]) for k in ("x", "y", "u", "v", "m")), *grid)
    return grid


def jittered_particles(cells, ppc, rng: "JavaRandom") -> dict:
    """ScatterTest.Particles.jittered (test/particle/ScatterTest.java:460-478): per cell in row,
    column order, per particle nextFloat x, nextFloat y, nextGaussian u, nextGaussian v, nextFloat m;
    x kept below col + 1 in binary32."""
    count = cells * cells * ppc
    out = {k: np.empty(count, np.float32) for k in ("x", "y", "u", "v", "m")}
    k = 0
    for row in range(cells):
        for col in range(cells):
            for _ in range(ppc):
                below_x = np.nextafter(np.float32(col + 1), np.float32(0))
                below_y = np.nextafter(np.float32(row + 1), np.float32(0))
                out["x"][k] = min(np.float32(np.float32(col) + np.float32(rng.next_float())), below_x)
                out["y"][k] = min(np.float32(np.float32(row) + np.float32(rng.next_float())), below_y)
                out["u"][k] = np.float32(rng.next_gaussian())
                out["v"][k] = np.float32(rng.next_gaussian())
                out["m"][k] = np.float32(np.float32(0.5) + np.float32(rng.next_float()))
                k += 1
    return out


def shuffled(p: dict, rng: "JavaRandom") -> dict:
    """ScatterTest.Particles.shuffled (:481-493): Fisher-Yates with nextInt(k + 1), k from the end."""
    count = len(p["x"])
    order = list(range(count))
    for k in range(count - 1, 0, -1):
        j = rng.next_int(k + 1)
        order[k], order[j] = order[j], order[k]
    return {key: v[np.array(order)] for key, v in p.items()}


PARTICLE_DIAGNOSTICS = ("min_density", "max_density", "low_density", "high_density", "mach", "courant", "non_finite")


def particle_diagnostics(j, speed, sound, dt) -> dict:
    """ParticleDiagnostics.of (core/particle/ParticleDiagnostics.java:45-65)."""
    out = np.zeros(7, np.float64)
    j = _c(np.asarray(j, np.float32))
    lib.orc_particle_diagnostics(len(j), j, float(speed), float(sound), float(dt), out)
    d = dict(zip(PARTICLE_DIAGNOSTICS, out.tolist()))
    d["non_finite"] = int(d["non_finite"])
    return d
