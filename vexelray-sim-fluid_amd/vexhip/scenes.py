"""Seeded synthetic inputs for the hot path's workloads.

Seeds go through java.util.Random's generator (the 48-bit LCG of the Java SE specification)
so the reference tests' inputs are reproduced bit for bit, e.g. the jittered column of
test/particle/FlipTest.java:134-150 and the slab of test/particle/Flip3Test.java:144-163.
Bulk draws are vectorised by jumping the LCG ahead: state_{k+j} = A^j state_k + C_j mod 2^48.
"""
from __future__ import annotations

import math

import numpy as np

_MULT = 0x5DEECE66D
_ADD = 0xB
_MASK = (1 << 48) - 1
_CHUNK = 1 << 16
_JUMP = None


def _jump_tables():
    global _JUMP
    if _JUMP is None:
        a = np.empty(_CHUNK, np.uint64)
        c = np.empty(_CHUNK, np.uint64)
        aj, cj = _MULT, _ADD
        for j in range(_CHUNK):
            a[j], c[j] = aj, cj
            aj = (aj * _MULT) & _MASK
            cj = (cj * _MULT + _ADD) & _MASK
        _JUMP = (a, c)
    return _JUMP


class JavaRandom:
    """java.util.Random, scalar, for the few draws a scene's layout needs."""

    def __init__(self, seed: int):
        self.seed = (seed ^ _MULT) & _MASK

    def _next(self, bits: int) -> int:
        self.seed = (self.seed * _MULT + _ADD) & _MASK
        v = self.seed >> (48 - bits)
        return v - (1 << bits) if bits == 32 and v >= 1 << 31 else v

    def next_int(self, bound: int | None = None) -> int:
        if bound is None:
            return self._next(32)
        r = self._next(31)
        m = bound - 1
        if bound & m == 0:
            return (bound * r) >> 31
        u = r
        while True:
            r = u % bound
            if ((u - r + m) & 0xFFFFFFFF) < 1 << 31:
                return r
            u = self._next(31)

    def next_float(self) -> float:
        return self._next(24) * 2.0**-24

    def next_double(self) -> float:
        return ((self._next(26) << 27) + self._next(27)) * 2.0**-53


class JavaRandomFloats:
    """Many nextFloat() draws at once, in call order, as float32."""

    def __init__(self, seed: int):
        self.state = (seed ^ _MULT) & _MASK

    def next_floats(self, n: int) -> np.ndarray:
        a, c = _jump_tables()
        out = np.empty(n, np.float32)
        k = 0
        mask = np.uint64(_MASK)
        while k < n:
            m = min(_CHUNK, n - k)
            s = np.uint64(self.state)
            states = (a[:m] * s + c[:m]) & mask
            out[k:k + m] = (states >> np.uint64(24)).astype(np.float32) * np.float32(2.0**-24)
            self.state = int(states[m - 1])
            k += m
        return out


# ------------------------------------------------------------------ 3D MPM

def jittered_block3(origin, size, ppc, rng: JavaRandomFloats, limit=None, jitter=0.999):
    """``ppc`` particles jittered in each cell of a box of ``size`` cells from ``origin``, filled
    layer (z), row (y), column (x), particle -- Flip3Test's order (test/particle/Flip3Test.java:152-163):
    ``x = (float)(origin + col) + nextFloat() * 0.999f``. Stops after ``limit`` particles."""
    sx, sy, sz = size
    count = sx * sy * sz * ppc if limit is None else int(limit)
    draws = rng.next_floats(3 * count).reshape(count, 3)
    idx = np.arange(count, dtype=np.int64)
    cell = idx // ppc
    col = cell % sx
    row = (cell // sx) % sy
    layer = cell // (sx * sy)
    j = np.float32(jitter)
    x = (origin[0] + col).astype(np.float32) + draws[:, 0] * j
    y = (origin[1] + row).astype(np.float32) + draws[:, 1] * j
    z = (origin[2] + layer).astype(np.float32) + draws[:, 2] * j
    return x, y, z


def mpm3_units(n: int, gravity=-9.8, bulk=10.0, rho0=1.0):
    """Physical unit-box values in the kernels' node units: h = 1/(n-1), g/h, B/h^2."""
    h = 1.0 / (n - 1)
    return gravity / h, bulk / (h * h), rho0


def mpm3_two_cubes(n=256, particles=8_388_608, ppc=8, seed=1):
    """Two cubes of water at seeded, non-overlapping positions inside the walls of an n^3 box.

    Each cube is the smallest whole cube holding half the particles at ``ppc`` a cell, filled in
    raster order and stopped at exactly half, so its last z-layer is partial. Returns a dict of
    x, y, z, m (m = rho0 / ppc with rho0 = 1) and the layout chosen."""
    half = particles // 2
    side = math.ceil((half / ppc) ** (1.0 / 3.0))
    while side ** 3 * ppc < half:
        side += 1
    hi = n - 2 - side          # a cube's cells run origin .. origin+side-1, inside nodes 1 .. n-2
    if hi < 1:
        raise ValueError(f"{particles} particles at {ppc} a cell do not fit in {n}^3")
    layout = JavaRandom(seed)
    origins = []
    while len(origins) < 2:
        o = tuple(1 + layout.next_int(hi) for _ in range(3))
        if all(any(abs(o[a] - p[a]) >= side for a in range(3)) for p in origins):
            origins.append(o)
    rng = JavaRandomFloats(seed + 1)
    parts = [jittered_block3(o, (side, side, side), ppc, rng, limit=half) for o in origins]
    x = np.concatenate([p[0] for p in parts])
    y = np.concatenate([p[1] for p in parts])
    z = np.concatenate([p[2] for p in parts])
    m = np.full(x.shape, np.float32(1.0 / ppc), np.float32)
    return {"x": x, "y": y, "z": z, "m": m, "origins": origins, "side": side}


# ------------------------------------------------------------------ 2D MPM

def jittered_block2(origin, size, ppc, rng: JavaRandomFloats, jitter=0.999):
    """FlipTest.column's order (test/particle/FlipTest.java:134-150): row, column, particle."""
    sx, sy = size
    count = sx * sy * ppc
    draws = rng.next_floats(2 * count).reshape(count, 2)
    cell = np.arange(count, dtype=np.int64) // ppc
    col = cell % sx
    row = cell // sx
    j = np.float32(jitter)
    x = (origin[0] + col).astype(np.float32) + draws[:, 0] * j
    y = (origin[1] + row).astype(np.float32) + draws[:, 1] * j
    return x, y


def mpm2_dam_block(n=128, width=32, height=64, ppc=4, seed=71):
    """A block of water against the lower-left walls of an n^2 grid, ``ppc`` particles a cell."""
    x, y = jittered_block2((1, 1), (width, height), ppc, JavaRandomFloats(seed))
    m = np.full(x.shape, np.float32(1.0 / ppc), np.float32)
    return {"x": x, "y": y, "m": m}


def mpm2_batch_layouts(scenes=256, seed=3, lo=4096, hi=262144, ppc=4, sizes=(64, 128, 256, 512)):
    """The batched 2D workload: ``scenes`` scenes with particle counts log-uniform in [lo, hi]
    and a grid drawn from ``sizes`` among those whose lower half holds the scene's block of
    water at ``ppc`` a cell. Returns a list of dicts: n, count, origin, block (w, h)."""
    r = JavaRandom(seed)
    out = []
    for _ in range(scenes):
        count = int(round(math.exp(math.log(lo) + r.next_double() * (math.log(hi) - math.log(lo)))))
        cells = -(-count // ppc)
        w = max(4, math.ceil(math.sqrt(cells / 2.0)))
        h = -(-cells // w)
        fits = [n for n in sizes if 1 + w <= n - 2 and 1 + h <= (n - 2)]
        n = fits[r.next_int(len(fits))]
        out.append({"n": n, "count": count, "origin": (1, 1), "block": (w, h)})
    return out


def mpm2_batch_scene(layout, seed):
    """One scene of the batch: FlipTest.column's jittered fill of the layout's block, stopped at
    its count, mass rho0 / ppc with rho0 = 1 (4 a cell), and Flip.params words in node units
    (h = 1/(n-1): g = -9.8/h, B = 10/h^2, dt scaled so the acoustic Courant number is about 0.04
    on every grid, 1e-4 at 128^2)."""
    n, count = layout["n"], layout["count"]
    w, h = layout["block"]
    x, y = jittered_block2(layout["origin"], (w, h), 4, JavaRandomFloats(seed))
    x, y = x[:count].copy(), y[:count].copy()
    m = np.full(count, np.float32(0.25), np.float32)
    spacing = 1.0 / (n - 1)
    dt = 1e-4 * 127.0 / (n - 1)
    words = np.zeros(28, np.int32)
    words[:6] = np.array([dt, 0.0, -9.8 / spacing, 10.0 / spacing ** 2, 1.0, 0.0],
                         np.float64).astype(np.float32).view(np.int32)
    return {"x": x, "y": y, "m": m}, words


# ------------------------------------------------------------------ shallow water

def swe_dam_bumps(n=4096, bumps=32, seed=0, radius=0.2):
    """A circular dam of depth 2 (radius ``radius`` about the centre of the unit square) in water
    of depth 1, plus ``bumps`` Gaussian humps of amplitude U(0.05, 0.3), width U(0.01, 0.05) and
    centre U(0, 1)^2, drawn per hump in that order with nextDouble. Returns h, hu, hv (float32)."""
    rng = JavaRandom(seed)
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    xx = c[None, :]
    yy = c[:, None]
    h = np.where((xx - 0.5) ** 2 + (yy - 0.5) ** 2 < radius * radius, 2.0, 1.0)
    for _ in range(bumps):
        amp = 0.05 + 0.25 * rng.next_double()
        width = 0.01 + 0.04 * rng.next_double()
        cx = rng.next_double()
        cy = rng.next_double()
        h = h + amp * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (width * width))
    h = h.astype(np.float32).ravel()
    z = np.zeros_like(h)
    return h, z, z.copy()
