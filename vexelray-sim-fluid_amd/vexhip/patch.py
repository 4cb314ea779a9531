"""PatchSimulation: one shallow-water patch on resident device buffers, advanced by slices of
simulated time -- gui/PatchSimulation.java's API over libvexelhip.so.

The host keeps the clock and grants the device budgets from it; the device sizes its own steps
and writes back what is left (core/stencil/ShallowWater.java:96-109). The buffer rotation of
PatchSimulation.dispatch (:197-205) lives in the library, behind vxr_swe_dispatch.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native
from .clock import Clock, seconds

WALL, OPEN = "WALL", "OPEN"


def params(g, courant, dx, dy, dry=1e-6) -> np.ndarray:
    """ShallowWater.params (:248-254): Courant in (0, 1/2]."""
    if not (0 < courant <= 0.5):
        raise ValueError(f"a first-order 2D step needs a Courant number in (0, 1/2], got {courant}")
    return params_allowing_instability(g, courant, dx, dy, dry)


def params_allowing_instability(g, courant, dx, dy, dry=1e-6) -> np.ndarray:
    """ShallowWater.paramsAllowingInstability (:261-271)."""
    if not courant > 0:
        raise ValueError(f"a Courant number must be positive, got {courant}")
    return np.array([g, courant, dx, dy, dry], np.float64).astype(np.float32).view(np.int32)


def fastest_wave(h, hu, hv, g, dry) -> int:
    """ShallowWater.fastestWave (:287-295) as the SPEED word: float quotients, a double sqrt cast
    to float, a float max -- elementwise, so the order of the max does not matter."""
    h = np.asarray(h, np.float32)
    wet = h.astype(np.float64) > dry
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.where(wet, np.asarray(hu, np.float32) / h, np.float32(0)).astype(np.float32)
        w = np.where(wet, np.asarray(hv, np.float32) / h, np.float32(0)).astype(np.float32)
    c = np.sqrt(g * h.astype(np.float64)).astype(np.float32)
    s = np.maximum(np.abs(u), np.abs(w)) + c
    fastest = np.float32(max(0.0, float(s.max()))) if s.size else np.float32(0)
    return int(np.array([fastest], np.float32).view(np.int32)[0])


@dataclass
class Advance:
    """What one advance did (PatchSimulation.Advance, :33-39)."""
    steps: int
    spent_ticks: int
    behind_ticks: int

    def mean_step(self) -> float:
        return 0.0 if self.steps == 0 else seconds(self.spent_ticks) / self.steps


class PatchSimulation:
    """gui/PatchSimulation.java:30-245. Owning-thread only."""

    def __init__(self, nx: int, ny: int, edges=(WALL, WALL, WALL, WALL)):
        flags = 0
        bits = (_native.FLAG_OPEN_WEST, _native.FLAG_OPEN_EAST, _native.FLAG_OPEN_SOUTH, _native.FLAG_OPEN_NORTH)
        for edge, bit in zip(edges, bits):
            if edge not in (WALL, OPEN):
                raise ValueError(f"every side of a patch needs an edge, got {edge!r}")
            if edge == OPEN:
                flags |= bit
        self._nx, self._ny = nx, ny
        self.sim = _native.Sim(_native.KIND_SWE, (nx, ny, 1), 0, flags)
        self.clock = Clock()
        self.total_steps = 0
        self.last_step = 0

    def nx(self) -> int:
        return self._nx

    def ny(self) -> int:
        return self._ny

    def on_gpu(self) -> bool:
        return True

    def load(self, h, hu, hv, words) -> None:
        """Replaces the state and the parameters and starts the clock again (:92-109)."""
        self.total_steps = 0
        self.last_step = 0
        self.clock = Clock()
        self.sim.clear()   # every buffer zero and the rotation back at slot 0
        self.sim.write("h", np.asarray(h, np.float32))
        self.sim.write("hu", np.asarray(hu, np.float32))
        self.sim.write("hv", np.asarray(hv, np.float32))
        w = np.asarray(words, np.int32)
        g = float(w[:1].view(np.float32)[0])
        dry = float(w[4:5].view(np.float32)[0])
        self.sim.write("speed", np.array([fastest_wave(h, hu, hv, g, dry)], np.int32))
        self.sim.write("params", w)

    def params(self, words) -> None:
        self.sim.write("params", np.asarray(words, np.int32))

    def write_budget(self, ticks: int) -> None:
        self.sim.write("budget", np.array([ticks], np.int32))

    def read_budget(self) -> int:
        return int(self.sim.read_words("budget")[0])

    def advance(self, ticks: int, max_steps: int) -> Advance:
        """Advances by up to ``ticks`` of simulated time in at most ``max_steps`` steps (:120-145)."""
        before = self.clock.now()
        granted = self.clock.grant(before + ticks)
        self.write_budget(granted)
        left = granted
        steps = 0
        while left > 0 and steps < max_steps:
            batch = min(max_steps - steps, (left + self.last_step - 1) // self.last_step) if self.last_step > 0 else 1
            self.sim.swe_dispatch(batch)
            steps += batch
            now = self.read_budget()
            # learn the step only from a batch that left budget over (:133-139)
            if 0 < now < left:
                self.last_step = (left - now) // batch
            left = now
        self.clock.settle(left)
        self.total_steps += steps
        return Advance(steps, self.clock.now() - before, left)

    def dispatch(self, count: int) -> None:
        self.sim.swe_dispatch(count)

    def read(self):
        """{h, hu, hv} as they are now; waits for the steps dispatched so far (:148-152)."""
        return [self.sim.read("h"), self.sim.read("hu"), self.sim.read("hv")]

    def fastest_wave(self) -> float:
        return float(self.sim.read("speed")[0])

    def allowed_step(self, courant, dx, dy) -> float:
        f = self.fastest_wave()
        return courant * min(dx, dy) / f if f > 0 else 0.0

    def diagnostics(self, g, dry, dt, dx, dy) -> dict:
        """Diagnostics.of (core/stencil/Diagnostics.java:48-81) of the current state, reduced on the
        device (no readback of the state)."""
        return self.sim.swe_diagnostics(g, dry, dt, dx, dy)

    def clamped(self) -> int:
        return int(self.sim.read_words("clamped")[0])

    def now(self) -> int:
        return self.clock.now()

    def steps(self) -> int:
        return self.total_steps

    def close(self) -> None:
        self.sim.close()
