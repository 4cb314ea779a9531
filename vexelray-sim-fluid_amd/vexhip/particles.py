"""ParticleSimulation (2D) and ParticleSimulation3 (3D): MLS-MPM on resident device buffers --
gui/ParticleSimulation.java and gui/ParticleSimulation3.java's API over libvexelhip.so.
"""
from __future__ import annotations

import numpy as np

from . import _native

AFFINE2 = ("c00", "c01", "c10", "c11")
AFFINE3 = ("c00", "c01", "c02", "c10", "c11", "c12", "c20", "c21", "c22")


def params2(dt, gx, gy, bulk, rho0, sigma=0.0) -> np.ndarray:
    """Flip.params (core/particle/Flip.java:145-211): 28 words, the optional physics off."""
    if not (dt > 0) or not (bulk >= 0) or not (rho0 > 0):
        raise ValueError(f"dt > 0, bulk >= 0 and rho0 > 0, got dt {dt}, bulk {bulk}, rho0 {rho0}")
    w = np.zeros(28, np.int32)
    w[:6] = np.array([dt, gx, gy, bulk, rho0, sigma], np.float64).astype(np.float32).view(np.int32)
    return w


def params3(dt, gx, gy, gz, bulk, rho0, sigma=0.0) -> np.ndarray:
    """Flip3.params (core/particle/Flip3.java:89-103)."""
    if not (dt > 0) or not (bulk >= 0) or not (rho0 > 0):
        raise ValueError(f"dt > 0, bulk >= 0 and rho0 > 0, got dt {dt}, bulk {bulk}, rho0 {rho0}")
    return np.array([dt, gx, gy, gz, bulk, rho0, sigma], np.float64).astype(np.float32).view(np.int32)


def stable_step(bulk, rho0, speed, courant) -> float:
    """Flip.stableStep (:214-216): C / (sqrt(B/rho0) + |v|)."""
    return courant / (np.sqrt(bulk / rho0) + speed)


def _flags(tension=False, heat=False, convection=False, relax=False, pump=False, plastic=False) -> int:
    f = 0
    for on, bit in ((tension, _native.FLAG_TENSION), (heat, _native.FLAG_HEAT),
                    (convection, _native.FLAG_CONVECTION), (relax, _native.FLAG_RELAX),
                    (pump, _native.FLAG_PUMP), (plastic, _native.FLAG_PLASTIC)):
        if on:
            f |= bit
    return f


class _Runner:
    def _write(self, name, values):
        self.sim.write(name, np.asarray(values, np.float32))

    def _read(self, name):
        return self.sim.read(name)

    def on_gpu(self) -> bool:
        return True

    def params(self, words) -> None:
        self.sim.write("params", np.asarray(words, np.int32))

    def steps(self) -> int:
        return self._steps

    def finish(self) -> None:
        self.sim.finish()

    def device_ptr(self, name: str) -> int:
        return self.sim.device_ptr(name)

    def close(self) -> None:
        self.sim.close()

    def diagnostics(self, speed, sound, dt) -> dict:
        """ParticleDiagnostics.of (core/particle/ParticleDiagnostics.java:45-65) of the particles' J,
        reduced on the device (demo/FlatRun.java:189 and demo/BoxRun.java:365 read it back instead)."""
        return self.sim.particle_diagnostics(speed, sound, dt)

    def _advance(self, count: int, sort_every: int | None) -> None:
        self.sim.advance(count, sort_every or 0, self._steps)
        self._steps += count


class ParticleSimulation(_Runner):
    """gui/ParticleSimulation.java:40-239, over core/particle/FlipStep.java's passes."""

    def __init__(self, nx, ny, particles, tension=False, heat=False, convection=False, relax=False, pump=False):
        if convection and (not heat or tension):
            raise ValueError("convection needs heat and does not run with tension")
        self.sim = _native.Sim(_native.KIND_MPM2, (nx, ny, 1), particles,
                               _flags(tension, heat, convection, relax, pump))
        self._nx, self._ny, self._particles = nx, ny, particles
        self._steps = 0

    def nx(self):
        return self._nx

    def ny(self):
        return self._ny

    def particles(self):
        return self._particles

    def load(self, x, y, u, v, m, words) -> None:
        """Replaces the particles and the parameters; J = 1, C = 0 (:129-143)."""
        n = self._particles
        self._write("x", x)
        self._write("y", y)
        self._write("u", u)
        self._write("v", v)
        self._write("m", m)
        self._write("j", np.ones(n, np.float32))
        for c in AFFINE2:
            self._write(c, np.zeros(n, np.float32))
        self.params(words)
        self._steps = 0

    def advance(self, count: int, sort_every: int) -> None:
        """count steps, sorting first on every step whose number is a multiple of sort_every (:175-183)."""
        self._advance(count, sort_every)

    def grid(self):
        return [self._read("gm"), self._read("gmu"), self._read("gmv")]

    def masses(self):
        return self._read("m")

    def velocities(self):
        return [self._read("u"), self._read("v")]

    def positions(self):
        return [self._read("x"), self._read("y")]

    def compression(self, j=None):
        if j is None:
            return self._read("j")
        self._write("j", j)
        return None


class ParticleSimulation3(_Runner):
    """gui/ParticleSimulation3.java:110-408, over core/particle/Flip3Step.java's passes."""

    def __init__(self, nx, ny, nz, particles, tension=False, plastic=False):
        self.sim = _native.Sim(_native.KIND_MPM3, (nx, ny, nz), particles, _flags(tension=tension, plastic=plastic))
        self._n = (nx, ny, nz)
        self._particles = particles
        self._steps = 0
        self._nominal = np.float32(np.finfo(np.float32).max)

    def nx(self):
        return self._n[0]

    def ny(self):
        return self._n[1]

    def nz(self):
        return self._n[2]

    def particles(self):
        return self._particles

    def load(self, at, velocity, mass, words, kind=None) -> None:
        """Replaces the particles (node units) and the parameters; J = 1, C = 0 (:227-275)."""
        if kind is not None:
            raise ValueError("a simulation that is not plastic has no kinds of particle")
        n = self._particles
        for name, v in zip("xyz", at):
            self._write(name, v)
        for name, v in zip("uvw", velocity):
            self._write(name, v)
        m = np.asarray(mass, np.float32)
        self._write("m", m)
        # the level of detail starts full: every slot staying, every group at 8 (:266-273)
        groups = n // 8
        if groups:
            self._write("level", np.full(groups, 8, np.float32))
            self._write("target", np.full(groups, 8, np.float32))
        self._write("stay", np.ones(n, np.float32))
        self._write("view", np.array([0, 0, 0, 1e9, 1], np.float32))
        live = m[m > 0]
        self._nominal = live.min() if live.size else np.float32(np.finfo(np.float32).max)
        self._write("j", np.ones(n, np.float32))
        # every slot at its own position again: the level of detail's groups are the slots as loaded
        self.sim.write("slot", np.arange(n, dtype=np.int32))
        for c in AFFINE3:
            self._write(c, np.zeros(n, np.float32))
        self.params(words)
        self._steps = 0

    def advance(self, count: int, sort_every: int | None = None) -> None:
        """count steps (:291-296); with sort_every, a sort by cell first on every step whose number
        is a multiple of it (the 2D runner's cadence, which the 3D reference step lacks)."""
        self._advance(count, sort_every)

    def sort(self) -> None:
        self.sim.run(_native.GROUP_SORT)

    def refine(self, eye, near, floor) -> None:
        """Moves every group of 8 slots one step toward the level its distance from `eye` asks for
        (:346-349): lodTarget, then lod. Groups are the slots as loaded; a sort moves slots, and
        the passes reach a group's slots wherever the sort put them. The next sort moves the slots
        left without mass behind every active one, and P2G and G2P then stop at the last active."""
        self._write("view", np.array([eye[0], eye[1], eye[2], near, floor], np.float32))
        self.sim.run(_native.GROUP_LOD)

    def active(self) -> int:
        """Slots with mass (:352-360), a readback."""
        return int(np.count_nonzero(self._read("m") > 0))

    def grid(self):
        return [self._read(g) for g in ("gm", "gmu", "gmv", "gmw")]

    def slots(self):
        """The slot each particle position holds (the identity until a sort)."""
        return self.sim.read_words("slot")

    def compression(self):
        """J of the particles holding at least half the nominal mass (:312-333)."""
        j, m = self._read("j"), self._read("m")
        keep = m >= np.float32(0.5) * self._nominal
        return j if keep.all() else j[keep]

    def velocities(self):
        return [self._read(v) for v in "uvw"]

    def positions(self):
        return [self._read(v) for v in "xyz"]


class ParticleBatch:
    """Many 2D scenes stepped together: one vxr batch (vxr_create_batch), one launch per pass for
    every scene. Scene k is what a ParticleSimulation of the same size would hold, and steps to
    the same bits as one with the fixed-point P2G (option p2g = 2). No reference counterpart: the
    reference steps one gui/ParticleSimulation per scene."""

    def __init__(self, sizes, particles):
        """``sizes``: (nx, ny) per scene; ``particles``: the particle count per scene."""
        nx = [int(s[0]) for s in sizes]
        ny = [int(s[1]) for s in sizes]
        self.sim = _native.Sim.batch(nx, ny, particles)
        self._sizes = list(zip(nx, ny))
        self._counts = [int(c) for c in particles]
        self.poff, self.goff, self.soff = self.sim.batch_layout()
        self._steps = 0

    def scenes(self) -> int:
        return len(self._counts)

    def on_gpu(self) -> bool:
        return True

    def _gather(self, per_scene, offsets, length, fill=0.0):
        out = np.full(length, fill, np.float32)
        for k, values in enumerate(per_scene):
            v = np.asarray(values, np.float32)
            out[offsets[k]:offsets[k] + len(v)] = v
        return out

    def load(self, scenes, params) -> None:
        """``scenes[k]``: dict with x, y and m (u, v optional), node units; ``params[k]``: the 28
        Flip.params words of scene k. J = 1, C = 0, as ParticleSimulation.load sets them."""
        if len(scenes) != self.scenes() or len(params) != self.scenes():
            raise ValueError(f"{self.scenes()} scenes and their params, got {len(scenes)} and {len(params)}")
        cap = int(self.poff[-1])
        for k, sc in enumerate(scenes):
            if len(sc["x"]) != self._counts[k]:
                raise ValueError(f"scene {k} holds {self._counts[k]} particles, got {len(sc['x'])}")
        for name in ("x", "y", "u", "v", "m"):
            self.sim.write(name, self._gather([sc.get(name, np.zeros(len(sc["x"]))) for sc in scenes],
                                              self.poff, cap))
        self.sim.write("j", self._gather([np.ones(c) for c in self._counts], self.poff, cap))
        zero = np.zeros(cap, np.float32)
        for c in AFFINE2:
            self.sim.write(c, zero)
        self.sim.write("params", np.concatenate([np.asarray(w, np.int32) for w in params]))
        self._steps = 0

    def advance(self, count: int, sort_every: int) -> None:
        """ParticleSimulation.advance for every scene at once."""
        self.sim.advance(count, sort_every, self._steps)
        self._steps += count

    def steps(self) -> int:
        return self._steps

    def scene(self, k: int, names=("x", "y", "u", "v", "m", "j", "c00", "c01", "c10", "c11")) -> dict:
        """Scene k's particle fields."""
        lo, n = int(self.poff[k]), self._counts[k]
        return {name: self.sim.read(name)[lo:lo + n].copy() for name in names}

    def scene_grid(self, k: int, names=("gm", "gmu", "gmv", "gu", "gv")) -> dict:
        nx, ny = self._sizes[k]
        lo = int(self.goff[k])
        return {name: self.sim.read(name)[lo:lo + nx * ny].copy() for name in names}

    def scene_starts(self, k: int) -> np.ndarray:
        nx, ny = self._sizes[k]
        lo = int(self.soff[k])
        return self.sim.read_words("starts")[lo:lo + (nx - 1) * (ny - 1) + 1].copy()

    def finish(self) -> None:
        self.sim.finish()

    def close(self) -> None:
        self.sim.close()
