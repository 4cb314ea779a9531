"""The fixed-point bilinear scatter (core/particle/FixedScatter.java) over libvexelhip.so.

``FixedPoint`` is FixedScatter.FixedPoint (:91-148): the mass and velocity quanta and the rule for
the finest scale whose i32 registers read back exactly. ``scatter`` deposits particles held in
device memory (torch tensors) into three i32 grids with either schedule (vxr_fixed_scatter).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

from . import _native

OFFSET_BITS = 10        # FixedScatter.OFFSET_BITS
MAX_MASS = 1 << 21      # FixedScatter.MAX_MASS
DIRECT, SEGMENTED = 0, 1


@dataclass(frozen=True)
class FixedPoint:
    mass_bits: int
    velocity_bits: int

    def __post_init__(self):
        if not (0 <= self.mass_bits <= 30 and 0 <= self.velocity_bits <= 30):
            raise ValueError(f"bits out of range: {self.mass_bits}, {self.velocity_bits}")

    @staticmethod
    def for_range(deposits: int, max_mass: float, max_speed: float) -> "FixedPoint":
        """The finest scale whose registers read back exactly for nodes of at most `deposits`
        deposits; momentum binds, and the velocity takes the smaller half of the bits (:102-117)."""
        total = math.floor(math.log(2.0 ** 30 / (deposits * max_mass * max(max_speed, 1))) / math.log(2))
        if total < 0:
            raise ValueError(f"no i32 scale fits {deposits} deposits of mass {max_mass} at speed {max_speed}")
        velocity = min(total // 2, 30)
        mass = min(total - velocity, 30)
        while mass > 0 and not FixedPoint(mass, velocity).fits(deposits, max_mass, max_speed):
            mass -= 1
        return FixedPoint(mass, velocity)

    def mass_scale(self) -> float:
        return float(2.0 ** self.mass_bits)

    def velocity_scale(self) -> float:
        return float(2.0 ** self.velocity_bits)

    def fits(self, deposits: int, max_mass: float, max_speed: float) -> bool:
        """2 K D < 2^31 for the largest mass and momentum deposit D, and a particle's mass in quanta
        below MAX_MASS (:128-138)."""
        mass = math.floor(max_mass * self.mass_scale())
        momentum = mass * math.floor(max_speed * self.velocity_scale())
        limit = 2.0 ** 31
        return mass < MAX_MASS and 2.0 * deposits * mass < limit and 2.0 * deposits * momentum < limit

    def mass(self, quanta: int) -> float:
        return math.ldexp(float(quanta), -self.mass_bits)

    def momentum(self, quanta: int) -> float:
        return math.ldexp(float(quanta), -(self.mass_bits + self.velocity_bits))


def scatter(n, x, y, u, v, m, scale: FixedPoint, schedule=SEGMENTED, grid=None, stream=0):
    """Adds the particles' shares into `grid` (three int32 device tensors of n*n; made zero if
    None) and returns it. Particle fields are float32 device tensors."""
    import torch

    if grid is None:
        grid = [torch.zeros(n * n, dtype=torch.int32, device=x.device) for _ in range(3)]
    if stream == 0:
        stream = torch.cuda.current_stream().cuda_stream
    _native.check(_native.lib().vxr_fixed_scatter(
        ctypes.c_void_p(stream), n, x.numel(), x.data_ptr(), y.data_ptr(), u.data_ptr(), v.data_ptr(), m.data_ptr(),
        scale.mass_bits, scale.velocity_bits, schedule, *(g.data_ptr() for g in grid)))
    return grid
