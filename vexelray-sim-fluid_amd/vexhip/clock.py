"""The host's clock in integer microsecond ticks.

Restates the behaviour the reference needs from ``dev.vexelray.sim.core.time.Clock``
(vexelray-sim-core 0.1.3-SNAPSHOT, an external dependency that is not vendored), as its call
sites use it: ``ShallowWater.TICKS_PER_SECOND`` (core/stencil/ShallowWater.java:146-147),
``grant``/``settle``/``now``/``outstanding`` (gui/PatchSimulation.java:120-145,
test/stencil/Stepper.java:54-108). A grant is what the device may spend; it fits an i32
(core/stencil/ShallowWater.java:104-109), so a grant reaching further is capped there.
"""
from __future__ import annotations

TICKS_PER_SECOND = 1_000_000
I32_MAX = 2**31 - 1


def ticks(seconds: float) -> int:
    """``seconds`` as ticks, to the nearest microsecond (ShallowWater.ticks, :274-276)."""
    return int(_java_round(seconds * TICKS_PER_SECOND))


def seconds(t: int) -> float:
    """``t`` ticks as seconds (ShallowWater.seconds, :279-281)."""
    return t / TICKS_PER_SECOND


def _java_round(x: float) -> int:
    # Math.round: floor(x + 1/2)
    import math
    return math.floor(x + 0.5)


class Clock:
    """Settled time plus at most one outstanding grant."""

    def __init__(self) -> None:
        self._now = 0
        self._outstanding = 0

    def now(self) -> int:
        return self._now

    def outstanding(self) -> int:
        return self._outstanding

    def grant(self, until: int) -> int:
        """Grants the ticks from now up to ``until``, at most an i32's worth; returns them."""
        amount = max(0, min(int(until) - self._now, I32_MAX))
        self._outstanding = amount
        return amount

    def settle(self, left: int) -> None:
        """Closes the grant: what was spent of it, the grant less ``left``, becomes settled time."""
        self._now += self._outstanding - int(left)
        self._outstanding = 0
