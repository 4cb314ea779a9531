"""vexhip: the MI355X-native hot path of vexelray-sim-fluid behind its simulator API.

The runners mirror the reference's gui runners (gui/PatchSimulation.java,
gui/ParticleSimulation.java, gui/ParticleSimulation3.java) over ``libvexelhip.so``'s C ABI
(``include/vexel/vxr.h``). There is no CPU fallback: without the HIP library or a device the
runners raise.
"""
from .clock import Clock, TICKS_PER_SECOND, seconds, ticks  # noqa: F401


def __getattr__(name):
    if name in ("PatchSimulation", "Advance"):
        from . import patch
        return getattr(patch, name)
    if name in ("ParticleSimulation", "ParticleSimulation3"):
        from . import particles
        return getattr(particles, name)
    raise AttributeError(name)
