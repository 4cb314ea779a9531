"""Backends over libvexelhip.so for the shared physics cases (swe_cases, mpm_cases)."""
import numpy as np

from vexhip import _native


class GpuSwe:
    """The shallow-water backend of swe_cases.Harness, over a vxr SWE handle."""

    def __init__(self, nx, ny, edges=(0, 0, 0, 0), rows=None):
        flags = 0
        for e, bit in zip(edges, (_native.FLAG_OPEN_WEST, _native.FLAG_OPEN_EAST,
                                  _native.FLAG_OPEN_SOUTH, _native.FLAG_OPEN_NORTH)):
            if e == 1:
                flags |= bit
        self.sim = _native.Sim(_native.KIND_SWE, (nx, ny, 1), 0, flags)
        if rows is not None:
            self.sim.set_option("rows", rows)
        self.dispatched = 0

    def load(self, h, hu, hv, params, first_speed):
        self.sim.clear()
        self.sim.write("h", np.asarray(h, np.float32))
        self.sim.write("hu", np.asarray(hu, np.float32))
        self.sim.write("hv", np.asarray(hv, np.float32))
        self.sim.write("speed", np.array([first_speed], np.int32))
        self.sim.write("params", np.asarray(params, np.int32))
        self.dispatched = 0

    def step(self, n):
        self.sim.swe_dispatch(n)
        self.dispatched += n

    def write_budget(self, ticks):
        self.sim.write("budget", np.array([ticks], np.int32))

    def read_budget(self):
        return int(self.sim.read_words("budget")[0])

    def read(self):
        return [self.sim.read("h").copy(), self.sim.read("hu").copy(), self.sim.read("hv").copy()]

    def clamped(self):
        return int(self.sim.read_words("clamped")[0])

    def fastest_wave_bits(self):
        return int(self.sim.read_words("speed")[0])


class GpuRig:
    """mpm_cases' rig over a vxr MPM handle."""

    def __init__(self, kind, dims, count, p2g=1, g2p=1, run=4, tension=False, **options):
        self.sim = _native.Sim(kind, dims, count, _native.FLAG_TENSION if tension else 0)
        self.sim.set_option("p2g", p2g)
        self.sim.set_option("g2p", g2p)
        self.sim.set_option("run", run)
        for key, value in options.items():
            self.sim.set_option(key, value)

    def write(self, name, values):
        self.sim.write(name, np.asarray(values, np.float32))

    def params(self, words):
        self.sim.write("params", np.asarray(words, np.int32))

    def run(self, passes):
        if passes == "step":
            self.sim.run(_native.GROUP_STEP)
        else:
            for p in passes:
                self.sim.run_pass(p)

    def read(self, name):
        return self.sim.read(name).copy()


def maker3(**opts):
    return lambda nx, ny, nz, count: GpuRig(_native.KIND_MPM3, (nx, ny, nz), count, **opts)


def maker2(**opts):
    return lambda nx, ny, count: GpuRig(_native.KIND_MPM2, (nx, ny, 1), count, **opts)
