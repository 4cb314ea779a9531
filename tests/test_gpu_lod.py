"""The device Flip3.lod and Flip3.lodTarget: the reference's Flip3LodTest cases on the HIP rig, and
every call bit for bit the oracle's over long thinning and filling sequences."""
import numpy as np
import pytest

import lod_cases as cases
import oracle
from vexhip import _native, particles

pytestmark = pytest.mark.gpu


class GpuLodRig:
    def __init__(self, n=cases.N, slots=cases.SLOTS):
        self.sim = _native.Sim(_native.KIND_MPM3, (n, n, n), slots)
        self.sim.set_option("p2g", 0)
        self.sim.set_option("g2p", 0)

    def write(self, name, values):
        v = np.asarray(values)
        self.sim.write(name, v.astype(np.int32) if name == "params" else v.astype(np.float32))

    def read(self, name):
        return self.sim.read(name).copy()

    def run(self, what):
        if what == "step":
            self.sim.run(_native.GROUP_STEP)
        else:
            self.sim.run_pass(what)


class SortingGpuLodRig(GpuLodRig):
    """The device rig with a sort by cell before every lod and lod_target call: the slots move
    (the inactive ones behind every active one) and the passes reach each group's slots through
    the slot map. Particle fields are written and read in slot order."""

    def _slots(self):
        return self.sim.read_words("slot")

    def write(self, name, values):
        if name in oracle.FIELDS3:
            values = np.asarray(values, np.float32)[self._slots()]
        super().write(name, values)

    def read(self, name):
        v = super().read(name)
        if name in oracle.FIELDS3:
            out = np.empty_like(v)
            out[self._slots()] = v
            return out
        return v

    def run(self, what):
        if what in ("lod", "lod_target"):
            self.sim.run_pass("sort")
            m = super().read("m")
            active = int(np.count_nonzero(m > 0))
            # the sort put every slot with mass first, and counted them
            assert int(self.sim.read_words("active")[0]) == active
            assert np.all(m[:active] > 0) and not np.any(m[active:] > 0)
        super().run(what)


RIGS = [GpuLodRig, SortingGpuLodRig]


@pytest.mark.parametrize("rig", RIGS, ids=["slots-in-place", "sorted"])
@pytest.mark.parametrize("case", cases.CASES, ids=[c.__name__ for c in cases.CASES])
def test_lod_case_on_the_device(gpu, case, rig):
    print(f"[lod] {case.__name__} {rig.__name__}", flush=True)
    case(rig)


@pytest.mark.parametrize("rig", RIGS, ids=["slots-in-place", "sorted"])
def test_lod_target_on_the_device(gpu, rig):
    cases.case_lod_target(lambda slots: rig(cases.N, slots))


@pytest.mark.parametrize("rig", RIGS, ids=["slots-in-place", "sorted"])
def test_every_call_is_the_oracles(gpu, rig):
    """Thinning to 1, then 5, then filling to 8: after every call, every field, stay and level
    bit for bit the oracle's (and a fresh slot set down at the group's mean the same); sorted,
    the slots have moved before every call."""
    print(f"[lod] bit-exact sequence {rig.__name__}", flush=True)
    dev, ref = rig(), cases.OracleLodRig()
    for rig in (dev, ref):
        cases.seed(rig)
    for want, calls in ((1, 40), (5, 25), (8, 40)):
        for rig in (dev, ref):
            cases.target(rig, want)
        for k in range(calls):
            dev.run("lod")
            ref.run("lod")
            for name in oracle.FIELDS3 + ("stay", "level"):
                np.testing.assert_array_equal(dev.read(name).view(np.int32), ref.read(name).view(np.int32),
                                              err_msg=f"target {want}, call {k}: {name}")


def test_refine_through_the_runner(gpu):
    """ParticleSimulation3.refine over a seeded block: far from the eye the groups thin (active
    falls), near it they stay full; mass is kept."""
    n = 32
    x, y, z = [], [], []
    for g in range(64):  # 64 groups on a 4 x 4 x 4 lattice of cells, one octant each
        cx, cy, cz = 4 + 2 * (g % 4), 4 + 2 * ((g // 4) % 4), 4 + 2 * (g // 16)
        for s in range(8):
            x.append(cx + 0.25 + 0.5 * (s & 1))
            y.append(cy + 0.25 + 0.5 * ((s >> 1) & 1))
            z.append(cz + 0.25 + 0.5 * ((s >> 2) & 1))
    count = len(x)
    sim = particles.ParticleSimulation3(n, n, n, count)
    zero = np.zeros(count, np.float32)
    m = np.full(count, 0.125, np.float32)
    sim.load((np.array(x, np.float32), np.array(y, np.float32), np.array(z, np.float32)), (zero, zero, zero), m,
             particles.params3(1e-3, 0, 0, 0, 10, 1))
    assert sim.active() == count
    for _ in range(60):
        sim.refine((0.0, 0.0, 0.0), 3.0, 1)
    assert sim.active() < count
    mass = float(np.sum(sim.masses() if hasattr(sim, "masses") else sim.sim.read("m"), dtype=np.float64))
    assert abs(mass - float(np.sum(m, dtype=np.float64))) <= 1e-4
    sim.close()
