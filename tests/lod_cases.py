"""test/particle/Flip3LodTest.java's cases, written once over a rig: ``write(name, values)``,
``read(name)`` for the particle fields and "level", "target", "stay", "view", and
``run(pass)`` for "lod", "lod_target" and "step". The oracle supplies one rig (OracleLodRig) and
the HIP runner another; each is held to the test's own assertions."""
import numpy as np

import oracle

N = 14
CELLS = ((3, 3, 3), (6, 3, 3), (3, 6, 3), (6, 6, 8))
GROUPS = len(CELLS)
GROUP = 8
SLOTS = GROUPS * GROUP
M = np.float32(0.125)
AFFINE = oracle.FIELDS3[8:]


class OracleLodRig:
    def __init__(self, n=N, slots=SLOTS):
        self.s = oracle.Mpm3(n, n, n, {"x": np.zeros(slots, np.float32)}, np.zeros(7, np.int32))
        groups = slots // GROUP
        self.extra = {"level": np.zeros(groups, np.float32), "target": np.zeros(groups, np.float32),
                      "stay": np.zeros(slots, np.float32), "view": np.zeros(5, np.float32)}

    def write(self, name, values):
        v = np.asarray(values)
        if name == "params":
            self.s.params = v.astype(np.int32).copy()
        elif name in self.extra:
            self.extra[name][:] = v.astype(np.float32)
        elif name in self.s.p:
            self.s.p[name][:] = v.astype(np.float32)
        else:
            self.s.g[name][:] = v.astype(np.float32)

    def read(self, name):
        if name in self.extra:
            return self.extra[name].copy()
        return (self.s.p if name in self.s.p else self.s.g)[name].copy()

    def run(self, what):
        if what == "lod":
            self.s.lod(self.extra["level"], self.extra["target"], self.extra["stay"])
        elif what == "lod_target":
            self.s.lod_target(self.extra["level"], self.extra["target"], self.extra["view"])
        else:
            self.s.step()


def seed(rig):
    """Flip3LodTest.seed (:33-73): eight particles a group, one in each octant of its cell."""
    r = oracle.JavaRandom(7)
    at = np.zeros((3, SLOTS), np.float32)
    vel = np.zeros((3, SLOTS), np.float32)
    aff = np.zeros((9, SLOTS), np.float32)
    mass = np.full(SLOTS, M, np.float32)
    vol = np.zeros(SLOTS, np.float32)
    for g in range(GROUPS):
        for s in range(GROUP):
            k = g * GROUP + s
            for a in range(3):
                inner = np.float32(np.float32((s >> a) & 1) + np.float32(0.25)) + np.float32(0.5) * np.float32(r.next_float())
                at[a][k] = np.float32(CELLS[g][a]) + np.float32(inner / np.float32(2))
                vel[a][k] = np.float32(4) * np.float32(r.next_float()) - np.float32(2)
            for c in range(9):
                aff[c][k] = np.float32(r.next_float()) - np.float32(0.5)
            vol[k] = np.float32(0.9) + np.float32(0.2) * np.float32(r.next_float())
    for a, name in enumerate("xyz"):
        rig.write(name, at[a])
    for a, name in enumerate("uvw"):
        rig.write(name, vel[a])
    for c, name in enumerate(AFFINE):
        rig.write(name, aff[c])
    rig.write("m", mass)
    rig.write("j", vol)
    rig.write("stay", np.ones(SLOTS, np.float32))
    rig.write("level", np.full(GROUPS, GROUP, np.float32))
    rig.write("target", np.full(GROUPS, GROUP, np.float32))
    rig.write("params", oracle.mpm3_params(0.001, 0, -10, 0, 100, 1))


def target(rig, n):
    rig.write("target", np.full(GROUPS, n, np.float32))


def totals(rig):
    m = rig.read("m").astype(np.float64)
    at = [rig.read(a).astype(np.float64) for a in "xyz"]
    v = [rig.read(a).astype(np.float64) for a in "uvw"]
    j = rig.read("j").astype(np.float64)
    live = m > 0
    out = {"active": int(live.sum()), "mass": float(m[live].sum()), "volume": float((m * j)[live].sum()),
           "momentum": [float((m * v[a])[live].sum()) for a in range(3)]}
    centre = np.zeros((GROUPS, 3))
    for g in range(GROUPS):
        sl = slice(g * GROUP, (g + 1) * GROUP)
        w = m[sl] * live[sl]
        for a in range(3):
            centre[g, a] = float((w * at[a][sl]).sum() / w.sum())
    out["centre"] = centre
    return out


def same(before, after, when):
    assert abs(before["mass"] - after["mass"]) <= 1e-5, when
    assert abs(before["volume"] - after["volume"]) <= 1e-4, when
    for a in range(3):
        assert abs(before["momentum"][a] - after["momentum"][a]) <= 1e-4, when


def unequal(rig, n):
    m = rig.read("m").astype(np.float64)
    most = 0.0
    for g in range(GROUPS):
        sl = m[g * GROUP:(g + 1) * GROUP]
        s = sl.sum()
        for x in sl:
            most = max(most, abs(x - s / n) if x > 0 else 0.0)
    return most


def fade(rig, n, start):
    """Calls lod until every group has n active slots, the sums checked after each (:142-162)."""
    frames, last = 0, 1 << 30
    while totals(rig)["active"] != GROUPS * n or unequal(rig, n) > 1e-6 or frames == 0:
        rig.run("lod")
        frames += 1
        now = totals(rig)
        same(start, now, f"frame {frames}")
        if n < GROUP:
            assert now["active"] <= last
            last = now["active"]
        assert frames < 400, f"still {now['active']} active after {frames} calls"
    return frames


def case_thinning(make):
    """thinningTakesManyCallsAndKeepsMassMomentumAndVolumeThroughout (:165-189)."""
    rig = make()
    seed(rig)
    start = totals(rig)
    target(rig, 1)
    rig.run("lod")
    assert totals(rig)["active"] == SLOTS
    frames = fade(rig, 1, start)
    assert frames > 10
    end = totals(rig)
    assert end["active"] == GROUPS
    assert np.all(np.abs(start["centre"] - end["centre"]) <= 0.4)
    m = rig.read("m").astype(np.float64)
    for g in range(GROUPS):
        assert abs(m[g * GROUP:(g + 1) * GROUP].sum() - GROUP * float(M)) <= 1e-6
    return rig


def case_filling_back(make):
    """fillingBackUpGivesEveryoneAnEqualShareAndKeepsTheSums (:191-221)."""
    rig = make()
    seed(rig)
    start = totals(rig)
    target(rig, 1)
    fade(rig, 1, start)
    one = totals(rig)
    target(rig, GROUP)
    fade(rig, GROUP, start)
    end = totals(rig)
    assert end["active"] == SLOTS
    assert np.all(np.abs(rig.read("m") - M) <= 1e-5)
    assert np.all(np.abs(one["centre"] - end["centre"]) <= 0.2)
    at = [rig.read(a) for a in "xyz"]
    for g in range(GROUPS):
        for s in range(GROUP):
            k = g * GROUP + s
            assert all(abs(at[a][k] - end["centre"][g][a]) < 0.8 for a in range(3))
    return rig


def case_turn_round(make):
    """changingTheTargetPartWayTurnsTheGroupRoundWithoutAJump (:223-242)."""
    rig = make()
    seed(rig)
    start = totals(rig)
    target(rig, 1)
    for k in range(6):
        rig.run("lod")
        same(start, totals(rig), f"thinning, call {k}")
    target(rig, GROUP)
    fade(rig, GROUP, start)
    assert np.all(np.abs(rig.read("m") - M) <= 1e-5)
    return rig


def case_nearest_leaves_first(make):
    """theSlotNearestTheMeanIsTheFirstToLeave (:244-275)."""
    rig = make()
    seed(rig)
    x, y, z = rig.read("x"), rig.read("y"), rig.read("z")
    for s in range(GROUP):
        x[s] = 7 + (s & 1) * 2 - 1
        y[s] = 7 + ((s >> 1) & 1) * 2 - 1
        z[s] = 7 + ((s >> 2) & 1) * 2 - 1
    x[5] = y[5] = z[5] = 7
    rig.write("x", x)
    rig.write("y", y)
    rig.write("z", z)
    target(rig, 7)
    rig.run("lod")
    stay = rig.read("stay")
    for s in range(GROUP):
        assert stay[s] == (0 if s == 5 else 1)
    target(rig, 6)
    rig.run("lod")
    assert int((rig.read("stay")[:GROUP] > 0.5).sum()) == 6
    return rig


def case_empty_group(make):
    """anEmptyGroupIsLeftAlone (:286-308)."""
    rig = make()
    seed(rig)
    m = rig.read("m")
    m[3 * GROUP:] = 0
    rig.write("m", m)
    target(rig, 1)
    for _ in range(5):
        rig.run("lod")
    m, x = rig.read("m"), rig.read("x")
    assert np.all(m[3 * GROUP:] == 0) and np.all(np.isfinite(x[3 * GROUP:]))
    return rig


def case_arrived_stays_put(make):
    """aGroupThatHasArrivedStaysPut (:310-327)."""
    rig = make()
    seed(rig)
    start = totals(rig)
    target(rig, 2)
    fade(rig, 2, start)
    before_m, before_u = rig.read("m"), rig.read("u")
    for _ in range(5):
        rig.run("lod")
    np.testing.assert_array_equal(rig.read("m"), before_m)
    np.testing.assert_array_equal(rig.read("u"), before_u)
    return rig


def case_inactive_slots_add_nothing(make):
    """inactiveSlotsAddNothingToTheGrid (:329-353)."""
    rig = make()
    seed(rig)
    start = totals(rig)
    target(rig, 1)
    fade(rig, 1, start)
    merged = totals(rig)
    before = rig.read("y")
    rig.run("step")
    grid = float(rig.read("gm").astype(np.float64).sum())
    assert abs(merged["mass"] - grid) <= 1e-5
    m, y = rig.read("m"), rig.read("y")
    assert np.all(y[m == 0] == before[m == 0])
    return rig


def case_lod_target(make_sized):
    """distanceFromTheEyePicksTheLevelAndAGroupOnABoundaryKeepsIt (:599-626)."""
    near = 2.0
    distance = [0.5, 1.5, 3, 10, 1.1, 0.9, 0.7, 2.4]
    start = np.array([8, 8, 8, 8, 8, 4, 4, 2], np.float32)
    expect = [8, 4, 2, 1, 8, 4, 8, 2]
    groups = len(distance)
    rig = make_sized(groups * GROUP)
    x = np.zeros(groups * GROUP, np.float32)
    for g in range(groups):
        x[g * GROUP] = np.float32(np.float32(distance[g]) * np.float32(near))
    rig.write("x", x)
    rig.write("level", start)
    rig.write("view", np.array([0, 0, 0, near, 1], np.float32))
    rig.run("lod_target")
    t = rig.read("target")
    assert [float(v) for v in t] == [float(e) for e in expect]
    rig.write("view", np.array([0, 0, 0, near, 2], np.float32))
    rig.run("lod_target")
    assert rig.read("target")[3] == 2
    return rig


CASES = [case_thinning, case_filling_back, case_turn_round, case_nearest_leaves_first, case_empty_group,
         case_arrived_stays_put, case_inactive_slots_add_nothing]
