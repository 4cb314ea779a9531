"""The oracle's ParticleDiagnostics.of held to test/particle/ParticleDiagnosticsTest.java (all five
cases), as the shallow-water diagnostics are held to DiagnosticsTest in test_oracle_swe.py."""
import numpy as np

import oracle
from vexhip import particles


def test_density_is_the_inverse_of_j():
    d = oracle.particle_diagnostics([1.0, 0.5, 2.0], 0, 1, 1)
    assert abs(d["min_density"] - 0.5) <= 1e-9 and abs(d["max_density"] - 2.0) <= 1e-9 and d["non_finite"] == 0


def test_courant_and_mach_come_from_the_sound_speed_the_step_was_sized_by():
    sound, fall = 300.0, 100.0
    dt = particles.stable_step(sound * sound, 1, fall, 0.45)
    d = oracle.particle_diagnostics([1.0], fall, sound, dt)
    assert abs(d["courant"] - 0.45) <= 1e-9 and abs(d["mach"] - fall / sound) <= 1e-9


def test_a_few_squeezed_particles_move_the_maximum_but_not_the_percentiles():
    j = np.ones(1000, np.float32)
    j[0], j[1] = 0.5, 0.6
    d = oracle.particle_diagnostics(j, 0, 1, 1)
    assert abs(d["max_density"] - 2.0) <= 1e-9
    assert abs(d["high_density"] - 1.0) <= 1e-9 and abs(d["low_density"] - 1.0) <= 1e-9


def test_a_broken_j_is_counted_and_left_out_of_the_range():
    d = oracle.particle_diagnostics([1.0, np.nan, 0.0, -1.0, np.inf], 0, 1, 1)
    assert d["non_finite"] == 4 and abs(d["min_density"] - 1.0) <= 1e-9 and abs(d["max_density"] - 1.0) <= 1e-9


def test_no_particles_read_as_zero():
    d = oracle.particle_diagnostics(np.zeros(0, np.float32), 0, 1, 1)
    assert d["max_density"] == 0 and d["min_density"] == 0


def test_percentiles_are_order_statistics():
    """Against numpy's sort: the 1st and 99th percentile picks of a seeded spread of J."""
    r = np.random.default_rng(3)
    j = r.uniform(0.8, 1.2, 5000).astype(np.float32)
    d = oracle.particle_diagnostics(j, 0, 1, 1)
    s = np.sort(j)
    assert d["low_density"] == float(np.float32(1) / s[min(4999, int(0.99 * 5000))])
    assert d["high_density"] == float(np.float32(1) / s[min(4999, int(0.01 * 5000))])


def test_densities_are_binary32_quotients():
    """ParticleDiagnostics divides 1 by a float (ParticleDiagnostics.java:50-51): the binary32
    quotient, widened to double -- for J = 3 that is 0.3333333432674408, not 1/3 in binary64."""
    d = oracle.particle_diagnostics(np.array([3.0, 3.0], np.float32), 0, 1, 1)
    assert d["min_density"] == float(np.float32(1) / np.float32(3))
    assert d["min_density"] != 1.0 / 3.0
