"""The oracle's MLS-MPM passes against the reference's particle tests and MPM's closed forms."""
import numpy as np
import pytest

import mpm_cases as cases
import oracle


def make3(nx, ny, nz, count):
    return cases.OracleRig3(nx, ny, nz, count)


def make2(nx, ny, count):
    return cases.OracleRig2(nx, ny, count)


def test_lone_particle_falls_3d():
    cases.case_lone_particle_falls_3d(make3)


def test_scatter_conserves_3d():
    cases.case_scatter_conserves_3d(make3)


def test_affine_field_is_reproduced_3d():
    cases.case_affine_field_is_reproduced_3d(make3)


def test_compressed_particle_pushes_out_3d():
    cases.case_compressed_particle_pushes_out_3d(make3)


def test_walls_hold_3d():
    cases.case_walls_hold_3d(make3)


@pytest.mark.slow
def test_dam_break_slab_3d():
    cases.case_dam_break_slab_3d(make3)


def test_lone_particle_falls_2d():
    cases.case_lone_particle_falls_2d(make2)


def test_scatter_matches_fp64_splat_2d():
    cases.case_scatter_matches_fp64_splat_2d(make2)


def test_broken_runs_2d():
    cases.case_broken_runs_2d(make2)


def test_affine_field_is_reproduced_2d():
    cases.case_affine_field_is_reproduced_2d(make2)


def test_dam_break_2d():
    cases.case_dam_break_2d(make2)


def test_weights_closed_form():
    """The quadratic B-spline weights of the stencil: they sum to 1, their first moment is 0 and
    their second moment 1/4 at every offset; at f = 1 they are (1/8, 3/4, 1/8). Probed through
    a lone particle's deposit, which is w times its mass on each node."""
    for xpos in (3.0, 3.25, 3.5, 3.9, 4.4999):
        rig = make2(8, 8, 1)
        rig.write("x", [xpos])
        rig.write("y", [4.0])
        rig.write("m", [1.0])
        rig.write("j", [1.0])
        rig.params(oracle.mpm2_params(1e-3, 0, 0, 0, 1))
        rig.run(["clear", "scatter"])
        gm = rig.read("gm").reshape(8, 8).astype(np.float64)
        wx = gm.sum(axis=0)
        nodes = np.arange(8)
        assert abs(wx.sum() - 1) < 1e-6
        assert abs(np.sum(wx * (nodes - xpos))) < 2e-6
        assert abs(np.sum(wx * (nodes - xpos) ** 2) - 0.25) < 2e-6
        # y = 4 sits at offset f = 1 from its base node 3
        np.testing.assert_allclose(gm.sum(axis=1)[3:6], [0.125, 0.75, 0.125], atol=1e-7)


def test_threaded_step_matches_the_serial_one():
    """orc_mpm3_step_omp (the CPU baseline's threaded substep) against orc_mpm3_step: grids and
    particles within FP32 reassociation of the P2G sums after a few steps, at 1, 3 and 8 threads."""
    import numpy as np
    from vexhip import scenes
    n = 24
    x, y, z = scenes.jittered_block3((3, 4, 3), (10, 12, 9), 8, scenes.JavaRandomFloats(17))
    count = len(x)
    f = {"x": x, "y": y, "z": z, "m": np.full(count, 0.125, np.float32)}
    params = oracle.mpm3_params(5e-4, 0, -150.0, 0, 1500.0, 1)
    ref = oracle.Mpm3(n, n, n, f, params)
    ref.step(4)
    for threads in (1, 3, 8):
        s = oracle.Mpm3(n, n, n, f, params)
        s.step_omp(4, threads)
        for name in ("x", "y", "z", "u", "v", "w", "j"):
            np.testing.assert_allclose(s.p[name], ref.p[name], rtol=1e-5, atol=1e-5, err_msg=f"{threads}: {name}")
        np.testing.assert_allclose(s.g["gm"], ref.g["gm"], rtol=1e-5, atol=1e-6)
