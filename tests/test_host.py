"""Host-side logic: derived parameters and rigid-body kinematics."""
import math

import numpy as np

from fsilbm3d_amd import params as P
from fsilbm3d_amd.solid import RigidPlate, aoa_to_ttt


def test_relaxation_times():
    tau, om = P.srt(0.05 * 1.0 / 3.0, 1.0)
    assert abs(tau - 0.55) < 1e-15 and abs(om - 1 / 0.55) < 1e-15
    # TRT with magic parameter lambda: (1/om+ - 1/2)(1/om- - 1/2) = lambda
    lam = 0.1875
    om2 = P.trt_omega2(om, lam)
    assert abs((1 / om - 0.5) * (1 / om2 - 0.5) - lam) < 1e-14


def test_mrt_reduces_to_srt_when_all_rates_equal(monkeypatch):
    for name in ("S0", "S1", "S2", "S4", "S10", "S16"):
        monkeypatch.setattr(P, name, 1.3)
    Mc, Mf = P.mrt_matrices(1.3)
    assert np.allclose(Mc, 1.3 * np.eye(19), atol=1e-13)
    assert np.allclose(Mf, np.eye(19) - 0.65 * np.eye(19), atol=1e-13)


def test_mrt_conserves_mass_and_momentum_rows():
    Mc, _ = P.mrt_matrices(1 / 0.55)
    ee = P.EE
    # the relaxation leaves conserved moments untouched: 1^T Mc = 0, e^T Mc = 0
    assert np.allclose(np.ones(19) @ Mc, 0.0, atol=1e-13)
    for k in range(3):
        assert np.allclose(ee[:, k] @ Mc, 0.0, atol=1e-13)


def test_block_extent_periodic_adds_one_cell():
    ext, per = P.block_extent((10, 8, 6), 0.5, (1.0, 2.0, 3.0), (301, 301, 302, 302, 101, 103))
    assert ext == (1.0 + 0.5 * 10, 2.0 + 0.5 * 7, 3.0 + 0.5 * 5)
    assert per == (1, 0, 0)


def _plate(**kw):
    nodes = np.array([[-2.0 + i, 0, 0, 2.0, 2.0, 0, 0, 1.0] for i in range(5)])
    els = [(i + 1, i + 2, 4) for i in range(4)]
    return RigidPlate(nodes, els, **kw)


def test_static_plate_markers_closed_form():
    p = _plate(XYZo=(8.0, 6.0, 6.0))
    assert p.M == 16 and p.v_move == 0
    # element-centre x, span samples z = 6 + (s + 0.5) - 2
    xs = sorted(set(p.Exyz[:, 0]))
    assert xs == [6.5, 7.5, 8.5, 9.5]
    assert sorted(set(p.Exyz[:, 2])) == [4.5, 5.5, 6.5, 7.5]
    assert (p.Exyz[:, 1] == 6.0).all()
    assert (p.Evel == 0.0).all()
    assert (p.Ea == -2.0).all()          # dl * len * (-2 alpha rho)
    assert abs(p.Asfac - 4.0) < 1e-15


def test_heaving_plate_kinematics():
    A, f, phi = 1.5, 0.02, 30.0
    p = _plate(XYZo=(8.0, 6.0, 6.0), XYZAmpl=(0, A, 0), XYZPhi=(0, phi, 0), Freq=f)
    assert p.v_move == 1
    t = 3.0
    p.structure(t, 1, 1.0, 1.0)
    p.update_markers()
    w = 2 * math.pi * f
    y = 6.0 + A * math.cos(w * t + math.radians(phi))
    v = -w * A * math.sin(w * t + math.radians(phi))
    assert np.allclose(p.Exyz[:, 1], y, atol=1e-14)
    assert np.allclose(p.Evel[:, 1], v, atol=1e-15)
    assert np.allclose(p.Evel[:, [0, 2]], 0.0)


def test_pitch_rotation_matrix_is_orthonormal_and_rotates_marker_velocity():
    T = np.array(aoa_to_ttt([0.1, -0.2, 0.3]))
    assert np.allclose(T @ T.T, np.eye(3), atol=1e-14)
    # prescribed pitching: every marker moves with the rigid velocity W x r
    p = _plate(XYZo=(8.0, 6.0, 6.0), AoAAmpl=(0, 0, 10.0), Freq=0.05, AoAPhi=(0, 0, -90.0))
    p.structure(0.0, 1, 0.0, 1.0)
    p.update_markers()
    W = p.vel[0, 3:6]
    r = p.Exyz - np.array(p.XYZ)
    assert np.linalg.norm(W) > 0
    assert np.allclose(p.Evel, np.cross(W, r), atol=1e-12)
    # and every node shares the same angular velocity
    assert np.allclose(p.vel[:, 3:6], W, atol=0)


def test_compare_blocks_containment():
    """CompareBlocks (FluidDomain.f90:1845-1972): containment both ways,
    separated blocks, and the father-adjacent son with one BCfluid face."""
    from types import SimpleNamespace as NS
    from fsilbm3d_amd.tree import compare_blocks

    def blk(o, m, bc=(0,) * 6, dh=1.0):
        return NS(origin=o, xmax=m[0], ymax=m[1], zmax=m[2], bc=bc, dh=dh)
    big = blk((0, 0, 0), (15, 11, 11), bc=(101, 103, 302, 302, 302, 302))
    son = blk((4, 3, 3), (10, 8, 8), dh=0.5)
    far = blk((20, 20, 20), (25, 25, 25))
    assert compare_blocks(big, son) == 1
    assert compare_blocks(son, big) == -1
    assert compare_blocks(son, far) == 0
    # a son sharing the father's x-min plane is still contained (aligned side)
    edge = blk((0, 3, 3), (6, 8, 8), dh=0.5)
    assert compare_blocks(big, edge) == 1
    # a son whose single BCfluid face lies one father spacing inside a
    # BCfluid_father face is shrunk to that face and counts as contained
    fat = blk((0, 0, 0), (15, 11, 11), bc=(101, 1, 302, 302, 302, 302))
    over = blk((14, 3, 3), (20, 8, 8), bc=(0, 302, 302, 302, 302, 302), dh=0.5)
    assert compare_blocks(fat, over) == 1
    assert compare_blocks(big, over) == 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it exits with status 2 and says
    why, before touching the device."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    p = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=str(root))
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout.strip() == ""


def test_surface_body_markers_match_oracle():
    """The host mirror of a surface body (gmsh read, incenters, area weights,
    rigid placement: Solidbody.f90:648-720, 1073-1128) equals the oracle's
    markers at start-up and after sub-steps, bit for bit."""
    import numpy as np
    from tests.golden.cases import BY_NAME
    from tests.helpers import derived, make_bodies, oracle_bodies
    pc = BY_NAME["ibm_surface_sphere"]
    host = make_bodies(pc, pc.case.denIn)[0]
    orc = oracle_bodies(pc, derived(pc.case))[0]
    assert host.M == orc.M == 80
    for step in range(3):
        if step:
            t = float(step)
            for isub in (1, 2):
                host.structure(t, isub, 1.0, 0.5)
                orc.structure(t, isub, 1.0, 0.5)
            host.update_markers()
            orc.update_markers()
        assert np.array_equal(host.Exyz, orc.Exyz) and np.array_equal(host.Evel, orc.Evel)
        assert np.array_equal(host.Ea, orc.Ea)
        assert np.array_equal(host.element_centres(), orc.element_centres())
