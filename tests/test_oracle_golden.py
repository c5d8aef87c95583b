"""The CPU oracle against the reference program's own output.

Each fixture in tests/golden/ was written by the reference (compiled from its
sources by oracle/build_ref.sh) from a seeded state injected through its
restart file (tests/golden/make_golden.py).  The oracle must reproduce every
one bit for bit.
"""
import numpy as np
import pytest

from tests.golden.cases import CASES, SINGLE, TREE, load_fixture
from tests.helpers import mismatch, run_oracle, run_oracle_tree


@pytest.mark.parametrize("pc", SINGLE, ids=[c.name for c in SINGLE])
def test_oracle_matches_reference_fixture(pc):
    fx = load_fixture(pc.name)
    f = run_oracle(pc)
    n, m = mismatch(f, fx["f_final"])
    assert n == 0, f"{pc.name}: {n} values differ, max |diff| {m:.3e}"


def test_fixtures_depend_on_the_bodies():
    """A body-free run must NOT reproduce the immersed-boundary fixtures, so
    those fixtures pin the coupling, not only the fluid."""
    import dataclasses
    for pc in CASES:
        if not pc.plates:
            continue
        f = run_oracle(dataclasses.replace(pc, plates=[]))
        n, _ = mismatch(f, load_fixture(pc.name)["f_final"])
        assert n > 0, pc.name


def test_planted_phase_error_is_caught():
    """A 1-degree heave phase error in the host kinematics changes the state."""
    import dataclasses
    from tests.golden.cases import BY_NAME
    pc = BY_NAME["ibm_heaving_plate"]
    p = dataclasses.replace(pc.plates[0], XYZPhi=(0.0, 31.0, 0.0))
    f = run_oracle(dataclasses.replace(pc, plates=[p]))
    n, _ = mismatch(f, load_fixture(pc.name)["f_final"])
    assert n > 0


@pytest.mark.parametrize("pc", TREE, ids=[c.name for c in TREE])
def test_oracle_tree_matches_reference_fixture(pc):
    """Father and son block after the reference's recursive tree step
    (LBMBlockComm.f90:279-318), both bit for bit."""
    fx = load_fixture(pc.name)
    fs = run_oracle_tree(pc)
    for k, key in enumerate(["f_final", "f_final_son1"]):
        n, m = mismatch(fs[k], fx[key])
        assert n == 0, f"{pc.name} {key}: {n} values differ, max |diff| {m:.3e}"


def test_tree_fixtures_pin_the_coupling():
    """Planted mistakes in the halo exchange (no time average of the father
    layer, coefficient 1 in the grid transform, the other interpolation
    scheme) each change the two-block fixtures."""
    import oracle.oracle_tree as T
    fx = {pc.name: load_fixture(pc.name) for pc in TREE}

    def differs():
        out = []
        for pc in TREE:
            fs = run_oracle_tree(pc)
            out.append(mismatch(fs[0], fx[pc.name]["f_final"])[0] + mismatch(fs[1], fx[pc.name]["f_final_son1"])[0])
        return out

    gt, ex, ip = T.grid_transform, T.OracleTree.extract, T.interp_plane

    def no_average(self, k, ff, time):
        p = self.pairs[k]
        for face in range(6):
            if p["sds"][face]:
                self.planes[k][(face, time)] = np.array(self._face_plane(ff, face, p))
    try:
        T.OracleTree.extract = no_average
        assert all(n > 0 for n in differs())
        T.OracleTree.extract = ex
        T.grid_transform = lambda fq, c, VF, dh: gt(fq, 1.0, VF, dh)
        assert all(n > 0 for n in differs())
        T.grid_transform = gt
        T.interp_plane = lambda fF, bS, aS, s: ip(fF, bS, aS, 1 if s == 2 else 2)
        assert all(n > 0 for n in differs())
    finally:
        T.grid_transform, T.OracleTree.extract, T.interp_plane = gt, ex, ip
