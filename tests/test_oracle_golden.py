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


BODY = [c for c in CASES if c.plates]


@pytest.mark.parametrize("pc", BODY, ids=[c.name for c in BODY])
def test_oracle_solid_info_matches_reference_datinfo(pc):
    """Every DatInfo row the reference wrote for every body at every step
    (forces = Cd/Cl/Cz, first/last/centre node, node average, power, energy;
    Beam_write_solid_info, SolidSolver.f90:1729-1799) equals the oracle's
    value rounded to the 10 digits the reference prints."""
    from tests.golden.cases import INFO_FILES
    from tests.helpers import print_equal
    fx = load_fixture(pc.name)
    rec = []
    run_oracle(pc, record=rec)
    for g in range(len(pc.plates)):
        for name in INFO_FILES:
            ref = fx[f"{name}_g{g + 1}"]
            got = np.array([r["solid_info"][g][name] for r in rec])
            assert got.shape == ref.shape, (name, got.shape, ref.shape)
            ok = print_equal(got, ref)
            assert ok.all(), f"{pc.name} group {g + 1} {name}: {int((~ok).sum())} entries differ"


@pytest.mark.parametrize("name,nsteps", [("c1_heaving_plate", 20), ("c3_flexible_plate", 2)])
def test_oracle_matches_full_size_reference_runs(name, nsteps):
    """The full-size workloads (128x64x64 rigid heaving plate, 2000 reference
    steps kept; 384x192x192 flexible plate, 20 kept): the oracle's first steps
    reproduce every DatInfo row of the reference at its print precision."""
    import dataclasses
    from tests.golden.cases import INFO_FILES, LONG_BY_NAME
    from tests.helpers import print_equal
    lc = LONG_BY_NAME[name]
    fx = load_fixture(name)
    rec = []
    run_oracle(dataclasses.replace(lc.pc, nsteps=nsteps), f0=lc.uniform_state(), record=rec)
    for fname in INFO_FILES:
        ref = fx[f"{fname}_g1"][:nsteps]
        got = np.array([r["solid_info"][0][fname] for r in rec])
        ok = print_equal(got, ref)
        assert ok.all(), f"{name} {fname}: {int((~ok).sum())} entries differ"


def test_long_fixtures_cover_their_steps():
    """The full-size fixtures hold one row per step and the final windows."""
    from tests.golden.cases import INFO_FILES, LONG_CASES
    for lc in LONG_CASES:
        fx = load_fixture(lc.name)
        assert int(fx["step"]) == lc.pc.nsteps
        for fname in INFO_FILES:
            assert fx[f"{fname}_g1"].shape[0] == lc.pc.nsteps
        for k, (x0, x1, y0, y1, z0, z1) in enumerate(lc.windows):
            assert fx[f"window{k}"].shape == (19, x1 - x0, y1 - y0, z1 - z0)


def test_planted_datinfo_mistakes_are_caught():
    """The reference's DatInfo rows pin the reductions: force coefficients
    over the span-wise area instead of the chord sum (Fref, Solidbody.f90:278),
    or the kinetic energy without its factor 1/2, each break them."""
    from tests.golden.cases import BY_NAME
    from tests.helpers import print_equal
    pc = BY_NAME["fem_flexible_plate"]
    fx = load_fixture(pc.name)
    rec = []
    run_oracle(pc, record=rec)
    cd = np.array([r["solid_info"][0]["forces"] for r in rec])
    assert print_equal(cd, fx["forces_g1"]).all()
    wrong = cd.copy()
    wrong[:, 3:] = wrong[:, 3:] / (pc.plates[0].lspan + pc.plates[0].rspan)
    assert not print_equal(wrong[:, 3:], fx["forces_g1"][:, 3:]).any()
    en = np.array([r["solid_info"][0]["energy"] for r in rec])
    assert print_equal(en, fx["energy_g1"]).all()
    wrong = en.copy()
    wrong[:, 4] = 2.0 * wrong[:, 4]
    assert not print_equal(wrong[:, 4], fx["energy_g1"][:, 4]).any()


def test_oracle_output_files_match_reference_bytes():
    """The reference's own DatFlow file (real(4) p, u, v, w over the output
    window, write_flow_, FluidDomain.f90:1628-1720) and continue file
    (write_continue_blocks, FluidDomain.f90:268-285, 1770-1776) after the last
    step, byte for byte from the oracle's state."""
    import os
    import tempfile
    from oracle.oracle import datflow_bytes
    from oracle.refcase import write_continue
    from tests.golden.cases import BY_NAME
    from tests.helpers import derived
    pc = BY_NAME["output_flow_continue"]
    fx = load_fixture(pc.name)
    f = run_oracle(pc)
    case = pc.case
    d = derived(case)
    got = datflow_bytes(f, case.volumeForceIn, case.dh, case.denIn, d["Uref"], case.offsetOutput, 1, case.origin)
    assert got == fx["datflow"].tobytes()
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "c")
        write_continue(p, f, case, pc.nsteps, pc.nsteps * case.dh / d["Tref"])
        assert open(p, "rb").read() == fx["continue_bytes"].tobytes()
    assert str(fx["datflow_name"]) == "Flow0000015000_b001" and str(fx["continue_name"]) == "continue0000015000"
