"""The CPU oracle against the reference program's own output.

Each fixture in tests/golden/ was written by the reference (compiled from its
sources by oracle/build_ref.sh) from a seeded state injected through its
restart file (tests/golden/make_golden.py).  The oracle must reproduce every
one bit for bit.
"""
import numpy as np
import pytest

from tests.golden.cases import CASES, load_fixture
from tests.helpers import mismatch, run_oracle


@pytest.mark.parametrize("pc", CASES, ids=[c.name for c in CASES])
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
