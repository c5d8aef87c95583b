"""The reference program driving the library: its own Fortran sources built
with four call sites of main.f90 redirected through include/fsil_bridge.f90
to libfsil.so (tests/fortran_host) run the fixture cases on the GPU from the
same inputs and restart file as the reference program's own run that wrote
each fixture.  The fluid step and the coupling run on the device; the markers,
the structure solver and every writer are the reference's.  Its continue file
holds the same final state bit for bit and its DatInfo files (forces,
first/last/centre node, node average, power, energy) the same rows."""
import os
import shutil

import numpy as np
import pytest

from tests.golden.cases import BY_NAME, LONG_BY_NAME, load_fixture
from tests.golden.make_golden import read_info

pytestmark = pytest.mark.gpu

HOST_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran_host", "_build", "FSILBM3D_fsil")
CASES = ["ibm_static_plate", "ibm_heaving_plate", "ibm_pitching_two_plates", "fem_clamped_heaving_plate",
         "fem_flexible_plate", "ibm_surface_sphere"]


def _host():
    if not os.path.exists(HOST_BIN):
        pytest.fail(f"{HOST_BIN} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                    "where the reference sources are present")
    return HOST_BIN


def _compare_info(workdir, ngroups, fx, label):
    info = read_info(workdir, ngroups)
    assert info, f"{label}: no DatInfo rows written"
    for k, rows in info.items():
        assert k in fx, f"{label}: DatInfo {k} has no fixture"
        assert rows.shape == fx[k].shape, (label, k, rows.shape, fx[k].shape)
        bad = np.argwhere(rows != fx[k])
        assert bad.size == 0, f"{label}: DatInfo {k} differs first at row/col {bad[0]}: {rows[tuple(bad[0])]} " \
                              f"vs {fx[k][tuple(bad[0])]}"
    return len(info)


@pytest.mark.parametrize("name", CASES)
def test_fortran_host_matches_reference_program(name):
    from oracle.refcase import run_reference
    exe = _host()
    pc = BY_NAME[name]
    print(f"[gpu] fortran host {name} dims={pc.case.dims} steps={pc.nsteps}", flush=True)
    case = pc.full_case()
    case.npsize = 1
    if pc.plates:
        case.info_every = 1
    fx = load_fixture(name)
    r = run_reference(case, pc.initial_state(), pc.nsteps, keep=True, binary=exe)
    try:
        assert r.step == int(fx["step"])
        n = int(np.count_nonzero(r.f != fx["f_final"]))
        assert n == 0, f"{name}: {n} populations differ from the reference program's"
        files = _compare_info(r.workdir, len(pc.plates), fx, name)
        print(f"[gpu] fortran host {name}: continue state and {files} DatInfo files equal the reference's", flush=True)
    finally:
        shutil.rmtree(r.workdir, ignore_errors=True)


def test_fortran_host_heaving_plate_2000_steps():
    """The 128x64x64 heaving plate for 2000 steps: every step's DatInfo row
    (Cd/Cl history and body state) and windows of the final state equal the
    reference program's."""
    from oracle.refcase import run_reference
    exe = _host()
    lc = LONG_BY_NAME["c1_heaving_plate"]
    pc = lc.pc
    print(f"[gpu] fortran host {lc.name} dims={pc.case.dims} steps={pc.nsteps}", flush=True)
    case = pc.full_case()
    case.npsize = lc.npsize  # as the fixture's run (make_golden.make_long)
    case.info_every = 1
    fx = load_fixture(lc.name)
    r = run_reference(case, lc.uniform_state(), pc.nsteps, keep=True, binary=exe, timeout=900.0)
    try:
        assert r.step == int(fx["step"])
        for k in range(len(lc.windows)):
            assert np.array_equal(lc.window(r.f, k), fx[f"window{k}"]), f"window {k} differs"
        files = _compare_info(r.workdir, len(pc.plates), fx, lc.name)
        print(f"[gpu] fortran host {lc.name}: {files} DatInfo files x {pc.nsteps} rows equal the reference's",
              flush=True)
    finally:
        shutil.rmtree(r.workdir, ignore_errors=True)
