"""Device parity: the HIP path through the C-ABI against the reference
program's fixtures and the CPU oracle, bit for bit (FP64 plain form)."""
import dataclasses
import sys

import numpy as np
import pytest

from tests.golden.cases import BY_NAME, CASES, ParityCase, Plate, load_fixture, seeded_state
from tests.helpers import mismatch, run_device, run_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


RIGID = [c for c in CASES if not c.case.sons and not any(p.iBodyModel == 2 for p in c.plates)]


@pytest.mark.parametrize("pc", RIGID, ids=[c.name for c in RIGID])
def test_device_matches_reference_fixture(pc):
    """Every fluid / rigid-body fixture bit for bit (elastic bodies:
    test_gpu_fem.py).  The WALE model (14) raises to the powers 1.5, 2.5 and
    1.25 (FluidDomain.f90:1419) with the device's pow, which may differ from
    the host C library's in the last place: within 1e-12 relative there."""
    _say(f"[gpu] fixture case {pc.name} dims={pc.case.dims} steps={pc.nsteps}")
    fx = load_fixture(pc.name)
    f = run_device(pc)
    n, m = mismatch(f, fx["f_final"])
    if pc.case.collide == 14:
        _say(f"[gpu] {pc.name}: {n} values differ, max {m:.3e}")
        assert m <= 1e-12 * np.abs(fx["f_final"]).max()
    else:
        assert n == 0, f"{pc.name}: {n} values differ from the reference, max |diff| {m:.3e}"


@pytest.mark.parametrize("name", ["ibm_static_plate", "ibm_heaving_plate", "ibm_pitching_two_plates",
                                  "ibm_surface_sphere"])
def test_device_ibm_forces_match_oracle(name):
    """Marker forces v_Eforce, beam loads lodFlow and the iteration count of
    every step equal the oracle's exactly."""
    pc = BY_NAME[name]
    _say(f"[gpu] ibm forces {name}")
    ro, rd = [], []
    fo = run_oracle(pc, record=ro)
    fd = run_device(pc, record=rd)
    assert mismatch(fo, fd)[0] == 0
    for so, sd in zip(ro, rd):
        assert so["iters"] == sd["iters"]
        Eo = np.concatenate(so["Eforce"])
        assert np.array_equal(Eo, sd["Eforce"])
        for lo, ld in zip(so["lod"], sd["lod"]):
            assert np.array_equal(lo, ld)


def _c1_like(nsteps, dims=(128, 64, 64)):
    """The heaving-plate configuration at its full size: 128x64x64, inflow
    0.05, Re 100 on a 16-cell chord, 101/103/302 faces, 17 chord nodes x 16
    span samples = 256 markers, heave 8 cells at 0.25 U/L, 3 MDF iterations."""
    from oracle.refcase import Case
    case = Case(dims=dims, bc=(101, 103, 302, 302, 302, 302), Re=100.0, uvwIn=(0.05, 0.0, 0.0),
                UrefType=0, Uref=0.05, LrefType=1, Lref=16.0, TrefType=0, ntolLBM=3, dtolLBM=1e-14)
    plate = Plate(17, -8.0, 8.0, 8.0, 16, XYZo=(40.0, 32.0, 32.0), XYZAmpl=(0.0, 8.0, 0.0), Freq=7.8125e-4)
    return ParityCase("c1_heaving_plate", case, nsteps=nsteps, seed=41, ubase=(0.05, 0, 0), neq=0.0,
                      drho=0.0, du=0.0, plates=[plate])


def test_full_size_heaving_plate_matches_oracle():
    pc = _c1_like(4)
    _say("[gpu] 128x64x64 heaving plate, 4 steps vs oracle")
    ro, rd = [], []
    fo = run_oracle(pc, record=ro)
    fd = run_device(pc, record=rd)
    n, m = mismatch(fo, fd)
    assert n == 0, f"{n} differ, max {m:.3e}"
    for so, sd in zip(ro, rd):
        assert so["iters"] == sd["iters"] == 3
        assert np.array_equal(np.concatenate(so["Eforce"]), sd["Eforce"])


@pytest.mark.parametrize("model", [1, 2, 3])
def test_random_periodic_models_match_oracle(model):
    from oracle.refcase import Case
    case = Case(dims=(40, 24, 20), bc=(301,) * 6, collide=model, params=(0.15,) + (0.0,) * 9, Re=3.0,
                Uref=0.05, UrefType=9, Lref=1.0, LrefType=1, TrefType=0, volumeForceIn=(1e-5, 0.0, -2e-6))
    pc = ParityCase(f"rand_periodic_{model}", case, nsteps=5, seed=100 + model, neq=0.03)
    _say(f"[gpu] random periodic model {model}")
    n, m = mismatch(run_oracle(pc), run_device(pc))
    assert n == 0, f"{n} differ, max {m:.3e}"


@pytest.mark.parametrize("model", [11, 14, 15])
def test_les_models_match_oracle(model):
    """The LES models on a larger block with open faces and a volume force,
    state and per-cell tau_all against the oracle (WALE within 1e-12, device
    pow)."""
    from oracle.refcase import Case
    from tests.helpers import run_device_block, run_oracle_block
    case = Case(dims=(40, 24, 36), bc=(101, 104, 201, 302, 301, 301), collide=model, Re=8.0, uvwIn=(0.05, 0.0, 0.0),
                shearRateIn=(2e-3, -1e-3, 5e-4), UrefType=9, Uref=0.05, LrefType=1, Lref=1.0, TrefType=0,
                volumeForceIn=(1e-5, 0.0, -2e-6))
    pc = ParityCase(f"les_{model}", case, nsteps=4, seed=300 + model, ubase=(0.05, 0, 0), neq=0.05, du=0.04)
    _say(f"[gpu] LES model {model} 40x24x36 vs oracle")
    fo, to = run_oracle_block(pc)
    fd, td = run_device_block(pc)
    n, m = mismatch(fo, fd)
    nt, mt = mismatch(to, td)
    _say(f"[gpu] LES {model}: state {n} differ (max {m:.3e}), tau_all {nt} differ (max {mt:.3e})")
    if model == 14:
        assert m <= 1e-12 * np.abs(fo).max() and mt <= 1e-12 * np.abs(to).max()
    else:
        assert n == 0 and nt == 0
    assert np.ptp(to) > 0.0


def test_odd_sizes_and_all_open_faces():
    """Ragged sizes (not multiples of the 64-lane wave) and every face a
    non-periodic code."""
    from oracle.refcase import Case
    case = Case(dims=(13, 11, 70), bc=(102, 104, 203, 202, 201, 302), Re=5.0, uvwIn=(0.04, 0.0, 0.0),
                shearRateIn=(1e-3, -2e-3, 5e-4), UrefType=9, Uref=0.05, LrefType=1, Lref=1.0, TrefType=0)
    pc = ParityCase("ragged_open", case, nsteps=4, seed=7, ubase=(0.04, 0, 0), neq=0.01)
    _say("[gpu] ragged sizes 13x11x70 all open faces")
    n, m = mismatch(run_oracle(pc), run_device(pc))
    assert n == 0, f"{n} differ, max {m:.3e}"


def test_macro_quantities_match_oracle():
    from fsilbm3d_amd.block import LBMBlock
    from oracle import oracle as O
    dims = (20, 16, 12)
    f = seeded_state(dims, 9, neq=0.02)
    blk = LBMBlock(dims, bc=(301,) * 6, nu=0.05 / 3)
    blk.set_flow(volumeForceIn=(1e-4, 2e-5, -3e-5))
    blk.upload(f)
    blk.update_volume_force(0.0)
    den, uuu = blk.calculate_macro_quantities()
    od, ou = O.macro(f, (1e-4, 2e-5, -3e-5), 1.0)
    assert np.array_equal(den, od) and np.array_equal(uuu, ou)
    assert np.array_equal(blk.download(), f)
    blk.close()


def test_stencil_out_of_bounds_is_an_error():
    """A marker whose stencil leaves a non-periodic, non-mirrored face is an
    error of the coupling call (Solidbody.f90:850-862); the stencil kernel
    clamps such indices, so no kernel leaves the block meanwhile."""
    from fsilbm3d_amd._lib import FsilError
    from oracle.refcase import Case
    case = Case(dims=(12, 12, 12), bc=(101, 103, 302, 302, 302, 302), Re=10.0, uvwIn=(0.05, 0, 0), UrefType=9,
                Uref=0.05, LrefType=1, Lref=1.0, TrefType=0)
    pc = ParityCase("oob", case, nsteps=1, seed=3, ubase=(0.05, 0, 0),
                    plates=[Plate(3, -1.0, 1.0, 1.0, 2, XYZo=(0.6, 6.0, 6.0))])
    with pytest.raises(FsilError, match="out of min bound"):
        run_device(pc)
    # a coupling call that does not wait reports it at the next call
    pc2 = dataclasses.replace(pc, nsteps=2)
    with pytest.raises(FsilError, match="out of min bound"):
        run_device(pc2, device_fem=True, wait=False)


@pytest.mark.parametrize("name", ["ibm_heaving_plate", "ibm_pitching_two_plates"])
def test_band_size_and_box_match_oracle_stencils(name):
    """fsil_ibm_band_info after a coupling call: the band is the set of cells
    the markers' trimmed 4x4x4 stencils touch (UpdateElmtInterp_,
    Solidbody.f90:760-821), counted once each, and its bounding box."""
    from fsilbm3d_amd import _lib
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import FSIBlock
    from oracle import oracle as O
    from tests.helpers import derived, flow_kwargs, make_bodies
    pc = BY_NAME[name]
    case = pc.case
    d = derived(case)
    _say(f"[gpu] band info {name}")
    blk = LBMBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["nu"], case.params)
    blk.set_flow(**flow_kwargs(case))
    blk.upload(np.array(pc.initial_state(), dtype=np.float64))
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    bodies = make_bodies(pc, case.denIn)
    fsi = FSIBlock(blk, bodies, ntolLBM=case.ntolLBM, dtolLBM=case.dtolLBM, Uref=d["Uref"], denIn=case.denIn)
    fsi.interaction_force(case.dh)
    nb = np.zeros(1, np.int32)
    box = np.zeros(6, np.int32)
    _lib.check(blk.lib.fsil_ibm_band_info(blk.handle, _lib.iptr(nb), _lib.iptr(box)))
    cells = set()
    geom4 = (case.dh,) + tuple(case.origin)
    for b in bodies:
        Ei, _ = O.stencil(b.Exyz, case.dims, geom4, case.bc)
        Ei = Ei.reshape(-1, 3, 4) - 1
        for row in Ei:
            for x in row[0]:
                for y in row[1]:
                    for z in row[2]:
                        cells.add((int(x), int(y), int(z)))
    blk.close()
    c = np.array(sorted(cells))
    lo, hi = c.min(axis=0), c.max(axis=0)
    assert int(nb[0]) == len(cells)
    assert list(box) == [*lo.tolist(), *(hi - lo + 1).tolist()]
