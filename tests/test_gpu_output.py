"""Output and restart on the device: the DatFlow and continue files the
library writes equal the reference program's own files byte for byte, and a
coupled run restarted from a continue file plus a beam checkpoint continues
bit for bit."""
import os
import tempfile

import numpy as np
import pytest

from tests.golden.cases import BY_NAME, load_fixture
from tests.helpers import derived, flow_kwargs, plate_specs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _fluid_block(pc, f0=None):
    from fsilbm3d_amd.block import LBMBlock
    case = pc.case
    d = derived(case)
    blk = LBMBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["nu"], case.params)
    blk.set_flow(**flow_kwargs(case))
    blk.upload(np.array(pc.initial_state() if f0 is None else f0, dtype=np.float64))
    blk.update_volume_force(0.0)
    return blk, d


def test_flow_and_continue_files_match_reference_bytes():
    from fsilbm3d_amd import _lib
    pc = BY_NAME["output_flow_continue"]
    _say(f"[gpu] output files {pc.name}")
    fx = load_fixture(pc.name)
    case = pc.case
    blk, d = _fluid_block(pc)
    blk.set_boundary_conditions()
    t = 0.0
    for _ in range(pc.nsteps):
        t = t + case.dh
        blk.update_volume_force(t)
        blk.collision_streaming_bc()
    lib = blk.lib
    ref = fx["datflow"].tobytes()
    nw = (len(ref) - 48) // 16
    out = np.zeros(4 * nw, np.float32)
    _lib.check(lib.fsil_flow_pack(blk.handle, case.offsetOutput, case.denIn, d["Uref"], _lib.fptr(out)))
    assert out.tobytes() == ref[48:]
    with tempfile.TemporaryDirectory() as td:
        pf = os.path.join(td, str(fx["datflow_name"]))
        _lib.check(lib.fsil_write_flow(blk.handle, case.offsetOutput, case.denIn, d["Uref"], pf.encode()))
        _lib.check(lib.fsil_write_wait())
        assert open(pf, "rb").read() == ref
        pcn = os.path.join(td, str(fx["continue_name"]))
        h = np.array([blk.handle], np.int32)
        _lib.check(lib.fsil_write_continue(1, _lib.iptr(h), pc.nsteps, pc.nsteps * case.dh / d["Tref"], pcn.encode()))
        assert open(pcn, "rb").read() == fx["continue_bytes"].tobytes()
    blk.close()


def test_flow_pack_fp32_matches_its_macro_quantities():
    """The FP32-deviation form packs the moments its own macro read-back forms."""
    from fsilbm3d_amd import _lib
    from fsilbm3d_amd.block import LBMBlock
    from tests.golden.cases import seeded_state
    dims = (24, 16, 20)
    blk = LBMBlock(dims, bc=(101, 103, 302, 302, 301, 301), nu=0.01, precision=1)
    blk.set_flow(uvwIn=(0.05, 0.0, 0.0), volumeForceIn=(1e-5, 0.0, 0.0))
    blk.upload(seeded_state(dims, 77, ubase=(0.05, 0, 0), du=0.02))
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    blk.collision_streaming_bc()
    den, u = blk.calculate_macro_quantities()
    off = 2
    nw = (dims[0] - 2 * off) * (dims[1] - 2 * off) * (dims[2] - 2 * off)
    out = np.zeros(4 * nw, np.float32)
    _lib.check(blk.lib.fsil_flow_pack(blk.handle, off, 1.0, 0.05, _lib.fptr(out)))
    sl = (slice(off, -off),) * 3
    want = [((1.0 / 3.0) * (den[sl] - 1.0)).astype(np.float32)] + [(u[k][sl] * (1.0 / 0.05)).astype(np.float32)
                                                                     for k in range(3)]
    assert out.tobytes() == b"".join(np.ascontiguousarray(w).tobytes() for w in want)
    blk.close()


def test_coupled_restart_is_bit_exact():
    """fem_flexible_plate: 3 steps, write the continue file and the beam
    checkpoint, 2 more steps; a fresh block restarted from those files runs the
    same 2 steps to the same fluid and beam state, bit for bit."""
    import ctypes
    from fsilbm3d_amd import _lib
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    from oracle.refcase import read_continue
    pc = BY_NAME["fem_flexible_plate"]
    _say(f"[gpu] restart {pc.name}")
    case = pc.case

    def bodies(blk, d):
        return DeviceBodyBlock(blk, plate_specs(pc), ntolLBM=case.ntolLBM, dtolLBM=case.dtolLBM, Uref=d["Uref"],
                               denIn=case.denIn, Lref=case.Lref, numsubstep=case.numsubstep,
                               IBPenaltyAlpha=case.IBPenaltyAlpha, dampK=case.dampK, dampM=case.dampM,
                               GeoGamma=case.GeoGamma, NewmarkGamma=case.NewmarkGamma,
                               NewmarkBeta=case.NewmarkBeta, dtolFEM=case.dtolFEM, ntolFEM=case.ntolFEM)

    blk, d = _fluid_block(pc)
    blk.set_boundary_conditions()
    fsi = bodies(blk, d)
    t = 0.0
    for _ in range(3):
        t = t + case.dh
        fsi.step(t)
    with tempfile.TemporaryDirectory() as td:
        pcn = os.path.join(td, "continue")
        h = np.array([blk.handle], np.int32)
        _lib.check(blk.lib.fsil_write_continue(1, _lib.iptr(h), 3, t / d["Tref"], pcn.encode()))
        n = ctypes.c_longlong(0)
        _lib.check(blk.lib.fsil_fem_checkpoint(blk.handle, None, ctypes.byref(n)))
        ck = np.zeros(n.value, np.uint8)
        _lib.check(blk.lib.fsil_fem_checkpoint(blk.handle, ck.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
        step, tt, f3 = read_continue(pcn, case.dims)
    t_restart = t
    for _ in range(2):
        t = t + case.dh
        fsi.step(t)
    fa = blk.download()
    sa = fsi.state(0)
    blk.close()

    blk2, _ = _fluid_block(pc, f0=f3)
    fsi2 = bodies(blk2, d)
    _lib.check(blk2.lib.fsil_fem_restore(blk2.handle, ck.ctypes.data_as(ctypes.c_void_p), n.value))
    assert step == 3 and tt == t_restart / d["Tref"]
    t = t_restart
    for _ in range(2):
        t = t + case.dh
        fsi2.step(t)
    fb = blk2.download()
    sb = fsi2.state(0)
    blk2.close()
    assert np.array_equal(fa, fb)
    for a, b in zip(sa, sb):
        assert np.array_equal(a, b)
