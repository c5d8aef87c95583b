"""The school configuration (1024x384x384, 16 flexible plates on a 4x4
diamond lattice, each its own group and heave phase): the FP32-deviation
fluid against the FP64 plain form, per body, after a few steps from the
uniform inflow state; and the band against the markers' stencils."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS = 4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(w, precision):
    from fsilbm3d_amd import _lib
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    blk = LBMBlock(w.dims, 1.0, (0.0, 0.0, 0.0), w.bc, 1, w.nu, precision=precision)
    try:
        blk.set_flow(**w.flow_kwargs())
        blk.fill(w.equilibrium())
        blk.update_volume_force(0.0)
        blk.set_boundary_conditions()
        fsi = DeviceBodyBlock(blk, w.plates(), **w.solver_kwargs())
        t = 0.0
        for k in range(NSTEPS):
            t = t + 1.0
            fsi.step(t)
            _say(f"[gpu] school precision {precision} step {k + 1}: MDF rounds {fsi.iters}, residual {fsi.dmax:.3e}")
        loads = np.array([fsi.body_loads(i).reshape(-1, 6).sum(0)[:3] for i in range(len(fsi.nND))])
        nb = np.zeros(1, np.int32)
        box = np.zeros(6, np.int32)
        _lib.check(blk.lib.fsil_ibm_band_info(blk.handle, _lib.iptr(nb), _lib.iptr(box)))
        tips = np.array([fsi.state(i)[0][-1, :3] for i in range(len(fsi.nND))])
    finally:
        blk.close()
    return loads, int(nb[0]), box, tips


def test_school_fp32_matches_fp64_per_body():
    from fsilbm3d_amd import workloads as W
    w = W.school()
    assert len(w.plates()) == 16
    _say("[gpu] school 1024x384x384, FP64 plain form")
    l64, nb64, box64, tip64 = _run(w, 0)
    _say("[gpu] school 1024x384x384, FP32-deviation form")
    l32, nb32, box32, tip32 = _run(w, 1)
    assert np.all(np.isfinite(l64)) and np.all(np.isfinite(l32))
    # the same markers in both forms: the same band
    assert nb64 == nb32 and list(box64) == list(box32)
    # every body carries load (drag in the flow direction)
    assert np.all(np.abs(l64[:, 0]) > 0.0)
    scale = np.abs(l64[:, :2]).max(axis=0)
    err = np.abs(l32[:, :2] - l64[:, :2]).max(axis=0) / scale
    _say(f"[gpu] school per-body drag/lift error FP32 vs FP64 after {NSTEPS} steps: {err}")
    assert err.max() < 1e-3
    assert np.abs(tip32 - tip64).max() < 1e-6 * np.abs(tip64).max()
