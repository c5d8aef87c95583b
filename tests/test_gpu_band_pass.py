"""The band-cell sums of the coupling call (velocity corrections and
spreading) in their two forms: a workgroup's pass over 32 cells staged in LDS
(the default) and one thread per cell reading its entries itself (taken when
a pass's entries exceed the LDS area; FSIL_BAND_PASS=0 forces it).  A run of
coupled steps ends in the same fluid state, forces and beam state in both
forms, bit for bit, in both storage forms."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(precision, cap, nsteps=6):
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    from fsilbm3d_amd.workloads import FlexiblePlate
    w = FlexiblePlate(dims=(96, 64, 64), L=16.0, Re=100.0, le=(24.0, 32.0, 32.0))
    blk = LBMBlock(w.dims, 1.0, (0.0, 0.0, 0.0), w.bc, 1, w.nu, precision=precision)
    blk.set_flow(**w.flow_kwargs())
    blk.upload(w.initial_state())
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    old = os.environ.get("FSIL_BAND_PASS")
    if cap is None:
        os.environ.pop("FSIL_BAND_PASS", None)
    else:
        os.environ["FSIL_BAND_PASS"] = cap
    try:
        # read when the coupling's bodies are defined
        fsi = DeviceBodyBlock(blk, [w.plate()], **w.solver_kwargs())
    finally:
        if old is None:
            os.environ.pop("FSIL_BAND_PASS", None)
        else:
            os.environ["FSIL_BAND_PASS"] = old
    forces, iters = [], []
    t = 0.0
    for _ in range(nsteps):
        t = t + 1.0
        fsi.step(t)
        forces.append(fsi.body_loads(0).copy())
        iters.append(fsi.iters)
    out = dict(f=blk.download(), forces=np.array(forces), iters=iters, state=[a.copy() for a in fsi.state(0)])
    blk.close()
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_band_cell_sums_per_cell_form_matches_pass_form(precision):
    print(f"[gpu] band-cell sum forms precision {precision}", flush=True)
    a = _run(precision, None)
    b = _run(precision, "0")
    assert a["iters"] == b["iters"]
    assert np.array_equal(a["forces"], b["forces"])
    assert np.array_equal(a["f"], b["f"])
    for x, y in zip(a["state"], b["state"]):
        assert np.array_equal(x, y)
