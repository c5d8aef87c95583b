"""The beam sub-step kernel's two forms: with a flexible body's element fields
held in LDS for the whole launch (the default where the LDS holds them) and
with every element access in the global element array (FSIL_FEM_HOT=0, the
form taken for bodies too large for the LDS or with stiffness damping).  A run
of coupled steps ends in the same fluid state, forces, beam state and
iteration counts in both forms, bit for bit, in both storage forms."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(precision, hot, nsteps=8):
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    from fsilbm3d_amd.workloads import FlexiblePlate
    w = FlexiblePlate(dims=(96, 64, 64), L=16.0, Re=100.0, le=(24.0, 32.0, 32.0))
    blk = LBMBlock(w.dims, 1.0, (0.0, 0.0, 0.0), w.bc, 1, w.nu, precision=precision)
    blk.set_flow(**w.flow_kwargs())
    blk.upload(w.initial_state())
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    old = os.environ.get("FSIL_FEM_HOT")
    os.environ["FSIL_FEM_HOT"] = hot
    try:
        # read when the bodies are defined
        fsi = DeviceBodyBlock(blk, [w.plate()], **w.solver_kwargs())
    finally:
        if old is None:
            os.environ.pop("FSIL_FEM_HOT", None)
        else:
            os.environ["FSIL_FEM_HOT"] = old
    forces, iters, fem = [], [], []
    t = 0.0
    for _ in range(nsteps):
        t = t + 1.0
        fsi.step(t)
        forces.append(fsi.body_loads(0).copy())
        iters.append(fsi.iters)
        fem.append(tuple(fsi.info(0)))
    out = dict(f=blk.download(), forces=np.array(forces), iters=iters, fem=fem,
               state=[a.copy() for a in fsi.state(0)])
    blk.close()
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_lds_element_form_matches_global_form(precision):
    print(f"[gpu] beam kernel forms precision {precision}", flush=True)
    a = _run(precision, "0")
    b = _run(precision, "1")
    assert a["iters"] == b["iters"]
    assert a["fem"] == b["fem"]
    assert np.array_equal(a["forces"], b["forces"])
    assert np.array_equal(a["f"], b["f"])
    for x, y in zip(a["state"], b["state"]):
        assert np.array_equal(x, y)
