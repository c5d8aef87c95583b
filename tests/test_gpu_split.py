"""Split steps of fsil_fsi_step: the collision of the cells outside the
predicted band box runs beside the coupling call and the beam sub-steps, the
rows of the band box's hull after the call.  A run of split steps ends in the
same fluid state, forces and beam state as the same run of unsplit steps, bit
for bit, in both storage forms, also when the predicted box is smaller than
the band (the hull launch then collides again, with their force, the band
cells the bulk launch collided without it)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(precision, env, nsteps=8):
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    from fsilbm3d_amd.workloads import FlexiblePlate
    w = FlexiblePlate(dims=(96, 64, 64), L=16.0, Re=100.0, le=(24.0, 32.0, 32.0))
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        # the block reads FSIL_SPLIT and FSIL_SPLIT_MARGIN when it is created
        blk = LBMBlock(w.dims, 1.0, (0.0, 0.0, 0.0), w.bc, 1, w.nu, precision=precision)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    blk.set_flow(**w.flow_kwargs())
    blk.upload(w.initial_state())
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    fsi = DeviceBodyBlock(blk, [w.plate()], **w.solver_kwargs())
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
@pytest.mark.parametrize("margin", ["3", "-4"])
def test_split_steps_match_unsplit(precision, margin):
    print(f"[gpu] split steps precision {precision} margin {margin}", flush=True)
    a = _run(precision, {"FSIL_SPLIT": "0"})
    b = _run(precision, {"FSIL_SPLIT": "1", "FSIL_SPLIT_MARGIN": margin})
    assert a["iters"] == b["iters"]
    assert np.array_equal(a["forces"], b["forces"])
    assert np.array_equal(a["f"], b["f"])
    for x, y in zip(a["state"], b["state"]):
        assert np.array_equal(x, y)
