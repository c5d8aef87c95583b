"""FP32-deviation storage form beside the FP64 plain form: open boundary faces
and the immersed-boundary coupling of a flexible plate.  The forms are not
bit-identical by design; each test states its tolerance (the plain form is
the reference for the stored state, drag/lift within 1e-3 relative)."""
import numpy as np
import pytest

from tests.golden.cases import BY_NAME
from tests.helpers import derived, flow_kwargs, run_plate_fsi

pytestmark = pytest.mark.gpu


def _run_block(pc, precision, nsteps):
    from fsilbm3d_amd.block import LBMBlock
    case = pc.case
    d = derived(case)
    blk = LBMBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["nu"], case.params, precision=precision)
    blk.set_flow(**flow_kwargs(case))
    blk.upload(pc.initial_state())
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    t = 0.0
    for _ in range(nsteps):
        t = t + case.dh
        blk.update_volume_force(t)
        blk.collision_streaming_bc()
    f = blk.download()
    u = blk.calculate_macro_quantities()[1]
    blk.close()
    return f, u


@pytest.mark.parametrize("name", ["bc_inflow_outflow_slip", "bc_moving_walls", "periodic_srt_force", "periodic_trt",
                                  "periodic_trt_force", "periodic_mrt", "periodic_mrt_force", "bc_mixed_trt_force",
                                  "bc_neq_extrap2_walls", "bc_halfway_walls"])
def test_fp32_open_faces_track_fp64(name):
    """Inflow/outflow/slip, moving walls, forced periodic blocks, the TRT and
    MRT collisions with and without a volume force (also TRT with open
    faces), non-equilibrium extrapolation inflow (102) and halfway walls
    (203/204): the FP32 form's velocity stays within 1e-5 (relative L2) of the
    FP64 form after 50 steps, and its populations within 1e-6 absolute."""
    pc = BY_NAME[name]
    print(f"[gpu] fp32 faces {name}", flush=True)
    f64, u64 = _run_block(pc, 0, 50)
    f32, u32 = _run_block(pc, 1, 50)
    l2 = np.sqrt(((u32 - u64) ** 2).sum() / (u64 ** 2).sum())
    assert l2 < 1e-5, l2
    assert np.abs(f32 - f64).max() < 1e-6


def test_fp32_flexible_plate_forces_within_1e_3():
    """A flexible plate (L = 16 cells, 17 chord nodes x 16 span markers) in a
    96x64x64 channel, 60 steps: the force on the body and the node positions of
    the FP32 form stay within 1e-3 relative of the FP64 form at every step."""
    from fsilbm3d_amd.workloads import FlexiblePlate
    w = FlexiblePlate(dims=(96, 64, 64), L=16.0, Re=100.0, le=(24.0, 32.0, 32.0))
    print("[gpu] fp32 flexible plate", flush=True)
    r64 = run_plate_fsi(w, 0, 60)
    r32 = run_plate_fsi(w, 1, 60)
    F64, F32 = r64["forces"], r32["forces"]
    scale = np.abs(F64[:, :2]).max(axis=0)
    rel = np.abs(F32[:, :2] - F64[:, :2]).max(axis=0) / scale
    print(f"[gpu] fp32 plate: max rel force error drag {rel[0]:.2e} lift {rel[1]:.2e}", flush=True)
    assert (rel < 1e-3).all(), rel
    assert np.abs(r32["pos"] - r64["pos"]).max() < 1e-3 * w.L
