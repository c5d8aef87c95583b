"""The fused collide/push kernel streams the populations with plain or with
nontemporal cache hints, chosen per block (FSIL_STREAM_HINT overrides the
choice at block creation), and the FP32-deviation kernel exchanges the values
that move along z through LDS when its workgroups are whole z-rows
(FSIL_ZROW=0 turns that off).  Both change only how memory is written, so
every combination gives bit-identical states; the default ones are those the
other parity tests compare with the oracle."""
import numpy as np
import pytest

from tests.golden.cases import BY_NAME
from tests.helpers import derived, flow_kwargs

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
    blk.close()
    return f


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["periodic_srt_force", "bc_inflow_outflow_slip"])
def test_stream_hints_bit_identical(name, precision, monkeypatch):
    pc = BY_NAME[name]
    out = {}
    for hint in ("plain", "nt"):
        for zrow in ("1", "0"):
            print(f"[gpu] stream hint {hint}, z-row exchange {zrow}: {name} precision {precision}", flush=True)
            monkeypatch.setenv("FSIL_STREAM_HINT", hint)
            monkeypatch.setenv("FSIL_ZROW", zrow)
            out[hint, zrow] = _run_block(pc, precision, 20)
    monkeypatch.delenv("FSIL_STREAM_HINT")
    monkeypatch.delenv("FSIL_ZROW")
    ref = out["plain", "1"]
    for k, v in out.items():
        assert np.array_equal(v, ref), k


def test_block_size_caps():
    """The fused kernels address a population plane with 32-bit byte offsets:
    blocks beyond 2^29 cells (FP64) or 2^30 cells (FP32 deviations) are refused
    at creation, before any allocation."""
    from fsilbm3d_amd._lib import FsilError
    from fsilbm3d_amd.block import LBMBlock
    print("[gpu] block size caps", flush=True)
    with pytest.raises(FsilError, match=r"2\^29"):
        LBMBlock((1024, 1024, 513), precision=0)
    with pytest.raises(FsilError, match=r"2\^30"):
        LBMBlock((1024, 1024, 1025), precision=1)
