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


@pytest.mark.parametrize("dims", [(6, 4, 64), (4, 4, 192), (4, 4, 256), (4, 4, 320), (4, 4, 384)])
def test_whole_wave_row_exchange_bit_identical(dims, monkeypatch):
    """Rows of whole waves (z a multiple of 64) exchange the z-moving values by
    lane permutes; against the LDS form (FSIL_ZROW=1) and plain shifted
    stores (FSIL_ZROW=0) the FP32-deviation state after 12 periodic steps
    from a seeded random state is bit-identical."""
    from fsilbm3d_amd import params as P
    from fsilbm3d_amd.block import LBMBlock
    rng = np.random.default_rng(20261202)
    N = dims[0] * dims[1] * dims[2]
    rho = 1.0 + rng.uniform(-1e-3, 1e-3, N)
    u = rng.uniform(-0.05, 0.05, (3, N))
    f0 = np.stack([np.asarray(P.equilibrium_point(rho[n], u[:, n]), np.float64) for n in range(N)], axis=1)
    f0 = f0.reshape((19,) + tuple(dims))
    out = {}
    for mode in ("default", "1", "0"):
        print(f"[gpu] z-row exchange {mode}: {dims}", flush=True)
        if mode == "default":
            monkeypatch.delenv("FSIL_ZROW", raising=False)
        else:
            monkeypatch.setenv("FSIL_ZROW", mode)
        blk = LBMBlock(dims, bc=(301,) * 6, nu=0.02, params=(0.0,) * 10, precision=1)
        blk.set_flow(denIn=1.0, uvwIn=(0.0, 0.0, 0.0))
        blk.upload(f0)
        blk.steps(12, 0.0, 1.0)
        out[mode] = blk.download()
        blk.close()
    monkeypatch.delenv("FSIL_ZROW", raising=False)
    assert np.array_equal(out["default"], out["1"])
    assert np.array_equal(out["default"], out["0"])


def test_whole_wave_row_exchange_with_coupled_plate(monkeypatch):
    """The coupled flexible plate (96x64x64, FP32-deviation fluid: rows of one
    wave) gives the same force on the body at every step, bit for bit, with the
    lane-permute exchange and with plain shifted stores."""
    from fsilbm3d_amd.workloads import FlexiblePlate
    from tests.helpers import run_plate_fsi
    w = FlexiblePlate(dims=(96, 64, 64), L=16.0, Re=100.0, le=(24.0, 32.0, 32.0))
    print("[gpu] z-row exchange, coupled plate", flush=True)
    monkeypatch.delenv("FSIL_ZROW", raising=False)
    ra = run_plate_fsi(w, 1, 6)
    monkeypatch.setenv("FSIL_ZROW", "0")
    rb = run_plate_fsi(w, 1, 6)
    monkeypatch.delenv("FSIL_ZROW", raising=False)
    assert np.array_equal(np.asarray(ra["forces"]), np.asarray(rb["forces"]))
    assert ra["iters"] == rb["iters"]
