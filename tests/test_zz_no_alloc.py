"""fsil_solid_info_static (the DatInfo reductions in the library's static
device array, no allocation) against fsil_solid_info on the same bodies,
bit for bit: two calls in a row, the second on a body with fewer elements,
then the first again; a body beyond the static array's 256 elements is an
error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REFS = np.array([16.0, 0.05, 0.05 / 320.0, 0.0125, 0.000625, 0.2])


def _plate(n_nodes, x0, y, z, nspan, half, amp, model=2):
    from fsilbm3d_amd.fsi import PlateSpec
    nodes = np.array([[i - 0.5 * (n_nodes - 1), 0.0, 0.0, half, half, 0.0, 0.0, 1.0] for i in range(n_nodes)])
    els = np.array([(i + 1, i + 2, nspan) for i in range(n_nodes - 1)], np.int32)
    cons = np.zeros((n_nodes, 6), np.int32)
    cons[0] = 1
    return PlateSpec(nodes8=nodes, elements=els, constraints=cons, iBodyModel=model, XYZo=(x0, y, z),
                     XYZAmpl=(0.0, amp, 0.0), Freq=0.01)


def _block(dims, bc, plates, steps):
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    from fsilbm3d_amd.params import equilibrium_point
    blk = LBMBlock(dims, 1.0, (0.0, 0.0, 0.0), bc, 1, 0.05 * 16 / 100.0)
    blk.set_flow(denIn=1.0, uvwIn=(0.05, 0.0, 0.0))
    blk.fill(equilibrium_point(1.0, (0.05, 0.0, 0.0)))
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    fsi = DeviceBodyBlock(blk, plates, ntolLBM=3, dtolLBM=1e-14, Uref=0.05, denIn=1.0, Lref=16.0, numsubstep=2)
    t = 0.0
    for _ in range(steps):
        t += 1.0
        fsi.step(t)
    return blk, fsi


def _rows(fsi, fn, body):
    from fsilbm3d_amd import _lib
    out = np.full(78, np.nan)
    _lib.check(fn(fsi.block.handle, body, _lib.dptr(REFS), _lib.dptr(out)))
    return out


def test_static_scratch_matches_allocating_entry_point():
    print("[gpu] static-scratch solid info: two bodies (8 and 4 elements)", flush=True)
    blk, fsi = _block((48, 24, 24), (101, 103, 302, 302, 302, 302),
                      [_plate(9, 16.0, 12.0, 12.0, 4, 2.0, 0.5), _plate(5, 34.0, 12.5, 12.0, 4, 2.0, 0.3)], 3)
    lib = fsi.lib
    a = _rows(fsi, lib.fsil_solid_info, 0)
    a_s = _rows(fsi, lib.fsil_solid_info_static, 0)
    b_s = _rows(fsi, lib.fsil_solid_info_static, 1)      # second call, smaller body
    b = _rows(fsi, lib.fsil_solid_info, 1)
    a_s2 = _rows(fsi, lib.fsil_solid_info_static, 0)
    blk.close()
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert np.array_equal(a_s, a) and np.array_equal(b_s, b) and np.array_equal(a_s2, a)
    assert not np.array_equal(a[45:], b[45:])


def test_static_scratch_rejects_large_bodies():
    from fsilbm3d_amd._lib import FsilError
    print("[gpu] static-scratch solid info: a 300-element body is refused", flush=True)
    blk, fsi = _block((320, 12, 12), (301,) * 6, [_plate(301, 160.0, 6.0, 6.0, 1, 0.5, 0.0, model=1)], 1)
    full = _rows(fsi, fsi.lib.fsil_solid_info, 0)
    with pytest.raises(FsilError, match="at most 256"):
        _rows(fsi, fsi.lib.fsil_solid_info_static, 0)
    blk.close()
    assert np.isfinite(full).all()
