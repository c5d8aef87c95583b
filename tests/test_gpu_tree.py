"""Father/son block tree on the device (halo exchange kernels through the
C-ABI) against the reference's two-block fixtures and the oracle tree."""
import dataclasses

import numpy as np
import pytest

from oracle.refcase import Case, SubBlock
from tests.golden.cases import TREE, ParityCase, load_fixture
from tests.helpers import mismatch, run_device_tree, run_oracle_tree

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("per_call", [False, True], ids=["one_call", "host_recursion"])
@pytest.mark.parametrize("pc", TREE, ids=[c.name for c in TREE])
def test_tree_matches_reference_fixture(pc, per_call):
    """Both blocks bit for bit after the reference's tree steps, with the
    recursion in the library (fsil_tree_steps) or in the host mirror."""
    print(f"[gpu] tree {pc.name} per_call={per_call}", flush=True)
    fs, _ = run_device_tree(pc, per_call=per_call)
    fx = load_fixture(pc.name)
    for k, key in enumerate(["f_final", "f_final_son1"]):
        n, m = mismatch(fs[k], fx[key])
        assert n == 0, f"{key}: {n} values differ, max |diff| {m:.3e}"


def test_tree_comm_pair_matches_oracle():
    from oracle.oracle_tree import comm_pair
    for pc in TREE:
        print(f"[gpu] comm pair {pc.name}", flush=True)
        _, pairs = run_device_tree(pc, nsteps=0)
        c = pc.case
        sb = c.sons[0]
        ref = comm_pair((c.dims, c.dh, c.origin), (sb.dims, sb.dh, sb.origin), sb.bc,
                        [int(sb.bc[2 * a] == 301) for a in range(3)])
        assert list(pairs) == [1]
        got = pairs[1]
        for key in ("sds", "s", "f", "si", "fi"):
            assert got[key] == ref[key], (pc.name, key)
        assert got["dimF"] == [ref["xDimF"], ref["yDimF"], ref["zDimF"]]


# Trees beyond the fixtures: a son periodic along z (even node count, so the
# wrap-around column of the plane interpolation is used), time-dependent
# volume force (the son's two sub-step times enter its force), mixed models.
EXTRA = [
    ParityCase("tree_periodic_son_quartic",
               Case(dims=(24, 16, 12), bc=(101, 103, 302, 302, 301, 301), Re=3.0, Uref=0.05, UrefType=9, Lref=1.0,
                    LrefType=1, TrefType=0, uvwIn=(0.05, 0.0, 0.0), interpolateScheme=2,
                    sons=[SubBlock(dims=(17, 13, 24), dh=0.5, origin=(6.0, 2.0, 0.0), bc=(0, 0, 0, 0, 301, 301))]),
               nsteps=4, seed=51, ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("tree_trt_mrt_linear_forced",
               Case(dims=(20, 18, 16), bc=(301,) * 6, Re=3.0, Uref=0.05, UrefType=9, Lref=1.0, LrefType=1,
                    TrefType=0, collide=2, params=(0.25,) + (0.0,) * 9, interpolateScheme=1,
                    volumeForceIn=(2e-5, -1e-5, 0.0), volumeForceAmp=1e-5, volumeForceFreq=0.05,
                    volumeForcePhi=30.0,
                    sons=[SubBlock(dims=(15, 13, 11), dh=0.5, origin=(5.0, 4.0, 3.0), collide=3)]),
               nsteps=4, seed=52, neq=0.01),
    ParityCase("tree_two_sons",
               Case(dims=(28, 14, 14), bc=(101, 103, 201, 201, 302, 302), Re=3.0, Uref=0.05, UrefType=9, Lref=1.0,
                    LrefType=1, TrefType=0, uvwIn=(0.05, 0.0, 0.0), interpolateScheme=2,
                    sons=[SubBlock(dims=(13, 13, 13), dh=0.5, origin=(3.0, 3.0, 3.0)),
                          SubBlock(dims=(15, 11, 13), dh=0.5, origin=(14.0, 4.0, 3.0))]),
               nsteps=3, seed=53, ubase=(0.05, 0, 0), neq=0.01),
]


@pytest.mark.parametrize("pc", EXTRA, ids=[c.name for c in EXTRA])
def test_tree_matches_oracle(pc):
    print(f"[gpu] tree vs oracle {pc.name}", flush=True)
    want = run_oracle_tree(pc)
    got, _ = run_device_tree(pc)
    for k in range(len(want)):
        n, m = mismatch(got[k], want[k])
        assert n == 0, f"block {k}: {n} values differ, max |diff| {m:.3e}"


def test_tree_mid_size_matches_oracle():
    """A 64x48x48 father with a 65x49x49 son (scheme 2), 2 steps, bit for bit."""
    pc = ParityCase("tree_mid", Case(dims=(64, 48, 48), bc=(101, 103, 302, 302, 302, 302), Re=20.0, Uref=0.05,
                                     UrefType=9, Lref=8.0, LrefType=1, TrefType=0, uvwIn=(0.05, 0.0, 0.0),
                                     interpolateScheme=2,
                                     sons=[SubBlock(dims=(65, 49, 49), dh=0.5, origin=(16.0, 12.0, 12.0))]),
                    nsteps=2, seed=54, ubase=(0.05, 0, 0), neq=0.01)
    print("[gpu] tree mid size", flush=True)
    want = run_oracle_tree(pc)
    got, _ = run_device_tree(pc)
    for k in range(2):
        n, m = mismatch(got[k], want[k])
        assert n == 0, f"block {k}: {n} values differ, max |diff| {m:.3e}"


@pytest.mark.parametrize("son", [
    SubBlock(dims=(13, 11, 11), dh=0.5, origin=(4.25, 3.0, 3.0)),   # off the father grid
    SubBlock(dims=(12, 11, 11), dh=0.5, origin=(4.0, 3.0, 3.0)),    # even count on a non-periodic axis
    SubBlock(dims=(13, 11, 11), dh=0.25, origin=(4.0, 3.0, 3.0)),   # ratio 4
], ids=["misaligned", "even_dims", "ratio4"])
def test_tree_rejects_mismatched_son(son):
    """check_blocks_params (LBMBlockComm.f90:508-544) stops the reference on
    these; the library refuses the pair with its message."""
    from fsilbm3d_amd import _lib
    from fsilbm3d_amd.block import LBMBlock
    father = LBMBlock((16, 12, 12), 1.0, (0.0, 0.0, 0.0), (101, 103, 302, 302, 302, 302), 1, 0.05)
    s = LBMBlock(son.dims, son.dh, son.origin, son.bc, 1, 0.05, block_id=2)
    idx = np.zeros(1, np.int32)
    rc = father.lib.fsil_tree_add_son(father.handle, s.handle, 1, _lib.iptr(idx))
    msg = father.lib.fsil_last_error().decode()
    s.close()
    father.close()
    assert rc != 0 and "grid points do not match" in msg


def test_tree_quartic_rejects_small_faces():
    from fsilbm3d_amd import _lib
    from fsilbm3d_amd.block import LBMBlock
    father = LBMBlock((16, 12, 12), 1.0, (0.0, 0.0, 0.0), (101, 103, 302, 302, 302, 302), 1, 0.05)
    # a son periodic along z with 4 nodes: x faces of 7 x 4 nodes
    s = LBMBlock((7, 7, 4), 0.5, (4.0, 3.0, 3.0), (0, 0, 0, 0, 301, 301), 1, 0.05, block_id=2)
    idx = np.zeros(1, np.int32)
    rc = father.lib.fsil_tree_add_son(father.handle, s.handle, 2, _lib.iptr(idx))
    msg = father.lib.fsil_last_error().decode()
    rc1 = father.lib.fsil_tree_add_son(father.handle, s.handle, 1, _lib.iptr(idx))
    s.close()
    father.close()
    assert rc != 0 and "at least 5 nodes" in msg
    assert rc1 == 0
