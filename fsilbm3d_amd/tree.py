"""Host mirror of the reference's block tree (gaoak/FSILBM3D src/LBMBlockComm.f90):
the father/son relation found from the block extents, and the recursive tree
step with the halo exchange on the device.

  CompareBlocks                      FluidDomain.f90:1845-1972  -> compare_blocks()
  findremove_blockTreeRoot, array_to_tree, build_block_tree
                                     LBMBlockComm.f90:99-199    -> BlockTree.__init__
  build_blocks_comunication          LBMBlockComm.f90:33-97     -> fsil_tree_add_son
  tree_collision_streaming_IBM_FEM   LBMBlockComm.f90:279-318   -> BlockTree.step()
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .block import LBMBlock

BC_FLUID, BC_FLUID_FATHER = 0, 1
MACHINE_TOL = 1.0e-12   # ConstParams.f90:36


def _extent(b: LBMBlock):
    return [b.origin[0], b.xmax, b.origin[1], b.ymax, b.origin[2], b.zmax]


def compare_blocks(bi: LBMBlock, bj: LBMBlock) -> int:
    """1 if block i contains block j, -1 if j contains i, 0 otherwise
    (FluidDomain.f90:1845-1972; the adjustment for a son with one BCfluid face
    facing a BCfluid_father face included)."""
    iv, jv = _extent(bi), _extent(bj)
    for a, av, b in ((bi, iv, bj), (bj, jv, bi)):
        fl = [c == BC_FLUID for c in a.bc]
        if sum(fl) == 1:
            for p in range(6):
                if not fl[p]:
                    continue
                ov = jv if av is iv else iv
                if p % 2 == 0:
                    if b.bc[p + 1] == BC_FLUID_FATHER and abs(ov[p + 1] - av[p] - b.dh) < MACHINE_TOL:
                        av[p + 1] = av[p]
                else:
                    if b.bc[p - 1] == BC_FLUID_FATHER and abs(av[p] - ov[p - 1] - b.dh) < MACHINE_TOL:
                        av[p - 1] = av[p]
    cnt, align = 0, 0
    for k in range(3):
        i0, i1, j0, j1 = iv[2 * k], iv[2 * k + 1], jv[2 * k], jv[2 * k + 1]
        d1 = i0 < j0 or abs(i0 - j0) < MACHINE_TOL
        d2 = j1 < i1 or abs(j1 - i1) < MACHINE_TOL
        d3 = j0 < i0 or abs(j0 - i0) < MACHINE_TOL
        d4 = i1 < j1 or abs(i1 - j1) < MACHINE_TOL
        d5 = i1 < j0
        d6 = j1 < i0
        if d1 and d2 and not (d3 and d4):
            cnt += 1
        elif d3 and d4 and not (d1 and d2):
            cnt -= 1
        elif d1 and d2 and d3 and d4:
            align += 1
        elif d5 or d6:
            return 0
    if align > 0:
        if cnt < 0:
            cnt -= align
        if cnt > 0:
            cnt += align
    if cnt == 3:
        return 1
    if cnt == -3:
        return -1
    return 0


class BlockTree:
    """Fluid blocks arranged as the reference's blockTree.  blocks[0..] are the
    LBMBlock objects in input order; bodies maps a block index to the object
    that couples its carried bodies (fsi.DeviceBodyBlock or fsi.FSIBlock)."""

    def __init__(self, blocks: Sequence[LBMBlock], interpolateScheme: int = 1,
                 bodies: Optional[Dict[int, object]] = None):
        self.blocks = list(blocks)
        self.lib = self.blocks[0].lib
        self.scheme = int(interpolateScheme)
        self.bodies = dict(bodies or {})
        nb = len(self.blocks)
        self.sons: List[List[int]] = [[] for _ in range(nb)]
        self.father = [-1] * nb
        rest = list(range(nb))
        self.root = self._find_root(rest)
        self._array_to_tree(rest, self.root)
        self.index: Dict[int, int] = {}
        self._build_comm(self.root)

    def _find_root(self, iblocks: List[int]) -> int:
        """findremove_blockTreeRoot (LBMBlockComm.f90:99-133)."""
        fa = [-1] * len(iblocks)
        for i in range(len(iblocks) - 1):
            for j in range(i + 1, len(iblocks)):
                t = compare_blocks(self.blocks[iblocks[i]], self.blocks[iblocks[j]])
                if t == 1:
                    fa[j] = iblocks[i]
                elif t == -1:
                    fa[i] = iblocks[j]
        roots = [b for b, f in zip(iblocks, fa) if f < 0]
        if len(roots) != 1:
            raise ValueError(f"there must be exactly one block tree root, found {len(roots)}: {fa}")
        iblocks[:] = [b for b, f in zip(iblocks, fa) if f >= 0]
        return roots[0]

    def _array_to_tree(self, iblocks: List[int], rootnode: int):
        """array_to_tree (LBMBlockComm.f90:135-185)."""
        nb = len(iblocks)
        if nb == 0:
            return
        fa = [-1] * nb
        for i in range(nb - 1):
            for j in range(i + 1, nb):
                t = compare_blocks(self.blocks[iblocks[i]], self.blocks[iblocks[j]])
                if t == 1:
                    fa[j] = i
                elif t == -1:
                    fa[i] = j
        changed = True
        while changed:
            changed = False
            for i in range(nb):
                if fa[i] >= 0 and fa[fa[i]] >= 0 and fa[i] != fa[fa[i]]:
                    fa[i] = fa[fa[i]]
                    changed = True
        roots = [iblocks[i] for i in range(nb) if fa[i] < 0]
        self.sons[rootnode] = roots
        for r in roots:
            self.father[r] = rootnode
        for r in roots:
            sub = [iblocks[j] for j in range(nb) if fa[j] >= 0 and iblocks[fa[j]] == r]
            self._array_to_tree(sub, r)

    def _build_comm(self, node: int):
        for s in self.sons[node]:
            idx = np.zeros(1, np.int32)
            _lib.check(self.lib.fsil_tree_add_son(self.blocks[node].handle, self.blocks[s].handle, self.scheme,
                                                  _lib.iptr(idx)))
            self.index[s] = int(idx[0])
            self._build_comm(s)

    # -- the halo exchange ---------------------------------------------------
    def comm_pair(self, son: int) -> dict:
        """CommPair of a son block (1-based indices)."""
        c = np.zeros(36, np.int32)
        _lib.check(self.lib.fsil_tree_pair_info(self.blocks[self.father[son]].handle, self.index[son],
                                                _lib.iptr(c)))
        g = lambda k: [int(v) for v in c[6 * k:6 * k + 6]]
        return dict(sds=g(0), s=g(1), f=g(2), si=g(3), fi=g(4), dimS=[int(v) for v in c[30:33]],
                    dimF=[int(v) for v in c[33:36]])

    def extract_interpolate_layer(self, node: int, level: int):
        _lib.check(self.lib.fsil_extract_interpolate_layer(self.blocks[node].handle, int(level)))

    def interpolation_father_to_son(self, son: int, n_timeStep: int):
        _lib.check(self.lib.fsil_interpolation_father_to_son(self.blocks[self.father[son]].handle, self.index[son],
                                                             int(n_timeStep)))

    def deliver_son_to_father(self, son: int):
        _lib.check(self.lib.fsil_deliver_son_to_father(self.blocks[self.father[son]].handle, self.index[son]))

    # -- time stepping ---------------------------------------------------------
    def set_boundary_conditions(self):
        """tree_set_boundary_conditions_block (LBMBlockComm.f90:262-277)."""
        for b in self.blocks:
            b.update_volume_force(0.0)
            b.set_boundary_conditions()

    def _step(self, node: int):
        blk = self.blocks[node]
        blk.update_volume_force()
        coupler = self.bodies.get(node)
        if coupler is not None:
            coupler.ibm_fem(blk.blktime)
        self.extract_interpolate_layer(node, 1)
        blk.collision_streaming_bc()
        self.extract_interpolate_layer(node, 2)
        for s in self.sons[node]:
            sb = self.blocks[s]
            for n in range(2):
                sb.blktime = sb.blktime + float(n) * sb.dh
                self._step(s)
                self.interpolation_father_to_son(s, n)
            self.deliver_son_to_father(s)

    def step(self, time: float):
        """One root step at time ``time`` (main.f90:95-106)."""
        for b in self.blocks:
            b.blktime = float(time)
        self._step(self.root)

    def steps(self, nsteps: int, t0: float, dt: float):
        """nsteps root steps from t0; a tree without bodies runs the whole
        recursion in one library call."""
        if not self.bodies:
            _lib.check(self.lib.fsil_tree_steps(self.blocks[self.root].handle, int(nsteps), float(t0), float(dt)))
            return
        t = float(t0)
        for _ in range(int(nsteps)):
            t = t + dt
            self.step(t)

    def time_steps(self, nsteps: int, t0: float, dt: float) -> float:
        ms = np.zeros(1)
        _lib.check(self.lib.fsil_time_tree_steps(self.blocks[self.root].handle, int(nsteps), float(t0), float(dt),
                                                 _lib.dptr(ms)))
        return float(ms[0])
