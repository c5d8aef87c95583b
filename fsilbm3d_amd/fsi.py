"""One fluid block coupled to rigid immersed bodies: the per-step order of
tree_collision_streaming_IBM_FEM + IBM_FEM (LBMBlockComm.f90:279-338) for a
single block, with the fluid and the immersed-boundary forcing on the device
and the prescribed rigid kinematics on the host.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np

from . import _lib
from .block import LBMBlock
from .solid import RigidPlate


class FSIBlock:
    def __init__(self, block: LBMBlock, bodies: List[RigidPlate], *, ntolLBM: int, dtolLBM: float,
                 Uref: float, denIn: float, numsubstep: int = 1):
        self.block = block
        self.bodies = bodies
        self.numsubstep = int(numsubstep)
        lib = block.lib
        self.lib = lib
        nm = np.array([b.M for b in bodies], dtype=np.int32)
        nn = np.array([b.nND for b in bodies], dtype=np.int32)
        ne = np.array([b.nEL for b in bodies], dtype=np.int32)
        me = np.ascontiguousarray(np.concatenate([b.marker_elem for b in bodies]) if bodies else np.zeros(1, np.int32), dtype=np.int32)
        en = np.ascontiguousarray(np.concatenate([b.elem_nodes.reshape(-1) for b in bodies]) if bodies else np.zeros(2, np.int32), dtype=np.int32)
        _lib.check(lib.fsil_ibm_define_bodies(block.handle, len(bodies), _lib.iptr(nm if len(nm) else np.zeros(1, np.int32)),
                                              _lib.iptr(nn if len(nn) else np.zeros(1, np.int32)),
                                              _lib.iptr(ne if len(ne) else np.zeros(1, np.int32)),
                                              _lib.iptr(me), _lib.iptr(en)))
        _lib.check(lib.fsil_ibm_set_params(block.handle, int(ntolLBM), float(dtolLBM), float(Uref), float(denIn)))
        self.M = int(nm.sum()) if len(nm) else 0
        self.Eforce = np.zeros((max(self.M, 1), 3))
        self.lodFlow = np.zeros(max(6 * int(nn.sum()) if len(nn) else 0, 1))
        self.iters = 0
        self.dmax = 0.0

    def interaction_force(self, dt: float):
        """FSInteraction_force (Solidbody.f90:589-602)."""
        for b in self.bodies:
            if b.v_move == 1:
                b.update_markers()
        if not self.bodies:
            return
        Exyz = np.ascontiguousarray(np.concatenate([b.Exyz for b in self.bodies]))
        Evel = np.ascontiguousarray(np.concatenate([b.Evel for b in self.bodies]))
        Ea = np.ascontiguousarray(np.concatenate([b.Ea for b in self.bodies]))
        xc = np.ascontiguousarray(np.concatenate([b.element_centres() for b in self.bodies]))
        it = np.zeros(1, dtype=np.int32)
        dm = np.zeros(1)
        _lib.check(self.lib.fsil_ibm_interaction_force(self.block.handle, float(dt), _lib.dptr(Exyz), _lib.dptr(Evel),
                                                       _lib.dptr(Ea), _lib.dptr(xc), _lib.dptr(self.Eforce),
                                                       _lib.dptr(self.lodFlow), _lib.iptr(it), _lib.dptr(dm)))
        self.iters, self.dmax = int(it[0]), float(dm[0])

    def body_loads(self, i: int) -> np.ndarray:
        off = 6 * sum(b.nND for b in self.bodies[:i])
        return self.lodFlow[off:off + 6 * self.bodies[i].nND]

    def step(self, time: float):
        """One time step at block time ``time`` (main.f90:95-107)."""
        blk = self.block
        blk.update_volume_force(time)
        self.interaction_force(blk.dh)
        dt_solid = blk.dh / float(self.numsubstep)
        for isub in range(1, self.numsubstep + 1):
            for b in self.bodies:
                b.structure(time, isub, blk.dh, dt_solid)
        blk.collision_streaming_bc()
