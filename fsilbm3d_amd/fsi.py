"""One fluid block coupled to rigid immersed bodies: the per-step order of
tree_collision_streaming_IBM_FEM + IBM_FEM (LBMBlockComm.f90:279-338) for a
single block, with the fluid and the immersed-boundary forcing on the device
and the prescribed rigid kinematics on the host.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np

import dataclasses

from . import _lib
from .block import LBMBlock
from .solid import RigidPlate

# DatInfo files of Beam_write_solid_info (SolidSolver.f90:1729-1799) and their row widths
INFO_WIDTHS = (("firstNode", 15), ("lastNode", 15), ("centerNode", 15), ("nodeAverage", 12), ("forces", 6),
               ("power", 7), ("energy", 8))


@dataclasses.dataclass
class PlateSpec:
    """A plate body as the reference's inputs describe it: plate.dat rows
    (POINT, ELEMENT, CONSTRAINT) plus the SolidBody group parameters of
    inFlow.dat (Solidbody.f90:121-155)."""
    nodes8: np.ndarray                 # nND x 8: x y z Lspan Rspan dirX dirY dirZ
    elements: np.ndarray               # nEL x 3: i j Nspan (1-based node ids)
    constraints: np.ndarray            # nND x 6
    iBodyModel: int = 1
    isMotionGiven: tuple = (1, 1, 1, 1, 1, 1)
    denR: float = 1.0
    KB: float = 0.1
    KS: float = 1000.0
    psR: float = 0.3
    Freq: float = 0.0
    XYZo: tuple = (0.0, 0.0, 0.0)
    XYZAmpl: tuple = (0.0, 0.0, 0.0)
    XYZPhi: tuple = (0.0, 0.0, 0.0)
    AoAo: tuple = (0.0, 0.0, 0.0)
    AoAAmpl: tuple = (0.0, 0.0, 0.0)
    AoAPhi: tuple = (0.0, 0.0, 0.0)
    initXYZVel: tuple = (0.0, 0.0, 0.0)

    def kin(self):
        return [self.Freq, *self.XYZo, *self.XYZAmpl, *self.XYZPhi, *self.AoAo, *self.AoAAmpl, *self.AoAPhi,
                *self.initXYZVel]


class DeviceBodyBlock:
    """A fluid block whose plate bodies live on the device beam solver: the
    whole IBM_FEM (LBMBlockComm.f90:320-338) runs through the C-ABI."""

    def __init__(self, block: LBMBlock, plates: List[PlateSpec], *, ntolLBM: int, dtolLBM: float, Uref: float,
                 denIn: float, Lref: float, numsubstep: int = 1, IBPenaltyAlpha: float = 1.0, dampK: float = 0.0,
                 dampM: float = 0.0, GeoGamma: float = 1.0, NewmarkGamma: float = 0.5, NewmarkBeta: float = 0.25,
                 dtolFEM: float = 1e-10, ntolFEM: int = 20):
        self.block = block
        self.lib = lib = block.lib
        self.numsubstep = int(numsubstep)
        nb = len(plates)
        nND = np.array([len(p.nodes8) for p in plates], np.int32)
        nEL = np.array([len(p.elements) for p in plates], np.int32)
        xyz8 = np.ascontiguousarray(np.concatenate([np.asarray(p.nodes8, np.float64) for p in plates]))
        els = np.ascontiguousarray(np.concatenate([np.asarray(p.elements, np.int32) for p in plates]))
        cons = np.ascontiguousarray(np.concatenate([np.asarray(p.constraints, np.int32) for p in plates]))
        ibm = np.array([p.iBodyModel for p in plates], np.int32)
        img = np.ascontiguousarray(np.array([p.isMotionGiven for p in plates], np.int32).reshape(-1))
        mat = np.ascontiguousarray(np.array([[p.denR, p.KB, p.KS, p.psR] for p in plates], np.float64).reshape(-1))
        kin = np.ascontiguousarray(np.array([p.kin() for p in plates], np.float64).reshape(-1))
        _lib.check(lib.fsil_fem_define_bodies(block.handle, nb, _lib.iptr(nND), _lib.iptr(nEL), _lib.dptr(xyz8),
                                              _lib.iptr(els), _lib.iptr(cons), _lib.iptr(ibm), _lib.iptr(img),
                                              _lib.dptr(mat), _lib.dptr(kin), float(Lref), float(Uref),
                                              float(denIn), float(IBPenaltyAlpha)))
        _lib.check(lib.fsil_fem_set_solver(block.handle, dampK, dampM, GeoGamma, NewmarkGamma, NewmarkBeta,
                                           dtolFEM, int(ntolFEM)))
        _lib.check(lib.fsil_ibm_set_params(block.handle, int(ntolLBM), float(dtolLBM), float(Uref), float(denIn)))
        self.nND = nND
        self.M = int(sum(int(np.asarray(p.elements)[:, 2].sum()) for p in plates))
        self._Eforce = np.zeros((max(self.M, 1), 3))
        self._lodFlow = np.zeros(6 * int(nND.sum()))
        self._stale = False
        self.iters = 0
        self.rounds = 0
        self.dmax = 0.0
        self._it = np.zeros(1, np.int32)
        self._dm = np.zeros(1)

    def interaction_force(self, dt: float, wait: bool = True):
        """The coupling call; with wait=False nothing is read back (no host
        wait; iters/dmax stay None and an error surfaces at the next call)."""
        _lib.check(self.lib.fsil_fem_update_markers(self.block.handle))
        if wait:
            it = np.zeros(1, dtype=np.int32)
            dm = np.zeros(1)
            _lib.check(self.lib.fsil_ibm_interaction_force(self.block.handle, float(dt), None, None, None, None,
                                                           None, None, _lib.iptr(it), _lib.dptr(dm)))
            self.iters, self.dmax = int(it[0]), float(dm[0])
            _lib.check(self.lib.fsil_ibm_rounds(self.block.handle, _lib.iptr(it)))
            self.rounds = int(it[0])
        else:
            _lib.check(self.lib.fsil_ibm_interaction_force(self.block.handle, float(dt), None, None, None, None,
                                                           None, None, None, None))
            self.iters, self.dmax = None, None
        self._stale = True

    def _fetch(self):
        if self._stale:
            _lib.check(self.lib.fsil_ibm_forces(self.block.handle, _lib.dptr(self._Eforce), _lib.dptr(self._lodFlow)))
            self._stale = False

    @property
    def Eforce(self) -> np.ndarray:
        """Marker forces of the last coupling call (read from the device on first use)."""
        self._fetch()
        return self._Eforce

    @property
    def lodFlow(self) -> np.ndarray:
        """Beam loads of the last coupling call (read from the device on first use)."""
        self._fetch()
        return self._lodFlow

    def set_warm_start(self, net=None, record: bool = False, skip: int = 0):
        """Learned warm start of the direct-forcing loop (fsilbm3d_amd.mlp.WarmStartNet):
        the network's guess stands for the first ``skip`` forcing rounds."""
        _lib.check(self.lib.fsil_ibm_set_warm_start(self.block.handle, net.handle if net is not None else 0,
                                                    int(bool(record)), int(skip)))

    def warm_start_data(self):
        """(features [M,12], targets [M,3], guess used, fell back) of the last step."""
        X = np.zeros((self.M, 12), np.float32)
        Y = np.zeros((self.M, 3), np.float32)
        info = np.zeros(2, np.int32)
        _lib.check(self.lib.fsil_ibm_warm_start_data(self.block.handle, _lib.fptr(X), _lib.fptr(Y), _lib.iptr(info)))
        return X, Y, bool(info[0]), bool(info[1])

    def body_loads(self, i: int) -> np.ndarray:
        off = 6 * int(self.nND[:i].sum())
        return self.lodFlow[off:off + 6 * int(self.nND[i])]

    def state(self, i: int):
        n = 6 * int(self.nND[i])
        out = [np.zeros(n) for _ in range(4)]
        _lib.check(self.lib.fsil_fem_state(self.block.handle, i, *[_lib.dptr(a) for a in out]))
        return [a.reshape(-1, 6) for a in out]

    def solid_info(self, i: int, refs) -> dict:
        """One DatInfo row per file of body i (Beam_write_solid_info,
        SolidSolver.f90:1729-1799), computed on the device; refs = (Lref,
        Uref, Aref, Fref, Pref, Eref)."""
        rows = np.zeros(78)
        _lib.check(self.lib.fsil_solid_info(self.block.handle, int(i),
                                            _lib.dptr(np.ascontiguousarray(refs, dtype=np.float64)), _lib.dptr(rows)))
        out, o = {}, 0
        for name, w in INFO_WIDTHS:
            out[name] = rows[o:o + w].copy()
            o += w
        return out

    def info(self, i: int):
        a = np.zeros(1, np.int32)
        b = np.zeros(1, np.int32)
        c = np.zeros(1)
        _lib.check(self.lib.fsil_fem_info(self.block.handle, i, _lib.iptr(a), _lib.iptr(b), _lib.dptr(c)))
        return int(a[0]), int(b[0]), float(c[0])

    def report(self):
        """Wait for the last coupling call and take its iteration count and
        residual (its errors are raised here)."""
        it = np.zeros(1, dtype=np.int32)
        dm = np.zeros(1)
        _lib.check(self.lib.fsil_ibm_report(self.block.handle, _lib.iptr(it), _lib.dptr(dm)))
        self.iters, self.dmax = int(it[0]), float(dm[0])
        _lib.check(self.lib.fsil_ibm_rounds(self.block.handle, _lib.iptr(it)))
        self.rounds = int(it[0])

    def ibm_fem(self, time: float, wait: bool = True):
        """IBM_FEM (LBMBlockComm.f90:320-338) at block time ``time``.  The beam
        sub-steps are queued behind the coupling call on the device before the
        host waits for the call (wait=True), so they start as soon as the call
        ends and ahead of the fluid update that the caller launches next."""
        blk = self.block
        self.interaction_force(blk.dh, False)
        _lib.check(self.lib.fsil_fem_substeps(blk.handle, float(time), self.numsubstep, blk.dh))
        if wait:
            self.report()

    def step(self, time: float, wait: bool = True):
        """One time step (main.f90:95-107): with wait, the library's one-call
        step (fsil_fsi_step); without, the calls one by one with a coupling
        call that reads nothing back."""
        blk = self.block
        if wait:
            it = self._it
            dm = self._dm
            _lib.check(self.lib.fsil_fsi_step(blk.handle, float(time), self.numsubstep, _lib.iptr(it), _lib.dptr(dm)))
            blk.blktime = float(time)
            self.iters, self.dmax = int(it[0]), float(dm[0])
            _lib.check(self.lib.fsil_ibm_rounds(blk.handle, _lib.iptr(it)))
            self.rounds = int(it[0])
            self._stale = True
            return
        blk.update_volume_force(time)
        self.ibm_fem(time, wait)
        blk.collision_streaming_bc()


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

    def ibm_fem(self, time: float):
        """IBM_FEM (LBMBlockComm.f90:320-338) at block time ``time``."""
        blk = self.block
        self.interaction_force(blk.dh)
        dt_solid = blk.dh / float(self.numsubstep)
        for isub in range(1, self.numsubstep + 1):
            for b in self.bodies:
                b.structure(time, isub, blk.dh, dt_solid)

    def step(self, time: float):
        """One time step at block time ``time`` (main.f90:95-107)."""
        blk = self.block
        blk.update_volume_force(time)
        self.ibm_fem(time)
        blk.collision_streaming_bc()
