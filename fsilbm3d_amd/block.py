"""Host mirror of the reference's LBMBlock type (FluidDomain.f90:17-56) over
the C-ABI.  Method names follow the reference's type-bound procedures; the
device owns the distribution functions, the host only sees them on
download.

Array layout on the host side is the reference's: fIn(z,y,x,0:18) in Fortran
order, i.e. a C-contiguous numpy array of shape (19, xDim, yDim, zDim).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import _lib
from . import params as P


class LBMBlock:
    def __init__(self, dims: Sequence[int], dh: float = 1.0, origin: Sequence[float] = (0.0, 0.0, 0.0),
                 bc: Sequence[int] = (301,) * 6, iCollidModel: int = 1, nu: float = 1.0 / 60.0,
                 params: Sequence[float] = (0.0,) * 10, precision: int = 0, block_id: int = 1,
                 device: int = 0):
        lib = _lib.init(device)
        self.lib = lib
        self.dims = tuple(int(d) for d in dims)
        self.dh = float(dh)
        self.origin = tuple(float(v) for v in origin)
        self.bc = tuple(int(b) for b in bc)
        self.iCollidModel = int(iCollidModel)
        self.precision = int(precision)
        (self.xmax, self.ymax, self.zmax), self.periodic = P.block_extent(self.dims, self.dh, self.origin, self.bc)
        self.tau, self.Omega = P.srt(nu, self.dh)
        self.Omega2 = self.Omega
        if self.iCollidModel == 2:
            self.Omega2 = P.trt_omega2(self.Omega, float(params[0]))
        mc = mf = None
        if self.iCollidModel == 3:
            Mc, Mf = P.mrt_matrices(self.Omega)
            mc, mf = P.fortran_order(Mc), P.fortran_order(Mf)
        dims_a = np.array(self.dims, dtype=np.int32)
        geom = np.array([self.dh, *self.origin, self.xmax, self.ymax, self.zmax], dtype=np.float64)
        bc_a = np.array(self.bc, dtype=np.int32)
        relax = np.array([self.tau, self.Omega, self.Omega2], dtype=np.float64)
        h = np.zeros(1, dtype=np.int32)
        _lib.check(lib.fsil_block_create(block_id, _lib.iptr(dims_a), _lib.dptr(geom), _lib.iptr(bc_a),
                                         self.iCollidModel, _lib.dptr(relax),
                                         _lib.dptr(mc) if mc is not None else None,
                                         _lib.dptr(mf) if mf is not None else None,
                                         self.precision, _lib.iptr(h)))
        self.handle = int(h[0])
        self.blktime = 0.0
        self.nu = float(nu)
        if self.iCollidModel in (11, 14, 15):
            _lib.check(lib.fsil_set_viscosity(self.handle, self.nu))
        self.set_flow()

    @property
    def N(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]

    def close(self):
        if getattr(self, "handle", 0):
            _lib.check(self.lib.fsil_block_destroy(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- flow conditions (FlowCondition.f90:11-26) --------------------------
    def set_flow(self, denIn: float = 1.0, uvwIn=(0.0, 0.0, 0.0), shearRateIn=(0.0, 0.0, 0.0),
                 velocityKind: int = 0, volumeForceIn=(0.0, 0.0, 0.0), volumeForceAmp: float = 0.0,
                 volumeForceFreq: float = 0.0, volumeForcePhi: float = 0.0):
        u = np.array(uvwIn, dtype=np.float64)
        s = np.array(shearRateIn, dtype=np.float64)
        v = np.array(volumeForceIn, dtype=np.float64)
        _lib.check(self.lib.fsil_set_flow(self.handle, float(denIn), _lib.dptr(u), _lib.dptr(s), int(velocityKind),
                                          _lib.dptr(v), float(volumeForceAmp), float(volumeForceFreq),
                                          float(volumeForcePhi)))

    # -- state transfer -----------------------------------------------------
    def upload(self, fIn: np.ndarray):
        f = np.ascontiguousarray(fIn, dtype=np.float64)
        assert f.shape == (19,) + self.dims
        _lib.check(self.lib.fsil_f_upload(self.handle, _lib.dptr(f)))

    def tau_all(self) -> np.ndarray:
        """Per-cell relaxation time tau_all(z,y,x) as (xDim, yDim, zDim)."""
        out = np.empty(self.dims)
        _lib.check(self.lib.fsil_tau_download(self.handle, _lib.dptr(out)))
        return out

    def fill(self, f19) -> None:
        """Every cell set to the same 19 populations (a uniform initial state)."""
        v = np.ascontiguousarray(np.asarray(f19, dtype=np.float64).reshape(19))
        _lib.check(self.lib.fsil_f_fill(self.handle, _lib.dptr(v)))

    def download(self) -> np.ndarray:
        f = np.empty((19,) + self.dims, dtype=np.float64)
        _lib.check(self.lib.fsil_f_download(self.handle, _lib.dptr(f)))
        return f

    # -- type-bound procedures ---------------------------------------------
    def update_volume_force(self, blktime: Optional[float] = None):
        """update_volume_force_ at the block time (FluidDomain.f90:1174-1180)."""
        if blktime is not None:
            self.blktime = float(blktime)
        _lib.check(self.lib.fsil_set_time(self.handle, self.blktime))

    def volume_force(self):
        vf = np.zeros(3)
        _lib.check(self.lib.fsil_get_volume_force(self.handle, _lib.dptr(vf)))
        return vf

    def set_boundary_conditions(self):
        """set_boundary_conditions_ on the stored state, as called once at
        start-up (main.f90:63)."""
        _lib.check(self.lib.fsil_set_boundary_conditions(self.handle))

    def calculate_macro_quantities(self):
        """calculate_macro_quantities_ (FluidDomain.f90:1128-1145) -> (den, uuu);
        den (xDim,yDim,zDim), uuu (3,xDim,yDim,zDim) in Fortran memory order."""
        den = np.empty(self.dims, dtype=np.float64)
        uuu = np.empty((3,) + self.dims, dtype=np.float64)
        _lib.check(self.lib.fsil_macro_download(self.handle, _lib.dptr(den), _lib.dptr(uuu)))
        return den, uuu

    def collision_streaming_bc(self):
        """collision_ + halfwayBCset_ + streaming_ + set_boundary_conditions_
        of one step (LBMBlockComm.f90:293-304), fused on the device."""
        _lib.check(self.lib.fsil_collide_stream_bc(self.handle))

    def steps(self, nsteps: int, t0: float, dt: float):
        """nsteps pure-fluid steps (no immersed bodies), the time accumulated from t0 by dt."""
        _lib.check(self.lib.fsil_steps(self.handle, int(nsteps), float(t0), float(dt)))
        t = float(t0)
        for _ in range(int(nsteps)):
            t = t + dt
        self.blktime = t

    def time_collide(self, nlaunch: int) -> float:
        ms = np.zeros(1)
        _lib.check(self.lib.fsil_time_collide(self.handle, int(nlaunch), _lib.dptr(ms)))
        return float(ms[0])

    def time_steps(self, nsteps: int, t0: float, dt: float) -> float:
        ms = np.zeros(1)
        _lib.check(self.lib.fsil_time_steps(self.handle, int(nsteps), float(t0), float(dt), _lib.dptr(ms)))
        return float(ms[0])


def synchronize():
    _lib.check(_lib.load().fsil_synchronize())
