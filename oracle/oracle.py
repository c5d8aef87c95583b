"""CPU oracle of the FSILBM3D fluid/coupling step (test infrastructure only).

Wraps the C restatement in oracle/fsil_oracle.c (built by oracle/Makefile into
oracle/libfsil_oracle.so) and composes one time step in the order of
tree_collision_streaming_IBM_FEM (LBMBlockComm.f90:279-318):
update_volume_force -> calculate_macro_quantities -> ResetVolumeForce ->
IBM (FSInteraction_force, Solidbody.f90:589-602, 869-918) -> add_volume_force
-> collision -> halfwayBCset -> streaming -> set_boundary_conditions.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfsil_oracle.so")
PI = 3.141592653589793

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_pf = C.POINTER(C.c_float)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_tables.argtypes = [_pi, _pi, _pd]
        L.orc_macro.argtypes = [_pd, C.c_int, C.c_int, C.c_int, _pd, C.c_double, _pd, _pd]
        L.orc_collision.argtypes = [_pd, _pd, _pd, _pd, C.c_size_t, C.c_int, C.c_double, C.c_double, _pd, _pd, C.c_double]
        L.orc_collision_les.argtypes = [_pd, _pd, _pd, _pd, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, _pd]
        L.orc_streaming.argtypes = [_pd, _pd, C.c_int, C.c_int, C.c_int]
        L.orc_bc.argtypes = [_pd, _pd, _pd, C.c_int, C.c_int, C.c_int, _pd, _pi, _pd, C.c_int, C.c_double, C.POINTER(_pd)]
        L.orc_halfway_save.argtypes = [_pd, C.c_int, C.c_int, C.c_int, _pi, C.POINTER(_pd)]
        L.orc_peskin.argtypes = [C.c_double]
        L.orc_peskin.restype = C.c_double
        L.orc_stencil.argtypes = [C.c_int, _pd, C.c_int, C.c_int, C.c_int, _pd, _pi, _pi, _pf]
        L.orc_stencil.restype = C.c_int
        L.orc_penalty.argtypes = [C.c_int, _pi, _pf, _pd, _pd, _pd, _pd, C.c_int, C.c_int, C.c_int,
                                  C.c_double, C.c_double, C.c_double, _pd]
        L.orc_spread.argtypes = [C.c_int, _pi, _pf, _pd, _pd, C.c_int, C.c_int, C.c_int, C.c_double]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_pd)


def _i(a):
    return a.ctypes.data_as(_pi)


def _f(a):
    return a.ctypes.data_as(_pf)


def tables():
    ee = np.zeros(57, np.int32)
    op = np.zeros(19, np.int32)
    wt = np.zeros(19)
    lib().orc_tables(_i(ee), _i(op), _d(wt))
    return ee.reshape(19, 3), op, wt


def peskin(x: float) -> float:
    return lib().orc_peskin(float(x))


def macro(f: np.ndarray, VF, dh: float):
    q, xd, yd, zd = f.shape
    den = np.empty((xd, yd, zd))
    u = np.empty((3, xd, yd, zd))
    vf = np.array(VF, dtype=np.float64)
    lib().orc_macro(_d(f), xd, yd, zd, _d(vf), dh, _d(den), _d(u))
    return den, u


def stencil(Exyz: np.ndarray, dims, geom4, bc):
    M = len(Exyz)
    Ei = np.zeros(12 * M, np.int32)
    Ew = np.zeros(12 * M, np.float32)
    E = np.ascontiguousarray(Exyz, dtype=np.float64)
    rc = lib().orc_stencil(M, _d(E), dims[0], dims[1], dims[2], _d(np.array(geom4, dtype=np.float64)),
                           _i(np.array(bc, dtype=np.int32)), _i(Ei), _f(Ew))
    if rc != 0:
        raise RuntimeError(f"stencil out of bounds at marker {-rc}")
    return Ei, Ew


class OracleBlock:
    """Single fluid block stepped on the CPU with the reference's algorithm."""

    def __init__(self, dims, dh, origin, bc, model, Omega, Omega2=0.0, Mc=None, Mf=None, *,
                 denIn=1.0, uvwIn=(0, 0, 0), shearRateIn=(0, 0, 0), velocityKind=0,
                 volumeForceIn=(0, 0, 0), volumeForceAmp=0.0, volumeForceFreq=0.0, volumeForcePhi=0.0,
                 tau=None, nu=None):
        self.dims = tuple(int(d) for d in dims)
        self.dh = float(dh)
        self.origin = tuple(float(v) for v in origin)
        self.bc = np.array(bc, dtype=np.int32)
        ext = []
        for a in range(3):
            m = self.origin[a] + self.dh * (self.dims[a] - 1)
            if bc[2 * a] == 301:
                m = m + self.dh
            ext.append(m)
        self.geom = np.array([self.dh, *self.origin, *ext], dtype=np.float64)
        self.model = int(model)
        self.Omega, self.Omega2 = float(Omega), float(Omega2)
        z = np.zeros(361)
        self.Mc = np.ascontiguousarray(Mc if Mc is not None else z, dtype=np.float64)
        self.Mf = np.ascontiguousarray(Mf if Mf is not None else z, dtype=np.float64)
        self.flowp = np.array([denIn, *uvwIn, *shearRateIn], dtype=np.float64)
        self.kind = int(velocityKind)
        self.VFin = tuple(float(v) for v in volumeForceIn)
        self.VFamp, self.VFfreq, self.VFphi = float(volumeForceAmp), float(volumeForceFreq), float(volumeForcePhi)
        xd, yd, zd = self.dims
        self.N = xd * yd * zd
        # LES models 11/14/15: per-cell relaxation time tau_all, initialised to
        # tau (calculate_SRT_params, FluidDomain.f90:450-456), and flow%nu
        self.tau, self.nu = tau, nu
        self.tau_all = None
        if self.model in (11, 14, 15):
            if tau is None or nu is None:
                raise ValueError("the LES collision models need tau and nu")
            self.tau_all = np.full(self.N, float(tau))
        # halfway planes, allocated on the first boundary call (FluidDomain.f90:651-654)
        self._hw_alloc = [False] * 6
        self._hw = [None] * 6
        self.den = None
        self.u = None

    def vf(self, t):
        return (self.VFin[0] + self.VFamp * math.sin(2.0 * PI * self.VFfreq * t + self.VFphi / 180.0 * PI),
                self.VFin[1], self.VFin[2])

    def _plane_shape(self, face):
        xd, yd, zd = self.dims
        return [(yd, zd), (yd, zd), (xd, zd), (xd, zd), (xd, yd), (xd, yd)][face]

    def _hw_ptrs(self):
        arr = (_pd * 6)()
        for face in range(6):
            arr[face] = _d(self._hw[face]) if self._hw[face] is not None else None
        return arr

    def set_boundary_conditions(self, f, den, u, blktime):
        hw = None
        first = []
        for face in range(6):
            if self.bc[face] in (203, 204) and not self._hw_alloc[face]:
                self._hw_alloc[face] = True
                first.append(face)
        arr = self._hw_ptrs()
        for face in first:           # allocated now, applied from the next call
            arr[face] = None
        lib().orc_bc(_d(f), _d(den), _d(u), *self.dims, _d(self.geom), _i(self.bc), _d(self.flowp), self.kind,
                     float(blktime), arr)
        for face in first:
            a, b = self._plane_shape(face)
            self._hw[face] = np.zeros(19 * a * b)

    def initial_fields(self, t0=0.0):
        """den/uuu as initialise_ leaves them (FluidDomain.f90:533-545)."""
        xd, yd, zd = self.dims
        den = np.full((xd, yd, zd), self.flowp[0])
        u = np.zeros((3, xd, yd, zd))
        d, (x0, y0, z0) = self.dh, self.origin
        s = self.flowp[4:7]
        uin = self.flowp[1:4]
        if self.kind == 0:
            X = x0 + d * np.arange(xd)[:, None, None]
            Y = y0 + d * np.arange(yd)[None, :, None]
            Z = z0 + d * np.arange(zd)[None, None, :]
            u[0] = ((uin[0] + 0.0 * s[0]) + Y * s[1]) + Z * s[2] + 0 * X
            u[1] = ((uin[1] + X * s[0]) + 0.0 * s[1]) + Z * s[2] + 0 * Y
            u[2] = ((uin[2] + X * s[0]) + Y * s[1]) + 0.0 * s[2] + 0 * Z
        elif self.kind == 2:
            u[0] = uin[0] + s[0] * math.cos(2 * PI * s[1] * t0 + s[2] / 180.0 * PI)
            u[1] = uin[1]
            u[2] = uin[2]
        return den, u

    def init_boundary(self, f, t0=0.0):
        """start-up call of set_boundary_conditions (main.f90:63)."""
        den, u = self.initial_fields(t0)
        self.set_boundary_conditions(f, den, u, t0)

    def step(self, f: np.ndarray, time: float, bodies: Optional[List[dict]] = None, ibm: Optional[dict] = None):
        """Advance f (shape (19,x,y,z), modified in place) by one step.

        bodies: list of dicts with Exyz (M,3), Evel (M,3), Ea (M,) per body in
        carried order; ibm: dict(ntol, dtol, Uref, denIn).  Returns a dict with
        the IBM outputs (Eforce per body, iterations)."""
        VF = self.vf(time)
        den, u = macro(f, VF, self.dh)
        force = np.zeros((3,) + self.dims)
        out = {"iters": 0, "Eforce": []}
        if bodies:
            sts = []
            for b in bodies:
                Ei, Ew = stencil(b["Exyz"], self.dims, self.geom[:4], self.bc)
                sts.append((Ei, Ew))
            Ef = [np.zeros((len(b["Ea"]), 3)) for b in bodies]
            it, dmax = 0, 1e10
            while it < ibm["ntol"] and dmax > ibm["dtol"]:
                dmax, dsum = 0.0, 0.0
                for b, (Ei, Ew), E in zip(bodies, sts, Ef):
                    tol = np.zeros(1)
                    lib().orc_penalty(len(b["Ea"]), _i(Ei), _f(Ew), _d(np.ascontiguousarray(b["Evel"], dtype=np.float64)),
                                      _d(np.ascontiguousarray(b["Ea"], dtype=np.float64)), _d(E), _d(u), *self.dims,
                                      self.dh, self.dh, ibm["denIn"], _d(tol))
                    dmax = dmax + tol[0]
                    dsum = dsum + float(len(b["Ea"]))
                dmax = dmax / (dsum * ibm["Uref"])
                it += 1
            for b, (Ei, Ew), E in zip(bodies, sts, Ef):
                lib().orc_spread(len(b["Ea"]), _i(Ei), _f(Ew), _d(E), _d(force), *self.dims, self.dh)
            out["iters"] = it
            out["Eforce"] = Ef
            out["stencils"] = sts
        force[0] += VF[0]
        force[1] += VF[1]
        force[2] += VF[2]
        if self.tau_all is not None:
            lib().orc_collision_les(_d(f), _d(den), _d(u), _d(force), *self.dims, self.model, float(self.tau),
                                    float(self.nu), self.dh, _d(self.tau_all))
        else:
            lib().orc_collision(_d(f), _d(den), _d(u), _d(force), self.N, self.model, self.Omega, self.Omega2,
                                _d(self.Mc), _d(self.Mf), self.dh)
        if any(self._hw_alloc):
            lib().orc_halfway_save(_d(f), *self.dims, _i(self.bc), self._hw_ptrs())
        tmp = np.empty(self.N)
        lib().orc_streaming(_d(f), _d(tmp), *self.dims)
        self.set_boundary_conditions(f, den, u, time)
        self.den, self.u = den, u
        return out


def lod_flow(Eforce: np.ndarray, Exyz: np.ndarray, marker_elem: np.ndarray, elem_nodes: np.ndarray,
             elem_xc: np.ndarray, nnodes: int) -> np.ndarray:
    """lodFlow of one body, serial marker order (Solidbody.f90:933-967)."""
    lod = [0.0] * (6 * nnodes)
    for m in range(len(Eforce)):
        e = int(marker_elem[m])
        n0, n1 = int(elem_nodes[e][0]), int(elem_nodes[e][1])
        F = [float(v) for v in Eforce[m]]
        rr = [float(Exyz[m][k]) - float(elem_xc[e][k]) for k in range(3)]
        Mm = [rr[1] * F[2] - rr[2] * F[1], rr[2] * F[0] - rr[0] * F[2], rr[0] * F[1] - rr[1] * F[0]]
        for k in range(3):
            lod[6 * n0 + k] = lod[6 * n0 + k] + 0.5 * F[k]
        for k in range(3):
            lod[6 * n1 + k] = lod[6 * n1 + k] + 0.5 * F[k]
        for k in range(3):
            lod[6 * n0 + 3 + k] = lod[6 * n0 + 3 + k] + 0.5 * Mm[k]
        for k in range(3):
            lod[6 * n1 + 3 + k] = lod[6 * n1 + 3 + k] + 0.5 * Mm[k]
    return np.array(lod)


class OracleBeam:
    """A plate body on the C restatement of the beam solver (fsil_oracle_beam.c)."""

    _sig_done = False

    @classmethod
    def _sigs(cls):
        if cls._sig_done:
            return
        L = lib()
        L.orc_beam_create.restype = C.c_void_p
        L.orc_beam_create.argtypes = [C.c_int, C.c_int, _pd, _pi, _pi, C.c_int, _pi, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, _pd, _pd, _pd, _pd, _pd, _pd, _pd, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_int]
        L.orc_beam_free.argtypes = [C.c_void_p]
        L.orc_beam_structure.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double]
        L.orc_beam_structure.restype = C.c_int
        L.orc_beam_markers.argtypes = [C.c_void_p, C.c_double, C.c_double, _pd, _pd, _pd]
        L.orc_beam_element_centres.argtypes = [C.c_void_p, _pd]
        L.orc_beam_nmarkers.argtypes = [C.c_void_p]
        L.orc_beam_nmarkers.restype = C.c_int
        L.orc_beam_set_lodflow.argtypes = [C.c_void_p, _pd]
        L.orc_beam_state.argtypes = [C.c_void_p, _pd, _pd, _pd, _pd]
        L.orc_beam_info.argtypes = [C.c_void_p, _pi, _pi, _pd]
        L.orc_beam_asfac.argtypes = [C.c_void_p]
        L.orc_beam_asfac.restype = C.c_double
        L.orc_beam_marker_elem.argtypes = [C.c_void_p, _pi, _pi]
        L.orc_beam_solid_info.argtypes = [C.c_void_p, _pd, _pd, _pd]
        L.orc_beam_surface_markers.argtypes = [C.c_void_p, C.c_int, _pd, _pd, _pd]
        L.orc_surface_geometry.argtypes = [C.c_int, _pi, _pd, C.c_double, C.c_double, _pd, _pd]
        cls._sig_done = True

    def __init__(self, nodes8, elements, constraints, *, iBodyModel=2, isMotionGiven=(1, 1, 1, 1, 1, 1), denR=1.0,
                 KB=0.1, KS=1000.0, psR=0.3, Freq=0.0, XYZo=(0, 0, 0), XYZAmpl=(0, 0, 0), XYZPhi=(0, 0, 0),
                 AoAo=(0, 0, 0), AoAAmpl=(0, 0, 0), AoAPhi=(0, 0, 0), initXYZVel=(0, 0, 0), Lref=1.0, Uref=1.0,
                 denIn=1.0, dampK=0.0, dampM=0.0, GeoGamma=1.0, NewmarkGamma=0.5, NewmarkBeta=0.25, dtolFEM=1e-10,
                 ntolFEM=20, IBPenaltyAlpha=1.0):
        self._sigs()
        nodes8 = np.ascontiguousarray(nodes8, dtype=np.float64)
        el = np.ascontiguousarray(elements, dtype=np.int32)
        cons = np.ascontiguousarray(constraints, dtype=np.int32)
        v = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.nND, self.nEL = len(nodes8), len(el)
        self.h = lib().orc_beam_create(self.nND, self.nEL, _d(nodes8), _i(el), _i(cons), int(iBodyModel),
                                       _i(np.array(isMotionGiven, dtype=np.int32)), denR, KB, KS, psR, Freq,
                                       _d(v(XYZo)), _d(v(XYZAmpl)), _d(v(XYZPhi)), _d(v(AoAo)), _d(v(AoAAmpl)),
                                       _d(v(AoAPhi)), _d(v(initXYZVel)), Lref, Uref, denIn, dampK, dampM, GeoGamma,
                                       NewmarkGamma, NewmarkBeta, dtolFEM, int(ntolFEM))
        self.iBodyModel = int(iBodyModel)
        self.alpha, self.denIn = float(IBPenaltyAlpha), float(denIn)
        self.M = lib().orc_beam_nmarkers(self.h)
        self.marker_elem = np.zeros(self.M, np.int32)
        self.elem_nodes = np.zeros((self.nEL, 2), np.int32)
        lib().orc_beam_marker_elem(self.h, _i(self.marker_elem), _i(self.elem_nodes))
        self.Asfac = lib().orc_beam_asfac(self.h)
        self.v_move = int(sum(abs(x) for x in initXYZVel) > 1e-5 or sum(abs(x) for x in XYZAmpl) > 1e-5
                          or sum(abs(x) for x in AoAAmpl) > 1e-5 or sum(isMotionGiven) < 6)
        self.Exyz = np.zeros((self.M, 3))
        self.Evel = np.zeros((self.M, 3))
        self.Ea = np.zeros(self.M)
        self.update_markers()

    def __del__(self):
        try:
            lib().orc_beam_free(self.h)
        except Exception:
            pass

    def update_markers(self):
        lib().orc_beam_markers(self.h, self.alpha, self.denIn, _d(self.Exyz), _d(self.Evel), _d(self.Ea))

    def element_centres(self):
        xc = np.zeros((self.nEL, 3))
        lib().orc_beam_element_centres(self.h, _d(xc))
        return xc

    def set_lodflow(self, lod):
        lib().orc_beam_set_lodflow(self.h, _d(np.ascontiguousarray(lod, dtype=np.float64)))

    def structure(self, time, isubstep, deltat, subdeltat):
        rc = lib().orc_beam_structure(self.h, float(time), int(isubstep), float(deltat), float(subdeltat))
        if rc:
            raise RuntimeError("CG solver breakdown")

    def state(self):
        g = 6 * self.nND
        out = [np.zeros(g) for _ in range(4)]
        lib().orc_beam_state(self.h, *[_d(a) for a in out])
        return [a.reshape(self.nND, 6) for a in out]

    def surface_markers(self, E0):
        """Markers of a surface body on this (two-node) beam's rigid kinematics."""
        E0 = np.ascontiguousarray(E0, dtype=np.float64)
        n = len(E0)
        X, V = np.zeros((n, 3)), np.zeros((n, 3))
        lib().orc_beam_surface_markers(self.h, n, _d(E0), _d(X), _d(V))
        return X, V

    def solid_info(self, XYZo, refs):
        """One DatInfo row per file of Beam_write_solid_info (SolidSolver.f90:1729-1799)
        as a dict name -> values; refs = (Lref, Uref, Aref, Fref, Pref, Eref)."""
        rows = np.zeros(78)
        lib().orc_beam_solid_info(self.h, _d(np.ascontiguousarray(XYZo, dtype=np.float64)),
                                  _d(np.ascontiguousarray(refs, dtype=np.float64)), _d(rows))
        return split_info_rows(rows)

    def info(self):
        a = np.zeros(1, np.int32)
        b = np.zeros(1, np.int32)
        c = np.zeros(1)
        lib().orc_beam_info(self.h, _i(a), _i(b), _d(c))
        return int(a[0]), int(b[0]), float(c[0])


INFO_WIDTHS = (("firstNode", 15), ("lastNode", 15), ("centerNode", 15), ("nodeAverage", 12), ("forces", 6),
               ("power", 7), ("energy", 8))


def split_info_rows(rows):
    """The 78 values of one info step as DatInfo file name -> row."""
    out, o = {}, 0
    for name, w in INFO_WIDTHS:
        out[name] = np.asarray(rows[o:o + w])
        o += w
    return out


def reference_scales(UrefType, uvwIn, Uref, Lref, TrefType, Tref, denIn, Asfac):
    """Lref, Uref, Aref, Fref, Pref, Eref as calculate_reference_params derives
    them (Solidbody.f90:219-284), for the reference-type choices the tests use
    (LrefType 1; UrefType 0 or >= 8; TrefType 0 or >= 2)."""
    U = abs(uvwIn[0]) if UrefType == 0 else Uref
    T = Lref / U if TrefType == 0 else Tref
    A = U / T
    q = (0.5 * denIn) * (U * U)
    F = q * Asfac
    E = (q * Asfac) * Lref
    P = (q * Asfac) * U
    return (Lref, U, A, F, P, E)


def datflow_bytes(f, VF, dh, denIn, Uref, offset, block_id, origin):
    """The reference's instantaneous DatFlow file of one block (write_flow_ with
    outputtype 1, FluidDomain.f90:1628-1720): header nxe, nye, nze, ID (int32)
    and xmin, ymin, zmin, dh (real8) of the output window (offsetOutput cells
    left out on every face), then real(4) p = Cs2 (den - denIn) and u, v, w /
    Uref over the window, z fastest; den/uuu as calculate_macro_quantities_
    leaves them after the step (main.f90:107)."""
    den, u = macro(f, VF, dh)
    xd, yd, zd = f.shape[1:]
    o = int(offset)
    sl = (slice(o, xd - o), slice(o, yd - o), slice(o, zd - o))
    cs2 = 1.0 / 3.0
    inv = 1.0 / Uref
    p = (cs2 * (den[sl] - denIn)).astype(np.float32)
    uu = [(u[k][sl] * inv).astype(np.float32) for k in range(3)]
    nx, ny, nz = p.shape
    head = np.array([nx, ny, nz, block_id], "<i4").tobytes()
    head += np.array([origin[0] + o * dh, origin[1] + o * dh, origin[2] + o * dh, dh], "<f8").tobytes()
    return head + b"".join(np.ascontiguousarray(a).astype("<f4").tobytes() for a in [p] + uu)


def surface_geometry(tri, xyz, alpha, denIn):
    """Incenters and area weights of a triangulated surface (1-based tri)."""
    lib()
    OracleBeam._sigs()
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = len(tri)
    E0, Ea = np.zeros((n, 3)), np.zeros(n)
    lib().orc_surface_geometry(n, _i(tri), _d(xyz), float(alpha), float(denIn), _d(E0), _d(Ea))
    return E0, Ea


class OracleSurface:
    """A v_type = -1 body on the oracle: a rigid two-node beam (the one
    SurfacetoBeam_write makes of the first two mesh nodes, Solidbody.f90:1130-1170)
    carrying one marker per mesh triangle (Solidbody.f90:648-720)."""

    def __init__(self, beam, tri, xyz, alpha, denIn):
        self.beam = beam
        self.E0, self.Ea0 = surface_geometry(tri, xyz, alpha, denIn)
        self.M = len(self.E0)
        self.nND, self.nEL = beam.nND, beam.nEL
        self.iBodyModel, self.v_move = beam.iBodyModel, beam.v_move
        self.marker_elem = np.zeros(self.M, np.int32)
        self.elem_nodes = beam.elem_nodes
        self.Asfac = beam.Asfac
        self.Ea = self.Ea0.copy()
        self.update_markers()

    def update_markers(self):
        self.Exyz, self.Evel = self.beam.surface_markers(self.E0)

    def element_centres(self):
        return self.beam.element_centres()

    def set_lodflow(self, lod):
        self.beam.set_lodflow(lod)

    def structure(self, *a):
        self.beam.structure(*a)

    def state(self):
        return self.beam.state()

    def info(self):
        return self.beam.info()

    def solid_info(self, XYZo, refs):
        return self.beam.solid_info(XYZo, refs)
