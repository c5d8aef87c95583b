"""Parity cases shared by the golden-fixture script, the oracle tests and the
GPU parity tests.

Initial states are generated from a seed with splitmix64 (64-bit integer
arithmetic, independent of numpy's random streams), so fixtures only store
the reference's output state.
"""
from __future__ import annotations

import dataclasses
import math
import os
import sys
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle.refcase import Case, Group, SubBlock, plate_text  # noqa: E402

WT = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)
EE = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
               [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
               [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]]).T


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0,1): splitmix64 outputs >> 11, times 2^-53."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def equilibrium(rho: np.ndarray, u: np.ndarray) -> np.ndarray:
    """calculate_distribution_funcion (FluidDomain.f90:1827-1834) per cell."""
    uSqr = u[0] * u[0]
    uSqr = uSqr + u[1] * u[1]
    uSqr = uSqr + u[2] * u[2]
    f = np.empty((19,) + rho.shape)
    for q in range(19):
        ux = (u[0] * EE[q, 0] + u[1] * EE[q, 1]) + u[2] * EE[q, 2]
        f[q] = (WT[q] * rho) * (((1.0 + 3.0 * ux) + (4.5 * ux) * ux) - 1.5 * uSqr)
    return f


def seeded_state(dims, seed: int, drho: float = 1e-3, du: float = 0.05, neq: float = 0.0,
                 ubase=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Equilibrium of rho = 1 + U(-drho, drho), u_i = ubase_i + U(-du, du) per
    cell (x -> y -> z -> component order), plus an optional non-equilibrium
    part neq*w_q*U(-1,1)."""
    xd, yd, zd = dims
    n = xd * yd * zd
    r = splitmix64_uniform(seed, 4 * n).reshape(xd, yd, zd, 4)
    rho = 1.0 + drho * (2.0 * r[..., 0] - 1.0)
    u = np.stack([ubase[k] + du * (2.0 * r[..., 1 + k] - 1.0) for k in range(3)])
    f = equilibrium(rho, u)
    if neq:
        p = splitmix64_uniform(seed ^ 0x5DEECE66D, 19 * n).reshape(19, xd, yd, zd)
        f = f + neq * WT[:, None, None, None] * (2.0 * p - 1.0)
    return np.ascontiguousarray(f)


def plate_nodes(n_nodes: int, x_lo: float, lspan: float, rspan: float):
    """Nodes of a straight plate along x (chord), spanwise direction z."""
    return np.array([[x_lo + i, 0.0, 0.0, lspan, rspan, 0.0, 0.0, 1.0] for i in range(n_nodes)])


@dataclasses.dataclass
class Plate:
    n_nodes: int
    x_lo: float
    lspan: float
    rspan: float
    nspan: int
    XYZo: tuple
    XYZAmpl: tuple = (0.0, 0.0, 0.0)
    XYZPhi: tuple = (0.0, 0.0, 0.0)
    AoAo: tuple = (0.0, 0.0, 0.0)
    AoAAmpl: tuple = (0.0, 0.0, 0.0)
    AoAPhi: tuple = (0.0, 0.0, 0.0)
    Freq: float = 0.0
    initXYZVel: tuple = (0.0, 0.0, 0.0)
    iBodyModel: int = 1
    denR: float = 1.0
    KB: float = 0.1
    KS: float = 1000.0
    psR: float = 0.3
    isMotionGiven: tuple = (1, 1, 1, 1, 1, 1)

    def nodes(self):
        return plate_nodes(self.n_nodes, self.x_lo, self.lspan, self.rspan)

    def elements(self):
        return [(i + 1, i + 2, self.nspan) for i in range(self.n_nodes - 1)]

    def group(self) -> Group:
        txt = plate_text([row[:3] for row in self.nodes()], self.lspan, self.rspan, (0.0, 0.0, 1.0), self.nspan)
        return Group(plate=txt, iBodyModel=self.iBodyModel, denR=self.denR, KB=self.KB, KS=self.KS, psR=self.psR,
                     isMotionGiven=self.isMotionGiven, firstXYZ=self.XYZo, XYZAmpl=self.XYZAmpl, XYZPhi=self.XYZPhi,
                     AoAo=self.AoAo, AoAAmpl=self.AoAAmpl, AoAPhi=self.AoAPhi, freq=self.Freq,
                     initXYZVel=self.initXYZVel)


def icosphere(radius: float, subdivisions: int = 1):
    """Vertices and triangles (1-based) of a subdivided icosahedron."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, np.float64) / math.sqrt(sum(c * c for c in p)) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid, nf = {}, []

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v) * radius, np.array(f, np.int32) + 1


def gmsh_text(xyz, tri) -> str:
    """A gmsh 2 ASCII mesh of the triangles, in the form Read_gmsh expects."""
    L = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(xyz))]
    L += [f"{i + 1} {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}" for i, p in enumerate(xyz)]
    L += ["$EndNodes", "$Elements", str(len(tri))]
    L += [f"{i + 1} 2 2 0 1 {a} {b} {c}" for i, (a, b, c) in enumerate(tri)]
    L += ["$EndElements"]
    return "\n".join(L) + "\n"


@dataclasses.dataclass
class SurfacePlate(Plate):
    """A rigid surface body (iBodyType -1): a sphere of `radius` triangulated
    as a once-subdivided icosahedron (80 triangles), centred at XYZo."""
    radius: float = 3.0

    def mesh(self) -> str:
        return gmsh_text(*icosphere(self.radius))

    def group(self) -> Group:
        g = super().group()
        return dataclasses.replace(g, plate=self.mesh(), iBodyType=-1)


@dataclasses.dataclass
class ParityCase:
    name: str
    case: Case
    nsteps: int
    seed: int
    drho: float = 1e-3
    du: float = 0.05
    neq: float = 0.0
    ubase: tuple = (0.0, 0.0, 0.0)
    plates: List[Plate] = dataclasses.field(default_factory=list)

    def initial_state(self):
        f = seeded_state(self.case.dims, self.seed, self.drho, self.du, self.neq, self.ubase)
        if not self.case.sons:
            return f
        out = [f]
        for k, sb in enumerate(self.case.sons):
            fs = seeded_state(sb.dims, self.seed * 7919 + k + 1, self.drho, self.du, self.neq, self.ubase)
            # father nodes that coincide with son nodes carry the son's values, as
            # the reference's restart reads them from the finer block
            r = int(round(self.case.dh / sb.dh))
            o = [int(round((sb.origin[a] - self.case.origin[a]) / self.case.dh)) for a in range(3)]
            n = [(sb.dims[a] - 1) // r + 1 for a in range(3)]
            f[:, o[0]:o[0] + n[0], o[1]:o[1] + n[1], o[2]:o[2] + n[2]] = fs[:, ::r, ::r, ::r]
            out.append(fs)
        return out

    def full_case(self) -> Case:
        c = dataclasses.replace(self.case)
        c.groups = [p.group() for p in self.plates]
        return c


def _periodic(**kw):
    base = dict(dims=(12, 8, 6), bc=(301,) * 6, Re=3.0, Uref=0.05, UrefType=9, Lref=1.0, LrefType=1,
                TrefType=0)
    base.update(kw)
    return Case(**base)


def _open(**kw):
    base = dict(dims=(12, 8, 6), Re=3.0, Uref=0.05, UrefType=9, Lref=1.0, LrefType=1, TrefType=0,
                uvwIn=(0.05, 0.0, 0.0))
    base.update(kw)
    return Case(**base)


CASES: List[ParityCase] = [
    ParityCase("periodic_srt", _periodic(), nsteps=3, seed=11, neq=0.02),
    ParityCase("periodic_trt", _periodic(collide=2, params=(0.1875,) + (0.0,) * 9), nsteps=3, seed=12, neq=0.02),
    ParityCase("periodic_mrt", _periodic(collide=3), nsteps=3, seed=13, neq=0.02),
    ParityCase("periodic_srt_force", _periodic(volumeForceIn=(1e-5, -2e-6, 3e-6), volumeForceAmp=2e-5,
                                               volumeForceFreq=0.01, volumeForcePhi=30.0), nsteps=3, seed=14, neq=0.02),
    ParityCase("periodic_mrt_force", _periodic(collide=3, volumeForceIn=(1e-5, 2e-6, -3e-6)), nsteps=2, seed=15, neq=0.02),
    ParityCase("periodic_trt_force", _periodic(collide=2, params=(0.25,) + (0.0,) * 9, volumeForceIn=(-1e-5, 2e-6, 3e-6)),
               nsteps=2, seed=16, neq=0.02),
    ParityCase("bc_inflow_outflow_slip", _open(bc=(101, 103, 302, 302, 302, 302)), nsteps=3, seed=21,
               ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("bc_neq_extrap2_walls", _open(bc=(102, 104, 201, 201, 302, 302), shearRateIn=(0.0, 1e-3, 2e-4)),
               nsteps=3, seed=22, ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("bc_moving_walls", _open(bc=(101, 103, 202, 202, 201, 302), shearRateIn=(1e-3, 2e-3, -1e-3)),
               nsteps=3, seed=23, ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("bc_halfway_walls", _open(bc=(101, 104, 203, 204, 203, 204), shearRateIn=(2e-3, 1e-3, 5e-4)),
               nsteps=3, seed=24, ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("bc_oscillatory_inflow", _open(bc=(101, 102, 302, 302, 301, 301), velocityKind=2,
                                              shearRateIn=(0.01, 0.02, 45.0)), nsteps=3, seed=25,
               ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("bc_mixed_trt_force", _open(bc=(101, 103, 302, 201, 301, 301), collide=2, params=(0.2,) + (0.0,) * 9,
                                           volumeForceIn=(2e-6, 0.0, 1e-6)), nsteps=3, seed=26, ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("ibm_static_plate", _open(dims=(20, 12, 12), bc=(101, 103, 302, 302, 302, 302), ntolLBM=3, dtolLBM=1e-14),
               nsteps=2, seed=31, ubase=(0.05, 0, 0), neq=0.005,
               plates=[Plate(5, -2.0, 2.0, 2.0, 4, XYZo=(8.0, 6.0, 6.0))]),
    ParityCase("ibm_heaving_plate", _open(dims=(20, 12, 12), bc=(101, 103, 302, 302, 302, 302), ntolLBM=3, dtolLBM=1e-14),
               nsteps=3, seed=32, ubase=(0.05, 0, 0), neq=0.005,
               plates=[Plate(5, -2.0, 2.0, 2.0, 4, XYZo=(8.0, 6.0, 6.0), XYZAmpl=(0.0, 1.5, 0.0), XYZPhi=(0.0, 30.0, 0.0),
                             Freq=0.02)]),
    ParityCase("ibm_pitching_two_plates", _open(dims=(24, 12, 14), bc=(101, 103, 302, 302, 301, 301), ntolLBM=4,
                                                dtolLBM=1e-14),
               nsteps=2, seed=33, ubase=(0.05, 0, 0), neq=0.005,
               plates=[Plate(4, -1.5, 2.0, 2.0, 3, XYZo=(7.0, 6.0, 7.0), AoAo=(0.0, 0.0, 10.0), AoAAmpl=(0.0, 0.0, 15.0),
                             Freq=0.03, XYZAmpl=(0.0, 0.5, 0.0)),
                       Plate(4, -1.5, 1.5, 1.5, 3, XYZo=(15.0, 6.5, 7.0), initXYZVel=(0.01, 0.0, 0.0))]),
    ParityCase("fem_clamped_heaving_plate", _open(dims=(20, 12, 12), bc=(101, 103, 302, 302, 302, 302), ntolLBM=3,
                                                  dtolLBM=1e-14, numsubstep=2, uvwIn=(0.05, 0.0, 0.0)),
               nsteps=3, seed=34, ubase=(0.05, 0, 0), neq=0.005,
               plates=[Plate(5, -2.0, 2.0, 2.0, 4, XYZo=(8.0, 6.0, 6.0), XYZAmpl=(0.0, 1.0, 0.0), Freq=0.03,
                             iBodyModel=2, denR=1.0, KB=0.1, KS=1000.0)]),
    ParityCase("fem_flexible_plate", _open(dims=(20, 12, 12), bc=(101, 103, 302, 302, 302, 302), ntolLBM=10,
                                           dtolLBM=1e-6, numsubstep=4, uvwIn=(0.05, 0.0, 0.0), Lref=8.0, Re=20.0,
                                           UrefType=0),
               nsteps=5, seed=35, ubase=(0.05, 0, 0), neq=0.002,
               plates=[Plate(5, -2.0, 2.0, 2.0, 4, XYZo=(8.0, 6.0, 6.0), XYZAmpl=(0.0, 1.0, 0.0), Freq=0.01,
                             iBodyModel=2, denR=1.0, KB=0.1, KS=1000.0)]),
    ParityCase("les_smagorinsky_periodic", _periodic(collide=11, volumeForceIn=(1e-5, -2e-6, 3e-6)), nsteps=3, seed=51,
               neq=0.05),
    ParityCase("les_wale_open", _open(bc=(101, 103, 302, 201, 301, 301), collide=14, volumeForceIn=(2e-6, 0.0, 1e-6)),
               nsteps=3, seed=52, ubase=(0.05, 0, 0), neq=0.05, du=0.03),
    ParityCase("les_vreman_walls", _open(bc=(101, 104, 201, 202, 302, 203), collide=15, shearRateIn=(1e-3, 2e-3, -1e-3)),
               nsteps=3, seed=53, ubase=(0.05, 0, 0), neq=0.05, du=0.03),
    ParityCase("les_wale_heaving_plate", _open(dims=(20, 12, 12), bc=(101, 103, 302, 302, 302, 302), collide=14,
                                               ntolLBM=3, dtolLBM=1e-14),
               nsteps=3, seed=54, ubase=(0.05, 0, 0), neq=0.02,
               plates=[Plate(5, -2.0, 2.0, 2.0, 4, XYZo=(8.0, 6.0, 6.0), XYZAmpl=(0.0, 1.5, 0.0), XYZPhi=(0.0, 30.0, 0.0),
                             Freq=0.02)]),
    ParityCase("ibm_surface_sphere", _open(dims=(24, 16, 16), bc=(101, 103, 302, 302, 302, 302), ntolLBM=3,
                                           dtolLBM=1e-14),
               nsteps=3, seed=71, ubase=(0.05, 0, 0), neq=0.005,
               plates=[SurfacePlate(2, 0.0, 1.0, 1.0, 1, XYZo=(10.0, 8.0, 8.0), XYZAmpl=(0.0, 1.0, 0.0),
                                    XYZPhi=(0.0, 20.0, 0.0), AoAo=(0.0, 0.0, 10.0), AoAAmpl=(0.0, 0.0, 8.0),
                                    Freq=0.02, radius=3.0)]),
    ParityCase("output_flow_continue", _open(bc=(101, 103, 302, 201, 301, 301), volumeForceIn=(2e-6, 0.0, 1e-6),
                                             flow_every=3, offsetOutput=1, outputtype=1), nsteps=3, seed=61,
               ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("twoblock_linear", _open(dims=(16, 12, 12), bc=(101, 103, 302, 302, 302, 302),
                                        sons=[SubBlock(dims=(13, 11, 11), dh=0.5, origin=(4.0, 3.0, 3.0))]),
               nsteps=3, seed=41, ubase=(0.05, 0, 0), neq=0.01),
    ParityCase("twoblock_quartic_periodic", _periodic(dims=(16, 12, 12), interpolateScheme=2,
                                                      volumeForceIn=(1e-5, 0.0, 0.0),
                                                      sons=[SubBlock(dims=(13, 13, 11), dh=0.5, origin=(5.0, 2.0, 3.0))]),
               nsteps=3, seed=42, neq=0.01),
]

BY_NAME: Dict[str, ParityCase] = {c.name: c for c in CASES}
SINGLE = [c for c in CASES if not c.case.sons]      # one fluid block
TREE = [c for c in CASES if c.case.sons]            # a root block with refined sons

# DatInfo files the reference writes per body group at every info step
# (Beam_write_solid_info, SolidSolver.f90:1729-1799): columns as written there.
INFO_FILES = ("forces", "firstNode", "lastNode", "centerNode", "nodeAverage", "power", "energy")


def heaving_plate_128(nsteps: int = 2000) -> ParityCase:
    """The rigid heaving-plate workload at its full size: 128x64x64, uniform
    inflow 0.05 (101), outflow (103), slip faces (302), Re 100 on a 16-cell
    chord; one rigid plate of 17 chord nodes x 16 span samples (256 markers)
    centred at x = 40, heave 0.5 L at 0.25 U/L; 3 direct-forcing rounds.  The
    state starts as the uniform inflow equilibrium."""
    case = Case(dims=(128, 64, 64), bc=(101, 103, 302, 302, 302, 302), Re=100.0, uvwIn=(0.05, 0.0, 0.0),
                UrefType=0, Uref=0.05, LrefType=1, Lref=16.0, TrefType=0, ntolLBM=3, dtolLBM=1e-14)
    plate = Plate(17, -8.0, 8.0, 8.0, 16, XYZo=(40.0, 32.0, 32.0), XYZAmpl=(0.0, 8.0, 0.0), Freq=7.8125e-4)
    return ParityCase("c1_heaving_plate", case, nsteps=nsteps, seed=41, ubase=(0.05, 0, 0), neq=0.0,
                      drho=0.0, du=0.0, plates=[plate])


def flexible_plate_384(nsteps: int = 20) -> ParityCase:
    """The flexible-plate workload at its full size: 384x192x192, inflow 0.05,
    Re 500 on L = 48; one elastic plate of 49 chord nodes x 48 span samples
    (2304 markers, mass ratio 1, KB 0.1, KS 1000), leading edge clamped at
    x = 96 and heaving 0.3 L at 0.2 U/L; numsubstep 4, direct forcing to 1e-6
    with at most 10 rounds.  The state starts as the uniform inflow equilibrium."""
    case = Case(dims=(384, 192, 192), bc=(101, 103, 302, 302, 302, 302), Re=500.0, uvwIn=(0.05, 0.0, 0.0),
                UrefType=0, Uref=0.05, LrefType=1, Lref=48.0, TrefType=0, ntolLBM=10, dtolLBM=1e-6,
                numsubstep=4, dtolFEM=1e-10, ntolFEM=20)
    plate = Plate(49, -24.0, 24.0, 24.0, 48, XYZo=(120.0, 96.0, 96.0), XYZAmpl=(0.0, 14.4, 0.0),
                  Freq=0.2 * 0.05 / 48.0, iBodyModel=2, denR=1.0, KB=0.1, KS=1000.0)
    return ParityCase("c3_flexible_plate", case, nsteps=nsteps, seed=43, ubase=(0.05, 0, 0), neq=0.0,
                      drho=0.0, du=0.0, plates=[plate])


@dataclasses.dataclass
class LongCase:
    """A full-size workload run by the reference for many steps.  Its fixture
    keeps every DatInfo row of every step and a few small windows of the
    final state (the whole state is 80 MB - 2 GB)."""
    pc: ParityCase
    windows: List[tuple]        # (x0, x1, y0, y1, z0, z1), half-open, 0-based
    npsize: int = 8

    @property
    def name(self) -> str:
        return self.pc.name

    def uniform_state(self) -> np.ndarray:
        from fsilbm3d_amd.params import equilibrium_point
        f19 = equilibrium_point(1.0, self.pc.ubase)
        f = np.empty((19,) + tuple(self.pc.case.dims))
        for q in range(19):
            f[q].fill(f19[q])
        return f

    def window(self, f: np.ndarray, k: int) -> np.ndarray:
        x0, x1, y0, y1, z0, z1 = self.windows[k]
        return np.ascontiguousarray(f[:, x0:x1, y0:y1, z0:z1])


LONG_CASES: List[LongCase] = [
    LongCase(heaving_plate_128(2000), windows=[(36, 44, 28, 36, 28, 36), (30, 34, 20, 44, 30, 34),
                                                (0, 2, 0, 8, 0, 8), (124, 128, 28, 36, 28, 36)]),
    LongCase(flexible_plate_384(20), windows=[(116, 124, 92, 100, 92, 100), (94, 98, 90, 102, 94, 98),
                                               (142, 146, 90, 102, 94, 98), (0, 2, 0, 8, 0, 8),
                                               (380, 384, 92, 100, 92, 100)]),
]
LONG_BY_NAME: Dict[str, LongCase] = {c.name: c for c in LONG_CASES}


def info_rows(fx: dict, name: str, group: int = 1) -> np.ndarray:
    """Rows of DatInfo/Group<g>_<name>.dat kept in a fixture."""
    return fx[f"{name}_g{group}"]


def fixture_path(name: str) -> str:
    return os.path.join(HERE, f"{name}.npz")


def load_fixture(name: str):
    with np.load(fixture_path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
