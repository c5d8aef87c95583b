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


def fixture_path(name: str) -> str:
    return os.path.join(HERE, f"{name}.npz")


def load_fixture(name: str):
    with np.load(fixture_path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
