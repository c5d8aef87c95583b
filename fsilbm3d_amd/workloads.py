"""Benchmark workloads of the hot path, stated as concrete inputs.

stencil:          D3Q19 512x256x256, all faces periodic (301), tau = 0.55 (Re 3 on Lref 1, Uref 0.05),
                  equilibrium from a seeded density 1+U(-1e-3,1e-3) and velocity U(-0.05,0.05).
flexible_plate:   D3Q19 384x192x192, uniform inflow U = 0.05 (101) at x-min, outflow (103) at x-max,
                  slip (302) lateral faces; Re = 500 on L = 48 cells (nu = 0.0048, tau = 0.5144).  One
                  flexible plate of chord L: 49 chord nodes (48 beam elements) x Nspan 48 span markers
                  = 2304 markers at one-cell spacing (the 49x49-node plate on the reference's
                  co-rotational beam), mass ratio 1, KB = 0.1, KS = 1000, leading-edge node clamped
                  and heaving with amplitude 0.3 L at 0.2 U/L; numsubstep 4, multi-direct forcing to
                  1e-6 (at most 10 iterations).
school:           D3Q19 1024x384x384 (151 M cells), the flexible-plate flow and solver settings, 16 such
                  plates on a 4x4 diamond lattice (2 L streamwise, 1.5 L spanwise, odd rows shifted by 1 L),
                  leading edges from x = 256, rows centred in z at y = 192, each plate its own group with
                  heave phase U(0, 360) degrees from numpy's default generator seeded 20261202
                  (36,864 markers, 16 beam bodies).
"""
from __future__ import annotations

import dataclasses
from typing import Tuple

import numpy as np

from . import params as P
from .fsi import PlateSpec


@dataclasses.dataclass
class FlexiblePlate:
    dims: Tuple[int, int, int] = (384, 192, 192)
    U: float = 0.05
    L: float = 48.0
    Re: float = 500.0
    le: Tuple[float, float, float] = (96.0, 96.0, 96.0)  # leading edge at rest (mean heave position)
    amp: float = 0.3                                      # heave amplitude / L
    freq: float = 0.2                                     # heave frequency * L / U
    numsubstep: int = 4
    ntolLBM: int = 10
    dtolLBM: float = 1e-6
    KB: float = 0.1
    KS: float = 1000.0
    denR: float = 1.0
    body_model: int = 2                                    # 2 flexible beam, 1 rigid prescribed motion
    bc: Tuple[int, ...] = (101, 103, 302, 302, 302, 302)

    @property
    def nu(self) -> float:
        return P.viscosity(self.U, self.L, self.Re)

    @property
    def n_chord(self) -> int:
        return int(round(self.L)) + 1

    def plate(self) -> PlateSpec:
        n = self.n_chord
        half = 0.5 * self.L
        nodes = np.array([[i - half, 0.0, 0.0, half, half, 0.0, 0.0, 1.0] for i in range(n)], np.float64)
        elements = np.array([(i + 1, i + 2, int(round(self.L))) for i in range(n - 1)], np.int32)
        cons = np.zeros((n, 6), np.int32)
        cons[0] = 1                                        # clamped leading edge
        xyzo = (self.le[0] + half, self.le[1], self.le[2])
        return PlateSpec(nodes8=nodes, elements=elements, constraints=cons, iBodyModel=self.body_model,
                         isMotionGiven=(1, 1, 1, 1, 1, 1), denR=self.denR, KB=self.KB, KS=self.KS, psR=0.3,
                         Freq=self.freq * self.U / self.L, XYZo=xyzo, XYZAmpl=(0.0, self.amp * self.L, 0.0))

    def solver_kwargs(self) -> dict:
        return dict(ntolLBM=self.ntolLBM, dtolLBM=self.dtolLBM, Uref=self.U, denIn=1.0, Lref=self.L,
                    numsubstep=self.numsubstep, IBPenaltyAlpha=1.0, dampK=0.0, dampM=0.0, GeoGamma=1.0,
                    NewmarkGamma=0.5, NewmarkBeta=0.25, dtolFEM=1e-10, ntolFEM=20)

    def flow_kwargs(self) -> dict:
        return dict(denIn=1.0, uvwIn=(self.U, 0.0, 0.0))

    def equilibrium(self) -> np.ndarray:
        """The 19 populations of the uniform inflow state (initialise_flow,
        FluidDomain.f90:524-544); every cell holds them at the start."""
        return np.asarray(P.equilibrium_point(1.0, (self.U, 0.0, 0.0)), np.float64)

    def initial_state(self) -> np.ndarray:
        """Uniform-flow equilibrium, as initialise_flow leaves the block
        (FluidDomain.f90:524-544); 19 constants broadcast over the block."""
        f = np.empty((19,) + tuple(self.dims), np.float64)
        feq = self.equilibrium()
        for q in range(19):
            f[q].fill(feq[q])
        return f


def flexible_plate(dims=(384, 192, 192)) -> FlexiblePlate:
    return FlexiblePlate(dims=tuple(dims))


@dataclasses.dataclass
class School(FlexiblePlate):
    """Sixteen flexible plates of the flexible-plate case in one block."""
    dims: Tuple[int, int, int] = (1024, 384, 384)
    rows: int = 4
    cols: int = 4
    x0: float = 256.0
    seed: int = 20261202

    def leading_edges(self):
        dx, dz = 2.0 * self.L, 1.5 * self.L
        z0 = 0.5 * (self.dims[2] - (self.rows - 1) * dz)
        out = []
        for j in range(self.rows):
            for i in range(self.cols):
                out.append((self.x0 + i * dx + (0.5 * dx if j % 2 else 0.0), 0.5 * self.dims[1], z0 + j * dz))
        return out

    def phases_deg(self) -> np.ndarray:
        return np.random.default_rng(self.seed).uniform(0.0, 360.0, self.rows * self.cols)

    def plates(self):
        out = []
        for le, ph in zip(self.leading_edges(), self.phases_deg()):
            one = FlexiblePlate(dims=self.dims, U=self.U, L=self.L, Re=self.Re, le=le, amp=self.amp, freq=self.freq,
                                numsubstep=self.numsubstep, ntolLBM=self.ntolLBM, dtolLBM=self.dtolLBM, KB=self.KB,
                                KS=self.KS, denR=self.denR, body_model=self.body_model, bc=self.bc).plate()
            out.append(dataclasses.replace(one, XYZPhi=(0.0, float(ph), 0.0)))
        return out


def school(dims=(1024, 384, 384)) -> School:
    return School(dims=tuple(dims))
