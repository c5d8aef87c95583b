"""Host-side derived parameters of a fluid block, computed in the reference's
order of operations so the scalars handed to the device are bit-identical.

- block extents: read_fuild_blocks (FluidDomain.f90:93-104)
- reference scales and viscosity: calculate_reference_params (Solidbody.f90:219-284)
- relaxation: calculate_SRT_params / _TRT_params / _MRT_params (FluidDomain.f90:449-522)
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np

Q = 19
EE = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
               [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
               [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]], dtype=np.int64).T
OPPO = [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
WT = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)
CS2 = 1.0 / 3.0
PI = 3.141592653589793

# MRT relaxation rates (ConstParams.f90:28)
S0, S1, S2, S4, S10, S16 = 0.0, 1.19, 1.4, 1.2, 1.4, 1.98

BC_PERIODIC = 301


def block_extent(dims: Sequence[int], dh: float, origin: Sequence[float], bc: Sequence[int]):
    """(xmax, ymax, zmax) and the periodic flags, FluidDomain.f90:93-104."""
    ext, per = [], []
    for a in range(3):
        p = 1 if (bc[2 * a] == BC_PERIODIC or bc[2 * a + 1] == BC_PERIODIC) else 0
        if p and bc[2 * a] != bc[2 * a + 1]:
            raise ValueError(f"Periodic boundaries must appear in pairs: {bc[2*a]} {bc[2*a+1]}")
        m = origin[a] + dh * (dims[a] - 1)
        if p:
            m = m + dh
        ext.append(m)
        per.append(p)
    return tuple(ext), tuple(per)


def viscosity(Uref: float, Lref: float, Re: float) -> float:
    return Uref * Lref / Re


def srt(nu: float, dh: float) -> Tuple[float, float]:
    tau = nu / (dh * CS2) + 0.5
    return tau, 1.0 / tau


def trt_omega2(Omega: float, lam: float) -> float:
    tmp = (lam * 4.0 - 1.0) * Omega + 2.0
    return 2.0 * (2.0 - Omega) / tmp


def _matmul(A, B):
    """Fortran MATMUL with the k-sum in ascending order."""
    n, m, p = len(A), len(B), len(B[0])
    out = [[0.0] * p for _ in range(n)]
    for i in range(n):
        for j in range(p):
            s = 0.0
            for k in range(m):
                s = s + A[i][k] * B[k][j]
            out[i][j] = s
    return out


def mrt_matrices(Omega: float):
    """M_COLLID, M_FORCE (19x19) as calculate_MRT_params builds them.

    Returned as numpy arrays indexed [i, j]; the C-ABI takes them in Fortran
    column-major order (``.T.copy()`` flattening)."""
    M = [[0.0] * Q for _ in range(Q)]
    for I in range(Q):
        e = [int(EE[I, k]) for k in range(3)]
        s2 = e[0] ** 2 + e[1] ** 2 + e[2] ** 2
        M[0][I] = 1.0
        M[1][I] = 19.0 * s2 - 30.0
        M[2][I] = (21.0 * s2 ** 2 - 53.0 * s2 + 24.0) / 2.0
        M[3][I] = float(e[0])
        M[5][I] = float(e[1])
        M[7][I] = float(e[2])
        M[4][I] = (5.0 * s2 - 9.0) * e[0]
        M[6][I] = (5.0 * s2 - 9.0) * e[1]
        M[8][I] = (5.0 * s2 - 9.0) * e[2]
        M[9][I] = 3.0 * e[0] ** 2 - s2
        M[10][I] = (3.0 * s2 - 5.0) * (3.0 * e[0] ** 2 - s2)
        M[11][I] = float(e[1] ** 2 - e[2] ** 2)
        M[12][I] = (3.0 * s2 - 5.0) * (e[1] ** 2 - e[2] ** 2)
        M[13][I] = float(e[0] * e[1])
        M[14][I] = float(e[1] * e[2])
        M[15][I] = float(e[2] * e[0])
        M[16][I] = float((e[1] ** 2 - e[2] ** 2) * e[0])
        M[17][I] = float((e[2] ** 2 - e[0] ** 2) * e[1])
        M[18][I] = float((e[0] ** 2 - e[1] ** 2) * e[2])
    MI = [[M[j][i] for j in range(Q)] for i in range(Q)]   # TRANSPOSE
    MM = _matmul(M, MI)
    for I in range(Q):
        for r in range(Q):
            MI[r][I] = MI[r][I] / MM[I][I]
    S = [S0, S1, S2, S0, S4, S0, S4, S0, S4, Omega, S10, Omega, S10, Omega, Omega, Omega, S16, S16, S16]
    SD = [[S[i] if i == j else 0.0 for j in range(Q)] for i in range(Q)]
    Mc = _matmul(_matmul(MI, SD), M)
    Mf = [[(1.0 if i == j else 0.0) - 0.5 * Mc[i][j] for j in range(Q)] for i in range(Q)]
    return np.array(Mc), np.array(Mf)


def fortran_order(m: np.ndarray) -> np.ndarray:
    """19x19 [i, j] -> flat column-major (element (i,j) at j*19 + i)."""
    return np.ascontiguousarray(m.T).reshape(-1).astype(np.float64)


def volume_force(VFin, amp, freq, phi, t):
    """update_volume_force_ (FluidDomain.f90:1174-1180)."""
    return (VFin[0] + amp * math.sin(2.0 * PI * freq * t + phi / 180.0 * PI), VFin[1], VFin[2])


def equilibrium_point(rho: float, u) -> np.ndarray:
    """calculate_distribution_funcion (FluidDomain.f90:1827-1834) of one node,
    in its association order: wt*rho*(((1+3ux)+4.5ux*ux)-1.5u^2)."""
    u1, u2, u3 = (float(v) for v in u)
    uSqr = u1 * u1
    uSqr = uSqr + u2 * u2
    uSqr = uSqr + u3 * u3
    out = np.empty(19)
    for q in range(19):
        ux = (u1 * EE[q, 0] + u2 * EE[q, 1]) + u3 * EE[q, 2]
        out[q] = (WT[q] * rho) * (((1.0 + 3.0 * ux) + (4.5 * ux) * ux) - 1.5 * uSqr)
    return out
