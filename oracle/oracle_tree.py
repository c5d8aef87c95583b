"""CPU oracle of the multi-block (father/son, refinement ratio 2) coupling of
the reference (gaoak/FSILBM3D src/LBMBlockComm.f90), test infrastructure only.

Plane work runs in numpy with elementwise IEEE operations written in the
reference's association order (no reductions), so it is bit-exact; block
steps use oracle.OracleBlock.

  build_blocks_comunication  LBMBlockComm.f90:33-107  -> comm_pair()
  extract_interpolate_layer  LBMBlockComm.f90:340-505 -> extract()
  interpolation_father_to_son LBMBlockComm.f90:655-806 -> father_to_son()
  interpolate_fIn / _tau     LBMBlockComm.f90:808-956 -> interp_plane()
  fIn_GridTransform          LBMBlockComm.f90:958-979 -> grid_transform()
  deliver_son_to_father      LBMBlockComm.f90:546-653 -> son_to_father()
  tree_collision_streaming_IBM_FEM LBMBlockComm.f90:279-318 -> OracleTree.step()
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from .oracle import OracleBlock

EE = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
               [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
               [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]], dtype=np.float64).T
WT = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)
GRID_DELTA = 2


def comm_pair(father, son, son_bc, son_periodic):
    """CommPair of a father/son pair (LBMBlockComm.f90:47-95); indices 1-based.
    father/son: (dims, dh, origin)."""
    (fd, fdh, fo), (sd, sdh, so) = father, son
    sds = [0] * 6
    for j in range(6):
        if son_bc[j] == 0:
            sds[j] = 1 if j % 2 == 0 else -1
    sxD, syD, szD = sd
    if son_periodic[0]:
        sxD -= 1
    if son_periodic[1]:
        syD -= 1
    if son_periodic[2]:
        szD -= 1
    s = [1, sxD, 1, syD, 1, szD]
    f = [0] * 6
    f[0] = math.floor((so[0] - fo[0]) / fdh + 1.5)
    f[2] = math.floor((so[1] - fo[1]) / fdh + 1.5)
    f[4] = math.floor((so[2] - fo[2]) / fdh + 1.5)
    ratio = math.floor(fdh / sdh + 0.5)
    f[1] = f[0] + (sxD - 1) // ratio
    f[3] = f[2] + (syD - 1) // ratio
    f[5] = f[4] + (szD - 1) // ratio
    si = [s[j] + sds[j] * ratio for j in range(6)]
    fi = [f[j] + sds[j] for j in range(6)]
    return dict(sds=sds, s=s, f=f, si=si, fi=fi, xDimS=sd[0], yDimS=sd[1], zDimS=sd[2],
                xDimF=f[1] - f[0] + 1, yDimF=f[3] - f[2] + 1, zDimF=f[5] - f[4] + 1)


def grid_transform(fq: np.ndarray, coeff: float, VF, dh: float) -> np.ndarray:
    """fIn_GridTransform for an array of nodes fq (19, ...)."""
    den, u = _moments(fq, VF, dh)
    uSqr = u[0] * u[0]
    uSqr = uSqr + u[1] * u[1]
    uSqr = uSqr + u[2] * u[2]
    out = np.empty_like(fq)
    for q in range(19):
        ux = (u[0] * EE[q, 0] + u[1] * EE[q, 1]) + u[2] * EE[q, 2]
        feq = (WT[q] * den) * ((1.0 - 1.5 * uSqr) + ux * (3.0 + 4.5 * ux))
        out[q] = feq + coeff * (fq[q] - feq)
    return out


def _moments(fq, VF, dh):
    s = np.zeros(fq.shape[1:])
    for q in range(19):
        s = s + fq[q]
    u = []
    for k in range(3):
        m = np.zeros(fq.shape[1:])
        for q in range(19):
            m = m + fq[q] * EE[q, k]
        u.append((m + 0.5 * VF[k] * dh) / s)
    return s, u


def interp_plane(fF: np.ndarray, bS: int, aS: int, scheme: int) -> np.ndarray:
    """interpolate_fIn (LBMBlockComm.f90:808-914): fF (19, aF, bF) coarse plane
    -> (19, aS, bS) fine plane (a the slow, b the fast in-plane index)."""
    q = fF.shape[0]
    fS = np.zeros((q, aS, bS))
    bT, aT = bS, aS
    r2 = r1 = 0
    if bS % 2 == 0:
        bT, r2 = bS - 1, 1
    if aS % 2 == 0:
        aT, r1 = aS - 1, 1
    F = lambda b, a: fF[:, a - 1, b - 1]  # 1-based access
    if scheme == 2:
        for b in range(1, bT + 1, 2):
            b1 = b // 2 + 1
            for a in range(1, aT + 1, 2):
                a1 = a // 2 + 1
                fS[:, a - 1, b - 1] = F(b1, a1)
                if b == 1:
                    fS[:, a - 1, b] = 0.375 * F(b1, a1) + 0.75 * F(b1 + 1, a1) - 0.125 * F(b1 + 2, a1)
                elif b == bT - 2:
                    fS[:, a - 1, b] = 0.375 * F(b1 + 1, a1) + 0.75 * F(b1, a1) - 0.125 * F(b1 - 1, a1)
                elif b != bT:
                    fS[:, a - 1, b] = (-(0.0625 * F(b1 - 1, a1)) + 0.5625 * F(b1, a1) + 0.5625 * F(b1 + 1, a1)
                                       - 0.0625 * F(b1 + 2, a1))
        S = lambda b, a: fS[:, a - 1, b - 1]
        for b in range(1, bT + 1):
            for a in range(2, aT + 1, 2):
                if a == 2:
                    v = 0.375 * S(b, a - 1) + 0.75 * S(b, a + 1) - 0.125 * S(b, a + 3)
                elif a == aT - 1:
                    v = 0.375 * S(b, a + 1) + 0.75 * S(b, a - 1) - 0.125 * S(b, a - 3)
                else:
                    v = -(0.0625 * S(b, a - 3)) + 0.5625 * S(b, a - 1) + 0.5625 * S(b, a + 1) - 0.0625 * S(b, a + 3)
                fS[:, a - 1, b - 1] = v
        if r2:
            for a in range(1, aT + 1):
                fS[:, a - 1, bT] = -(0.0625 * S(bT - 2, a)) + 0.5625 * S(bT, a) + 0.5625 * S(1, a) - 0.0625 * S(3, a)
        if r1:
            for b in range(1, bT + 1):
                fS[:, aT, b - 1] = -(0.0625 * S(b, aT - 2)) + 0.5625 * S(b, aT) + 0.5625 * S(b, 1) - 0.0625 * S(b, 3)
            if r2:
                fS[:, aT, bT] = (-(0.0625 * S(bT + 1, aT - 2)) + 0.5625 * S(bT + 1, aT) + 0.5625 * S(bT + 1, 1)
                                 - 0.0625 * S(bT + 1, 3))
    else:
        for b in range(1, bT + 1, 2):
            b1 = b // 2 + 1
            for a in range(1, aT + 1, 2):
                a1 = a // 2 + 1
                fS[:, a - 1, b - 1] = F(b1, a1)
                if b < bT:
                    fS[:, a - 1, b] = (F(b1, a1) + F(b1 + 1, a1)) * 0.5
        S = lambda b, a: fS[:, a - 1, b - 1]
        for b in range(1, bT + 1):
            for a in range(2, aT + 1, 2):
                fS[:, a - 1, b - 1] = (S(b, a - 1) + S(b, a + 1)) * 0.5
        if r2:
            for a in range(1, aT + 1):
                fS[:, a - 1, bT] = (S(bT, a) + S(1, a)) * 0.5
        if r1:
            for b in range(1, bT + 1):
                fS[:, aT, b - 1] = (S(b, aT) + S(b, 1)) * 0.5
            if r2:
                fS[:, aT, bT] = (S(bT + 1, aT) + S(bT + 1, 1)) * 0.5
    return fS


class OracleTree:
    """A root block and refined son blocks (one level), stepped as the
    reference's recursive tree step."""

    def __init__(self, root: OracleBlock, sons: List[OracleBlock], pairs: List[dict], taus: List[float],
                 scheme: int = 1):
        self.root, self.sons, self.pairs, self.scheme = root, sons, pairs, scheme
        self.tau_root, self.tau_sons = taus[0], taus[1:]
        self.planes = [dict() for _ in sons]

    @staticmethod
    def _face_plane(f, face, p):
        """father fIn plane region of a face (19, a, b)."""
        fx = p["f"]
        x0, x1, y0, y1, z0, z1 = fx[0] - 1, fx[1] - 1, fx[2] - 1, fx[3] - 1, fx[4] - 1, fx[5] - 1
        if face == 0 or face == 1:
            xF = fx[face] - 1
            return f[:, xF, y0:y1 + 1, z0:z1 + 1]
        if face == 2 or face == 3:
            yF = fx[face] - 1
            return f[:, x0:x1 + 1, yF, z0:z1 + 1]
        zF = fx[face] - 1
        return f[:, x0:x1 + 1, y0:y1 + 1, zF]

    def extract(self, k, ff, time):
        p = self.pairs[k]
        for face in range(6):
            if p["sds"][face] == 0:
                continue
            pl = np.array(self._face_plane(ff, face, p))
            if time == 1:
                self.planes[k][(face, 1)] = pl
            else:
                self.planes[k][(face, 2)] = pl
                self.planes[k][(face, 1)] = 0.5 * (self.planes[k][(face, 1)] + pl)

    def father_to_son(self, k, fs, n_timeStep):
        p = self.pairs[k]
        VF = self.root.vf(self.root_time)
        dh = self.root.dh
        coeff = (self.tau_sons[k] / self.tau_root) / float(GRID_DELTA)
        sd = (p["xDimS"], p["yDimS"], p["zDimS"])
        for face in range(6):
            if p["sds"][face] == 0:
                continue
            pl = self.planes[k][(face, 1 if n_timeStep == 0 else 2)]
            axis = face // 2
            if axis == 0:
                aS, bS = sd[1], sd[2]
            elif axis == 1:
                aS, bS = sd[0], sd[2]
            else:
                aS, bS = sd[0], sd[1]
            t = interp_plane(pl, bS, aS, self.scheme)
            t = grid_transform(t, coeff, VF, dh)
            c = p["s"][face] - 1
            if axis == 0:
                fs[:, c, :, :] = t
            elif axis == 1:
                fs[:, :, c, :] = t
            else:
                fs[:, :, :, c] = t

    def son_to_father(self, k, ff, fs, son_vf, son_dh):
        p = self.pairs[k]
        coeff = (self.tau_root / self.tau_sons[k]) * float(GRID_DELTA)
        si, fi = p["si"], p["fi"]
        for face in range(6):
            if p["sds"][face] == 0:
                continue
            axis = face // 2
            rng = [slice(si[0] - 1, si[1], 2), slice(si[2] - 1, si[3], 2), slice(si[4] - 1, si[5], 2)]
            frng = [slice(fi[0] - 1, fi[0] - 1 + len(range(si[0], si[1] + 1, 2))),
                    slice(fi[2] - 1, fi[2] - 1 + len(range(si[2], si[3] + 1, 2))),
                    slice(fi[4] - 1, fi[4] - 1 + len(range(si[4], si[5] + 1, 2)))]
            rng[axis] = si[face] - 1
            frng[axis] = fi[face] - 1
            src = np.array(fs[(slice(None),) + tuple(rng)])
            ff[(slice(None),) + tuple(frng)] = grid_transform(src, coeff, son_vf, son_dh)

    def step(self, fs_all, time, sons_first=None):
        """One root step at block time `time` (main.f90:95-107)."""
        ff = fs_all[0]
        self.root_time = time
        for k in range(len(self.sons)):
            self.extract(k, ff, 1)
        self.root.step(ff, time)
        for k in range(len(self.sons)):
            self.extract(k, ff, 2)
        for k, sb in enumerate(self.sons):
            fs = fs_all[1 + k]
            st = time
            for n in range(GRID_DELTA):
                st = st + float(n) * sb.dh
                sb.step(fs, st)
                self.father_to_son(k, fs, n)
            self.son_to_father(k, ff, fs, sb.vf(st), sb.dh)
