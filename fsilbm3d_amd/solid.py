"""Rigid plate bodies on the host: prescribed kinematics and the immersed-
boundary markers that the device coupling consumes.

This is the host half of the coupling step, restated in the reference's
association order so the marker arrays are bit-identical:
  - Segment_Build / Init / cptdxyz1        SolidSolver.f90:60-129
  - BuildAxisDirTriad / RigidUpdateTriad_D SolidSolver.f90:602-697
  - get_angle_triad                        SolidSolver.f90:909-952
  - Map*ToCurrent                          SolidSolver.f90:1126-1145
  - Beam_Initialise (rigid part)           SolidSolver.f90:1401-1446
  - UpdateVelFromPosAngular                SolidSolver.f90:1478-1494
  - Beam_structure, iBodyModel = 1         SolidSolver.f90:1826-1858
  - AoAtoTTT (with its pi constant)        SolidSolver.f90:2404-2472
  - calculate_angle_material (unit part)   SolidSolver.f90:1543-1547
  - PlateBuild_ / PlateUpdatePosVelArea_   Solidbody.f90:1051-1071, 604-646
Elastic bodies (iBodyModel = 2) are not handled here.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence

import numpy as np

M_PI = 3.141592653589793          # m_pi (SolidSolver.f90:1261)
AOA_PI = 3.141562653589793         # the constant AoAtoTTT uses (SolidSolver.f90:2406)
AOA_EPS = 1.0e-5


def _sum(xs):
    s = 0.0
    for v in xs:
        s = s + v
    return s


def _dot(a, b):
    return _sum(a[i] * b[i] for i in range(len(a)))


def _matvec(T, v):
    return [_sum(T[i][k] * v[k] for k in range(3)) for i in range(3)]


def _matmul(A, B):
    return [[_sum(A[i][k] * B[k][j] for k in range(3)) for j in range(3)] for i in range(3)]


def aoa_to_ttt(aoa):
    def cs(a):
        c, s = math.cos(a), math.sin(a)
        if abs(a) < AOA_EPS:
            c, s = 1.0, 0.0
        if abs(a - 0.5 * AOA_PI) < AOA_EPS:
            c, s = 0.0, 1.0
        if abs(a + 0.5 * AOA_PI) < AOA_EPS:
            c, s = 0.0, -1.0
        return c, s
    c, s = cs(aoa[0])
    rrx = [[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]]
    c, s = cs(aoa[1])
    rry = [[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]]
    c, s = cs(aoa[2])
    rrz = [[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]]
    T = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    T = _matmul(rrz, T)
    T = _matmul(rry, T)
    T = _matmul(rrx, T)
    return T


def angle_triad(t11, t22):
    rr = [[_sum(t22[i][k] * t11[j][k] for k in range(3)) for j in range(3)] for i in range(3)]
    dtx = (rr[2][1] - rr[1][2]) / 2.0
    dty = (rr[0][2] - rr[2][0]) / 2.0
    dtz = (rr[1][0] - rr[0][1]) / 2.0
    tr = rr[0][0] + rr[1][1] + rr[2][2]
    tr = (tr - 1.0) / 2.0
    tr = min(max(tr, -1.0), 1.0)
    sint = math.sqrt(dtx * dtx + dty * dty + dtz * dtz)
    theta = math.acos(tr)
    if sint < 1.0e-10 or theta < 1.0e-10:
        return [dtx, dty, dtz]
    f = theta / sint
    return [f * dtx, f * dty, f * dtz]


def axis_dir_triad(l, m, n, d):
    ex = [l, m, n]
    dd = math.sqrt(_dot(ex, ex))
    if dd > 1.0e-14:
        ex = [v / dd for v in ex]
    dr = list(d)
    dd = math.sqrt(_dot(dr, dr))
    if dd > 1.0e-14:
        dr = [v / dd for v in dr]
    proj = _dot(dr, ex)
    ey = [dr[i] - proj * ex[i] for i in range(3)]
    dd = math.sqrt(_dot(ey, ey))
    if dd <= 1.0e-10:
        if abs(ex[2]) > 0.995:
            ey = [0.0, 1.0, 0.0]
            ez = [-ex[2], 0.0, 0.0]
        else:
            dd = math.sqrt(ex[0] * ex[0] + ex[1] * ex[1])
            ey = [-ex[1] / dd, ex[0] / dd, 0.0]
            ez = [-ex[0] * ex[2] / dd, -ex[1] * ex[2] / dd, dd]
    else:
        ey = [v / dd for v in ey]
        ez = [ex[1] * ey[2] - ex[2] * ey[1], ex[2] * ey[0] - ex[0] * ey[2], ex[0] * ey[1] - ex[1] * ey[0]]
        dd = math.sqrt(_dot(ez, ez))
        if dd > 1.0e-14:
            ez = [v / dd for v in ez]
    # triad(:,1)=ex, (:,2)=ey, (:,3)=ez ; returned by column
    return [ex, ey, ez]


@dataclasses.dataclass
class Element:
    node0: int
    node1: int
    Nspan: int
    x00: List[float]
    Lspan: float
    spanlen: float
    dirc00: List[float]
    area00: float
    x1: List[float] = None
    dirc1: List[float] = None
    len1: float = 0.0
    l1: float = 0.0
    m1: float = 0.0
    n1: float = 0.0
    triad_y: List[float] = None   # column 2 of triad_n1 (= triad_n2 = triad_ee for rigid)


class RigidPlate:
    """A v_type = 1 body (plate on a line mesh) with iBodyModel = 1."""

    def __init__(self, nodes: np.ndarray, elements: Sequence[Sequence[int]], *, XYZo, XYZAmpl=(0, 0, 0),
                 XYZPhi=(0, 0, 0), AoAo=(0, 0, 0), AoAAmpl=(0, 0, 0), AoAPhi=(0, 0, 0), Freq=0.0,
                 initXYZVel=(0, 0, 0), isMotionGiven=(1, 1, 1, 1, 1, 1), denIn=1.0, IBPenaltyAlpha=1.0):
        """nodes: (nND, 8) rows x y z Lspan Rspan dirX dirY dirZ (plate.dat POINT);
        elements: (i, j, Nspan) with 1-based node ids (plate.dat ELEMENT)."""
        self.nodes = np.asarray(nodes, dtype=np.float64)
        self.nND = len(self.nodes)
        self.XYZo = [float(v) for v in XYZo]
        self.XYZAmpl = [float(v) for v in XYZAmpl]
        self.AoAo = [float(v) for v in AoAo]
        self.AoAAmpl = [float(v) for v in AoAAmpl]
        self.initXYZVel = [float(v) for v in initXYZVel]
        self.Freq = float(Freq)
        self.denIn = float(denIn)
        self.alpha = float(IBPenaltyAlpha)
        self.v_move = int(sum(abs(v) for v in initXYZVel) > 1e-5 or sum(abs(v) for v in XYZAmpl) > 1e-5
                          or sum(abs(v) for v in AoAAmpl) > 1e-5 or sum(isMotionGiven) < 6)
        # degrees -> radians (SolidSolver.f90:1543-1547)
        self.AoAo = [v / 180.0 * M_PI for v in self.AoAo]
        self.AoAAmpl = [v / 180.0 * M_PI for v in self.AoAAmpl]
        self.AoAPhi = [float(v) / 180.0 * M_PI for v in AoAPhi]
        self.XYZPhi = [float(v) / 180.0 * M_PI for v in XYZPhi]
        self.elements: List[Element] = []
        xyz = self.nodes
        for (i, j, ns) in elements:
            p0, p1 = i - 1, j - 1
            x00 = [xyz[p0, 0], xyz[p0, 1], xyz[p0, 2], 0.0, 0.0, 0.0, xyz[p1, 0], xyz[p1, 1], xyz[p1, 2], 0.0, 0.0, 0.0]
            x00 = [float(v) for v in x00]
            area = math.sqrt((x00[6] - x00[0]) ** 2 + (x00[7] - x00[1]) ** 2 + (x00[8] - x00[2]) ** 2)
            Lspan = 0.5 * (float(xyz[p0, 3]) + float(xyz[p1, 3]))
            spanlen = 0.5 * (float(xyz[p0, 4]) + float(xyz[p1, 4])) + Lspan
            d = [0.5 * (float(xyz[p0, 5 + k]) + float(xyz[p1, 5 + k])) for k in range(3)]
            nrm = math.sqrt(_sum(v * v for v in d))
            if not nrm > 1e-10:
                raise ValueError("element node directions are opposite; no unique bisector exists")
            d = [v / nrm for v in d]
            self.elements.append(Element(p0, p1, int(ns), x00, Lspan, spanlen, d, area))
        self.nEL = len(self.elements)
        self.Asfac = _sum(e.area00 for e in self.elements)
        self.rtov = [0]
        for e in self.elements:
            self.rtov.append(self.rtov[-1] + e.Nspan)
        self.M = self.rtov[-1]
        self.marker_elem = np.repeat(np.arange(self.nEL, dtype=np.int32),
                                     [e.Nspan for e in self.elements]).astype(np.int32)
        self.elem_nodes = np.array([[e.node0, e.node1] for e in self.elements], dtype=np.int32)
        self.pos = np.zeros((self.nND, 6))
        self.vel = np.zeros((self.nND, 6))
        self._initialise(0.0)
        self.Exyz = np.zeros((self.M, 3))
        self.Evel = np.zeros((self.M, 3))
        self.Ea = np.zeros(self.M)
        self.update_markers()

    # -- kinematics ---------------------------------------------------------
    def _place(self, t: float):
        c = 2.0 * M_PI * self.Freq
        XYZ = [self.XYZo[k] + self.XYZAmpl[k] * math.cos(c * t + self.XYZPhi[k]) + self.initXYZVel[k] * t for k in range(3)]
        AoA = [self.AoAo[k] + self.AoAAmpl[k] * math.cos(c * t + self.AoAPhi[k]) for k in range(3)]
        return XYZ, AoA

    def _velocity(self, t: float, AoA, TTT):
        c = 2.0 * M_PI * self.Freq
        UVW = [-(c * self.XYZAmpl[k] * math.sin(c * t + self.XYZPhi[k])) + self.initXYZVel[k] for k in range(3)]
        W1 = [-(c * self.AoAAmpl[k] * math.sin(c * t + self.AoAPhi[k])) for k in range(3)]
        W2 = [W1[0] * math.cos(AoA[1]) + W1[2],
              W1[0] * math.sin(AoA[1]) * math.sin(AoA[2]) + W1[1] * math.cos(AoA[2]),
              W1[0] * math.sin(AoA[1]) * math.cos(AoA[2]) - W1[1] * math.sin(AoA[2])]
        W3 = _matvec(TTT, W2)
        self.W3, self.UVW = W3, UVW
        for n in range(self.nND):
            rel = [self.pos[n, k] - self.XYZ[k] for k in range(3)]
            w = [W3[1] * rel[2] - W3[2] * rel[1], W3[2] * rel[0] - W3[0] * rel[2], W3[0] * rel[1] - W3[1] * rel[0]]
            for k in range(3):
                self.vel[n, k] = w[k] + UVW[k]
                self.vel[n, 3 + k] = W3[k]

    def _map_elements(self, TTT, XYZ, AoAd):
        for e in self.elements:
            a = _matvec(TTT, e.x00[0:3])
            b = _matvec(TTT, e.x00[6:9])
            e.x1 = [a[k] + XYZ[k] for k in range(3)] + list(AoAd) + [b[k] + XYZ[k] for k in range(3)] + list(AoAd)
            e.dirc1 = _matvec(TTT, e.dirc00)
            self.pos[e.node0, :] = e.x1[0:6]
            self.pos[e.node1, :] = e.x1[6:12]
            dx, dy, dz = e.x1[6] - e.x1[0], e.x1[7] - e.x1[1], e.x1[8] - e.x1[2]
            e.len1 = math.sqrt(dx * dx + dy * dy + dz * dz)
            e.l1, e.m1, e.n1 = dx / e.len1, dy / e.len1, dz / e.len1
            e.triad_y = axis_dir_triad(e.l1, e.m1, e.n1, e.dirc1)[1]

    def _initialise(self, t: float):
        self.TTT00 = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
        self.XYZ, AoA = self._place(t)
        self.TTT0 = aoa_to_ttt(AoA)
        self.TTTnxt = self.TTT0
        AoAd = angle_triad(self.TTT0, self.TTT0)
        self._map_elements(self.TTT0, self.XYZ, AoAd)
        self._velocity(t, AoA, self.TTT0)

    def structure(self, time: float, isubstep: int, deltat: float, subdeltat: float):
        """Beam_structure for iBodyModel = 1 (SolidSolver.f90:1826-1858)."""
        t = (time - deltat) + float(isubstep) * subdeltat
        self.XYZ, AoA = self._place(t)
        TTTnxt = aoa_to_ttt(AoA)
        self.TTTnxt = TTTnxt
        AoAd = angle_triad(self.TTT0, TTTnxt)
        self._map_elements(TTTnxt, self.XYZ, AoAd)
        self._velocity(t, AoA, TTTnxt)

    # -- markers ------------------------------------------------------------
    def update_markers(self):
        """PlateUpdatePosVelArea_ (Solidbody.f90:604-646)."""
        beta = -(self.alpha * 2.0 * self.denIn)
        for i, e in enumerate(self.elements):
            i1, i2 = e.node0, e.node1
            tmpxyz = [0.5 * (self.pos[i1, k] + self.pos[i2, k]) for k in range(3)]
            tmpvel = [0.5 * (self.vel[i1, k] + self.vel[i2, k]) for k in range(3)]
            om = [0.5 * (self.vel[i1, 3 + k] + self.vel[i2, 3 + k]) for k in range(3)]
            dl = e.spanlen / float(e.Nspan)
            area = dl * e.len1 * beta
            dirc = [0.5 * (e.triad_y[k] + e.triad_y[k]) for k in range(3)]
            nrm = math.sqrt(_dot(dirc, dirc))
            dirc = [v / nrm for v in dirc] if nrm > 1.0e-12 else list(e.triad_y)
            c0 = self.rtov[i]
            for s in range(e.Nspan):
                ls = dl * (0.5 + float(s)) - e.Lspan
                r = [dirc[k] * ls for k in range(3)]
                w = [om[1] * r[2] - om[2] * r[1], om[2] * r[0] - om[0] * r[2], om[0] * r[1] - om[1] * r[0]]
                for k in range(3):
                    self.Exyz[c0 + s, k] = tmpxyz[k] + r[k]
                    self.Evel[c0 + s, k] = tmpvel[k] + w[k]
                self.Ea[c0 + s] = area

    def element_centres(self) -> np.ndarray:
        """Beam axis midpoints of the current element coordinates, the moment
        arm origin of FluidVolumeForce_ (Solidbody.f90:954)."""
        return np.array([[0.5 * (e.x1[k] + e.x1[6 + k]) for k in range(3)] for e in self.elements])

    def force_coefficients(self, lodFlow: np.ndarray, Fref: float):
        """sum(lodFlow(1:gEQ:6))/Fref etc. (SolidSolver.f90:1761-1765)."""
        return tuple(_sum(lodFlow[k::6]) / Fref for k in range(3))


def read_gmsh(text: str):
    """Read_gmsh (Solidbody.f90:1088-1128): the nodes and triangles of a gmsh
    2 ASCII mesh.  Four header lines, the node count and "id x y z" rows, two
    lines, the element count, then every element from the first triangle
    (type 2 with two tags, first tag 0) on is a triangle "id 2 2 0 t a b c"."""
    lines = text.splitlines()
    n = int(lines[4].split()[0])
    xyz = np.array([[float(v) for v in lines[5 + i].split()[1:4]] for i in range(n)])
    i = 5 + n + 2
    ne = int(lines[i].split()[0])
    i += 1
    for j in range(ne):
        parts = lines[i + j].split()
        p = [int(v) for v in parts[1:5]]
        if p[0] == 2 and p[1] == 2 and p[2] == 0:
            ntri = ne - int(parts[0]) + 1
            tri = np.array([[int(v) for v in lines[i + j + k].split()[5:8]] for k in range(ntri)], np.int32)
            return xyz, tri
    raise ValueError("no triangles in the mesh")


def surface_geometry(xyz, tri, alpha: float, denIn: float):
    """SurfaceBuildPosVelArea_ (Solidbody.f90:648-705): per triangle its
    incenter (body frame) and area times -2 alpha denIn."""
    beta = -(alpha * 2.0 * denIn)
    E0 = np.zeros((len(tri), 3))
    Ea = np.zeros(len(tri))
    for i, (a, b, c) in enumerate(tri):
        x1, y1, z1 = (float(v) for v in xyz[a - 1])
        x2, y2, z2 = (float(v) for v in xyz[b - 1])
        x3, y3, z3 = (float(v) for v in xyz[c - 1])
        la = math.sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3) + (z2 - z3) * (z2 - z3))
        lb = math.sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1) + (z3 - z1) * (z3 - z1))
        lc = math.sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2))
        inv = 1.0 / (la + lb + lc)
        E0[i] = [(la * x1 + lb * x2 + lc * x3) * inv, (la * y1 + lb * y2 + lc * y3) * inv,
                 (la * z1 + lb * z2 + lc * z3) * inv]
        ax = (z1 - z2) * (y3 - y2) + (y2 - y1) * (z3 - z2)
        ay = (x1 - x2) * (z3 - z2) + (z2 - z1) * (x3 - x2)
        az = (y1 - y2) * (x3 - x2) + (x2 - x1) * (y3 - y2)
        Ea[i] = math.sqrt(ax * ax + ay * ay + az * az) * 0.5 * beta
    return E0, Ea


def surface_beam_nodes(xyz) -> np.ndarray:
    """The two-node beam SurfacetoBeam_write makes of a surface mesh
    (Solidbody.f90:1130-1170): its first two nodes written with F20.10, Lspan
    = Rspan = 1, direction z."""
    return np.array([[float("%.10f" % float(v)) for v in xyz[i]] + [1.0, 1.0, 0.0, 0.0, 1.0] for i in range(2)])


class SurfaceBody(RigidPlate):
    """A v_type = -1 body (triangulated surface from a gmsh mesh) with
    iBodyModel = 1: one marker per triangle at its incenter, moved rigidly with
    the body's beam (SurfaceUpdatePosVel_, Solidbody.f90:707-720) and loading
    that beam's single element (vtor = 1)."""

    def __init__(self, msh_text: str, **kw):
        xyz, tri = read_gmsh(msh_text)
        self.E0, self.Ea0 = surface_geometry(xyz, tri, kw.get("IBPenaltyAlpha", 1.0), kw.get("denIn", 1.0))
        super().__init__(surface_beam_nodes(xyz), [(1, 2, 1)], **kw)
        self.M = len(self.E0)
        self.marker_elem = np.zeros(self.M, np.int32)
        self.rtov = [0, self.M]

    def update_markers(self):
        T, W, U = self.TTTnxt, self.W3, self.UVW
        n = len(self.E0)
        X, V = np.zeros((n, 3)), np.zeros((n, 3))
        for i in range(n):
            rel = _matvec(T, [float(v) for v in self.E0[i]])
            w = [W[1] * rel[2] - W[2] * rel[1], W[2] * rel[0] - W[0] * rel[2], W[0] * rel[1] - W[1] * rel[0]]
            for k in range(3):
                X[i, k] = rel[k] + self.XYZ[k]
                V[i, k] = w[k] + U[k]
        self.Exyz, self.Evel, self.Ea = X, V, self.Ea0.copy()
