// D3Q19 fluid kernels for gfx950 (FP64 plain form).
//
// Exactness: this file is compiled with -ffp-contract=off, so no multiply-add
// is fused; every expression below is written in the association order of the
// reference's Fortran (gaoak/FSILBM3D src/FluidDomain.f90).  Terms whose lattice
// component e is 0 are skipped: in IEEE arithmetic x + (+-0) == x for every
// nonzero x, and a running sum that starts at +0 never becomes -0, so skipping
// them changes at most the sign of an intermediate zero, never a stored value.
#pragma once
#include "fsil_common.h"

namespace fsil {

// ---------------------------------------------------------------------------
// per-direction helpers, fully unrolled at compile time
template <int E>
__device__ __forceinline__ double sgn(double v) {
  return E > 0 ? v : -v;
}

// uxyz(q) = u1*e1 + u2*e2 + u3*e3   (FluidDomain.f90:1219)
template <int q>
__device__ __forceinline__ double dotu(double u1, double u2, double u3) {
  constexpr int a = ex(q), b = ey(q), c = ez(q);
  if constexpr (a == 0 && b == 0 && c == 0) return 0.0;
  else if constexpr (a != 0 && b != 0) return sgn<a>(u1) + sgn<b>(u2);
  else if constexpr (a != 0 && c != 0) return sgn<a>(u1) + sgn<c>(u3);
  else if constexpr (b != 0 && c != 0) return sgn<b>(u2) + sgn<c>(u3);
  else if constexpr (a != 0) return sgn<a>(u1);
  else if constexpr (b != 0) return sgn<b>(u2);
  else return sgn<c>(u3);
}

// (e_i - u_i + 3*uxyz*e_i)   (FluidDomain.f90:1222-1224)
template <int E>
__device__ __forceinline__ double guo_a(double ui, double t3) {
  if constexpr (E == 0) return -ui;
  else if constexpr (E > 0) return (1.0 - ui) + t3;
  else return (-1.0 - ui) - t3;
}

// Flb(q) = dt3*wt*(A1*F1 + A2*F2 + A3*F3)
template <int q>
__device__ __forceinline__ double guo_force(double dt3, double ux, double u1, double u2,
                                            double u3, double F1, double F2, double F3) {
  const double t3 = 3.0 * ux;
  const double s = (guo_a<ex(q)>(u1, t3) * F1 + guo_a<ey(q)>(u2, t3) * F2) +
                   guo_a<ez(q)>(u3, t3) * F3;
  return (dt3 * wt(q)) * s;
}

// Sequential moments of one cell (FluidDomain.f90:1136-1139).
__device__ __forceinline__ void moments(const double (&f)[Q], double &rho, double &m1,
                                        double &m2, double &m3) {
  // the reference's sums start from +0; 0 + a == a for every a but -0, and a
  // population is never -0, so each sum starts from its first term
  double s = f[0];
#pragma unroll
  for (int q = 1; q < Q; ++q) s = s + f[q];
  rho = s;
  // e_x != 0: 1,2,7,8,9,10,11,12,13,14 ; e_y: 3,4,7,8,9,10,15,16,17,18 ;
  // e_z: 5,6,11,12,13,14,15,16,17,18 (q ascending)
  m1 = f[1]; m1 = m1 - f[2]; m1 = m1 + f[7]; m1 = m1 - f[8]; m1 = m1 + f[9];
  m1 = m1 - f[10]; m1 = m1 + f[11]; m1 = m1 - f[12]; m1 = m1 + f[13]; m1 = m1 - f[14];
  m2 = f[3]; m2 = m2 - f[4]; m2 = m2 + f[7]; m2 = m2 + f[8]; m2 = m2 - f[9];
  m2 = m2 - f[10]; m2 = m2 + f[15]; m2 = m2 - f[16]; m2 = m2 + f[17]; m2 = m2 - f[18];
  m3 = f[5]; m3 = m3 - f[6]; m3 = m3 + f[11]; m3 = m3 + f[12]; m3 = m3 - f[13];
  m3 = m3 - f[14]; m3 = m3 + f[15]; m3 = m3 + f[16]; m3 = m3 - f[17]; m3 = m3 - f[18];
}

// The coupling call's error word in its control words (IbmCtl::kErr): while it
// holds an error (a call that did not wait, whose error the host reports at
// its next call), the fused kernels leave the band overlay out.
constexpr int kBandErrWord = 7;
constexpr int kBandNoErr = 0x7fffffff;

// Cells a collide launch covers.  kRegionAll: the whole grid.  kRegionBulk:
// the cells outside the rows x in [hole[0], hole[1]], y in [hole[2], hole[3]]
// (a predicted band box of whole z-rows).  kRegionHull: the rows of the
// smallest box holding that hole and the band's box (ctl[0..5]), by a
// grid-stride loop over tiles of the workgroup's shape with y aligned to the
// workgroup's rows.  Both read only the pre-collision state, so a cell
// collided by a bulk and then a hull launch of one step ends with the hull
// launch's values, and the hull may cover more than the hole.
enum CollideRegion { kRegionAll = 0, kRegionBulk = 1, kRegionHull = 2 };

// the rows of a hull launch: x in [x0, x1], y in [y0, y1] with y0 a multiple of by
__device__ inline void region_hull(const int *ctl, const int *hole, int xd, int yd, int by, int &x0, int &x1,
                                   int &y0, int &y1) {
  x0 = max(0, min(hole[0], ctl[0]));
  x1 = min(xd - 1, max(hole[1], ctl[3]));
  y0 = max(0, min(hole[2], ctl[1]));
  y1 = min(yd - 1, max(hole[3], ctl[4]));
  y0 = (y0 / by) * by;
}

// grid-stride tiles of a hull launch: calls cell(x, y, z) for every cell of
// the hull rows (every thread runs the same number of iterations)
template <typename F>
__device__ __forceinline__ void for_hull_cells(const int *ctl, const int *hole, int xd, int yd, int zd, F cell) {
  int x0, x1, y0, y1;
  const int by = (int)blockDim.y, bz = (int)blockDim.x;
  region_hull(ctl, hole, xd, yd, by, x0, x1, y0, y1);
  if (x1 < x0 || y1 < y0) return;
  const int ny = (y1 - y0 + by) / by, nz = (zd + bz - 1) / bz;
  const int ntile = (x1 - x0 + 1) * ny * nz;
  for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
    const int tz = t % nz, r = t / nz, ty = r % ny, x = x0 + r / ny;
    cell(x, y0 + ty * by + (int)threadIdx.y, tz * bz + (int)threadIdx.x);
  }
}

__device__ __forceinline__ bool in_hole(const int *hole, int x, int y) {
  return x >= hole[0] && x <= hole[1] && y >= hole[2] && y <= hole[3];
}

// ---------------------------------------------------------------------------
// Arguments of the fused collide + push-stream kernel.
struct CollideArgs {
  const double *__restrict__ fin;  // pre-collision state, SoA [q][x][y][z]
  double *__restrict__ fout;       // post-stream state (before boundary faces)
  int xd, yd, zd;
  long long N;
  double Omega, Omega2, dh;
  double c1;        // 1 - 0.5*Omega
  double hO, cT1;   // TRT: 0.5*Omega, 0.5-0.25*Omega
  double hO2, cT2;  // TRT: 0.5*Omega2, 0.5-0.25*Omega2
  double hv[3];     // 0.5*VF*dh   (numerator shift of the velocity moment)
  double VF[3];     // uniform volume force
  const double *__restrict__ Mc;  // MRT M_COLLID, Fortran column-major 19x19
  const double *__restrict__ Mf;  // MRT M_FORCE
  // immersed-boundary band overlay (cells whose velocity/force come from IBM)
  const int *__restrict__ ctl;        // band bounding box in ctl[0..5] (device memory)
  const int *__restrict__ cellmap;    // cell -> band index, -1 outside the band
  const double *__restrict__ band_u;  // [3][cap] corrected velocity
  const double *__restrict__ band_F;  // [3][cap] spread force (+VF)
  int cap;
  // auxiliary plane outputs for boundary codes 102 (den/u of the first
  // interior plane) and 203/204 (post-collision boundary plane)
  int aux_face_mask102, aux_face_mask_hw;
  double *nb_den[6];
  double *nb_u[6];   // [3][plane]
  double *hw[6];     // [q][plane]
  int nt;            // nontemporal population loads and stores (selects the kernel instantiation)
  // LES collision models 11 (Smagorinsky), 14 (WALE), 15 (Vreman)
  const double *__restrict__ uf;  // [3][N] velocity the collision reads (models 14, 15: neighbour gradients)
  double *__restrict__ tau_all;   // [N] per-cell relaxation time (WALE reads the previous step's)
  double tau, nu;                 // block tau, flow%nu
  double les_c;                   // CsmagConst, CWALEConst or CvremConst
  int region;                     // CollideRegion
  int hole[4];                    // kRegionBulk / kRegionHull: x0, x1, y0, y1
};

// velocity of one cell as the collision reads it: the moments of the stored
// state with the uniform volume force, or the immersed-boundary band's
// corrected value (FluidDomain.f90:1128-1145, Solidbody.f90:1034-1048)
template <bool BAND>
__device__ __forceinline__ void cell_velocity(const CollideArgs &a, const double (&f)[Q], unsigned n, int x, int y,
                                              int z, double &rho, double &u1, double &u2, double &u3, bool &in_band,
                                              int &b) {
  double m1, m2, m3;
  moments(f, rho, m1, m2, m3);
  u1 = (m1 + a.hv[0]) / rho;
  u2 = (m2 + a.hv[1]) / rho;
  u3 = (m3 + a.hv[2]) / rho;
  in_band = false;
  b = -1;
  if constexpr (BAND) {
    if (x >= a.ctl[0] && x <= a.ctl[3] && y >= a.ctl[1] && y <= a.ctl[4] && z >= a.ctl[2] && z <= a.ctl[5]) {
      b = a.cellmap[n];
      if (b >= 0) {
        u1 = a.band_u[b];
        u2 = a.band_u[a.cap + b];
        u3 = a.band_u[2 * a.cap + b];
        in_band = true;
      }
    }
  }
}

// derivative of velocity component c along axis ax at (x,y,z) from the full
// field uf: central inside, one-sided second order at the block's first and
// last cell, times s (center_diff / onesid_diff, FluidDomain.f90:1425-1436)
__device__ __forceinline__ double les_diff(const CollideArgs &a, int x, int y, int z, int ax, int c, double s) {
  const double *u = a.uf + (size_t)c * (size_t)a.N;
  const int i = ax == 0 ? x : (ax == 1 ? y : z);
  const int nd = ax == 0 ? a.xd : (ax == 1 ? a.yd : a.zd);
  const long long st = ax == 0 ? (long long)a.yd * a.zd : (ax == 1 ? (long long)a.zd : 1LL);
  const long long n = ((long long)x * a.yd + y) * a.zd + z;
  if (i > 0 && i < nd - 1) return (u[n + st] - u[n - st]) * s;
  const long long d = i == 0 ? st : -st;
  return (-3.0 * u[n] + 4.0 * u[n + d] - u[n + 2 * d]) * s;
}

// plane index of a cell on face `face` in the reference's (outer, inner) loop order
__device__ __forceinline__ int face_plane_index(int face, int x, int y, int z, int yd, int zd) {
  const int axis = face >> 1;
  if (axis == 0) return y * zd + z;
  if (axis == 1) return x * zd + z;
  return x * yd + y;
}

// One cell: load its 19 pre-collision populations, collide, push the 19
// post-collision values to the neighbours (x+e_q, periodic wrap).
template <int MODEL, bool FORCE, bool BAND, bool AUX, bool NT>
__device__ __forceinline__ void collide_push_cell(const CollideArgs &a, int x, int y, int z) {
  // 32-bit cell index and byte offsets (N < 2^29 for every block this library
  // accepts), so each access is a scalar plane base plus one 32-bit offset
  const unsigned n = ((unsigned)x * (unsigned)a.yd + (unsigned)y) * (unsigned)a.zd + (unsigned)z;
  const size_t plane = (size_t)a.N * sizeof(double);
  const unsigned nb = n << 3;

  double f[Q];
  load_cell(f, a.fin, plane, nb, NT);

  double rho, m1, m2, m3;
  moments(f, rho, m1, m2, m3);
  double u1 = (m1 + a.hv[0]) / rho;
  double u2 = (m2 + a.hv[1]) / rho;
  double u3 = (m3 + a.hv[2]) / rho;
  bool forced = FORCE;
  double F1 = a.VF[0], F2 = a.VF[1], F3 = a.VF[2];
  if constexpr (BAND) {
    if (a.ctl[kBandErrWord] == kBandNoErr && x >= a.ctl[0] && x <= a.ctl[3] && y >= a.ctl[1] && y <= a.ctl[4] && z >= a.ctl[2] &&
        z <= a.ctl[5]) {
      const int b = a.cellmap[n];
      if (b >= 0) {
        u1 = a.band_u[b];
        u2 = a.band_u[a.cap + b];
        u3 = a.band_u[2 * a.cap + b];
        F1 = a.band_F[b];
        F2 = a.band_F[a.cap + b];
        F3 = a.band_F[2 * a.cap + b];
        forced = true;
      }
    }
  }
  if constexpr (AUX) {
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      if (!((a.aux_face_mask102 >> face) & 1)) continue;
      const int axis = face >> 1, hi = face & 1;
      const int c = axis == 0 ? x : (axis == 1 ? y : z);
      const int nd = axis == 0 ? a.xd : (axis == 1 ? a.yd : a.zd);
      if (c == (hi ? nd - 2 : 1)) {
        const int p = face_plane_index(face, x, y, z, a.yd, a.zd);
        const int np = axis == 0 ? a.yd * a.zd : (axis == 1 ? a.xd * a.zd : a.xd * a.yd);
        a.nb_den[face][p] = rho;
        a.nb_u[face][p] = u1;
        a.nb_u[face][np + p] = u2;
        a.nb_u[face][2 * np + p] = u3;
      }
    }
  }

  // equilibrium difference fEq - f  (FluidDomain.f90:1217-1220)
  double uSqr = u1 * u1;
  uSqr = uSqr + u2 * u2;
  uSqr = uSqr + u3 * u3;
  const double base = 1.0 - 1.5 * uSqr;
  const double dt3 = 3.0 * a.dh;
  double post[Q];
#define FSIL_FEQ(q)                                                       \
  const double ux##q = dotu<q>(u1, u2, u3);                               \
  const double fe##q = (wt(q) * rho) * (base + ux##q * (3.0 + 4.5 * ux##q)) - f[q]; \
  const double fl##q = forced ? guo_force<q>(dt3, ux##q, u1, u2, u3, F1, F2, F3) : 0.0;
  FSIL_FEQ(0) FSIL_FEQ(1) FSIL_FEQ(2) FSIL_FEQ(3) FSIL_FEQ(4) FSIL_FEQ(5) FSIL_FEQ(6)
  FSIL_FEQ(7) FSIL_FEQ(8) FSIL_FEQ(9) FSIL_FEQ(10) FSIL_FEQ(11) FSIL_FEQ(12) FSIL_FEQ(13)
  FSIL_FEQ(14) FSIL_FEQ(15) FSIL_FEQ(16) FSIL_FEQ(17) FSIL_FEQ(18)
#undef FSIL_FEQ
  const double fe[Q] = {fe0, fe1, fe2, fe3, fe4, fe5, fe6, fe7, fe8, fe9,
                        fe10, fe11, fe12, fe13, fe14, fe15, fe16, fe17, fe18};
  const double fl[Q] = {fl0, fl1, fl2, fl3, fl4, fl5, fl6, fl7, fl8, fl9,
                        fl10, fl11, fl12, fl13, fl14, fl15, fl16, fl17, fl18};
  if constexpr (MODEL == 1) {
    // SRT  (FluidDomain.f90:1227)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double t = f[q] + a.Omega * fe[q];
      post[q] = forced ? t + a.c1 * fl[q] : t;
    }
  } else if constexpr (MODEL == 2) {
    // TRT  (FluidDomain.f90:1230-1235)
    constexpr int P[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
    constexpr int M[9] = {2, 4, 6, 10, 9, 14, 13, 18, 17};
    {
      const double t = a.Omega * fe[0];
      post[0] = f[0] + (forced ? t + a.c1 * fl[0] : t);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int p = P[k], m = M[k];
      double tp = a.hO * (fe[p] + fe[m]);
      double tm = a.hO2 * (fe[p] - fe[m]);
      if (forced) {
        tp = tp + a.cT1 * (fl[p] + fl[m]);
        tm = tm + a.cT2 * (fl[p] - fl[m]);
      }
      post[p] = f[p] + (tp + tm);
      post[m] = f[m] + (tp - tm);
    }
  } else if constexpr (MODEL == 11 || MODEL == 14 || MODEL == 15) {
    // LES: SRT with a per-cell omega (FluidDomain.f90:1239-1258); the
    // non-equilibrium moments from fneq = -(fEq - f) (smag/WALE, :1271-1277)
    double fn[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) fn[q] = -fe[q];
    const double Q11 = fn[1] + fn[2] + fn[7] + fn[8] + fn[9] + fn[10] + fn[11] + fn[12] + fn[13] + fn[14];
    const double Q22 = fn[3] + fn[4] + fn[7] + fn[8] + fn[9] + fn[10] + fn[15] + fn[16] + fn[17] + fn[18];
    const double Q33 = fn[5] + fn[6] + fn[11] + fn[12] + fn[13] + fn[14] + fn[15] + fn[16] + fn[17] + fn[18];
    const double Q12 = fn[7] - fn[8] - fn[9] + fn[10];
    const double Q13 = fn[11] - fn[12] - fn[13] + fn[14];
    const double Q23 = fn[15] - fn[16] - fn[17] + fn[18];
    double omega;
    if constexpr (MODEL == 11) {
      // smag (:1265-1280)
      const double Qs = Q11 * Q11 + Q22 * Q22 + Q33 * Q33 + 2.0 * (Q12 * Q12 + Q13 * Q13 + Q23 * Q23);
      const double tau_t = sqrt(a.tau * a.tau + a.les_c * sqrt(Qs) / rho);
      a.tau_all[n] = 0.5 * (a.tau + tau_t);
      omega = 2.0 / (a.tau + tau_t);
    } else if constexpr (MODEL == 14) {
      // WALE (:1308-1424); the reference's `invdh` holds dh
      const double s = a.dh, tp = a.tau_all[n];
      const double S11 = -1.5 * s * Q11 / (rho * tp), S22 = -1.5 * s * Q22 / (rho * tp);
      const double S33 = -1.5 * s * Q33 / (rho * tp), S12 = -1.5 * s * Q12 / (rho * tp);
      const double S13 = -1.5 * s * Q13 / (rho * tp), S23 = -1.5 * s * Q23 / (rho * tp);
      const double S = S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23);
      double ox = les_diff(a, x, y, z, 1, 2, s);
      ox = 0.5 * (ox - les_diff(a, x, y, z, 2, 1, s));
      double oy = les_diff(a, x, y, z, 2, 0, s);
      oy = 0.5 * (oy - les_diff(a, x, y, z, 0, 2, s));
      double oz = les_diff(a, x, y, z, 0, 1, s);
      oz = 0.5 * (oz - les_diff(a, x, y, z, 1, 0, s));
      const double O12 = -0.5 * oz, O13 = 0.5 * oy, O23 = -0.5 * ox;
      const double O = 2.0 * (O12 * O12 + O23 * O23 + O13 * O13);
      const double SO11 = -(0.0 + S11 * S11 * O12 * O12 + S11 * S11 * O13 * O13 + 0.0 + S12 * S12 * O12 * O12 +
                            S12 * S12 * O13 * O13 + 0.0 + S13 * S13 * O12 * O12 + S13 * S13 * O13 * O13);
      const double SO22 = -(S12 * S12 * O12 * O12 + 0.0 + S12 * S12 * O23 * O23 + S22 * S22 * O12 * O12 + 0.0 +
                            S22 * S22 * O23 * O23 + S23 * S23 * O12 * O12 + 0.0 + S23 * S23 * O23 * O23);
      const double SO33 = -(S13 * S13 * O13 * O13 + S13 * S13 * O23 * O23 + 0.0 + S23 * S23 * O13 * O13 +
                            S23 * S23 * O23 * O23 + 0.0 + S33 * S33 * O13 * O13 + S33 * S33 * O23 * O23 + 0.0);
      const double SO12 = -(0.0 + 0.0 + S11 * S12 * O13 * O23 + 0.0 + 0.0 + S12 * S22 * O13 * O23 + 0.0 + 0.0 +
                            S13 * S23 * O13 * O23);
      const double SO13 = (0.0 + S11 * S13 * O12 * O23 + 0.0 + 0.0 + S12 * S23 * O12 * O23 + 0.0 + 0.0 +
                           S13 * S33 * O12 * O23 + 0.0);
      const double SO23 = -(S12 * S13 * O12 * O13 + 0.0 + 0.0 + S22 * S23 * O12 * O13 + 0.0 + 0.0 +
                            S23 * S33 * O12 * O13 + 0.0 + 0.0);
      const double SO = SO11 + SO22 + SO33 + 2.0 * (SO12 + SO13 + SO23);
      const double SdSd = (S * S + O * O) / 6.0 + 2.0 * S * O / 3.0 + 2.0 * SO;
      double OP = pow(SdSd, 1.5) / (pow(S, 2.5) + pow(SdSd, 1.25));
      if (!isfinite(OP) || OP < 0.0) OP = 0.0;
      const double tn = (a.nu + a.les_c * OP * a.dh * a.dh) / (a.dh * (1.0 / 3.0)) + 0.5;
      a.tau_all[n] = tn;
      omega = 1.0 / tn;
    } else {
      // vrem (:1437-1507)
      const double s = a.dh;
      const double a11 = 0.5 * les_diff(a, x, y, z, 0, 0, s), a21 = 0.5 * les_diff(a, x, y, z, 0, 1, s);
      const double a31 = 0.5 * les_diff(a, x, y, z, 0, 2, s), a12 = 0.5 * les_diff(a, x, y, z, 1, 0, s);
      const double a22 = 0.5 * les_diff(a, x, y, z, 1, 1, s), a32 = 0.5 * les_diff(a, x, y, z, 1, 2, s);
      const double a13 = 0.5 * les_diff(a, x, y, z, 2, 0, s), a23 = 0.5 * les_diff(a, x, y, z, 2, 1, s);
      const double a33 = 0.5 * les_diff(a, x, y, z, 2, 2, s);
      const double b11 = a11 * a11, b12 = a12 * a12, b13 = a13 * a13, b21 = a21 * a21, b22 = a22 * a22;
      const double b23 = a23 * a23, b31 = a31 * a31, b32 = a32 * a32, b33 = a33 * a33;
      const double aa = b11 + b12 + b13 + b21 + b22 + b23 + b31 + b32 + b33;
      const double c12 = a11 * a21 + a12 * a22 + a13 * a23, c13 = a11 * a31 + a12 * a32 + a13 * a33;
      const double c23 = a21 * a31 + a22 * a32 + a23 * a33;
      const double bb = (b11 + b12 + b13) * (b21 + b22 + b23) - c12 * c12 + (b11 + b12 + b13) * (b31 + b32 + b33) -
                        c13 * c13 + (b21 + b22 + b23) * (b31 + b32 + b33) - c23 * c23;
      double OP = sqrt(bb / aa);
      if (!isfinite(OP)) OP = 0.0;
      const double tn = (a.nu + a.les_c * OP * a.dh * a.dh) / (a.dh * (1.0 / 3.0)) + 0.5;
      a.tau_all[n] = tn;
      omega = 1.0 / tn;
    }
    const double cf = 1.0 - 0.5 * omega;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const double t = f[q] + omega * fe[q];
      post[q] = forced ? t + cf * fl[q] : t;
    }
  } else {
    // MRT  (FluidDomain.f90:1238): f + MATMUL(M_COLLID,fEq) + MATMUL(M_FORCE,Flb)
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      double s = 0.0;
      for (int j = 0; j < Q; ++j) s = s + a.Mc[j * Q + i] * fe[j];
      double t = f[i] + s;
      if (forced) {
        double r = 0.0;
        for (int j = 0; j < Q; ++j) r = r + a.Mf[j * Q + i] * fl[j];
        t = t + r;
      }
      post[i] = t;
    }
  }

  if constexpr (AUX) {
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      if (!((a.aux_face_mask_hw >> face) & 1)) continue;
      const int axis = face >> 1, hi = face & 1;
      const int c = axis == 0 ? x : (axis == 1 ? y : z);
      const int nd = axis == 0 ? a.xd : (axis == 1 ? a.yd : a.zd);
      if (c == (hi ? nd - 1 : 0)) {
        const int p = face_plane_index(face, x, y, z, a.yd, a.zd);
        const int np = axis == 0 ? a.yd * a.zd : (axis == 1 ? a.xd * a.zd : a.xd * a.yd);
        for (int q = 0; q < Q; ++q) a.hw[face][(long long)q * np + p] = post[q];
      }
    }
  }

  // push-stream with periodic wrap in every direction (FluidDomain.f90:1514-1625):
  // destination offset = n + per-axis step, the step wrapping at the block edge
  const unsigned yz = (unsigned)a.yd * (unsigned)a.zd, zd = (unsigned)a.zd;
  const unsigned sxp = x + 1 == a.xd ? 0u - (unsigned)(a.xd - 1) * yz : yz;
  const unsigned sxm = x == 0 ? (unsigned)(a.xd - 1) * yz : 0u - yz;
  const unsigned syp = y + 1 == a.yd ? 0u - (unsigned)(a.yd - 1) * zd : zd;
  const unsigned sym = y == 0 ? (unsigned)(a.yd - 1) * zd : 0u - zd;
  const unsigned szp = z + 1 == a.zd ? 0u - (unsigned)(a.zd - 1) : 1u;
  const unsigned szm = z == 0 ? (unsigned)(a.zd - 1) : 0u - 1u;
  const unsigned st[6] = {sxp, sxm, syp, sym, szp, szm};
  store_cell(post, a.fout, plane, n, st, NT);
}

template <int MODEL, bool FORCE, bool BAND, bool AUX, bool NT>
__global__ __launch_bounds__(256) void k_collide_push(CollideArgs a) {
  if (a.region == kRegionHull) {
    for_hull_cells(a.ctl, a.hole, a.xd, a.yd, a.zd, [&](int x, int y, int z) {
      if (z < a.zd && y < a.yd) collide_push_cell<MODEL, FORCE, BAND, AUX, NT>(a, x, y, z);
    });
    return;
  }
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int x = blockIdx.z;
  if (z >= a.zd || y >= a.yd) return;
  if (a.region == kRegionBulk && in_hole(a.hole, x, y)) return;
  collide_push_cell<MODEL, FORCE, BAND, AUX, NT>(a, x, y, z);
}

// Velocity field the LES models 14 and 15 differentiate: every cell's u as
// the collision reads it (uuu after calculate_macro_quantities_ and the
// immersed-boundary correction), into a.uf [3][N] before the fused kernel runs.
template <bool BAND>
__global__ __launch_bounds__(256) void k_les_velocity(CollideArgs a) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int x = blockIdx.z;
  if (z >= a.zd || y >= a.yd) return;
  const unsigned n = ((unsigned)x * (unsigned)a.yd + (unsigned)y) * (unsigned)a.zd + (unsigned)z;
  double f[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = a.fin[(size_t)q * a.N + n];
  double rho, u1, u2, u3;
  bool in_band;
  int b;
  cell_velocity<BAND>(a, f, n, x, y, z, rho, u1, u2, u3, in_band, b);
  double *uf = const_cast<double *>(a.uf);
  uf[n] = u1;
  uf[(size_t)a.N + n] = u2;
  uf[2 * (size_t)a.N + n] = u3;
}

// every cell set to v
__global__ __launch_bounds__(256) void k_fill_value(double *__restrict__ p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// macroscopic fields of the stored state (FluidDomain.f90:1128-1145)
__global__ __launch_bounds__(256) void k_macro(const double *__restrict__ f, long long N,
                                               double hv0, double hv1, double hv2,
                                               double *__restrict__ den, double *__restrict__ u) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) g[q] = f[q * N + n];
  double rho, m1, m2, m3;
  moments(g, rho, m1, m2, m3);
  den[n] = rho;
  u[n] = (m1 + hv0) / rho;
  u[N + n] = (m2 + hv1) / rho;
  u[2 * N + n] = (m3 + hv2) / rho;
}

// ---------------------------------------------------------------------------
// Boundary faces (FluidDomain.f90:616-1126), one launch per face in the order
// x-, x+, y-, y+, z-, z+ so that edges and corners see earlier faces' values.
struct FaceArgs {
  double *__restrict__ f;   // FP64 plain state, or
  float *__restrict__ g;    // FP32-deviation state (g = f - w)
  int xd, yd, zd;
  long long N;
  int face, code;
  double dh, xmin, ymin, zmin, wallc, hwc;
  double denIn, uvw[3], shear[3];
  int vkind;        // 0 shear profile, 2 oscillatory (velocity given in uvw_osc)
  double uvw_osc[3];
  const double *__restrict__ nb_den;  // 102: den of first interior plane
  const double *__restrict__ nb_u;    // 102: u of first interior plane
  const double *__restrict__ hw;      // 203/204: saved post-collision plane (nullptr: skip)
};

// evaluate_velocity (FluidDomain.f90:1791-1824) at node (xc,yc,zc)
__device__ __forceinline__ void eval_velocity(const FaceArgs &a, double xc, double yc,
                                              double zc, double v[3]) {
  if (a.vkind == 0) {
    v[0] = ((a.uvw[0] + 0.0 * a.shear[0]) + yc * a.shear[1]) + zc * a.shear[2];
    v[1] = ((a.uvw[1] + xc * a.shear[0]) + 0.0 * a.shear[1]) + zc * a.shear[2];
    v[2] = ((a.uvw[2] + xc * a.shear[0]) + yc * a.shear[1]) + 0.0 * a.shear[2];
  } else {
    v[0] = a.uvw_osc[0];
    v[1] = a.uvw_osc[1];
    v[2] = a.uvw_osc[2];
  }
}

// calculate_distribution_funcion (FluidDomain.f90:1827-1834), one direction
template <int q>
__device__ __forceinline__ double feq_bc(double rho, double v1, double v2, double v3,
                                         double uSqr) {
  const double ux = dotu<q>(v1, v2, v3);
  return (wt(q) * rho) * (((1.0 + 3.0 * ux) + (4.5 * ux) * ux) - 1.5 * uSqr);
}

template <int q>
__device__ __forceinline__ double mw_term(double rho, double v1, double v2, double v3) {
  // 2.0*wt*density*uxyz*3.0  (FluidDomain.f90:1842)
  const double ux = dotu<q>(v1, v2, v3);
  return (((2.0 * wt(q)) * rho) * ux) * 3.0;
}

template <int FACE>
struct FaceSets;
template <> struct FaceSets<0> { static constexpr int in[5] = {1, 7, 9, 11, 13}; static constexpr int sym[5] = {2, 8, 10, 12, 14}; };
template <> struct FaceSets<1> { static constexpr int in[5] = {2, 8, 10, 12, 14}; static constexpr int sym[5] = {1, 7, 9, 11, 13}; };
template <> struct FaceSets<2> { static constexpr int in[5] = {3, 7, 8, 15, 17}; static constexpr int sym[5] = {4, 9, 10, 16, 18}; };
template <> struct FaceSets<3> { static constexpr int in[5] = {4, 9, 10, 16, 18}; static constexpr int sym[5] = {3, 7, 8, 15, 17}; };
template <> struct FaceSets<4> { static constexpr int in[5] = {5, 11, 12, 15, 16}; static constexpr int sym[5] = {6, 13, 14, 17, 18}; };
template <> struct FaceSets<5> { static constexpr int in[5] = {6, 13, 14, 17, 18}; static constexpr int sym[5] = {5, 11, 12, 15, 16}; };

template <int q>
__device__ __forceinline__ double feq_dir(double rho, double v1, double v2, double v3, double uSqr) {
  return feq_bc<q>(rho, v1, v2, v3, uSqr);
}

// runtime-q dispatch over the compile-time feq helpers
__device__ __forceinline__ double feq_q(int q, double rho, double v1, double v2, double v3,
                                        double uSqr) {
  switch (q) {
#define C(k) case k: return feq_bc<k>(rho, v1, v2, v3, uSqr);
    C(0) C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16) C(17) C(18)
#undef C
  }
  return 0.0;
}
__device__ __forceinline__ double mw_q(int q, double rho, double v1, double v2, double v3) {
  switch (q) {
#define C(k) case k: return mw_term<k>(rho, v1, v2, v3);
    C(0) C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16) C(17) C(18)
#undef C
  }
  return 0.0;
}

// Population access of the face kernel in either storage form; the
// deviation form converts through f = g + w, so the copies of the wall,
// symmetry and extrapolation codes return the stored deviations unchanged.
template <bool F32>
struct FaceIO {
  __device__ static __forceinline__ double ld(const FaceArgs &a, int q, long long i) {
    if constexpr (F32) return (double)a.g[i] + wt(q);
    else return a.f[i];
  }
  __device__ static __forceinline__ void st(const FaceArgs &a, int q, long long i, double v) {
    if constexpr (F32) a.g[i] = (float)(v - wt(q));
    else a.f[i] = v;
  }
};

// one boundary node t of face FACE
template <int FACE, bool F32>
__device__ __forceinline__ void face_node(const FaceArgs &a, int t) {
  using IO = FaceIO<F32>;
  constexpr int axis = FACE >> 1, hi = FACE & 1;
  const int nd0 = axis == 0 ? a.yd : a.xd;             // outer in-plane extent
  const int nd1 = axis == 2 ? a.yd : a.zd;             // inner in-plane extent
  if (t >= nd0 * nd1) return;
  const int A = t / nd1, B = t - A * nd1;
  const int nda = axis == 0 ? a.xd : (axis == 1 ? a.yd : a.zd);
  const int pb = hi ? nda - 1 : 0, p2 = hi ? nda - 2 : 1, p3 = hi ? nda - 3 : 2;
  int c[3], c2[3], c3[3];
  c[axis] = pb; c2[axis] = p2; c3[axis] = p3;
  const int oa = axis == 0 ? 1 : 0, ia = axis == 2 ? 1 : 2;
  c[oa] = c2[oa] = c3[oa] = A;
  c[ia] = c2[ia] = c3[ia] = B;
  const long long N = a.N;
  const long long n = ((long long)c[0] * a.yd + c[1]) * a.zd + c[2];
  const long long n2 = ((long long)c2[0] * a.yd + c2[1]) * a.zd + c2[2];
  const long long n3 = ((long long)c3[0] * a.yd + c3[1]) * a.zd + c3[2];
  double co[3] = {a.xmin + a.dh * (double)c[0], a.ymin + a.dh * (double)c[1],
                  a.zmin + a.dh * (double)c[2]};
  co[axis] = a.wallc;
  constexpr const int *S = FaceSets<FACE>::in;
  switch (a.code) {
    case BC_EQ_DIRICHLET: {
      double v[3];
      eval_velocity(a, co[0], co[1], co[2], v);
      double uSqr = v[0] * v[0];
      uSqr = uSqr + v[1] * v[1];
      uSqr = uSqr + v[2] * v[2];
#pragma unroll
      for (int q = 0; q < Q; ++q) IO::st(a, q, q * N + n, feq_q(q, a.denIn, v[0], v[1], v[2], uSqr));
      break;
    }
    case BC_NEQ_DIRICHLET: {
      double v[3];
      eval_velocity(a, co[0], co[1], co[2], v);
      double uSqr = v[0] * v[0];
      uSqr = uSqr + v[1] * v[1];
      uSqr = uSqr + v[2] * v[2];
      const int np = nd0 * nd1;
      const double r2 = a.nb_den[t];
      const double w1 = a.nb_u[t], w2 = a.nb_u[np + t], w3 = a.nb_u[2 * np + t];
      double wSqr = w1 * w1;
      wSqr = wSqr + w2 * w2;
      wSqr = wSqr + w3 * w3;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int q = S[k];
        const double fe = feq_q(q, a.denIn, v[0], v[1], v[2], uSqr);
        const double fi = feq_q(q, r2, w1, w2, w3, wSqr);
        IO::st(a, q, q * N + n, fe + (IO::ld(a, q, q * N + n2) - fi));
      }
      break;
    }
    case BC_EXTRAP1:
#pragma unroll
      for (int k = 0; k < 5; ++k) IO::st(a, S[k], S[k] * N + n, IO::ld(a, S[k], S[k] * N + n2));
      break;
    case BC_EXTRAP2:
#pragma unroll
      for (int k = 0; k < 5; ++k)
        IO::st(a, S[k], S[k] * N + n, 2.0 * IO::ld(a, S[k], S[k] * N + n2) - IO::ld(a, S[k], S[k] * N + n3));
      break;
    case BC_WALL: {
      double tmp[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) tmp[k] = IO::ld(a, oppo(S[k]), oppo(S[k]) * N + n);
#pragma unroll
      for (int k = 0; k < 5; ++k) IO::st(a, S[k], S[k] * N + n, tmp[k]);
      break;
    }
    case BC_WALL_HALFWAY: {
      if (!a.hw) break;
      const int np = nd0 * nd1;
      double tmp[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) tmp[k] = a.hw[(long long)oppo(S[k]) * np + t];
#pragma unroll
      for (int k = 0; k < 5; ++k) IO::st(a, S[k], S[k] * N + n, tmp[k]);
      break;
    }
    case BC_MOVING_WALL: {
      double v[3];
      eval_velocity(a, co[0], co[1], co[2], v);
      double tmp[5];
#pragma unroll
      for (int k = 0; k < 5; ++k)
        tmp[k] = IO::ld(a, oppo(S[k]), oppo(S[k]) * N + n) + mw_q(S[k], a.denIn, v[0], v[1], v[2]);
#pragma unroll
      for (int k = 0; k < 5; ++k) IO::st(a, S[k], S[k] * N + n, tmp[k]);
      break;
    }
    case BC_MOVING_WALL_HALFWAY: {
      if (!a.hw) break;
      co[axis] = a.hwc;
      double v[3];
      eval_velocity(a, co[0], co[1], co[2], v);
      const int np = nd0 * nd1;
      double tmp[5];
#pragma unroll
      for (int k = 0; k < 5; ++k)
        tmp[k] = a.hw[(long long)oppo(S[k]) * np + t] + mw_q(S[k], a.denIn, v[0], v[1], v[2]);
#pragma unroll
      for (int k = 0; k < 5; ++k) IO::st(a, S[k], S[k] * N + n, tmp[k]);
      break;
    }
    case BC_SYMMETRIC: {
      constexpr const int *M = FaceSets<FACE>::sym;
      double tmp[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) tmp[k] = IO::ld(a, M[k], M[k] * N + n);
#pragma unroll
      for (int k = 0; k < 5; ++k) IO::st(a, S[k], S[k] * N + n, tmp[k]);
      break;
    }
    default:
      break;
  }
}

template <int FACE, bool F32>
__global__ __launch_bounds__(256) void k_face(FaceArgs a) {
  face_node<FACE, F32>(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// Both faces of one axis in one launch (workgroups [0, nblk) the low face,
// the rest the high face).  A face writes only its boundary plane and reads
// that plane and the next two inward, so the two faces of an axis at least
// six nodes long touch disjoint planes and may run together; the axes stay
// in order (edges belong to two axes).
template <int AXIS, bool F32>
__global__ __launch_bounds__(256) void k_face_pair(FaceArgs lo, FaceArgs hi, int nblk) {
  if ((int)blockIdx.x < nblk) face_node<2 * AXIS, F32>(lo, blockIdx.x * blockDim.x + threadIdx.x);
  else face_node<2 * AXIS + 1, F32>(hi, ((int)blockIdx.x - nblk) * blockDim.x + threadIdx.x);
}

}  // namespace fsil
