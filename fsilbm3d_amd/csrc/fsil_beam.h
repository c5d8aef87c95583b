// Co-rotational 3-D Timoshenko beam (plate bodies) for the device FEM, FP64.
//
// Element routines are __host__ __device__: the host runs the body set-up
// (material, triads, rotated mass matrix) with the C library's libm, the
// device runs the Newmark / Newton / block-Jacobi PCG sub-steps.  Every
// expression keeps the association order of gaoak/FSILBM3D
// src/SolidSolver.f90 (routine citations at each function).  Integer powers
// x**n are evaluated by binary exponentiation, as the reference binary does.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>

namespace fsil {

constexpr int NED = 12;  // element degrees of freedom

struct BeamElem {
  int node0, node1, Nspan, pad;
  int bc[NED];
  double x00[NED], x0[NED], x1[NED], xnxt[NED];
  double dx0, dy0, dz0, len0, dx1, dy1, dz1, len1;
  double xll0, xmm0, xnn0, xll1, xmm1, xnn1;
  double Lspan, spanlen, area00, geoFRM;
  double dirc00[3], dirc0[3];
  double tee[9], tn1[9], tn2[9];  // [row*3 + col]
  double prop[8];
  double mas[NED * NED], stf[NED * NED], geo[NED * NED], coef[NED * NED];
  double stl[NED * NED];  // local stiffness: depends on the constant properties and len0 only
  double rot[9];
  double fint[NED], fload[NED], fdamp[NED];  // per-element contributions (internal force, loads)
  double gsc[9];  // geometric-stiffness coefficients gy1..gy4, gz1..gz4, gt (el_geom_scalars)
  double ksc[10];  // the distinct entries of stl (StiffPattern): EA/L, ky1..ky4, kz1..kz4, GJ/L
  static constexpr bool kFull = true;
};

// The fields of an element that a flexible body's Newton iterations read and
// write, kept in LDS by the device sub-step kernel for the whole launch (the
// element routines take either struct).  Without stiffness damping (which
// reads the full rotated stiffness) and without the rotated mass matrix's
// zero blocks: masb holds its diagonal 3x3 blocks 0 and 1, which blocks 2 and
// 3 repeat (the same diagonal rotated by the same triad).  An odd number of
// 8-byte words, so that one element per lane reads distinct LDS banks.
struct BeamHot {
  int node0, node1;
  double dx0, dy0, dz0, len0, dx1, dy1, dz1, len1;
  double xll1, xmm1, xnn1, geoFRM;
  double prop[8];
  double tee[9], tn1[9], tn2[9];
  double gsc[9], ksc[10];
  double masb[18];
  double x1[NED], fint[NED], fload[NED];
  static constexpr bool kFull = false;
};
constexpr int kHotWords = (int)(sizeof(BeamHot) / sizeof(double));
static_assert(sizeof(BeamHot) % sizeof(double) == 0 && (kHotWords & 1) == 1, "odd word count");

// entry (i, j), in a diagonal 3x3 block, of the rotated mass matrix
__host__ __device__ inline double mas_entry(const BeamElem &e, int i, int j) { return e.mas[i * NED + j]; }
__host__ __device__ inline double mas_entry(const BeamHot &e, int i, int j) {
  return e.masb[((i / 3) & 1) * 9 + (i % 3) * 3 + j % 3];
}

// word w of BeamHot <- word src[w] of BeamElem (masb from the mas blocks 0, 1)
struct HotMap {
  short src[kHotWords];
  constexpr HotMap() : src() {
    put(offsetof(BeamHot, node0), offsetof(BeamElem, node0), 1);
    put(offsetof(BeamHot, dx0), offsetof(BeamElem, dx0), 8);
    put(offsetof(BeamHot, xll1), offsetof(BeamElem, xll1), 3);
    put(offsetof(BeamHot, geoFRM), offsetof(BeamElem, geoFRM), 1);
    put(offsetof(BeamHot, prop), offsetof(BeamElem, prop), 8);
    put(offsetof(BeamHot, tee), offsetof(BeamElem, tee), 9);
    put(offsetof(BeamHot, tn1), offsetof(BeamElem, tn1), 9);
    put(offsetof(BeamHot, tn2), offsetof(BeamElem, tn2), 9);
    put(offsetof(BeamHot, gsc), offsetof(BeamElem, gsc), 9);
    put(offsetof(BeamHot, ksc), offsetof(BeamElem, ksc), 10);
    put(offsetof(BeamHot, x1), offsetof(BeamElem, x1), NED);
    put(offsetof(BeamHot, fint), offsetof(BeamElem, fint), NED);
    put(offsetof(BeamHot, fload), offsetof(BeamElem, fload), NED);
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          src[offsetof(BeamHot, masb) / 8 + b * 9 + r * 3 + c] =
              (short)(offsetof(BeamElem, mas) / 8 + (3 * b + r) * NED + 3 * b + c);
  }
  constexpr void put(size_t h, size_t e, int n) {
    for (int k = 0; k < n; ++k) src[h / 8 + k] = (short)(e / 8 + k);
  }
};

// 0 when the rotated mass matrix's diagonal blocks 2 and 3 equal blocks 0 and
// 1 bit for bit, else 1 + the first entry that differs
__host__ inline int el_mass_blocks_repeat(const BeamElem &e) {
  for (int b = 0; b < 2; ++b)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const double *u = &e.mas[(3 * b + r) * NED + 3 * b + c], *v = &e.mas[(3 * b + 6 + r) * NED + 3 * b + 6 + c];
        if (__builtin_memcmp(u, v, sizeof(double)) != 0) return 1 + (3 * b + 6 + r) * NED + 3 * b + 6 + c;
      }
  return 0;
}

__host__ __device__ inline double bpow(double x, int n) {
  double res = 0.0, cur = x;
  bool have = false;
  while (n) {
    if (n & 1) {
      if (have) res = res * cur;
      else { res = cur; have = true; }
    }
    n >>= 1;
    if (n) cur = cur * cur;
  }
  return have ? res : 1.0;
}

__host__ __device__ inline double bdot3(const double *a, const double *b) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s = s + a[k] * b[k];
  return s;
}

// C = A*B (3x3, k ascending); C may alias A or B
__host__ __device__ inline void bmm3(const double *A, const double *B, double *C) {
  double T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + A[i * 3 + k] * B[k * 3 + j];
      T[i * 3 + j] = s;
    }
  for (int k = 0; k < 9; ++k) C[k] = T[k];
}

__host__ __device__ inline void bmv3(const double *A, const double *v, double *out) {
  double t[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + A[i * 3 + k] * v[k];
    t[i] = s;
  }
  out[0] = t[0]; out[1] = t[1]; out[2] = t[2];
}

// get_angle_triad (SolidSolver.f90:909-952)
__host__ __device__ inline void angle_triad(const double *t11, const double *t22, double *t) {
  double rr[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + t22[i * 3 + k] * t11[j * 3 + k];
      rr[i * 3 + j] = s;
    }
  const double dtx = (rr[7] - rr[5]) / 2.0;
  const double dty = (rr[2] - rr[6]) / 2.0;
  const double dtz = (rr[3] - rr[1]) / 2.0;
  double tr = rr[0] + rr[4] + rr[8];
  tr = (tr - 1.0) / 2.0;
  if (tr > 1.0) tr = 1.0;
  if (tr < -1.0) tr = -1.0;
  const double sint = sqrt(dtx * dtx + dty * dty + dtz * dtz);
  const double theta = acos(tr);
  if (sint < 1.0e-10 || theta < 1.0e-10) {
    t[0] = dtx; t[1] = dty; t[2] = dtz;
  } else {
    const double f = theta / sint;
    t[0] = f * dtx; t[1] = f * dty; t[2] = f * dtz;
  }
}

// AoAtoTTT with its pi constant (SolidSolver.f90:2404-2472)
__host__ __device__ inline void aoa_to_ttt(const double *aoa, double *T) {
  const double pi_ = 3.141562653589793;
  double R[3][9];
  for (int a = 0; a < 3; ++a) {
    double c = cos(aoa[a]), s = sin(aoa[a]);
    if (fabs(aoa[a]) < 1.0e-5) { c = 1.0; s = 0.0; }
    if (fabs(aoa[a] - 0.5 * pi_) < 1.0e-5) { c = 0.0; s = 1.0; }
    if (fabs(aoa[a] + 0.5 * pi_) < 1.0e-5) { c = 0.0; s = -1.0; }
    for (int k = 0; k < 9; ++k) R[a][k] = 0.0;
    if (a == 0) { R[0][0] = 1.0; R[0][4] = c; R[0][7] = s; R[0][5] = -s; R[0][8] = c; }
    else if (a == 1) { R[1][0] = c; R[1][6] = -s; R[1][4] = 1.0; R[1][2] = s; R[1][8] = c; }
    else { R[2][0] = c; R[2][3] = s; R[2][1] = -s; R[2][4] = c; R[2][8] = 1.0; }
  }
  const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  bmm3(R[2], I, T);
  bmm3(R[1], T, T);
  bmm3(R[0], T, T);
}

// Segment_FiniteRot (SolidSolver.f90:1066-1124)
__host__ __device__ inline void finite_rot(double t1, double t2, double t3, double *rr) {
  const double tt = sqrt(bpow(t1, 2) + bpow(t2, 2) + bpow(t3, 2));
  const double ss = sin(tt), cc = cos(tt);
  const double r2[9] = {0.0, -t3, t2, t3, 0.0, -t1, -t2, t1, 0.0};
  double r3[9];
  r3[0] = -t3 * t3 - t2 * t2; r3[1] = t2 * t1; r3[2] = t3 * t1;
  r3[3] = t1 * t2; r3[4] = -t3 * t3 - t1 * t1; r3[5] = t3 * t2;
  r3[6] = t1 * t3; r3[7] = t2 * t3; r3[8] = -t2 * t2 - t1 * t1;
  double c1, c2;
  if (tt < 1.0e-10) { c1 = 1.0; c2 = 0.5; }
  else { c1 = ss / tt; c2 = (1.0 - cc) / bpow(tt, 2); }
  for (int k = 0; k < 9; ++k) {
    const double r1 = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
    rr[k] = r1 + r2[k] * c1 + r3[k] * c2;
  }
}

// Segment_BuildAxisDirTriad (SolidSolver.f90:624-697)
__host__ __device__ inline void axis_dir_triad(double l, double m, double n, const double *d, double *triad) {
  double ex[3] = {l, m, n}, ey[3], ez[3], dir[3] = {d[0], d[1], d[2]};
  double dd = sqrt(bdot3(ex, ex));
  if (dd > 1.0e-14) for (int k = 0; k < 3; ++k) ex[k] = ex[k] / dd;
  dd = sqrt(bdot3(dir, dir));
  if (dd > 1.0e-14) for (int k = 0; k < 3; ++k) dir[k] = dir[k] / dd;
  const double proj = bdot3(dir, ex);
  for (int k = 0; k < 3; ++k) ey[k] = dir[k] - proj * ex[k];
  dd = sqrt(bdot3(ey, ey));
  if (dd <= 1.0e-10) {
    if (fabs(ex[2]) > 0.995) {
      ey[0] = 0.0; ey[1] = 1.0; ey[2] = 0.0;
      ez[0] = -ex[2]; ez[1] = 0.0; ez[2] = 0.0;
    } else {
      dd = sqrt(ex[0] * ex[0] + ex[1] * ex[1]);
      ey[0] = -ex[1] / dd; ey[1] = ex[0] / dd; ey[2] = 0.0;
      ez[0] = -ex[0] * ex[2] / dd; ez[1] = -ex[1] * ex[2] / dd; ez[2] = dd;
    }
  } else {
    for (int k = 0; k < 3; ++k) ey[k] = ey[k] / dd;
    ez[0] = ex[1] * ey[2] - ex[2] * ey[1];
    ez[1] = ex[2] * ey[0] - ex[0] * ey[2];
    ez[2] = ex[0] * ey[1] - ex[1] * ey[0];
    dd = sqrt(bdot3(ez, ez));
    if (dd > 1.0e-14) for (int k = 0; k < 3; ++k) ez[k] = ez[k] / dd;
  }
  for (int k = 0; k < 3; ++k) {
    triad[k * 3 + 0] = ex[k];
    triad[k * 3 + 1] = ey[k];
    triad[k * 3 + 2] = ez[k];
  }
}

// Segment_cptdxyz1 (SolidSolver.f90:119-129)
template <class ET>
__host__ __device__ inline void el_cptdxyz1(ET &e) {
  e.dx1 = e.x1[6] - e.x1[0];
  e.dy1 = e.x1[7] - e.x1[1];
  e.dz1 = e.x1[8] - e.x1[2];
  e.len1 = sqrt(e.dx1 * e.dx1 + e.dy1 * e.dy1 + e.dz1 * e.dz1);
  e.xll1 = e.dx1 / e.len1;
  e.xmm1 = e.dy1 / e.len1;
  e.xnn1 = e.dz1 / e.len1;
}

// Segment_MapReferencePosToCurrent (SolidSolver.f90:1126-1137)
__host__ __device__ inline void el_map_pos(const BeamElem &e, double *out, const double *T, const double *XYZ,
                                           const double *AoA) {
  double a[3], b[3];
  bmv3(T, e.x00, a);
  bmv3(T, e.x00 + 6, b);
  for (int k = 0; k < 3; ++k) {
    out[k] = a[k] + XYZ[k];
    out[3 + k] = AoA[k];
    out[6 + k] = b[k] + XYZ[k];
    out[9 + k] = AoA[k];
  }
}

// Segment_FormMassMatrix (SolidSolver.f90:318-363)
__host__ __device__ inline void el_form_mass(BeamElem &e) {
  const double area = e.prop[2], rho = e.prop[3], ziy = e.prop[6], ziz = e.prop[7];
  for (int k = 0; k < NED * NED; ++k) e.mas[k] = 0.0;
  const double roal = rho * area * e.len0 / 2.0;
  e.mas[0 * NED + 0] = roal; e.mas[1 * NED + 1] = roal; e.mas[2 * NED + 2] = roal;
  e.mas[3 * NED + 3] = roal * (ziy + ziz) / area;
  e.mas[4 * NED + 4] = roal * ziy / area;
  e.mas[5 * NED + 5] = roal * ziz / area;
  for (int k = 0; k < 6; ++k) e.mas[(6 + k) * NED + 6 + k] = e.mas[k * NED + k];
}

// Segment_FormStiffMatrix (SolidSolver.f90:365-478)
__host__ __device__ inline void el_form_stiff(BeamElem &e) {
  const double emod = e.prop[0], gmod = e.prop[1], area = e.prop[2], zix = e.prop[5], ziy = e.prop[6],
               ziz = e.prop[7];
  const double L = e.len0;
  double *K = e.stf;
  for (int k = 0; k < NED * NED; ++k) K[k] = 0.0;
  const double Inv = 1.0 / L;
  const double ksy = 5.0 / 6.0, ksz = 5.0 / 6.0;
  const double phiy = 12.0 * emod * ziz / (ksy * gmod * area * L * L);
  const double phiz = 12.0 * emod * ziy / (ksz * gmod * area * L * L);
  const double ky1 = 12.0 * emod * ziz / (bpow(L, 3) * (1.0 + phiy));
  const double ky2 = 6.0 * emod * ziz / (bpow(L, 2) * (1.0 + phiy));
  const double ky3 = (4.0 + phiy) * emod * ziz / (L * (1.0 + phiy));
  const double ky4 = (2.0 - phiy) * emod * ziz / (L * (1.0 + phiy));
  const double kz1 = 12.0 * emod * ziy / (bpow(L, 3) * (1.0 + phiz));
  const double kz2 = 6.0 * emod * ziy / (bpow(L, 2) * (1.0 + phiz));
  const double kz3 = (4.0 + phiz) * emod * ziy / (L * (1.0 + phiz));
  const double kz4 = (2.0 - phiz) * emod * ziy / (L * (1.0 + phiz));
#define KK(i, j) K[(i) * NED + (j)]
  KK(0, 0) = area * emod * Inv; KK(1, 1) = ky1; KK(2, 2) = kz1; KK(3, 3) = gmod * zix * Inv; KK(4, 4) = kz3;
  KK(5, 5) = ky3;
  for (int k = 0; k < 6; ++k) KK(6 + k, 6 + k) = KK(k, k);
  KK(0, 6) = -KK(0, 0);
  KK(1, 5) = ky2; KK(1, 7) = -ky1; KK(1, 11) = ky2; KK(5, 7) = -ky2; KK(5, 11) = ky4; KK(7, 11) = -ky2;
  KK(2, 4) = -kz2; KK(2, 8) = -kz1; KK(2, 10) = -kz2; KK(4, 8) = kz2; KK(4, 10) = kz4; KK(8, 10) = kz2;
  KK(3, 9) = -KK(3, 3);
  KK(6, 0) = KK(0, 6);
  KK(5, 1) = KK(1, 5); KK(7, 1) = KK(1, 7); KK(11, 1) = KK(1, 11); KK(7, 5) = KK(5, 7); KK(11, 5) = KK(5, 11);
  KK(11, 7) = KK(7, 11);
  KK(4, 2) = KK(2, 4); KK(8, 2) = KK(2, 8); KK(10, 2) = KK(2, 10); KK(8, 4) = KK(4, 8); KK(10, 4) = KK(4, 10);
  KK(10, 8) = KK(8, 10);
  KK(9, 3) = KK(3, 9);
#undef KK
}

// Pattern of Segment_FormStiffMatrix (SolidSolver.f90:365-478): entry (i, j)
// of the local stiffness is 0 (the formed matrix's exact +0), k (ksc[k-1]) or
// -k (its negative, formed by negation); symmetric.  ksc order: EA/L, ky1,
// ky2, ky3, ky4, kz1, kz2, kz3, kz4, GJ/L.
struct StiffPattern {
  signed char c[NED * NED];
  constexpr StiffPattern() : c() {
    const int e[26][3] = {{0, 0, 1},  {6, 6, 1},   {0, 6, -1},  {1, 1, 2},   {7, 7, 2},   {1, 7, -2},  {1, 5, 3},
                          {1, 11, 3}, {5, 7, -3},  {7, 11, -3}, {5, 5, 4},   {11, 11, 4}, {5, 11, 5},  {2, 2, 6},
                          {8, 8, 6},  {2, 8, -6},  {2, 4, -7},  {2, 10, -7}, {4, 8, 7},   {8, 10, 7},  {4, 4, 8},
                          {10, 10, 8}, {4, 10, 9}, {3, 3, 10},  {9, 9, 10},  {3, 9, -10}};
    for (int q = 0; q < 26; ++q) {
      c[e[q][0] * NED + e[q][1]] = (signed char)e[q][2];
      c[e[q][1] * NED + e[q][0]] = (signed char)e[q][2];
    }
  }
};
__device__ constexpr StiffPattern kStiffPatternD{};
constexpr StiffPattern kStiffPatternH{};

__host__ __device__ inline int stiff_code(int i, int j) {
#ifdef __HIP_DEVICE_COMPILE__
  return kStiffPatternD.c[i * NED + j];
#else
  return kStiffPatternH.c[i * NED + j];
#endif
}

// entry (i, j) of the local stiffness from its scalars
__host__ __device__ inline double stiff_entry(const double *ksc, int i, int j) {
  const int k = stiff_code(i, j);
  return k == 0 ? 0.0 : (k > 0 ? ksc[k - 1] : -ksc[-k - 1]);
}

// ksc from the formed stl; 0 when the pattern reproduces every entry of stl
// bit for bit, else 1 + the first entry that differs
__host__ inline int el_stiff_scalars(BeamElem &e) {
  const int at[10][2] = {{0, 0}, {1, 1}, {1, 5}, {5, 5}, {5, 11}, {2, 2}, {4, 8}, {4, 4}, {4, 10}, {3, 3}};
  for (int k = 0; k < 10; ++k) e.ksc[k] = e.stl[at[k][0] * NED + at[k][1]];
  for (int k = 0; k < NED * NED; ++k) {
    const double v = stiff_entry(e.ksc, k / NED, k % NED);
    if (__builtin_memcmp(&v, &e.stl[k], sizeof(double)) != 0) return 1 + k;
  }
  return 0;
}

// Segment_FormGeomMatrix (SolidSolver.f90:480-600): the nine distinct
// coefficients of the geometric stiffness, into e.gsc
template <class ET>
__host__ __device__ inline void el_geom_scalars(ET &e) {
  const double s = e.geoFRM;
  const double emod = e.prop[0], gmod = e.prop[1], area = e.prop[2], zix = e.prop[5], ziy = e.prop[6],
               ziz = e.prop[7];
  const double L = e.len0;
  const double ksy = 5.0 / 6.0, ksz = 5.0 / 6.0;
  const double phiy = 12.0 * emod * ziz / (ksy * gmod * area * L * L);
  const double phiz = 12.0 * emod * ziy / (ksz * gmod * area * L * L);
  const double py2 = bpow(1.0 + phiy, 2), pz2 = bpow(1.0 + phiz, 2);
  double *g = e.gsc;
  g[0] = s / L * (6.0 / 5.0 + 2.0 * phiy + phiy * phiy) / py2;
  g[1] = s / L * (L / 10.0) / py2;
  g[2] = s / L * (2.0 * L * L / 15.0 + phiy * L * L / 6.0 + phiy * phiy * L * L / 12.0) / py2;
  g[3] = s / L * (-(L * L / 30.0) - phiy * L * L / 6.0 - phiy * phiy * L * L / 12.0) / py2;
  g[4] = s / L * (6.0 / 5.0 + 2.0 * phiz + phiz * phiz) / pz2;
  g[5] = s / L * (L / 10.0) / pz2;
  g[6] = s / L * (2.0 * L * L / 15.0 + phiz * L * L / 6.0 + phiz * phiz * L * L / 12.0) / pz2;
  g[7] = s / L * (-(L * L / 30.0) - phiz * L * L / 6.0 - phiz * phiz * L * L / 12.0) / pz2;
  g[8] = s * zix / (area * L);
}

// Pattern of Segment_FormGeomMatrix: entry (i, j) is 0 (zero), k (coefficient
// k-1) or -k (its negative); symmetric.
struct GeoPattern {
  signed char c[NED * NED];
  constexpr GeoPattern() : c() {
    const int e[23][3] = {{1, 1, 1},   {7, 7, 1},   {5, 5, 3},  {11, 11, 3}, {1, 5, 2},  {1, 11, 2},
                          {1, 7, -1},  {5, 7, -2},  {7, 11, -2}, {5, 11, 4}, {2, 2, 5},  {8, 8, 5},
                          {4, 4, 7},   {10, 10, 7}, {2, 4, -6},  {2, 10, -6}, {2, 8, -5}, {4, 8, 6},
                          {8, 10, 6},  {4, 10, 8},  {3, 3, 9},   {9, 9, 9},   {3, 9, -9}};
    for (int q = 0; q < 23; ++q) {
      c[e[q][0] * NED + e[q][1]] = (signed char)e[q][2];
      c[e[q][1] * NED + e[q][0]] = (signed char)e[q][2];
    }
  }
};
__device__ constexpr GeoPattern kGeoPatternD{};
constexpr GeoPattern kGeoPatternH{};

// entry (i, j) of the geometric stiffness from its coefficients g
__host__ __device__ inline double geo_entry(const double *g, int i, int j) {
#ifdef __HIP_DEVICE_COMPILE__
  const int k = kGeoPatternD.c[i * NED + j];
#else
  const int k = kGeoPatternH.c[i * NED + j];
#endif
  return k == 0 ? 0.0 : (k > 0 ? g[k - 1] : -g[-k - 1]);
}

__host__ __device__ inline void el_form_geom(BeamElem &e) {
  el_geom_scalars(e);
  for (int i = 0; i < NED; ++i)
    for (int j = 0; j < NED; ++j) e.geo[i * NED + j] = geo_entry(e.gsc, i, j);
}

// RotateMatrix + RKR (SolidSolver.f90:699-748): ek <- R^T ek R by 3x3 blocks
__host__ __device__ inline void el_rotate(BeamElem &e) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) e.rot[i * 3 + j] = e.tee[j * 3 + i];
}
__host__ __device__ inline void el_rkr(const BeamElem &e, double *ek) {
  const double *r = e.rot;
  double rt[9];
  for (int in = 0; in < 3; ++in)
    for (int jn = 0; jn < 3; ++jn) rt[jn * 3 + in] = r[in * 3 + jn];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int j1 = i * 3, j2 = j * 3;
      double kt[9];
      for (int k = 0; k < 3; ++k)
        for (int ii = 0; ii < 3; ++ii) {
          double s = 0.0;
          for (int jj = 0; jj < 3; ++jj) s = s + ek[(j1 + k) * NED + j2 + jj] * r[jj * 3 + ii];
          kt[k * 3 + ii] = s;
        }
      for (int k = 0; k < 3; ++k)
        for (int ii = 0; ii < 3; ++ii) {
          double s = 0.0;
          for (int jj = 0; jj < 3; ++jj) s = s + rt[k * 3 + jj] * kt[jj * 3 + ii];
          ek[(j1 + k) * NED + j2 + ii] = s;
        }
    }
}

// entry (I, J) of R^T ek R computed as el_rkr computes it (3x3 blocks,
// kt = ek_block * r, then rt * kt, each sum in ascending order)
__host__ __device__ inline double rkr_entry(const double *ek, const double *r, int I, int J) {
  const int j1 = (I / 3) * 3, k = I - j1, j2 = (J / 3) * 3, ii = J - j2;
  double s = 0.0;
  for (int jj = 0; jj < 3; ++jj) {
    double kt = 0.0;
    for (int l = 0; l < 3; ++l) kt = kt + ek[(j1 + jj) * NED + j2 + l] * r[l * 3 + ii];
    s = s + r[jj * 3 + k] * kt;
  }
  return s;
}

// Segment_BodyStress_D (SolidSolver.f90:750-839): local internal force -> e.fint,
// with the local stiffness of Segment_FormStiffMatrix (e.stl, formed once)
template <class ET>
__host__ __device__ inline double el_stress_dl(const ET &e) {
  const double du = e.dx1 - e.dx0, dv = e.dy1 - e.dy0, dw = e.dz1 - e.dz0;
  return ((e.dx0 + e.dx1) * du + (e.dy0 + e.dy1) * dv + (e.dz0 + e.dz1) * dw) / (e.len0 + e.len1);
}
// the local rotation of one end (0: tn1, 1: tn2) relative to the element triad
template <class ET>
__host__ __device__ inline void el_stress_end(const ET &e, int end, double *out) {
  double t[3];
  const double *T = e.tee;
  angle_triad(e.tee, end ? e.tn2 : e.tn1, t);
  out[0] = T[0] * t[0] + T[3] * t[1] + T[6] * t[2];
  out[1] = T[1] * t[0] + T[4] * t[1] + T[7] * t[2];
  out[2] = T[2] * t[0] + T[5] * t[1] + T[8] * t[2];
}
__host__ __device__ inline double el_geo_frm(double dl, const double *prop, double len0) {
  return dl * prop[0] * prop[2] / len0;
}
// the internal force from the local deformation (dl, end rotations a, b)
template <class ET>
__host__ __device__ inline void el_stress_force(ET &e, double dl, const double *a, const double *b) {
  const double *T = e.tee;
  const double ub[NED] = {0.0, 0.0, 0.0, a[0], a[1], a[2], dl, 0.0, 0.0, b[0], b[1], b[2]};
  // fb = stl ub in ascending k.  The terms with a zero stiffness entry, and
  // those with k in {0, 1, 2, 7, 8} where ub is zero, add a signed zero to a
  // sum that starts at +0 and so change nothing; they are left out, and the
  // entries come from the ten scalars rather than the stored matrix.
  double fb[NED];
#pragma unroll
  for (int i = 0; i < NED; ++i) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NED; ++k) {
      if (k == 0 || k == 1 || k == 2 || k == 7 || k == 8) continue;
      if (stiff_code(i, k) == 0) continue;
      s = s + stiff_entry(e.ksc, i, k) * ub[k];
    }
    fb[i] = s;
  }
  for (int i = 0; i < NED; ++i) e.fint[i] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      e.fint[0 + i] = e.fint[0 + i] + T[i * 3 + j] * fb[0 + j];
      e.fint[3 + i] = e.fint[3 + i] + T[i * 3 + j] * fb[3 + j];
      e.fint[6 + i] = e.fint[6 + i] + T[i * 3 + j] * fb[6 + j];
      e.fint[9 + i] = e.fint[9 + i] + T[i * 3 + j] * fb[9 + j];
    }
}
template <class ET>
__host__ __device__ inline void el_body_stress(ET &e) {
  const double dl = el_stress_dl(e);
  double a[3], b[3];
  el_stress_end(e, 0, a);
  el_stress_end(e, 1, b);
  e.geoFRM = el_geo_frm(dl, e.prop, e.len0);
  el_stress_force(e, dl, a, b);
}

// Segment_StrainEnergy_D after FormStiffMatrix (SolidSolver.f90:841-907,
// 2340-2348): stretching (stiffness entries (1,1), (1,7), (7,1), (7,7)) and
// bending/torsion (all others) energy 0.5 ub^T K ub of the local deformation
__host__ __device__ inline void el_strain_energy(const BeamElem &e, double *es, double *eb) {
  const double du = e.dx1 - e.dx0, dv = e.dy1 - e.dy0, dw = e.dz1 - e.dz0;
  const double dl = ((e.dx0 + e.dx1) * du + (e.dy0 + e.dy1) * dv + (e.dz0 + e.dz1) * dw) / (e.len0 + e.len1);
  double t[3], a[3], b[3];
  const double *T = e.tee;
  angle_triad(e.tee, e.tn1, t);
  a[0] = T[0] * t[0] + T[3] * t[1] + T[6] * t[2];
  a[1] = T[1] * t[0] + T[4] * t[1] + T[7] * t[2];
  a[2] = T[2] * t[0] + T[5] * t[1] + T[8] * t[2];
  angle_triad(e.tee, e.tn2, t);
  b[0] = T[0] * t[0] + T[3] * t[1] + T[6] * t[2];
  b[1] = T[1] * t[0] + T[4] * t[1] + T[7] * t[2];
  b[2] = T[2] * t[0] + T[5] * t[1] + T[8] * t[2];
  const double ub[NED] = {0.0, 0.0, 0.0, a[0], a[1], a[2], dl, 0.0, 0.0, b[0], b[1], b[2]};
  double s1 = 0.0, s2 = 0.0;
  for (int i = 0; i < NED; ++i) {
    double x = 0.0, y = 0.0;
    for (int k = 0; k < NED; ++k) {
      const double K = e.stl[i * NED + k];
      const bool stretch = (i == 0 || i == 6) && (k == 0 || k == 6);
      x = x + (stretch ? K : 0.0) * ub[k];
      y = y + (stretch ? 0.0 : K) * ub[k];
    }
    s1 = s1 + x * ub[i];
    s2 = s2 + y * ub[i];
  }
  *es = 0.5 * s1;
  *eb = 0.5 * s2;
}

// Segment_UpdateMatrix (SolidSolver.f90:131-157)
__host__ __device__ inline void el_update_matrix(BeamElem &e, const double *c, double gamma, double dampM,
                                                 double dampK) {
  el_form_geom(e);
  el_rotate(e);
  el_rkr(e, e.stf);
  el_rkr(e, e.geo);
  for (int k = 0; k < NED * NED; ++k) {
    const double tanv = e.stf[k] + gamma * e.geo[k];
    e.coef[k] = tanv + c[0] * e.mas[k] + c[1] * dampM * e.mas[k];
  }
  if (dampK > 0.0)
    for (int k = 0; k < NED * NED; ++k) e.coef[k] = e.coef[k] + c[1] * dampK * e.stf[k];
}

// Segment_UpdateLoad (SolidSolver.f90:159-193) -> e.fload (mass load, then the
// stiffness-damping load added in the same order as two LocToGlobal calls)
template <class ET>
__host__ __device__ inline void el_update_load(ET &e, const double *c, double dampM, double dampK,
                                               const double *dspO, const double *dsp, const double *vel,
                                               const double *acc) {
  double qM[NED], qC[NED], qMC[NED];
  const int n0 = e.node0, n1 = e.node1;
  for (int i = 0; i < 6; ++i) {
    qM[i] = c[0] * (dspO[6 * n0 + i] - dsp[6 * n0 + i]) + c[2] * vel[6 * n0 + i] + c[3] * acc[6 * n0 + i];
    qM[i + 6] = c[0] * (dspO[6 * n1 + i] - dsp[6 * n1 + i]) + c[2] * vel[6 * n1 + i] + c[3] * acc[6 * n1 + i];
    qC[i] = c[1] * (dspO[6 * n0 + i] - dsp[6 * n0 + i]) + c[4] * vel[6 * n0 + i] + c[5] * acc[6 * n0 + i];
    qC[i + 6] = c[1] * (dspO[6 * n1 + i] - dsp[6 * n1 + i]) + c[4] * vel[6 * n1 + i] + c[5] * acc[6 * n1 + i];
  }
  for (int i = 0; i < NED; ++i) qMC[i] = qM[i] + dampM * qC[i];
  for (int i = 0; i < NED; ++i) e.fload[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    e.fload[i] = mas_entry(e, i, i) * qMC[i];
    e.fload[i + 6] = mas_entry(e, i + 6, i + 6) * qMC[i + 6];
  }
  for (int i = 0; i < 3; ++i) {
    double s = 0.0, t = 0.0;
    for (int k = 0; k < 3; ++k) s = s + mas_entry(e, 3 + i, 3 + k) * qMC[3 + k];
    for (int k = 0; k < 3; ++k) t = t + mas_entry(e, 9 + i, 9 + k) * qMC[9 + k];
    e.fload[3 + i] = s;
    e.fload[9 + i] = t;
  }
  if constexpr (ET::kFull) {
    if (dampK > 0.0)
      for (int i = 0; i < NED; ++i) {
        double s = 0.0;
        for (int k = 0; k < NED; ++k) s = s + e.stf[i * NED + k] * qC[k];
        e.fdamp[i] = dampK * s;
      }
  }
}

// Segment_UpdateTriad_D (SolidSolver.f90:966-997) for one end of the
// element: the end's triad rotated by its node's rotation increment
template <class ET>
__host__ __device__ inline void el_update_node_triad(ET &e, const double *dspnn, int end) {
  double rr[9];
  const int n = end ? e.node1 : e.node0;
  finite_rot(dspnn[6 * n + 3], dspnn[6 * n + 4], dspnn[6 * n + 5], rr);
  if (end) bmm3(rr, e.tn2, e.tn2);
  else bmm3(rr, e.tn1, e.tn1);
}
// Segment_MakeTriad_ee (SolidSolver.f90:999-1064) from the updated end triads
template <class ET>
__host__ __device__ inline void el_make_triad_ee(ET &e) {
  double rr[9];
  double *E = e.tee;
  for (int k = 0; k < 9; ++k) E[k] = 0.0;
  E[0] = e.xll1; E[3] = e.xmm1; E[6] = e.xnn1;
  double t[3], aa[9];
  angle_triad(e.tn1, e.tn2, t);
  finite_rot(t[0] / 2.0, t[1] / 2.0, t[2] / 2.0, rr);
  bmm3(rr, e.tn1, aa);
  double r2e1 = 0.0, r3e1 = 0.0;
  for (int k = 0; k < 3; ++k) {
    r2e1 = r2e1 + aa[k * 3 + 1] * E[k * 3 + 0];
    r3e1 = r3e1 + aa[k * 3 + 2] * E[k * 3 + 0];
  }
  for (int j = 0; j < 3; ++j) {
    E[j * 3 + 1] = aa[j * 3 + 1] - r2e1 * (aa[j * 3 + 0] + E[j * 3 + 0]) / 2.0;
    E[j * 3 + 2] = aa[j * 3 + 2] - r3e1 * (aa[j * 3 + 0] + E[j * 3 + 0]) / 2.0;
  }
  const double c1[3] = {E[0], E[3], E[6]};
  double c2[3] = {E[1], E[4], E[7]};
  const double p = bdot3(c2, c1);
  for (int k = 0; k < 3; ++k) E[k * 3 + 1] = E[k * 3 + 1] - p * E[k * 3 + 0];
  for (int k = 0; k < 3; ++k) c2[k] = E[k * 3 + 1];
  const double dd = sqrt(bdot3(c2, c2));
  if (dd > 1.0e-14)
    for (int k = 0; k < 3; ++k) E[k * 3 + 1] = E[k * 3 + 1] / dd;
  const double a0 = E[3] * E[7] - E[6] * E[4];
  const double a1 = E[6] * E[1] - E[0] * E[7];
  const double a2 = E[0] * E[4] - E[3] * E[1];
  E[2] = a0; E[5] = a1; E[8] = a2;
}

// Segment_UpdateTriad_D + Segment_MakeTriad_ee (SolidSolver.f90:966-1064)
template <class ET>
__host__ __device__ inline void el_update_triads(ET &e, const double *dspnn) {
  el_update_node_triad(e, dspnn, 0);
  el_update_node_triad(e, dspnn, 1);
  el_make_triad_ee(e);
}

// Invert6x6 (SolidSolver.f90:2234-2279), Gauss-Jordan with partial pivoting
__host__ __device__ inline int invert6(const double *A, double *Ai) {
  double aug[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) { aug[i][j] = A[i * 6 + j]; aug[i][6 + j] = 0.0; }
  for (int i = 0; i < 6; ++i) aug[i][6 + i] = 1.0;
  for (int i = 0; i < 6; ++i) {
    int p = i;
    double pivot = fabs(aug[i][i]);
    for (int k = i + 1; k < 6; ++k)
      if (fabs(aug[k][i]) > pivot) { pivot = fabs(aug[k][i]); p = k; }
    if (pivot <= 1.0e-30) {
      for (int k = 0; k < 36; ++k) Ai[k] = 0.0;
      return i + 1;
    }
    if (p != i)
      for (int j = 0; j < 12; ++j) { const double t = aug[i][j]; aug[i][j] = aug[p][j]; aug[p][j] = t; }
    pivot = aug[i][i];
    for (int j = 0; j < 12; ++j) aug[i][j] = aug[i][j] / pivot;
    for (int k = 0; k < 6; ++k)
      if (k != i) {
        const double factor = aug[k][i];
        for (int j = 0; j < 12; ++j) aug[k][j] = aug[k][j] - factor * aug[i][j];
      }
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) Ai[i * 6 + j] = aug[i][6 + j];
  return 0;
}

// Invert6x6 as invert6, on a caller-provided 6x12 work area (LDS) whose left
// 6x6 half holds the matrix; the inverse is left in the right half.
__device__ inline int invert6_inplace(double *aug) {
#define AUG(i, j) aug[(i) * 12 + (j)]
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) AUG(i, 6 + j) = 0.0;
  for (int i = 0; i < 6; ++i) AUG(i, 6 + i) = 1.0;
  for (int i = 0; i < 6; ++i) {
    int p = i;
    double pivot = fabs(AUG(i, i));
    for (int k = i + 1; k < 6; ++k)
      if (fabs(AUG(k, i)) > pivot) { pivot = fabs(AUG(k, i)); p = k; }
    if (pivot <= 1.0e-30) return i + 1;
    if (p != i)
      for (int j = 0; j < 12; ++j) { const double t = AUG(i, j); AUG(i, j) = AUG(p, j); AUG(p, j) = t; }
    pivot = AUG(i, i);
    for (int j = 0; j < 12; ++j) AUG(i, j) = AUG(i, j) / pivot;
    for (int k = 0; k < 6; ++k)
      if (k != i) {
        const double factor = AUG(k, i);
        for (int j = 0; j < 12; ++j) AUG(k, j) = AUG(k, j) - factor * AUG(i, j);
      }
  }
#undef AUG
  return 0;
}

}  // namespace fsil
