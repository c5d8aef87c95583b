/*
 * CPU restatement of the reference's co-rotational 3-D Timoshenko beam solver
 * (gaoak/FSILBM3D src/SolidSolver.f90) for plate bodies, in FP64, used ONLY as
 * test infrastructure (the checker of the device FEM and the body kinematics
 * of the oracle's coupled runs).
 *
 * Structure and expression order follow the reference routine by routine:
 *   Segment_Build / Init / cptdxyz1            SolidSolver.f90:60-129
 *   FormMassMatrix / FormStiffMatrix / Geom    :318-600
 *   BuildAxisDirTriad / RotateMatrix / RKR     :624-748
 *   BodyStress_D                               :750-839
 *   get_angle_triad / global_to_local          :909-964
 *   UpdateTriad_D / MakeTriad_ee / FiniteRot   :966-1124
 *   Map*ToCurrent                              :1126-1145
 *   calculate_angle_material (isKB = 1)        :1533-1616
 *   Beam_Initialise, InitTriadANDFormMass      :1401-1510
 *   Beam_structure / UpdateNewmarkCoeffs       :1820-1892
 *   Beam_Solver .. UpdateVelAcc                :1894-2359 (block-Jacobi PCG)
 *   AoAtoTTT                                   :2404-2472
 * Sums run in the reference's loop order; integer powers x**n are evaluated
 * by binary exponentiation (x**3 = x*(x*x), x**5 = x*((x*x)*(x*x))).
 * The fixture tests/golden/fem_clamped_heaving_plate.npz (written by the
 * reference program) pins this file bit for bit through the coupled flow.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NE 12

typedef struct {
  int node0, node1, l2g[NE], bc[NE], Nspan;
  double x00[NE], x0[NE], x1[NE], xnxt[NE];
  double dx0, dy0, dz0, dx1, dy1, dz1, xll0, xmm0, xnn0, xll1, xmm1, xnn1, len0, len1;
  double Lspan, spanlen, dirc00[3], dirc0[3], dirc1[3], dircnxt[3];
  double geoFRM, area00;
  double tee[3][3], tn1[3][3], tn2[3][3]; /* [row][col] = Fortran (i,j) */
  double prop[8];
  double coef[NE][NE], tan_[NE][NE], stf[NE][NE], mas[NE][NE], geo[NE][NE], rot[3][3];
} Seg;

typedef struct {
  int nND, nEL, gEQ, iBodyModel;
  Seg *el;
  double *pos, *dsp, *vel, *acc, *mss; /* [node][6] */
  double *lodInte, *lodExte, *lodEffe, *lodFlow, *lodRepl, *lodGrav, *vBC;
  double coeffs[8];
  double Freq, denR, KB, KS, psR, EmR, tcR;
  double XYZ[3], XYZo[3], initXYZVel[3], XYZAmpl[3], XYZPhi[3], AoA[3], AoAo[3], AoAAmpl[3], AoAPhi[3], AoAd[3];
  double UVW[3], WWW1[3], WWW2[3], WWW3[3], TTT00[3][3], TTT0[3][3], TTTnxt[3][3];
  int isMotionGiven[6];
  /* global solver parameters (Set_SolidSolver_Params) */
  double dampK, dampM, GeoGamma, NewmarkGamma, NewmarkBeta, dtolFEM;
  int ntolFEM;
  /* diagnostics of the last solve */
  int last_newton_iters, last_cg_iters_total;
  double last_dnorm;
} Beam;

static const double M_PI_REF = 3.141592653589793;
static const double AOA_PI = 3.141562653589793;

static double ipow(double x, int n) { /* binary exponentiation, low bit first */
  double res = 0.0, cur = x;
  int have = 0;
  while (n) {
    if (n & 1) {
      if (have) res = res * cur;
      else { res = cur; have = 1; }
    }
    n >>= 1;
    if (n) cur = cur * cur;
  }
  return have ? res : 1.0;
}

static double dot3(const double *a, const double *b) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) s = s + a[k] * b[k];
  return s;
}

/* C = A*B for 3x3, k ascending from 0 */
static void mm3(double A[3][3], double B[3][3], double C[3][3]) {
  double T[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + A[i][k] * B[k][j];
      T[i][j] = s;
    }
  memcpy(C, T, sizeof(T));
}

static void mv3(double A[3][3], const double *v, double *out) {
  double t[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + A[i][k] * v[k];
    t[i] = s;
  }
  memcpy(out, t, sizeof(t));
}

void orc_aoa_to_ttt(const double *aoa, double T[3][3]) {
  double rr[3][3][3];
  for (int a = 0; a < 3; ++a) {
    double c = cos(aoa[a]), s = sin(aoa[a]);
    if (fabs(aoa[a]) < 1.0e-5) { c = 1.0; s = 0.0; }
    if (fabs(aoa[a] - 0.5 * AOA_PI) < 1.0e-5) { c = 0.0; s = 1.0; }
    if (fabs(aoa[a] + 0.5 * AOA_PI) < 1.0e-5) { c = 0.0; s = -1.0; }
    memset(rr[a], 0, sizeof(rr[a]));
    if (a == 0) {
      rr[0][0][0] = 1.0; rr[0][1][1] = c; rr[0][2][1] = s; rr[0][1][2] = -s; rr[0][2][2] = c;
    } else if (a == 1) {
      rr[1][0][0] = c; rr[1][2][0] = -s; rr[1][1][1] = 1.0; rr[1][0][2] = s; rr[1][2][2] = c;
    } else {
      rr[2][0][0] = c; rr[2][1][0] = s; rr[2][0][1] = -s; rr[2][1][1] = c; rr[2][2][2] = 1.0;
    }
  }
  double I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  mm3(rr[2], I, T);
  mm3(rr[1], T, T);
  mm3(rr[0], T, T);
}

static void get_angle_triad(double t11[3][3], double t22[3][3], double *tx, double *ty, double *tz) {
  double rr[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + t22[i][k] * t11[j][k];
      rr[i][j] = s;
    }
  const double dtx = (rr[2][1] - rr[1][2]) / 2.0;
  const double dty = (rr[0][2] - rr[2][0]) / 2.0;
  const double dtz = (rr[1][0] - rr[0][1]) / 2.0;
  double tr = rr[0][0] + rr[1][1] + rr[2][2];
  tr = (tr - 1.0) / 2.0;
  if (tr > 1.0) tr = 1.0;
  if (tr < -1.0) tr = -1.0;
  const double sint = sqrt(dtx * dtx + dty * dty + dtz * dtz);
  const double theta = acos(tr);
  if (sint < 1.0e-10 || theta < 1.0e-10) {
    *tx = dtx; *ty = dty; *tz = dtz;
  } else {
    const double f = theta / sint;
    *tx = f * dtx; *ty = f * dty; *tz = f * dtz;
  }
}

void orc_angle_triad(const double *t11, const double *t22, double *out) {
  double a[3][3], b[3][3];
  memcpy(a, t11, sizeof(a));
  memcpy(b, t22, sizeof(b));
  get_angle_triad(a, b, out, out + 1, out + 2);
}

static void global_to_local(double t[3][3], double tx, double ty, double tz, double *a, double *b, double *c) {
  *a = t[0][0] * tx + t[1][0] * ty + t[2][0] * tz;
  *b = t[0][1] * tx + t[1][1] * ty + t[2][1] * tz;
  *c = t[0][2] * tx + t[1][2] * ty + t[2][2] * tz;
}

static void finite_rot(double t1, double t2, double t3, double rr[3][3]) {
  const double tt = sqrt(ipow(t1, 2) + ipow(t2, 2) + ipow(t3, 2));
  const double ss = sin(tt), cc = cos(tt);
  double r1[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double r2[3][3] = {{0.0, -t3, t2}, {t3, 0.0, -t1}, {-t2, t1, 0.0}};
  double r3[3][3];
  r3[0][0] = -t3 * t3 - t2 * t2;
  r3[0][1] = t2 * t1;
  r3[0][2] = t3 * t1;
  r3[1][0] = t1 * t2;
  r3[1][1] = -t3 * t3 - t1 * t1;
  r3[1][2] = t3 * t2;
  r3[2][0] = t1 * t3;
  r3[2][1] = t2 * t3;
  r3[2][2] = -t2 * t2 - t1 * t1;
  double c1, c2;
  if (tt < 1.0e-10) { c1 = 1.0; c2 = 0.5; }
  else { c1 = ss / tt; c2 = (1.0 - cc) / ipow(tt, 2); }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) rr[i][j] = r1[i][j] + r2[i][j] * c1 + r3[i][j] * c2;
}

static void axis_dir_triad(double l, double m, double n, const double *d, double triad[3][3]) {
  double ex[3] = {l, m, n}, ey[3], ez[3], dir[3] = {d[0], d[1], d[2]};
  double dd = sqrt(dot3(ex, ex));
  if (dd > 1.0e-14) for (int k = 0; k < 3; ++k) ex[k] = ex[k] / dd;
  dd = sqrt(dot3(dir, dir));
  if (dd > 1.0e-14) for (int k = 0; k < 3; ++k) dir[k] = dir[k] / dd;
  const double proj = dot3(dir, ex);
  for (int k = 0; k < 3; ++k) ey[k] = dir[k] - proj * ex[k];
  dd = sqrt(dot3(ey, ey));
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
    dd = sqrt(dot3(ez, ez));
    if (dd > 1.0e-14) for (int k = 0; k < 3; ++k) ez[k] = ez[k] / dd;
  }
  for (int k = 0; k < 3; ++k) {
    triad[k][0] = ex[k];
    triad[k][1] = ey[k];
    triad[k][2] = ez[k];
  }
}

static void seg_cptdxyz1(Seg *e) {
  e->dx1 = e->x1[6] - e->x1[0];
  e->dy1 = e->x1[7] - e->x1[1];
  e->dz1 = e->x1[8] - e->x1[2];
  e->len1 = sqrt(e->dx1 * e->dx1 + e->dy1 * e->dy1 + e->dz1 * e->dz1);
  e->xll1 = e->dx1 / e->len1;
  e->xmm1 = e->dy1 / e->len1;
  e->xnn1 = e->dz1 / e->len1;
}

static void seg_init(Seg *e) {
  memcpy(e->x1, e->x0, sizeof(e->x1));
  memcpy(e->xnxt, e->x1, sizeof(e->xnxt));
  memcpy(e->dirc1, e->dirc0, sizeof(e->dirc1));
  memcpy(e->dircnxt, e->dirc1, sizeof(e->dircnxt));
  e->dx0 = e->x0[6] - e->x0[0];
  e->dy0 = e->x0[7] - e->x0[1];
  e->dz0 = e->x0[8] - e->x0[2];
  e->len0 = sqrt(e->dx0 * e->dx0 + e->dy0 * e->dy0 + e->dz0 * e->dz0);
  e->xll0 = e->dx0 / e->len0;
  e->xmm0 = e->dy0 / e->len0;
  e->xnn0 = e->dz0 / e->len0;
  seg_cptdxyz1(e);
}

static void map_pos(Seg *e, double *out, double T[3][3], const double *XYZ, const double *AoA) {
  double a[3], b[3];
  mv3(T, e->x00, a);
  mv3(T, e->x00 + 6, b);
  for (int k = 0; k < 3; ++k) {
    out[k] = a[k] + XYZ[k];
    out[3 + k] = AoA[k];
    out[6 + k] = b[k] + XYZ[k];
    out[9 + k] = AoA[k];
  }
}

static void form_mass(Seg *e) {
  const double area = e->prop[2], rho = e->prop[3], ziy = e->prop[6], ziz = e->prop[7];
  const double length = e->len0;
  memset(e->mas, 0, sizeof(e->mas));
  const double roal = rho * area * length / 2.0;
  e->mas[0][0] = roal; e->mas[1][1] = roal; e->mas[2][2] = roal;
  e->mas[3][3] = roal * (ziy + ziz) / area;
  e->mas[4][4] = roal * ziy / area;
  e->mas[5][5] = roal * ziz / area;
  for (int k = 0; k < 6; ++k) e->mas[6 + k][6 + k] = e->mas[k][k];
}

static void form_stiff(Seg *e) {
  const double emod = e->prop[0], gmod = e->prop[1], area = e->prop[2], zix = e->prop[5],
               ziy = e->prop[6], ziz = e->prop[7];
  const double length = e->len0;
  memset(e->stf, 0, sizeof(e->stf));
  const double Inv = 1.0 / length;
  const double ksy = 5.0 / 6.0, ksz = 5.0 / 6.0;
  const double phiy = 12.0 * emod * ziz / (ksy * gmod * area * length * length);
  const double phiz = 12.0 * emod * ziy / (ksz * gmod * area * length * length);
  const double ky1 = 12.0 * emod * ziz / (ipow(length, 3) * (1.0 + phiy));
  const double ky2 = 6.0 * emod * ziz / (ipow(length, 2) * (1.0 + phiy));
  const double ky3 = (4.0 + phiy) * emod * ziz / (length * (1.0 + phiy));
  const double ky4 = (2.0 - phiy) * emod * ziz / (length * (1.0 + phiy));
  const double kz1 = 12.0 * emod * ziy / (ipow(length, 3) * (1.0 + phiz));
  const double kz2 = 6.0 * emod * ziy / (ipow(length, 2) * (1.0 + phiz));
  const double kz3 = (4.0 + phiz) * emod * ziy / (length * (1.0 + phiz));
  const double kz4 = (2.0 - phiz) * emod * ziy / (length * (1.0 + phiz));
  double (*K)[NE] = e->stf;
  K[0][0] = area * emod * Inv; K[1][1] = ky1; K[2][2] = kz1; K[3][3] = gmod * zix * Inv; K[4][4] = kz3; K[5][5] = ky3;
  for (int k = 0; k < 6; ++k) K[6 + k][6 + k] = K[k][k];
  K[0][6] = -K[0][0];
  K[1][5] = ky2; K[1][7] = -ky1; K[1][11] = ky2; K[5][7] = -ky2; K[5][11] = ky4; K[7][11] = -ky2;
  K[2][4] = -kz2; K[2][8] = -kz1; K[2][10] = -kz2; K[4][8] = kz2; K[4][10] = kz4; K[8][10] = kz2;
  K[3][9] = -K[3][3];
  K[6][0] = K[0][6];
  K[5][1] = K[1][5]; K[7][1] = K[1][7]; K[11][1] = K[1][11]; K[7][5] = K[5][7]; K[11][5] = K[5][11]; K[11][7] = K[7][11];
  K[4][2] = K[2][4]; K[8][2] = K[2][8]; K[10][2] = K[2][10]; K[8][4] = K[4][8]; K[10][4] = K[4][10]; K[10][8] = K[8][10];
  K[9][3] = K[3][9];
}

static void form_geom(Seg *e) {
  const double s = e->geoFRM;
  const double emod = e->prop[0], gmod = e->prop[1], area = e->prop[2], zix = e->prop[5],
               ziy = e->prop[6], ziz = e->prop[7];
  const double L = e->len0;
  memset(e->geo, 0, sizeof(e->geo));
  const double ksy = 5.0 / 6.0, ksz = 5.0 / 6.0;
  const double phiy = 12.0 * emod * ziz / (ksy * gmod * area * L * L);
  const double phiz = 12.0 * emod * ziy / (ksz * gmod * area * L * L);
  const double py2 = ipow(1.0 + phiy, 2), pz2 = ipow(1.0 + phiz, 2);
  const double gy1 = s / L * (6.0 / 5.0 + 2.0 * phiy + phiy * phiy) / py2;
  const double gy2 = s / L * (L / 10.0) / py2;
  const double gy3 = s / L * (2.0 * L * L / 15.0 + phiy * L * L / 6.0 + phiy * phiy * L * L / 12.0) / py2;
  const double gy4 = s / L * (-(L * L / 30.0) - phiy * L * L / 6.0 - phiy * phiy * L * L / 12.0) / py2;
  double (*G)[NE] = e->geo;
  G[1][1] = gy1; G[5][5] = gy3; G[7][7] = gy1; G[11][11] = gy3;
  G[1][5] = gy2; G[1][7] = -gy1; G[1][11] = gy2; G[5][7] = -gy2; G[5][11] = gy4; G[7][11] = -gy2;
  const double gz1 = s / L * (6.0 / 5.0 + 2.0 * phiz + phiz * phiz) / pz2;
  const double gz2 = s / L * (L / 10.0) / pz2;
  const double gz3 = s / L * (2.0 * L * L / 15.0 + phiz * L * L / 6.0 + phiz * phiz * L * L / 12.0) / pz2;
  const double gz4 = s / L * (-(L * L / 30.0) - phiz * L * L / 6.0 - phiz * phiz * L * L / 12.0) / pz2;
  G[2][2] = gz1; G[4][4] = gz3; G[8][8] = gz1; G[10][10] = gz3;
  G[2][4] = -gz2; G[2][8] = -gz1; G[2][10] = -gz2; G[4][8] = gz2; G[4][10] = gz4; G[8][10] = gz2;
  const double gt = s * zix / (area * L);
  G[3][3] = gt; G[3][9] = -gt; G[9][9] = gt;
  G[5][1] = G[1][5]; G[7][1] = G[1][7]; G[11][1] = G[1][11]; G[7][5] = G[5][7]; G[11][5] = G[5][11]; G[11][7] = G[7][11];
  G[4][2] = G[2][4]; G[8][2] = G[2][8]; G[10][2] = G[2][10]; G[8][4] = G[4][8]; G[10][4] = G[4][10]; G[10][8] = G[8][10];
  G[9][3] = G[3][9];
}

static void rotate_matrix(Seg *e) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) e->rot[i][j] = e->tee[j][i];
}

static void rkr(Seg *e, double ek[NE][NE]) {
  double r[3][3], rt[3][3], kt[NE][NE];
  memcpy(r, e->rot, sizeof(r));
  for (int in = 0; in < 3; ++in)
    for (int jn = 0; jn < 3; ++jn) rt[jn][in] = e->rot[in][jn];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int j1 = i * 3, j2 = j * 3;
      for (int k = 0; k < 3; ++k)
        for (int ii = 0; ii < 3; ++ii) {
          double s = 0.0;
          for (int jj = 0; jj < 3; ++jj) s = s + ek[j1 + k][j2 + jj] * r[jj][ii];
          kt[j1 + k][j2 + ii] = s;
        }
      for (int k = 0; k < 3; ++k)
        for (int ii = 0; ii < 3; ++ii) {
          double s = 0.0;
          for (int jj = 0; jj < 3; ++jj) s = s + rt[k][jj] * kt[j1 + jj][j2 + ii];
          ek[j1 + k][j2 + ii] = s;
        }
    }
}

static void loc_to_global(const Seg *e, const double *lx, double *x) {
  for (int i = 0; i < NE; ++i) x[e->l2g[i]] = x[e->l2g[i]] + lx[i];
}

static void body_stress(Seg *e, double *lodInte) {
  const double du = e->dx1 - e->dx0, dv = e->dy1 - e->dy0, dw = e->dz1 - e->dz0;
  const double dl = ((e->dx0 + e->dx1) * du + (e->dy0 + e->dy1) * dv + (e->dz0 + e->dz1) * dw) / (e->len0 + e->len1);
  double tx, ty, tz, tx1, ty1, tz1, tx2, ty2, tz2;
  get_angle_triad(e->tee, e->tn1, &tx, &ty, &tz);
  global_to_local(e->tee, tx, ty, tz, &tx1, &ty1, &tz1);
  get_angle_triad(e->tee, e->tn2, &tx, &ty, &tz);
  global_to_local(e->tee, tx, ty, tz, &tx2, &ty2, &tz2);
  double ub[NE] = {0.0, 0.0, 0.0, tx1, ty1, tz1, dl, 0.0, 0.0, tx2, ty2, tz2};
  const double emod = e->prop[0], area = e->prop[2];
  e->geoFRM = dl * emod * area / e->len0;
  double fb[NE], f[NE];
  for (int i = 0; i < NE; ++i) {
    double s = 0.0;
    for (int k = 0; k < NE; ++k) s = s + e->stf[i][k] * ub[k];
    fb[i] = s;
  }
  for (int i = 0; i < NE; ++i) f[i] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      f[0 + i] = f[0 + i] + e->tee[i][j] * fb[0 + j];
      f[3 + i] = f[3 + i] + e->tee[i][j] * fb[3 + j];
      f[6 + i] = f[6 + i] + e->tee[i][j] * fb[6 + j];
      f[9 + i] = f[9 + i] + e->tee[i][j] * fb[9 + j];
    }
  loc_to_global(e, f, lodInte);
}

static void update_matrix(Seg *e, const double *c, double gamma, double dampM, double dampK) {
  form_geom(e);
  rotate_matrix(e);
  rkr(e, e->stf);
  rkr(e, e->geo);
  for (int i = 0; i < NE; ++i)
    for (int j = 0; j < NE; ++j) {
      e->tan_[i][j] = e->stf[i][j] + gamma * e->geo[i][j];
      e->coef[i][j] = e->tan_[i][j] + c[0] * e->mas[i][j] + c[1] * dampM * e->mas[i][j];
    }
  if (dampK > 0.0)
    for (int i = 0; i < NE; ++i)
      for (int j = 0; j < NE; ++j) e->coef[i][j] = e->coef[i][j] + c[1] * dampK * e->stf[i][j];
}

static void mass_multiply(const Seg *e, const double *q, double *mq) {
  for (int i = 0; i < NE; ++i) mq[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    mq[i] = e->mas[i][i] * q[i];
    mq[i + 6] = e->mas[i + 6][i + 6] * q[i + 6];
  }
  for (int i = 0; i < 3; ++i) {
    double s = 0.0, t = 0.0;
    for (int k = 0; k < 3; ++k) s = s + e->mas[3 + i][3 + k] * q[3 + k];
    for (int k = 0; k < 3; ++k) t = t + e->mas[9 + i][9 + k] * q[9 + k];
    mq[3 + i] = s;
    mq[9 + i] = t;
  }
}

static void update_load(const Seg *e, const double *c, double dampM, double dampK, const double *dspO,
                        const double *dsp, const double *vel, const double *acc, double *lodEffe) {
  double qM[NE], qC[NE], qMC[NE], ml[NE];
  const int n0 = e->node0, n1 = e->node1;
  for (int i = 0; i < 6; ++i) {
    qM[i] = c[0] * (dspO[6 * n0 + i] - dsp[6 * n0 + i]) + c[2] * vel[6 * n0 + i] + c[3] * acc[6 * n0 + i];
    qM[i + 6] = c[0] * (dspO[6 * n1 + i] - dsp[6 * n1 + i]) + c[2] * vel[6 * n1 + i] + c[3] * acc[6 * n1 + i];
    qC[i] = c[1] * (dspO[6 * n0 + i] - dsp[6 * n0 + i]) + c[4] * vel[6 * n0 + i] + c[5] * acc[6 * n0 + i];
    qC[i + 6] = c[1] * (dspO[6 * n1 + i] - dsp[6 * n1 + i]) + c[4] * vel[6 * n1 + i] + c[5] * acc[6 * n1 + i];
  }
  for (int i = 0; i < NE; ++i) qMC[i] = qM[i] + dampM * qC[i];
  mass_multiply(e, qMC, ml);
  loc_to_global(e, ml, lodEffe);
  if (dampK > 0.0) {
    double dl[NE];
    for (int i = 0; i < NE; ++i) {
      double s = 0.0;
      for (int k = 0; k < NE; ++k) s = s + e->stf[i][k] * qC[k];
      dl[i] = dampK * s;
    }
    loc_to_global(e, dl, lodEffe);
  }
}

static void update_triad(Seg *e, const double *dspnn) {
  double rr[3][3];
  finite_rot(dspnn[6 * e->node0 + 3], dspnn[6 * e->node0 + 4], dspnn[6 * e->node0 + 5], rr);
  mm3(rr, e->tn1, e->tn1);
  finite_rot(dspnn[6 * e->node1 + 3], dspnn[6 * e->node1 + 4], dspnn[6 * e->node1 + 5], rr);
  mm3(rr, e->tn2, e->tn2);
}

static void make_triad_ee(Seg *e) {
  const double xll = e->xll1, xmm = e->xmm1, xnn = e->xnn1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) e->tee[i][j] = 0.0;
  e->tee[0][0] = xll; e->tee[1][0] = xmm; e->tee[2][0] = xnn;
  double tx, ty, tz;
  get_angle_triad(e->tn1, e->tn2, &tx, &ty, &tz);
  tx = tx / 2.0; ty = ty / 2.0; tz = tz / 2.0;
  double rr[3][3], aa[3][3];
  finite_rot(tx, ty, tz, rr);
  mm3(rr, e->tn1, aa);
  double r2e1 = 0.0, r3e1 = 0.0;
  for (int k = 0; k < 3; ++k) {
    r2e1 = r2e1 + aa[k][1] * e->tee[k][0];
    r3e1 = r3e1 + aa[k][2] * e->tee[k][0];
  }
  for (int j = 0; j < 3; ++j) {
    e->tee[j][1] = aa[j][1] - r2e1 * (aa[j][0] + e->tee[j][0]) / 2.0;
    e->tee[j][2] = aa[j][2] - r3e1 * (aa[j][0] + e->tee[j][0]) / 2.0;
  }
  double c2[3] = {e->tee[0][1], e->tee[1][1], e->tee[2][1]}, c1[3] = {e->tee[0][0], e->tee[1][0], e->tee[2][0]};
  const double p = dot3(c2, c1);
  for (int k = 0; k < 3; ++k) e->tee[k][1] = e->tee[k][1] - p * e->tee[k][0];
  for (int k = 0; k < 3; ++k) c2[k] = e->tee[k][1];
  const double dd = sqrt(dot3(c2, c2));
  if (dd > 1.0e-14)
    for (int k = 0; k < 3; ++k) e->tee[k][1] = e->tee[k][1] / dd;
  const double a0 = e->tee[1][0] * e->tee[2][1] - e->tee[2][0] * e->tee[1][1];
  const double a1 = e->tee[2][0] * e->tee[0][1] - e->tee[0][0] * e->tee[2][1];
  const double a2 = e->tee[0][0] * e->tee[1][1] - e->tee[1][0] * e->tee[0][1];
  e->tee[0][2] = a0; e->tee[1][2] = a1; e->tee[2][2] = a2;
}

/* ------------------------------------------------------------------ */
/* body lifetime */

Beam *orc_beam_create(int nND, int nEL, const double *xyz8 /* nND x 8 */, const int *elems /* nEL x 3: i j Nspan (1-based) */,
                      const int *cons /* nND x 6 */, int iBodyModel, const int *isMotionGiven,
                      double denR, double KB, double KS, double psR, double Freq,
                      const double *XYZo, const double *XYZAmpl, const double *XYZPhi_deg, const double *AoAo_deg,
                      const double *AoAAmpl_deg, const double *AoAPhi_deg, const double *initXYZVel,
                      double Lref, double Uref, double denIn,
                      double dampK, double dampM, double GeoGamma, double NewmarkGamma, double NewmarkBeta,
                      double dtolFEM, int ntolFEM) {
  Beam *b = (Beam *)calloc(1, sizeof(Beam));
  b->nND = nND; b->nEL = nEL; b->gEQ = 6 * nND; b->iBodyModel = iBodyModel;
  b->el = (Seg *)calloc(nEL, sizeof(Seg));
  const int g = b->gEQ;
  b->pos = calloc(g, 8); b->dsp = calloc(g, 8); b->vel = calloc(g, 8); b->acc = calloc(g, 8); b->mss = calloc(3 * nND, 8);
  b->lodInte = calloc(g, 8); b->lodExte = calloc(g, 8); b->lodEffe = calloc(g, 8); b->lodFlow = calloc(g, 8);
  b->lodRepl = calloc(g, 8); b->lodGrav = calloc(g, 8); b->vBC = calloc(g, 8);
  b->dampK = dampK; b->dampM = dampM; b->GeoGamma = GeoGamma; b->NewmarkGamma = NewmarkGamma;
  b->NewmarkBeta = NewmarkBeta; b->dtolFEM = dtolFEM; b->ntolFEM = ntolFEM;
  b->denR = denR; b->KB = KB; b->KS = KS; b->psR = psR; b->Freq = Freq;
  for (int k = 0; k < 6; ++k) b->isMotionGiven[k] = isMotionGiven[k];
  for (int k = 0; k < 3; ++k) {
    b->XYZo[k] = XYZo[k]; b->XYZAmpl[k] = XYZAmpl[k]; b->initXYZVel[k] = initXYZVel[k];
    /* calculate_angle_material: degrees -> radians (SolidSolver.f90:1543-1547) */
    b->AoAo[k] = AoAo_deg[k] / 180.0 * M_PI_REF;
    b->AoAAmpl[k] = AoAAmpl_deg[k] / 180.0 * M_PI_REF;
    b->AoAPhi[k] = AoAPhi_deg[k] / 180.0 * M_PI_REF;
    b->XYZPhi[k] = XYZPhi_deg[k] / 180.0 * M_PI_REF;
  }
  /* Segment_Build */
  for (int n = 0; n < nEL; ++n) {
    Seg *e = &b->el[n];
    const int p0 = elems[3 * n] - 1, p1 = elems[3 * n + 1] - 1;
    e->node0 = p0; e->node1 = p1; e->Nspan = elems[3 * n + 2];
    for (int i = 0; i < 6; ++i) { e->l2g[i] = 6 * p0 + i; e->l2g[i + 6] = 6 * p1 + i; }
    for (int k = 0; k < 3; ++k) {
      e->x00[k] = xyz8[8 * p0 + k];
      e->x00[6 + k] = xyz8[8 * p1 + k];
      e->x00[3 + k] = 0.0;
      e->x00[9 + k] = 0.0;
    }
    e->area00 = sqrt(ipow(e->x00[6] - e->x00[0], 2) + ipow(e->x00[7] - e->x00[1], 2) + ipow(e->x00[8] - e->x00[2], 2));
    for (int k = 0; k < 6; ++k) { e->bc[k] = cons[6 * p0 + k]; e->bc[6 + k] = cons[6 * p1 + k]; }
    e->Lspan = 0.5 * (xyz8[8 * p0 + 3] + xyz8[8 * p1 + 3]);
    e->spanlen = 0.5 * (xyz8[8 * p0 + 4] + xyz8[8 * p1 + 4]) + e->Lspan;
    for (int k = 0; k < 3; ++k) e->dirc00[k] = 0.5 * (xyz8[8 * p0 + 5 + k] + xyz8[8 * p1 + 5 + k]);
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + ipow(e->dirc00[k], 2);
    const double nrm = sqrt(s);
    for (int k = 0; k < 3; ++k) e->dirc00[k] = e->dirc00[k] / nrm;
    /* Beam_adjustBC */
    if (e->bc[0] == 1) for (int k = 0; k < 6; ++k) e->bc[k] = isMotionGiven[k];
    if (e->bc[6] == 1) for (int k = 0; k < 6; ++k) e->bc[6 + k] = isMotionGiven[k];
  }
  /* calculate_angle_material, isKB = 1 */
  for (int n = 0; n < nEL; ++n) {
    Seg *e = &b->el[n];
    const double len = e->spanlen;
    const double th = sqrt(KB / KS * 12.0) * Lref;
    e->prop[2] = len * th;
    e->prop[0] = KS * denIn * ipow(Uref, 2) * Lref * len / e->prop[2];
    e->prop[1] = e->prop[0] / (2.0 * (1.0 + psR));
    e->prop[3] = denR * len * Lref * denIn / e->prop[2];
    const double ratio = th / len;
    e->prop[5] = len * ipow(th, 3) / 3.0 * (1.0 - 0.63 * ratio + 0.052 * ipow(ratio, 5));
    e->prop[6] = len * ipow(th, 3) / 12.0;
    e->prop[7] = th * ipow(len, 3) / 12.0;
  }
  /* Beam_Initialise(time = 0) */
  const double t = 0.0;
  for (int k = 0; k < 3; ++k) {
    b->XYZ[k] = b->XYZo[k] + b->XYZAmpl[k] * cos(2.0 * M_PI_REF * b->Freq * t + b->XYZPhi[k]) + b->initXYZVel[k] * t;
    b->AoA[k] = b->AoAo[k] + b->AoAAmpl[k] * cos(2.0 * M_PI_REF * b->Freq * t + b->AoAPhi[k]);
  }
  orc_aoa_to_ttt(b->AoA, b->TTT0);
  orc_aoa_to_ttt(b->AoA, b->TTTnxt);
  get_angle_triad(b->TTT0, b->TTTnxt, &b->AoAd[0], &b->AoAd[1], &b->AoAd[2]);
  for (int n = 0; n < nEL; ++n) {
    Seg *e = &b->el[n];
    map_pos(e, e->x0, b->TTT0, b->XYZ, b->AoAd);
    mv3(b->TTT0, e->dirc00, e->dirc0);
    seg_init(e);
  }
  for (int n = 0; n < nEL; ++n) {
    Seg *e = &b->el[n];
    for (int k = 0; k < 6; ++k) { b->pos[6 * e->node0 + k] = e->x1[k]; b->pos[6 * e->node1 + k] = e->x1[6 + k]; }
  }
  if (iBodyModel == 1) {
    for (int k = 0; k < 3; ++k) {
      b->UVW[k] = -(2.0 * M_PI_REF * b->Freq * b->XYZAmpl[k] * sin(2.0 * M_PI_REF * b->Freq * t + b->XYZPhi[k])) + b->initXYZVel[k];
      b->WWW1[k] = -(2.0 * M_PI_REF * b->Freq * b->AoAAmpl[k] * sin(2.0 * M_PI_REF * b->Freq * t + b->AoAPhi[k]));
    }
    b->WWW2[0] = b->WWW1[0] * cos(b->AoA[1]) + b->WWW1[2];
    b->WWW2[1] = b->WWW1[0] * sin(b->AoA[1]) * sin(b->AoA[2]) + b->WWW1[1] * cos(b->AoA[2]);
    b->WWW2[2] = b->WWW1[0] * sin(b->AoA[1]) * cos(b->AoA[2]) - b->WWW1[1] * sin(b->AoA[2]);
    mv3(b->TTT0, b->WWW2, b->WWW3);
    for (int n = 0; n < nND; ++n) {
      double rel[3];
      for (int k = 0; k < 3; ++k) rel[k] = b->pos[6 * n + k] - b->XYZ[k];
      const double *W = b->WWW3;
      b->vel[6 * n + 0] = (W[1] * rel[2] - W[2] * rel[1]) + b->UVW[0];
      b->vel[6 * n + 1] = (W[2] * rel[0] - W[0] * rel[2]) + b->UVW[1];
      b->vel[6 * n + 2] = (W[0] * rel[1] - W[1] * rel[0]) + b->UVW[2];
      for (int k = 0; k < 3; ++k) b->vel[6 * n + 3 + k] = W[k];
    }
  }
  for (int n = 0; n < nEL; ++n) {
    Seg *e = &b->el[n];
    axis_dir_triad(e->xll0, e->xmm0, e->xnn0, e->dirc0, e->tn1);
    memcpy(e->tn2, e->tn1, sizeof(e->tn1));
    memcpy(e->tee, e->tn1, sizeof(e->tn1));
    form_mass(e);
    rotate_matrix(e);
    rkr(e, e->mas);
  }
  for (int n = 0; n < nEL; ++n) {
    Seg *e = &b->el[n];
    for (int j = 0; j < 3; ++j) {
      b->mss[3 * e->node0 + j] = b->mss[3 * e->node0 + j] + e->mas[j][j];
      b->mss[3 * e->node1 + j] = b->mss[3 * e->node1 + j] + e->mas[j + 6][j + 6];
    }
  }
  return b;
}

void orc_beam_free(Beam *b) {
  if (!b) return;
  free(b->el); free(b->pos); free(b->dsp); free(b->vel); free(b->acc); free(b->mss);
  free(b->lodInte); free(b->lodExte); free(b->lodEffe); free(b->lodFlow); free(b->lodRepl);
  free(b->lodGrav); free(b->vBC);
  free(b);
}

/* ---- solver ---- */

static void matvec(Beam *b, const double *x, double *out) {
  for (int i = 0; i < b->gEQ; ++i) out[i] = 0.0;
  for (int n = 0; n < b->nEL; ++n) {
    Seg *e = &b->el[n];
    double lx[NE], lb[NE];
    for (int i = 0; i < NE; ++i) lx[i] = x[e->l2g[i]];
    for (int i = 0; i < NE; ++i) {
      double s = 0.0;
      for (int k = 0; k < NE; ++k) s = s + e->coef[i][k] * lx[k];
      lb[i] = s;
    }
    loc_to_global(e, lb, out);
  }
}

static double dotg(const double *a, const double *c, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + a[i] * c[i];
  return s;
}

static int invert6(double A[6][6], double Ai[6][6]) {
  double aug[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) { aug[i][j] = A[i][j]; aug[i][6 + j] = 0.0; }
  for (int i = 0; i < 6; ++i) aug[i][6 + i] = 1.0;
  for (int i = 0; i < 6; ++i) {
    int p = i;
    double pivot = fabs(aug[i][i]);
    for (int k = i + 1; k < 6; ++k)
      if (fabs(aug[k][i]) > pivot) { pivot = fabs(aug[k][i]); p = k; }
    if (pivot <= 1.0e-30) {
      memset(Ai, 0, sizeof(double) * 36);
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
    for (int j = 0; j < 6; ++j) Ai[i][j] = aug[i][6 + j];
  return 0;
}

/* returns 0, or 2 on CG breakdown */
static int cg_solve(Beam *b, double *x, const double *rhs, int iterNR, int *cg_iters) {
  const int g = b->gEQ, nN = b->nND;
  double *r = calloc(g, 8), *p = calloc(g, 8), *Ap = calloc(g, 8), *z = calloc(g, 8), *xF = calloc(g, 8);
  int *fixed = calloc(g, sizeof(int));
  double (*blk)[6][6] = calloc(nN, sizeof(double[6][6]));
  const double err = 1.0e-6;
  int rc = 0;
  *cg_iters = 0;
  for (int n = 0; n < b->nEL; ++n) {
    Seg *e = &b->el[n];
    for (int i = 0; i < NE; ++i)
      if (e->bc[i] > 0) {
        const int gid = e->l2g[i];
        fixed[gid] = 1;
        xF[gid] = (iterNR == 1) ? b->vBC[gid] : 0.0;
      }
  }
  memcpy(x, xF, 8 * g);
  matvec(b, x, Ap);
  for (int i = 0; i < g; ++i) r[i] = rhs[i] - Ap[i];
  for (int i = 0; i < g; ++i) if (fixed[i]) r[i] = 0.0;
  for (int n = 0; n < b->nEL; ++n) {
    Seg *e = &b->el[n];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        blk[e->node0][i][j] = blk[e->node0][i][j] + e->coef[i][j];
        blk[e->node1][i][j] = blk[e->node1][i][j] + e->coef[6 + i][6 + j];
      }
  }
  for (int node = 0; node < nN; ++node) {
    for (int i = 0; i < 6; ++i)
      if (fixed[6 * node + i]) {
        for (int j = 0; j < 6; ++j) { blk[node][i][j] = 0.0; blk[node][j][i] = 0.0; }
        blk[node][i][i] = 1.0;
      }
    double inv[6][6];
    if (invert6(blk[node], inv)) {
      memset(blk[node], 0, sizeof(blk[node]));
      for (int i = 0; i < 6; ++i) blk[node][i][i] = 1.0;
    } else {
      memcpy(blk[node], inv, sizeof(inv));
    }
  }
#define PRECOND()                                                            \
  for (int node = 0; node < nN; ++node)                                      \
    for (int i = 0; i < 6; ++i) {                                            \
      double s = 0.0;                                                        \
      for (int k = 0; k < 6; ++k) s = s + blk[node][i][k] * r[6 * node + k]; \
      z[6 * node + i] = fixed[6 * node + i] ? 0.0 : s;                       \
    }
  PRECOND();
  double rn = sqrt(dotg(r, r, g));
  if (!(rn <= err)) {
    memcpy(p, z, 8 * g);
    double rsold = dotg(r, z, g);
    if (fabs(rsold) <= 1.0e-30) { rc = 2; goto done; }
    for (int it = 1; it <= 10000; ++it) {
      *cg_iters = it;
      matvec(b, p, Ap);
      for (int i = 0; i < g; ++i) if (fixed[i]) Ap[i] = 0.0;
      const double pAp = dotg(p, Ap, g);
      if (fabs(pAp) <= 1.0e-30) { rc = 2; goto done; }
      const double alpha = rsold / pAp;
      for (int i = 0; i < g; ++i) x[i] = x[i] + alpha * p[i];
      for (int i = 0; i < g; ++i) if (fixed[i]) x[i] = xF[i];
      for (int i = 0; i < g; ++i) r[i] = r[i] - alpha * Ap[i];
      for (int i = 0; i < g; ++i) if (fixed[i]) r[i] = 0.0;
      rn = sqrt(dotg(r, r, g));
      if (rn <= err) break;
      PRECOND();
      const double rsnew = dotg(r, z, g);
      if (fabs(rsold) <= 1.0e-30) { rc = 2; goto done; }
      const double beta = rsnew / rsold;
      for (int i = 0; i < g; ++i) p[i] = z[i] + beta * p[i];
      for (int i = 0; i < g; ++i) if (fixed[i]) p[i] = 0.0;
      rsold = rsnew;
    }
  }
#undef PRECOND
done:
  free(r); free(p); free(Ap); free(z); free(xF); free(fixed); free(blk);
  return rc;
}

static void update_newmark(Beam *b, double dt) {
  double *c = b->coeffs;
  c[2] = 1.0 / (b->NewmarkBeta * dt);
  c[1] = b->NewmarkGamma * c[2];
  c[0] = c[2] / dt;
  c[3] = 0.5 / b->NewmarkBeta - 1.0;
  c[4] = b->NewmarkGamma / b->NewmarkBeta - 1.0;
  c[5] = dt * (0.5 * b->NewmarkGamma / b->NewmarkBeta - 1.0);
  c[6] = dt * (1.0 - b->NewmarkGamma);
  c[7] = dt * b->NewmarkGamma;
}

static int beam_solver(Beam *b) {
  const int g = b->gEQ;
  double *dspO = malloc(8 * g), *velO = malloc(8 * g), *accO = malloc(8 * g), *dspn = calloc(g, 8);
  int rc = 0;
  /* InitDspVelAccATTimeT */
  for (int i = 0; i < g; ++i) b->vBC[i] = 0.0;
  for (int n = 0; n < b->nEL; ++n) {
    Seg *e = &b->el[n];
    for (int i = 0; i < NE; ++i) b->vBC[e->l2g[i]] = e->xnxt[i] - e->x1[i];
  }
  for (int i = 0; i < g; ++i) b->lodExte[i] = b->lodFlow[i] + b->lodGrav[i] + b->lodRepl[i];
  memcpy(dspO, b->dsp, 8 * g);
  memcpy(velO, b->vel, 8 * g);
  memcpy(accO, b->acc, 8 * g);
  double dnorm = 1.0;
  int iter;
  b->last_cg_iters_total = 0;
  for (iter = 1; iter <= b->ntolFEM; ++iter) {
    /* UpdateMatrixANDLoad */
    for (int i = 0; i < g; ++i) b->lodInte[i] = 0.0;
    for (int n = 0; n < b->nEL; ++n) {
      Seg *e = &b->el[n];
      form_stiff(e);
      body_stress(e, b->lodInte);
      update_matrix(e, b->coeffs, b->GeoGamma, b->dampM, b->dampK);
    }
    for (int i = 0; i < g; ++i) b->lodEffe[i] = b->lodExte[i] - b->lodInte[i];
    for (int n = 0; n < b->nEL; ++n)
      update_load(&b->el[n], b->coeffs, b->dampM, b->dampK, dspO, b->dsp, b->vel, b->acc, b->lodEffe);
    int cgi = 0;
    if (cg_solve(b, dspn, b->lodEffe, iter, &cgi)) { rc = 2; break; }
    b->last_cg_iters_total += cgi;
    /* UpdateDspANDTride (beta = 1 for every iteration since maxramp = 1) */
    const double zi = ipow(2.0, iter <= 1 ? iter : 1), z0 = 2.0;
    const double beta = iter <= 1 ? zi / z0 * 1.0 : 1.0 * 1.0;
    double *dspnn = malloc(8 * g);
    for (int i = 0; i < g; ++i) dspnn[i] = beta * dspn[i];
    for (int i = 0; i < g; ++i) b->dsp[i] = b->dsp[i] + dspnn[i];
    for (int n = 0; n < b->nEL; ++n) {
      Seg *e = &b->el[n];
      for (int k = 0; k < 6; ++k) {
        e->x1[k] = e->x0[k] + b->dsp[6 * e->node0 + k];
        e->x1[6 + k] = e->x0[6 + k] + b->dsp[6 * e->node1 + k];
      }
      for (int k = 0; k < 6; ++k) { b->pos[6 * e->node0 + k] = e->x1[k]; b->pos[6 * e->node1 + k] = e->x1[6 + k]; }
      seg_cptdxyz1(e);
      update_triad(e, dspnn);
      make_triad_ee(e);
    }
    double mx = -INFINITY;
    for (int i = 0; i < g; ++i) { const double v = ipow(beta * dspn[i], 2); if (v > mx) mx = v; }
    dnorm = fabs(mx);
    free(dspnn);
    if (dnorm <= b->dtolFEM) break;
  }
  b->last_newton_iters = iter > b->ntolFEM ? b->ntolFEM : iter;
  b->last_dnorm = dnorm;
  /* UpdateVelAcc */
  const double *c = b->coeffs;
  for (int i = 0; i < g; ++i) {
    b->acc[i] = c[0] * (b->dsp[i] - dspO[i]) - c[2] * velO[i] - c[3] * accO[i];
    b->vel[i] = velO[i] + c[6] * accO[i] + c[7] * b->acc[i];
  }
  free(dspO); free(velO); free(accO); free(dspn);
  return rc;
}

/* Beam_structure (SolidSolver.f90:1820-1877) */
int orc_beam_structure(Beam *b, double time, int isubstep, double deltat, double subdeltat) {
  const double tt = time - deltat + (double)isubstep * subdeltat;
  for (int k = 0; k < 3; ++k) {
    b->XYZ[k] = b->XYZo[k] + b->XYZAmpl[k] * cos(2.0 * M_PI_REF * b->Freq * tt + b->XYZPhi[k]) + b->initXYZVel[k] * tt;
    b->AoA[k] = b->AoAo[k] + b->AoAAmpl[k] * cos(2.0 * M_PI_REF * b->Freq * tt + b->AoAPhi[k]);
  }
  orc_aoa_to_ttt(b->AoA, b->TTTnxt);
  get_angle_triad(b->TTT0, b->TTTnxt, &b->AoAd[0], &b->AoAd[1], &b->AoAd[2]);
  if (b->iBodyModel == 1) {
    for (int n = 0; n < b->nEL; ++n) {
      Seg *e = &b->el[n];
      map_pos(e, e->xnxt, b->TTTnxt, b->XYZ, b->AoAd);
      mv3(b->TTTnxt, e->dirc00, e->dircnxt);
      memcpy(e->x1, e->xnxt, sizeof(e->x1));
      memcpy(e->dirc1, e->dircnxt, sizeof(e->dirc1));
      for (int k = 0; k < 6; ++k) { b->pos[6 * e->node0 + k] = e->x1[k]; b->pos[6 * e->node1 + k] = e->x1[6 + k]; }
      seg_cptdxyz1(e);
      axis_dir_triad(e->xll1, e->xmm1, e->xnn1, e->dirc1, e->tn1);
      memcpy(e->tn2, e->tn1, sizeof(e->tn1));
      memcpy(e->tee, e->tn1, sizeof(e->tn1));
    }
    for (int k = 0; k < 3; ++k) {
      b->UVW[k] = -(2.0 * M_PI_REF * b->Freq * b->XYZAmpl[k] * sin(2.0 * M_PI_REF * b->Freq * tt + b->XYZPhi[k])) + b->initXYZVel[k];
      b->WWW1[k] = -(2.0 * M_PI_REF * b->Freq * b->AoAAmpl[k] * sin(2.0 * M_PI_REF * b->Freq * tt + b->AoAPhi[k]));
    }
    b->WWW2[0] = b->WWW1[0] * cos(b->AoA[1]) + b->WWW1[2];
    b->WWW2[1] = b->WWW1[0] * sin(b->AoA[1]) * sin(b->AoA[2]) + b->WWW1[1] * cos(b->AoA[2]);
    b->WWW2[2] = b->WWW1[0] * sin(b->AoA[1]) * cos(b->AoA[2]) - b->WWW1[1] * sin(b->AoA[2]);
    mv3(b->TTTnxt, b->WWW2, b->WWW3);
    for (int n = 0; n < b->nND; ++n) {
      double rel[3];
      for (int k = 0; k < 3; ++k) rel[k] = b->pos[6 * n + k] - b->XYZ[k];
      const double *W = b->WWW3;
      b->vel[6 * n + 0] = (W[1] * rel[2] - W[2] * rel[1]) + b->UVW[0];
      b->vel[6 * n + 1] = (W[2] * rel[0] - W[0] * rel[2]) + b->UVW[1];
      b->vel[6 * n + 2] = (W[0] * rel[1] - W[1] * rel[0]) + b->UVW[2];
      for (int k = 0; k < 3; ++k) b->vel[6 * n + 3 + k] = W[k];
    }
    return 0;
  }
  for (int n = 0; n < b->nEL; ++n) map_pos(&b->el[n], b->el[n].xnxt, b->TTTnxt, b->XYZ, b->AoAd);
  update_newmark(b, subdeltat);
  return beam_solver(b);
}

/* PlateUpdatePosVelArea_ (Solidbody.f90:604-646) for this body */
void orc_beam_markers(const Beam *b, double alpha, double denIn, double *Exyz, double *Evel, double *Ea) {
  const double beta = -(alpha * 2.0 * denIn);
  int cnt = 0;
  for (int i = 0; i < b->nEL; ++i) {
    const Seg *e = &b->el[i];
    const int i1 = e->node0, i2 = e->node1;
    double txyz[3], tvel[3], om[3], dirc[3];
    for (int k = 0; k < 3; ++k) {
      txyz[k] = 0.5 * (b->pos[6 * i1 + k] + b->pos[6 * i2 + k]);
      tvel[k] = 0.5 * (b->vel[6 * i1 + k] + b->vel[6 * i2 + k]);
      om[k] = 0.5 * (b->vel[6 * i1 + 3 + k] + b->vel[6 * i2 + 3 + k]);
    }
    const double left = e->Lspan, len = e->spanlen;
    const double dl = len / (double)e->Nspan;
    const double area = dl * e->len1 * beta;
    for (int k = 0; k < 3; ++k) dirc[k] = 0.5 * (e->tn1[k][1] + e->tn2[k][1]);
    const double nrm = sqrt(dot3(dirc, dirc));
    if (nrm > 1.0e-12) for (int k = 0; k < 3; ++k) dirc[k] = dirc[k] / nrm;
    else for (int k = 0; k < 3; ++k) dirc[k] = e->tee[k][1];
    for (int s = 0; s < e->Nspan; ++s) {
      const double ls = dl * (0.5 + (double)s) - left;
      double r[3], w[3];
      for (int k = 0; k < 3; ++k) r[k] = dirc[k] * ls;
      w[0] = om[1] * r[2] - om[2] * r[1];
      w[1] = om[2] * r[0] - om[0] * r[2];
      w[2] = om[0] * r[1] - om[1] * r[0];
      for (int k = 0; k < 3; ++k) {
        Exyz[3 * cnt + k] = txyz[k] + r[k];
        Evel[3 * cnt + k] = tvel[k] + w[k];
      }
      Ea[cnt] = area;
      ++cnt;
    }
  }
}

void orc_beam_element_centres(const Beam *b, double *xc) {
  for (int i = 0; i < b->nEL; ++i)
    for (int k = 0; k < 3; ++k) xc[3 * i + k] = 0.5 * (b->el[i].x1[k] + b->el[i].x1[6 + k]);
}

int orc_beam_nmarkers(const Beam *b) {
  int m = 0;
  for (int i = 0; i < b->nEL; ++i) m += b->el[i].Nspan;
  return m;
}

void orc_beam_set_lodflow(Beam *b, const double *lod) { memcpy(b->lodFlow, lod, 8 * b->gEQ); }

void orc_beam_state(const Beam *b, double *pos, double *vel, double *acc, double *dsp) {
  if (pos) memcpy(pos, b->pos, 8 * b->gEQ);
  if (vel) memcpy(vel, b->vel, 8 * b->gEQ);
  if (acc) memcpy(acc, b->acc, 8 * b->gEQ);
  if (dsp) memcpy(dsp, b->dsp, 8 * b->gEQ);
}

void orc_beam_info(const Beam *b, int *newton_iters, int *cg_iters, double *dnorm) {
  *newton_iters = b->last_newton_iters;
  *cg_iters = b->last_cg_iters_total;
  *dnorm = b->last_dnorm;
}

double orc_beam_asfac(const Beam *b) {
  double s = 0.0;
  for (int i = 0; i < b->nEL; ++i) s = s + b->el[i].area00;
  return s;
}

void orc_beam_marker_elem(const Beam *b, int *me, int *en) {
  int cnt = 0;
  for (int i = 0; i < b->nEL; ++i) {
    for (int s = 0; s < b->el[i].Nspan; ++s) me[cnt++] = i;
    en[2 * i] = b->el[i].node0;
    en[2 * i + 1] = b->el[i].node1;
  }
}

/* ------------------------------------------------------------------ */
/* Beam_write_solid_info (SolidSolver.f90:1729-1799): the values of one DatInfo
 * row per file, in the reference's arithmetic order.  refs = Lref, Uref, Aref,
 * Fref, Pref, Eref (Solidbody.f90:276-280); XYZo the body's reference point.
 * rows (78 values): firstNode[15] lastNode[15] centerNode[15] nodeAverage[12]
 * forces[6] power[7] energy[8]. */

/* Segment_StrainEnergy_D after FormStiffMatrix (SolidSolver.f90:841-907,
 * 2340-2348): stretching and bending/torsion energy of the local deformation */
static void strain_energy(Seg *e, double *es, double *eb) {
  form_stiff(e);
  const double du = e->dx1 - e->dx0, dv = e->dy1 - e->dy0, dw = e->dz1 - e->dz0;
  const double dl = ((e->dx0 + e->dx1) * du + (e->dy0 + e->dy1) * dv + (e->dz0 + e->dz1) * dw) / (e->len0 + e->len1);
  double tx, ty, tz, tx1, ty1, tz1, tx2, ty2, tz2;
  get_angle_triad(e->tee, e->tn1, &tx, &ty, &tz);
  global_to_local(e->tee, tx, ty, tz, &tx1, &ty1, &tz1);
  get_angle_triad(e->tee, e->tn2, &tx, &ty, &tz);
  global_to_local(e->tee, tx, ty, tz, &tx2, &ty2, &tz2);
  const double ub[NE] = {0.0, 0.0, 0.0, tx1, ty1, tz1, dl, 0.0, 0.0, tx2, ty2, tz2};
  double S[NE][NE], B[NE][NE];
  for (int i = 0; i < NE; ++i)
    for (int j = 0; j < NE; ++j) { S[i][j] = 0.0; B[i][j] = e->stf[i][j]; }
  S[0][0] = e->stf[0][0]; S[0][6] = e->stf[0][6]; S[6][6] = e->stf[6][6]; S[6][0] = e->stf[6][0];
  B[0][0] = 0.0; B[0][6] = 0.0; B[6][6] = 0.0; B[6][0] = 0.0;
  double s1 = 0.0, s2 = 0.0;
  for (int i = 0; i < NE; ++i) {
    double a = 0.0, c = 0.0;
    for (int k = 0; k < NE; ++k) { a = a + S[i][k] * ub[k]; c = c + B[i][k] * ub[k]; }
    s1 = s1 + a * ub[i];
    s2 = s2 + c * ub[i];
  }
  *es = 0.5 * s1;
  *eb = 0.5 * s2;
}

static void node_row(const Beam *b, int n, const double *XYZo, const double *refs, double *r) {
  const double Lref = refs[0], Uref = refs[1], Aref = refs[2];
  for (int k = 0; k < 3; ++k) {
    r[k] = XYZo[k] / Lref;
    r[3 + k] = (b->pos[6 * n + k] - XYZo[k]) / Lref;
    r[6 + k] = b->pos[6 * n + 3 + k];
    r[9 + k] = b->vel[6 * n + k] / Uref;
    r[12 + k] = b->acc[6 * n + k] / Aref;
  }
}

void orc_beam_solid_info(Beam *b, const double *XYZo, const double *refs, double *rows) {
  const double Lref = refs[0], Uref = refs[1], Aref = refs[2], Fref = refs[3], Pref = refs[4], Eref = refs[5];
  const int nND = b->nND;
  node_row(b, 0, XYZo, refs, rows);
  node_row(b, nND - 1, XYZo, refs, rows + 15);
  node_row(b, (nND + 1) / 2 - 1, XYZo, refs, rows + 30);
  double *r = rows + 45;
  for (int k = 0; k < 3; ++k) {
    double ms = 0.0, xs = 0.0, vs = 0.0, as = 0.0;
    for (int n = 0; n < nND; ++n) ms = ms + b->mss[3 * n + k];
    for (int n = 0; n < nND; ++n) xs = xs + b->pos[6 * n + k] * b->mss[3 * n + k];
    for (int n = 0; n < nND; ++n) vs = vs + b->vel[6 * n + k] * b->mss[3 * n + k];
    for (int n = 0; n < nND; ++n) as = as + b->acc[6 * n + k] * b->mss[3 * n + k];
    r[k] = XYZo[k] / Lref;
    r[3 + k] = (xs / ms - XYZo[k]) / Lref;
    r[6 + k] = (vs / ms) / Uref;
    r[9 + k] = (as / ms) / Aref;
  }
  r = rows + 57;
  double P[3];
  for (int k = 0; k < 3; ++k) {
    double f = 0.0, p = 0.0;
    for (int n = 0; n < nND; ++n) f = f + b->lodFlow[6 * n + k];
    for (int n = 0; n < nND; ++n) p = p + b->lodFlow[6 * n + k] * b->vel[6 * n + k];
    r[k] = XYZo[k] / Lref;
    r[3 + k] = f / Fref;
    P[k] = p / Pref;
  }
  double *pw = rows + 63;
  for (int k = 0; k < 3; ++k) pw[k] = XYZo[k] / Lref;
  pw[3] = (P[0] + P[1]) + P[2];
  pw[4] = P[0]; pw[5] = P[1]; pw[6] = P[2];
  double Es = 0.0, Eb = 0.0, Ev = 0.0;
  for (int i = 0; i < b->nEL; ++i) {
    double es, eb;
    strain_energy(&b->el[i], &es, &eb);
    Es = Es + es;
    Eb = Eb + eb;
  }
  Es = Es / Eref;
  Eb = Eb / Eref;
  const double Ep = Es + Eb;
  for (int i = 0; i < b->nEL; ++i) {
    const Seg *e = &b->el[i];
    double ve[NE];
    for (int k = 0; k < 6; ++k) { ve[k] = b->vel[6 * e->node0 + k]; ve[6 + k] = b->vel[6 * e->node1 + k]; }
    double d = 0.0;
    for (int j = 0; j < NE; ++j) {
      double mv = 0.0;
      for (int k = 0; k < NE; ++k) mv = mv + e->mas[j][k] * ve[k];
      d = d + ve[j] * mv;
    }
    Ev = Ev + 0.5 * d;
  }
  Ev = Ev / Eref;
  double *en = rows + 70;
  for (int k = 0; k < 3; ++k) en[k] = XYZo[k] / Lref;
  en[3] = Ev + Ep; en[4] = Ev; en[5] = Ep; en[6] = Es; en[7] = Eb;
}

/* SurfaceUpdatePosVel_ (Solidbody.f90:707-720): markers of a v_type = -1 body
 * (one per surface triangle) from their body-frame positions E0 and the rigid
 * kinematics of the body's beam: x = TTTnxt E0 + XYZ, v = WWW3 x rel + UVW. */
void orc_beam_surface_markers(const Beam *b, int n, const double *E0, double *Exyz, double *Evel) {
  for (int i = 0; i < n; ++i) {
    double rel[3];
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s = s + b->TTTnxt[r][k] * E0[3 * i + k];
      rel[r] = s;
    }
    const double *w = b->WWW3;
    const double v[3] = {w[1] * rel[2] - w[2] * rel[1], w[2] * rel[0] - w[0] * rel[2], w[0] * rel[1] - w[1] * rel[0]};
    for (int k = 0; k < 3; ++k) {
      Exyz[3 * i + k] = rel[k] + b->XYZ[k];
      Evel[3 * i + k] = v[k] + b->UVW[k];
    }
  }
}

/* SurfaceBuildPosVelArea_ (Solidbody.f90:648-705): per triangle (1-based node
 * ids tri[3i..3i+2]) its incenter E0 and area times -2 alpha denIn. */
void orc_surface_geometry(int ntri, const int *tri, const double *xyz, double alpha, double denIn, double *E0,
                          double *Ea) {
  const double beta = -(alpha * 2.0 * denIn);
  for (int i = 0; i < ntri; ++i) {
    const double *A = xyz + 3 * (tri[3 * i] - 1), *B = xyz + 3 * (tri[3 * i + 1] - 1), *C = xyz + 3 * (tri[3 * i + 2] - 1);
    const double x1 = A[0], x2 = B[0], x3 = C[0], y1 = A[1], y2 = B[1], y3 = C[1], z1 = A[2], z2 = B[2], z3 = C[2];
    const double la = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3) + (z2 - z3) * (z2 - z3));
    const double lb = sqrt((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1) + (z3 - z1) * (z3 - z1));
    const double lc = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
    const double invC = 1.0 / (la + lb + lc);
    E0[3 * i] = (la * x1 + lb * x2 + lc * x3) * invC;
    E0[3 * i + 1] = (la * y1 + lb * y2 + lc * y3) * invC;
    E0[3 * i + 2] = (la * z1 + lb * z2 + lc * z3) * invC;
    const double ax = ((z1 - z2) * (y3 - y2) + (y2 - y1) * (z3 - z2));
    const double ay = ((x1 - x2) * (z3 - z2) + (z2 - z1) * (x3 - x2));
    const double az = ((y1 - y2) * (x3 - x2) + (x2 - x1) * (y3 - y2));
    Ea[i] = sqrt(ax * ax + ay * ay + az * az) * 0.5 * beta;
  }
}
