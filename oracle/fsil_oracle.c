/*
 * CPU restatement of the FSILBM3D fluid/coupling hot path, in FP64, used ONLY
 * as test infrastructure: the parity tests, smoke() and bench.py's
 * cpu_baseline leg load it as the checker.  Nothing in the product links it.
 *
 * Every routine follows the reference's own structure and expression order
 * (cited file:line under gaoak/FSILBM3D src/), NOT the GPU kernels' fused
 * design, so the two are independent restatements:
 *   - collision is in place on the pre-collision state (FluidDomain.f90:1208-1263),
 *   - streaming is a periodic shift of the whole field (FluidDomain.f90:1514-1625),
 *   - boundary conditions run face by face x-,x+,y-,y+,z-,z+ on the streamed
 *     field (FluidDomain.f90:616-1126).
 * Layout: Fortran fIn(z,y,x,0:18) == C f[q][x][y][z], z fastest
 * (FluidDomain.f90:384); uuu/force(z,y,x,1:3) == C u[d][x][y][z].
 *
 * Arithmetic: compiled with -O2 -ffp-contract=off so each operation rounds as
 * written; sums run in the reference's loop order (the amdflang build of the
 * reference sums SUM() sequentially, which the golden fixtures confirm).
 *
 * Parity pins (tests/test_oracle_*.py): D3Q19 table identities; moments of the
 * equilibrium; mass/momentum conservation; the Peskin kernel identities; and
 * bit-for-bit agreement with the compiled reference program on the committed
 * golden fixtures (periodic SRT/TRT/MRT, every boundary code, volume force,
 * immersed-boundary plate).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define Q 19

/* ConstParams.f90:11-25 */
static const int EE[Q][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1},
    {-1, 0, -1}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
static const int OPPO[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
static const int POSD[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
static const int NEGD[9] = {2, 4, 6, 10, 9, 14, 13, 18, 17};
static double WT[Q];
static void init_wt(void) {
  WT[0] = 1.0 / 3.0;
  for (int q = 1; q <= 6; ++q) WT[q] = 1.0 / 18.0;
  for (int q = 7; q < Q; ++q) WT[q] = 1.0 / 36.0;
}

typedef struct {
  int xd, yd, zd;
} dims_t;

static inline size_t IX(const dims_t *d, int x, int y, int z) {
  return ((size_t)x * d->yd + y) * d->zd + z;
}

void orc_tables(int *ee, int *oppo, double *wt) {
  init_wt();
  for (int q = 0; q < Q; ++q) {
    for (int k = 0; k < 3; ++k) ee[q * 3 + k] = EE[q][k];
    oppo[q] = OPPO[q];
    wt[q] = WT[q];
  }
}

/* calculate_macro_quantities_  FluidDomain.f90:1128-1145 */
void orc_macro(const double *f, int xd, int yd, int zd, const double *VF, double dh,
               double *den, double *u) {
  dims_t d = {xd, yd, zd};
  size_t N = (size_t)xd * yd * zd;
  for (size_t n = 0; n < N; ++n) {
    double s = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    for (int q = 0; q < Q; ++q) s = s + f[q * N + n];
    for (int q = 0; q < Q; ++q) m1 = m1 + f[q * N + n] * (double)EE[q][0];
    for (int q = 0; q < Q; ++q) m2 = m2 + f[q * N + n] * (double)EE[q][1];
    for (int q = 0; q < Q; ++q) m3 = m3 + f[q * N + n] * (double)EE[q][2];
    den[n] = s;
    u[0 * N + n] = (m1 + 0.5 * VF[0] * dh) / s;
    u[1 * N + n] = (m2 + 0.5 * VF[1] * dh) / s;
    u[2 * N + n] = (m3 + 0.5 * VF[2] * dh) / s;
  }
  (void)d;
}

/* collision_  FluidDomain.f90:1208-1238 (models 1 SRT, 2 TRT, 3 MRT).
 * force: full field (ResetVolumeForce + IBM spread + add_volume_force). */
void orc_collision(double *f, const double *den, const double *u, const double *force,
                   size_t N, int model, double Omega, double Omega2, const double *Mc,
                   const double *Mf, double dh) {
  init_wt();
  const double dt3 = 3.0 * dh;
  for (size_t n = 0; n < N; ++n) {
    const double u1 = u[n], u2 = u[N + n], u3 = u[2 * N + n];
    const double F1 = force[n], F2 = force[N + n], F3 = force[2 * N + n];
    const double rho = den[n];
    double uSqr = 0.0;
    uSqr = uSqr + u1 * u1;
    uSqr = uSqr + u2 * u2;
    uSqr = uSqr + u3 * u3;
    double uxyz[Q], fEq[Q], Flb[Q], fin[Q];
    for (int q = 0; q < Q; ++q) {
      const double e1 = EE[q][0], e2 = EE[q][1], e3 = EE[q][2];
      fin[q] = f[q * N + n];
      uxyz[q] = u1 * e1 + u2 * e2 + u3 * e3;
      fEq[q] = WT[q] * rho * ((1.0 - 1.5 * uSqr) + uxyz[q] * (3.0 + 4.5 * uxyz[q])) - fin[q];
      Flb[q] = dt3 * WT[q] *
               ((e1 - u1 + 3.0 * uxyz[q] * e1) * F1 + (e2 - u2 + 3.0 * uxyz[q] * e2) * F2 +
                (e3 - u3 + 3.0 * uxyz[q] * e3) * F3);
    }
    if (model == 1) {
      for (int q = 0; q < Q; ++q)
        f[q * N + n] = fin[q] + Omega * fEq[q] + (1.0 - 0.5 * Omega) * Flb[q];
    } else if (model == 2) {
      double t[Q];
      fEq[0] = Omega * fEq[0] + (1.0 - 0.5 * Omega) * Flb[0];
      for (int k = 0; k < 9; ++k) {
        const int p = POSD[k], m = NEGD[k];
        t[p] = 0.5 * Omega * (fEq[p] + fEq[m]) + (0.5 - 0.25 * Omega) * (Flb[p] + Flb[m]);
      }
      for (int k = 0; k < 9; ++k) {
        const int p = POSD[k], m = NEGD[k];
        t[m] = 0.5 * Omega2 * (fEq[p] - fEq[m]) + (0.5 - 0.25 * Omega2) * (Flb[p] - Flb[m]);
      }
      for (int k = 0; k < 9; ++k) {
        const int p = POSD[k], m = NEGD[k];
        const double a = t[p] + t[m], b = t[p] - t[m];
        fEq[p] = a;
        fEq[m] = b;
      }
      for (int q = 0; q < Q; ++q) f[q * N + n] = fin[q] + fEq[q];
    } else if (model == 3) {
      /* MATMUL(M, v)(i) = sum_j M(i,j) v(j), j ascending; Fortran column-major
       * M(i,j) is stored at Mc[j*19 + i]. */
      for (int i = 0; i < Q; ++i) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < Q; ++j) a = a + Mc[j * Q + i] * fEq[j];
        for (int j = 0; j < Q; ++j) b = b + Mf[j * Q + i] * Flb[j];
        f[i * N + n] = fin[i] + a + b;
      }
    }
  }
}

/* streaming_  FluidDomain.f90:1514-1625: every population shifts by e_q with
 * periodic wrap in all three directions. */
void orc_streaming(double *f, double *tmp, int xd, int yd, int zd) {
  dims_t d = {xd, yd, zd};
  size_t N = (size_t)xd * yd * zd;
  for (int q = 0; q < Q; ++q) {
    double *fq = f + q * N;
    for (int x = 0; x < xd; ++x)
      for (int y = 0; y < yd; ++y)
        for (int z = 0; z < zd; ++z) {
          int xs = (x + EE[q][0] + xd) % xd, ys = (y + EE[q][1] + yd) % yd,
              zs = (z + EE[q][2] + zd) % zd;
          tmp[IX(&d, xs, ys, zs)] = fq[IX(&d, x, y, z)];
        }
    memcpy(fq, tmp, N * sizeof(double));
  }
}

/* calculate_distribution_funcion  FluidDomain.f90:1827-1834 */
static void feq_bc(double rho, const double v[3], double out[Q]) {
  double uSqr = 0.0;
  uSqr = uSqr + v[0] * v[0];
  uSqr = uSqr + v[1] * v[1];
  uSqr = uSqr + v[2] * v[2];
  for (int q = 0; q < Q; ++q) {
    const double ux = v[0] * EE[q][0] + v[1] * EE[q][1] + v[2] * EE[q][2];
    out[q] = WT[q] * rho * (1.0 + 3.0 * ux + 4.5 * ux * ux - 1.5 * uSqr);
  }
}

/* evaluate_velocity  FluidDomain.f90:1791-1824.  kind 0: shear profile;
 * kind 2: oscillatory inflow (cos evaluated here with libm, like dcos). */
typedef struct {
  double denIn, uvwIn[3], shear[3];
  int kind;
} flowbc_t;

static const double PI_ = 3.141592653589793;

static void eval_vel(const flowbc_t *fl, double t, double zc, double yc, double xc, double v[3]) {
  if (fl->kind == 0) {
    v[0] = fl->uvwIn[0] + 0 * fl->shear[0] + yc * fl->shear[1] + zc * fl->shear[2];
    v[1] = fl->uvwIn[1] + xc * fl->shear[0] + 0 * fl->shear[1] + zc * fl->shear[2];
    v[2] = fl->uvwIn[2] + xc * fl->shear[0] + yc * fl->shear[1] + 0 * fl->shear[2];
  } else if (fl->kind == 2) {
    v[0] = fl->uvwIn[0] + fl->shear[0] * cos(2 * PI_ * fl->shear[1] * t + fl->shear[2] / 180.0 * PI_);
    v[1] = fl->uvwIn[1];
    v[2] = fl->uvwIn[2];
  } else {
    v[0] = v[1] = v[2] = 0.0; /* the reference leaves velocityOut unset */
  }
}

/* evaluate_moving_wall  FluidDomain.f90:1837-1843 */
static void moving_wall(double rho, const double v[3], const double in[Q], double out[Q]) {
  for (int q = 0; q < Q; ++q) {
    const double ux = v[0] * EE[q][0] + v[1] * EE[q][1] + v[2] * EE[q][2];
    out[q] = in[OPPO[q]] + 2.0 * WT[q] * rho * ux * 3.0;
  }
}

/* Incoming sets per face (FluidDomain.f90:622-1125) and the outgoing set each
 * symmetric face mirrors from. */
static const int IN_SET[6][5] = {{1, 7, 9, 11, 13},  {2, 8, 10, 12, 14}, {3, 7, 8, 15, 17},
                                 {4, 9, 10, 16, 18}, {5, 11, 12, 15, 16}, {6, 13, 14, 17, 18}};
static const int SYM_SRC[6][5] = {{2, 8, 10, 12, 14}, {1, 7, 9, 11, 13},  {4, 9, 10, 16, 18},
                                  {3, 7, 8, 15, 17},  {6, 13, 14, 17, 18}, {5, 11, 12, 15, 16}};

/* set_boundary_conditions_  FluidDomain.f90:616-1126.
 * den/u: the macro fields of this step (used by 102).  hw: six post-collision
 * planes (fIn_hw*, laid out [q][a][b] with (a,b) the face's two in-plane
 * indices in the reference's loop order) or NULL when the halfway planes are
 * not yet allocated (first call, FluidDomain.f90:651-654). */
void orc_bc(double *f, const double *den, const double *u, int xd, int yd, int zd,
            const double *geom /* dh,xmin,ymin,zmin,xmax,ymax,zmax */, const int *bc,
            const double *flowp /* denIn,u1,u2,u3,s1,s2,s3 */, int kind, double blktime,
            double *const *hw) {
  init_wt();
  dims_t d = {xd, yd, zd};
  size_t N = (size_t)xd * yd * zd;
  flowbc_t fl;
  fl.denIn = flowp[0];
  for (int k = 0; k < 3; ++k) {
    fl.uvwIn[k] = flowp[1 + k];
    fl.shear[k] = flowp[4 + k];
  }
  fl.kind = kind;
  const double dh = geom[0], xmin = geom[1], ymin = geom[2], zmin = geom[3];
  const double xmax = geom[4], ymax = geom[5], zmax = geom[6];
  for (int face = 0; face < 6; ++face) {
    const int code = bc[face];
    if (code == 301 || code == 0 || code == 1) continue;
    const int axis = face / 2, hi = face & 1;
    const int nd[3] = {xd, yd, zd};
    /* the two in-plane axes in the reference's loop nesting (outer, inner) */
    int oa, ia;
    if (axis == 0) { oa = 1; ia = 2; }
    else if (axis == 1) { oa = 0; ia = 2; }
    else { oa = 0; ia = 1; }
    const int pb = hi ? nd[axis] - 1 : 0;          /* boundary plane */
    const int p2 = hi ? nd[axis] - 2 : 1;          /* first interior plane */
    const int p3 = hi ? nd[axis] - 3 : 2;
    const double wallc = (axis == 0) ? (hi ? xmax : xmin) : (axis == 1) ? (hi ? ymax : ymin) : (hi ? zmax : zmin);
    const double hwc = hi ? wallc + dh * 0.5 : wallc - dh * 0.5;
    const int *S = IN_SET[face];
    for (int a = 0; a < nd[oa]; ++a)
      for (int b = 0; b < nd[ia]; ++b) {
        int c[3], c2[3], c3[3];
        c[axis] = pb; c2[axis] = p2; c3[axis] = p3;
        c[oa] = c2[oa] = c3[oa] = a;
        c[ia] = c2[ia] = c3[ia] = b;
        const size_t n = IX(&d, c[0], c[1], c[2]);
        const size_t n2 = IX(&d, c2[0], c2[1], c2[2]);
        double co[3];
        co[0] = xmin + dh * c[0];
        co[1] = ymin + dh * c[1];
        co[2] = zmin + dh * c[2];
        co[axis] = wallc;
        double v[3], fe[Q], fi[Q], tmp[Q], cur[Q];
        switch (code) {
          case 101:
            eval_vel(&fl, blktime, co[2], co[1], co[0], v);
            feq_bc(fl.denIn, v, fe);
            for (int q = 0; q < Q; ++q) f[q * N + n] = fe[q];
            break;
          case 102: {
            eval_vel(&fl, blktime, co[2], co[1], co[0], v);
            feq_bc(fl.denIn, v, fe);
            double v2[3] = {u[n2], u[N + n2], u[2 * N + n2]};
            feq_bc(den[n2], v2, fi);
            for (int k = 0; k < 5; ++k) {
              const int q = S[k];
              f[q * N + n] = fe[q] + (f[q * N + n2] - fi[q]);
            }
            break;
          }
          case 103:
            for (int k = 0; k < 5; ++k) f[S[k] * N + n] = f[S[k] * N + n2];
            break;
          case 104: {
            const size_t n3 = IX(&d, c3[0], c3[1], c3[2]);
            for (int k = 0; k < 5; ++k) f[S[k] * N + n] = 2.0 * f[S[k] * N + n2] - f[S[k] * N + n3];
            break;
          }
          case 201:
            for (int k = 0; k < 5; ++k) tmp[k] = f[OPPO[S[k]] * N + n];
            for (int k = 0; k < 5; ++k) f[S[k] * N + n] = tmp[k];
            break;
          case 203:
            if (!hw || !hw[face]) break;
            for (int k = 0; k < 5; ++k) tmp[k] = hw[face][((size_t)OPPO[S[k]] * nd[oa] + a) * nd[ia] + b];
            for (int k = 0; k < 5; ++k) f[S[k] * N + n] = tmp[k];
            break;
          case 202:
            eval_vel(&fl, blktime, co[2], co[1], co[0], v);
            for (int q = 0; q < Q; ++q) cur[q] = f[q * N + n];
            moving_wall(fl.denIn, v, cur, tmp);
            for (int k = 0; k < 5; ++k) f[S[k] * N + n] = tmp[S[k]];
            break;
          case 204:
            if (!hw || !hw[face]) break;
            co[axis] = hwc;
            eval_vel(&fl, blktime, co[2], co[1], co[0], v);
            for (int q = 0; q < Q; ++q) cur[q] = hw[face][((size_t)q * nd[oa] + a) * nd[ia] + b];
            moving_wall(fl.denIn, v, cur, tmp);
            for (int k = 0; k < 5; ++k) f[S[k] * N + n] = tmp[S[k]];
            break;
          case 302:
            for (int k = 0; k < 5; ++k) tmp[k] = f[SYM_SRC[face][k] * N + n];
            for (int k = 0; k < 5; ++k) f[S[k] * N + n] = tmp[k];
            break;
          default:
            break;
        }
      }
  }
}

/* halfwayBCset_  FluidDomain.f90:567-614: copy post-collision boundary planes */
void orc_halfway_save(const double *f, int xd, int yd, int zd, const int *bc, double *const *hw) {
  dims_t d = {xd, yd, zd};
  size_t N = (size_t)xd * yd * zd;
  const int nd[3] = {xd, yd, zd};
  for (int face = 0; face < 6; ++face) {
    if (!(bc[face] == 203 || bc[face] == 204) || !hw[face]) continue;
    const int axis = face / 2, hi = face & 1;
    int oa, ia;
    if (axis == 0) { oa = 1; ia = 2; }
    else if (axis == 1) { oa = 0; ia = 2; }
    else { oa = 0; ia = 1; }
    const int pb = hi ? nd[axis] - 1 : 0;
    for (int q = 0; q < Q; ++q)
      for (int a = 0; a < nd[oa]; ++a)
        for (int b = 0; b < nd[ia]; ++b) {
          int c[3];
          c[axis] = pb; c[oa] = a; c[ia] = b;
          hw[face][((size_t)q * nd[oa] + a) * nd[ia] + b] = f[q * N + IX(&d, c[0], c[1], c[2])];
        }
  }
}

/* ---------------- immersed boundary ---------------- */

/* Phi  Solidbody.f90:822-833 (Peskin 4-point) */
static double peskin(double x) {
  const double r = fabs(x);
  if (r < 1.0) return (3.0 - 2.0 * r + sqrt(1.0 + 4.0 * r * (1.0 - r))) * 0.125;
  if (r < 2.0) return (5.0 - 2.0 * r - sqrt(-7.0 + 4.0 * r * (3.0 - r))) * 0.125;
  return 0.0;
}

double orc_peskin(double x) { return peskin(x); }

/* trimedindex  Solidbody.f90:834-866; returns 0, or -1 when out of bounds */
static int trim_index(int i, int nd, int *out, int bclo, int bchi) {
  for (int k = -1; k <= 2; ++k) {
    int v = i + k;
    if (v < 1) {
      if (bclo == 301) v += nd;
      else if ((bclo == 302 || bclo == 201) && v == 0) v = 2;
      else if (bclo == 203 && v == 0) v = 1;
      else return -1;
    } else if (v > nd) {
      if (bchi == 301) v -= nd;
      else if ((bchi == 302 || bchi == 201) && v == nd + 1) v = nd - 1;
      else if (bchi == 203 && v == nd + 1) v = nd;
      else return -1;
    }
    out[k + 1] = v;
  }
  return 0;
}

/* UpdateElmtInterp_  Solidbody.f90:760-821: per body, anchor on marker 1,
 * 1-based grid indices (as the reference stores them) and FP32 weights. */
int orc_stencil(int nmark, const double *Exyz /* 3*M, marker-major */, int xd, int yd, int zd,
                const double *geom /* dh,xmin,ymin,zmin */, const int *bc, int *Ei /* 12*M */,
                float *Ew /* 12*M */) {
  const double dh = geom[0];
  const double invdh = 1.0 / dh;
  const int nd[3] = {xd, yd, zd};
  int i0[3];
  double x0[3];
  for (int k = 0; k < 3; ++k) {
    int ii = (int)floor((Exyz[k] - geom[1 + k]) * invdh);
    x0[k] = geom[1 + k] + (double)ii * dh;
    i0[k] = ii + 1;
  }
  for (int m = 0; m < nmark; ++m) {
    for (int k = 0; k < 3; ++k) {
      double off = (Exyz[3 * m + k] - x0[k]) * invdh;
      int idx = (int)floor(off);
      off = off - (double)idx;
      idx = idx + i0[k];
      if (trim_index(idx, nd[k], Ei + 12 * m + 4 * k, bc[2 * k], bc[2 * k + 1])) return -1 - m;
      for (int j = -1; j <= 2; ++j) Ew[12 * m + 4 * k + j + 1] = (float)peskin((double)j - off);
    }
  }
  return 0;
}

/* PenaltyForce_  Solidbody.f90:981-1049 for one body.  uuu is modified in
 * place (serial marker order); Eforce accumulates; returns tol (sum |dV|_1
 * in marker order) through *tol. */
void orc_penalty(int nmark, const int *Ei, const float *Ew, const double *Evel, const double *Ea,
                 double *Eforce, double *uuu, int xd, int yd, int zd, double dt, double dh,
                 double denIn, double *tol) {
  dims_t d = {xd, yd, zd};
  size_t N = (size_t)xd * yd * zd;
  const double invh3 = 0.5 * dt * ((1.0 / dh) * (1.0 / dh) * (1.0 / dh)) / denIn;
  double *fT = (double *)malloc(sizeof(double) * 3 * (size_t)(nmark > 0 ? nmark : 1));
  double t = 0.0;
  for (int m = 0; m < nmark; ++m) {
    const int *I = Ei + 12 * m;
    const float *W = Ew + 12 * m;
    double v[3] = {0.0, 0.0, 0.0};
    for (int x = 0; x < 4; ++x)
      for (int y = 0; y < 4; ++y)
        for (int z = 0; z < 4; ++z) {
          const size_t n = IX(&d, I[x] - 1, I[4 + y] - 1, I[8 + z] - 1);
          const double rx = W[x], ry = W[4 + y], rz = W[8 + z];
          for (int c = 0; c < 3; ++c) v[c] = v[c] + uuu[c * N + n] * rx * ry * rz;
        }
    for (int c = 0; c < 3; ++c) v[c] = Evel[3 * m + c] - v[c];
    double F[3];
    for (int c = 0; c < 3; ++c) F[c] = v[c] * Ea[m];
    t = t + fabs(v[0]) + fabs(v[1]) + fabs(v[2]);
    for (int c = 0; c < 3; ++c) {
      Eforce[3 * m + c] = Eforce[3 * m + c] + F[c];
      fT[3 * m + c] = F[c] * invh3;
    }
  }
  for (int m = 0; m < nmark; ++m) {
    const int *I = Ei + 12 * m;
    const float *W = Ew + 12 * m;
    for (int x = 0; x < 4; ++x)
      for (int y = 0; y < 4; ++y)
        for (int z = 0; z < 4; ++z) {
          const size_t n = IX(&d, I[x] - 1, I[4 + y] - 1, I[8 + z] - 1);
          const double rx = W[x], ry = W[4 + y], rz = W[8 + z];
          for (int c = 0; c < 3; ++c) uuu[c * N + n] = uuu[c * N + n] - fT[3 * m + c] * rx * ry * rz;
        }
  }
  free(fT);
  *tol = t;
}

/* The fluid part of FluidVolumeForce_  Solidbody.f90:968-977 for one body:
 * force -= Eforce/h^3 * delta, serial marker order. */
void orc_spread(int nmark, const int *Ei, const float *Ew, const double *Eforce, double *force,
                int xd, int yd, int zd, double dh) {
  dims_t d = {xd, yd, zd};
  size_t N = (size_t)xd * yd * zd;
  const double invh3 = (1.0 / dh) * (1.0 / dh) * (1.0 / dh);
  for (int m = 0; m < nmark; ++m) {
    const int *I = Ei + 12 * m;
    const float *W = Ew + 12 * m;
    double Fh[3];
    for (int c = 0; c < 3; ++c) Fh[c] = Eforce[3 * m + c] * invh3;
    for (int x = 0; x < 4; ++x)
      for (int y = 0; y < 4; ++y)
        for (int z = 0; z < 4; ++z) {
          const size_t n = IX(&d, I[x] - 1, I[4 + y] - 1, I[8 + z] - 1);
          const double rx = W[x], ry = W[4 + y], rz = W[8 + z];
          for (int c = 0; c < 3; ++c) force[c * N + n] = force[c * N + n] + (-(Fh[c] * rx * ry * rz));
        }
  }
}

/* ---------------- LES collision models ---------------- */
/* collision_ with iCollidModel 11 (Smagorinsky), 14 (WALE) and 15 (Vreman):
 * FluidDomain.f90:1239-1258 (dispatch), 1265-1280 (smag), 1308-1424 (WALE),
 * 1425-1436 (center/one-sided differences), 1437-1507 (vrem); constants
 * ConstParams.f90:39-45.  u is the full velocity field the collision reads
 * (macro quantities with the immersed-boundary correction), tau_all the
 * per-cell relaxation time (read by WALE from the previous step, written by
 * all three). */
static double cdiff(double g1, double g2, double s) { return (g1 - g2) * s; }
static double odiff(double g1, double g2, double g3, double s) { return (-3.0 * g1 + 4.0 * g2 - g3) * s; }

/* derivative of velocity component c along axis a at cell (x,y,z), the
 * reference's interior/edge choice (no wrap) */
static double dudx(const double *u, const dims_t *d, int x, int y, int z, int a, int c, double s) {
  const size_t N = (size_t)d->xd * d->yd * d->zd;
  const double *uc = u + (size_t)c * N;
  int p[3] = {x, y, z};
  const int nd = a == 0 ? d->xd : (a == 1 ? d->yd : d->zd);
  const int i = p[a];
#define AT(k) (p[a] = (k), uc[IX(d, p[0], p[1], p[2])])
  double r;
  if (i > 0 && i < nd - 1) {
    const double g1 = AT(i + 1), g2 = AT(i - 1);
    r = cdiff(g1, g2, s);
  } else if (i == 0) {
    const double g1 = AT(i), g2 = AT(i + 1), g3 = AT(i + 2);
    r = odiff(g1, g2, g3, s);
  } else {
    const double g1 = AT(i), g2 = AT(i - 1), g3 = AT(i - 2);
    r = odiff(g1, g2, g3, s);
  }
#undef AT
  return r;
}

void orc_collision_les(double *f, const double *den, const double *u, const double *force, int xd, int yd, int zd,
                       int model, double tau, double nu, double dh, double *tau_all) {
  init_wt();
  const dims_t d = {xd, yd, zd};
  const size_t N = (size_t)xd * yd * zd;
  const double dt3 = 3.0 * dh;
  const double Cs2 = 1.0 / 3.0, Csmag = 0.17, CWALE = 0.50;
  const double CsmagConst = 2.0 * Csmag * Csmag * sqrt(2.0) * 9.0;
  const double CWALEConst = CWALE * CWALE;
  const double CvremConst = 2.5 * Csmag * Csmag;
  for (int x = 0; x < xd; ++x)
    for (int y = 0; y < yd; ++y)
      for (int z = 0; z < zd; ++z) {
        const size_t n = IX(&d, x, y, z);
        const double u1 = u[n], u2 = u[N + n], u3 = u[2 * N + n];
        const double F1 = force[n], F2 = force[N + n], F3 = force[2 * N + n];
        const double rho = den[n];
        double uSqr = 0.0;
        uSqr = uSqr + u1 * u1;
        uSqr = uSqr + u2 * u2;
        uSqr = uSqr + u3 * u3;
        double uxyz[Q], fEq[Q], Flb[Q], fin[Q], fn[Q];
        for (int q = 0; q < Q; ++q) {
          const double e1 = EE[q][0], e2 = EE[q][1], e3 = EE[q][2];
          fin[q] = f[q * N + n];
          uxyz[q] = u1 * e1 + u2 * e2 + u3 * e3;
          fEq[q] = WT[q] * rho * ((1.0 - 1.5 * uSqr) + uxyz[q] * (3.0 + 4.5 * uxyz[q])) - fin[q];
          Flb[q] = dt3 * WT[q] *
                   ((e1 - u1 + 3.0 * uxyz[q] * e1) * F1 + (e2 - u2 + 3.0 * uxyz[q] * e2) * F2 +
                    (e3 - u3 + 3.0 * uxyz[q] * e3) * F3);
          fn[q] = -fEq[q];
        }
        const double Q11 = fn[1] + fn[2] + fn[7] + fn[8] + fn[9] + fn[10] + fn[11] + fn[12] + fn[13] + fn[14];
        const double Q22 = fn[3] + fn[4] + fn[7] + fn[8] + fn[9] + fn[10] + fn[15] + fn[16] + fn[17] + fn[18];
        const double Q33 = fn[5] + fn[6] + fn[11] + fn[12] + fn[13] + fn[14] + fn[15] + fn[16] + fn[17] + fn[18];
        const double Q12 = fn[7] - fn[8] - fn[9] + fn[10];
        const double Q13 = fn[11] - fn[12] - fn[13] + fn[14];
        const double Q23 = fn[15] - fn[16] - fn[17] + fn[18];
        double omega;
        if (model == 11) {
          const double Qs = Q11 * Q11 + Q22 * Q22 + Q33 * Q33 + 2.0 * (Q12 * Q12 + Q13 * Q13 + Q23 * Q23);
          const double tau_t = sqrt(tau * tau + CsmagConst * sqrt(Qs) / rho);
          tau_all[n] = 0.5 * (tau + tau_t);
          omega = 2.0 / (tau + tau_t);
        } else if (model == 14) {
          const double s = dh; /* the reference's `invdh` holds dh */
          const double tp = tau_all[n];
          const double S11 = -1.5 * s * Q11 / (rho * tp), S22 = -1.5 * s * Q22 / (rho * tp);
          const double S33 = -1.5 * s * Q33 / (rho * tp), S12 = -1.5 * s * Q12 / (rho * tp);
          const double S13 = -1.5 * s * Q13 / (rho * tp), S23 = -1.5 * s * Q23 / (rho * tp);
          const double S = S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23);
          double ox = dudx(u, &d, x, y, z, 1, 2, s);
          ox = 0.5 * (ox - dudx(u, &d, x, y, z, 2, 1, s));
          double oy = dudx(u, &d, x, y, z, 2, 0, s);
          oy = 0.5 * (oy - dudx(u, &d, x, y, z, 0, 2, s));
          double oz = dudx(u, &d, x, y, z, 0, 1, s);
          oz = 0.5 * (oz - dudx(u, &d, x, y, z, 1, 0, s));
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
          const double tn = (nu + CWALEConst * OP * dh * dh) / (dh * Cs2) + 0.5;
          tau_all[n] = tn;
          omega = 1.0 / tn;
        } else {
          const double s = dh;
          const double a11 = 0.5 * dudx(u, &d, x, y, z, 0, 0, s), a21 = 0.5 * dudx(u, &d, x, y, z, 0, 1, s);
          const double a31 = 0.5 * dudx(u, &d, x, y, z, 0, 2, s), a12 = 0.5 * dudx(u, &d, x, y, z, 1, 0, s);
          const double a22 = 0.5 * dudx(u, &d, x, y, z, 1, 1, s), a32 = 0.5 * dudx(u, &d, x, y, z, 1, 2, s);
          const double a13 = 0.5 * dudx(u, &d, x, y, z, 2, 0, s), a23 = 0.5 * dudx(u, &d, x, y, z, 2, 1, s);
          const double a33 = 0.5 * dudx(u, &d, x, y, z, 2, 2, s);
          const double b11 = a11 * a11, b12 = a12 * a12, b13 = a13 * a13, b21 = a21 * a21, b22 = a22 * a22;
          const double b23 = a23 * a23, b31 = a31 * a31, b32 = a32 * a32, b33 = a33 * a33;
          const double aa = b11 + b12 + b13 + b21 + b22 + b23 + b31 + b32 + b33;
          const double c12 = a11 * a21 + a12 * a22 + a13 * a23, c13 = a11 * a31 + a12 * a32 + a13 * a33;
          const double c23 = a21 * a31 + a22 * a32 + a23 * a33;
          const double bb = (b11 + b12 + b13) * (b21 + b22 + b23) - c12 * c12 + (b11 + b12 + b13) * (b31 + b32 + b33) -
                            c13 * c13 + (b21 + b22 + b23) * (b31 + b32 + b33) - c23 * c23;
          double OP = sqrt(bb / aa);
          if (!isfinite(OP)) OP = 0.0;
          const double tn = (nu + CvremConst * OP * dh * dh) / (dh * Cs2) + 0.5;
          tau_all[n] = tn;
          omega = 1.0 / tn;
        }
        for (int q = 0; q < Q; ++q) f[q * N + n] = fin[q] + omega * fEq[q] + (1.0 - 0.5 * omega) * Flb[q];
      }
}
