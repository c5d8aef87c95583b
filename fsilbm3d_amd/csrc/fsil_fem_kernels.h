// Device FEM sub-step: one workgroup per body, FP64.
//
// Beam_structure for iBodyModel 2 (SolidSolver.f90:1859-1914): prescribed
// positions, Newmark coefficients, full Newton with a block-Jacobi
// preconditioned conjugate-gradient solve (SolidSolver.f90:1949-2020) and the
// displacement/triad update, then the velocity/acceleration update.  For
// iBodyModel 1 the same launch applies the prescribed rigid kinematics.
//
// Parallel structure inside a workgroup: element work (matrices, internal
// force, loads, triads) one thread per element; element-by-element
// mat-vec products one thread per (element, row); the 6x6 block inverses one
// thread per node; every assembly (LocToGlobal) is a per-DOF sequential sum
// over that DOF's elements in element order, and every dot product is a
// sequential sum in DOF order by one thread, so the arithmetic is the
// reference's (only the device's libm sin/cos/acos can differ in the last
// place from the host C library's).
#pragma once
#include "fsil_beam.h"

namespace fsil {

struct FemBodyDesc {
  int nND, nEL, eoff, goff, noff;   // element / dof / node offsets into the global arrays
  int iBodyModel, ntolFEM;
  int tree_sums;                    // 1: CG dot products summed in a tree (FP32-deviation blocks), 0: in DOF order
  double dtolFEM, dampK, dampM, GeoGamma;
  double coeffs[8];
  double XYZ[3], AoAd[3], T[9];     // prescribed placement of this sub-step (host libm)
  double UVW[3], W3[3];             // rigid velocities (iBodyModel 1)
  // outputs
  int newton_iters, cg_iters, status, pad2;
  double dnorm;
};

struct FemArrays {
  BeamElem *el;
  double *pos, *dsp, *vel, *acc, *lod;          // lod: lodFlow from the immersed-boundary coupling
  double *dspO, *velO, *accO, *lodInte, *lodEffe, *lodExte, *vBC;
  double *x, *r, *p, *Ap, *z, *xF, *lbuf, *blk; // lbuf: 12 per element; blk: 36 per node
  int *fixed;
  const int *dof_ptr, *dof_e, *dof_i;           // dof -> (element, local index) in element order
  const int *node_ptr, *node_e, *node_h;        // node -> (element, half 0/1) in element order
};

__device__ inline void fem_sync() { __syncthreads(); }

// s + t[0] + t[1] + ... + t[n-1], in that order.  The terms are read in
// batches of sixteen into registers, the next batch while the current one is
// added, so only the chain of additions is serial.  Every batch read is
// unconditional (the last full batch is read once more, unused) so that the
// wait before a batch's additions covers that batch alone and not the
// prefetch behind it.
__device__ inline double ordered_sum(double s, const double *t, int n) {
  constexpr int B = 16;
  const int nb = n / B;
  // drain every outstanding memory access first (generic accesses elsewhere
  // in the kernel count on the LDS counter too), so that inside the loop only
  // the in-order LDS reads are pending and a batch waits for itself alone
  __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) expcnt(7) lgkmcnt(0)
  if (nb > 0) {
    double a[B], b[B];
#pragma unroll
    for (int j = 0; j < B; ++j) a[j] = t[j];
    int k = 0;
    // the scheduling barriers keep each prefetch ahead of the additions it
    // is meant to overlap
    for (;;) {
      const double *nx = t + B * min(k + 1, nb - 1);
#pragma unroll
      for (int j = 0; j < B; ++j) b[j] = nx[j];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < B; ++j) s = s + a[j];
      __builtin_amdgcn_sched_barrier(0);
      if (++k == nb) break;
      const double *nx2 = t + B * min(k + 1, nb - 1);
#pragma unroll
      for (int j = 0; j < B; ++j) a[j] = nx2[j];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < B; ++j) s = s + b[j];
      __builtin_amdgcn_sched_barrier(0);
      if (++k == nb) break;
    }
  }
  for (int i = nb * B; i < n; ++i) s = s + t[i];
  return s;
}

// The conjugate-gradient vectors of one body: in LDS when they fit (the usual
// case), else in the global work arrays; indexed by the body's DOF number.
struct FemVecs {
  double *x, *r, *p, *z, *Ap, *blk;  // blk: the 6x6 inverse blocks, 36 per node
  double *xF;                        // lifted fixed-DOF values of the CG
  const int *fixed;                  // fixed-DOF flags
  const int *en;                     // node0, node1 of each element (null: read the elements)
};

// LDS copy of a body's assembly lists for the CG: per DOF the range of its
// (element, local index) entries, each packed as element*12 + index, and the
// element products of the mat-vec (lbuf)
// LDS stride (doubles) of one element's matrix: 144 entries padded so that
// the five elements a wave's mat-vec lanes touch fall on distinct LDS banks
// (172 doubles = 24 banks mod 64)
constexpr int kCoefStride = 172;

struct CgLds {
  double *lbuf;
  const int *ptr, *ent;
  bool on;
  const double *coef;  // LDS copy of the body's element matrices (transposed), or null
};

// element-by-element product out = A*v for one body (Beam_MatrixMultipy,
// SolidSolver.f90:2035-2049); optionally zeroes the fixed DOFs of the result
// and leaves the products v_g*out_g in prod (for the next dot product)
__device__ inline void fem_matvec(const FemArrays &A, const FemBodyDesc &D, const FemVecs &V, const CgLds &C,
                                  const double *v, double *out, bool zero_fixed = false, double *prod = nullptr) {
  const int nE = D.nEL, g0 = D.goff;
  for (int t = threadIdx.x; t < nE * NED; t += blockDim.x) {
    const int e = t / NED, i = t - e * NED;
    const BeamElem &E = A.el[D.eoff + e];
    const int n0 = 6 * (V.en ? V.en[2 * e] : E.node0), n1 = 6 * (V.en ? V.en[2 * e + 1] : E.node1);
    double vv[NED], cc[NED];
#pragma unroll
    for (int k = 0; k < NED; ++k) {
      vv[k] = v[k < 6 ? n0 + k : n1 + (k - 6)];
      cc[k] = C.coef ? C.coef[e * kCoefStride + k * NED + i] : E.coef[i * NED + k];
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NED; ++k) s = s + cc[k] * vv[k];
    if (C.on) C.lbuf[t] = s;
    else A.lbuf[(size_t)(D.eoff + e) * NED + i] = s;
  }
  fem_sync();
  for (int g = threadIdx.x; g < 6 * D.nND; g += blockDim.x) {
    double s = 0.0;
    if (C.on) {
      for (int k = C.ptr[g]; k < C.ptr[g + 1]; ++k) s = s + C.lbuf[C.ent[k]];
    } else {
      for (int k = A.dof_ptr[g0 + g]; k < A.dof_ptr[g0 + g + 1]; ++k)
        s = s + A.lbuf[(size_t)A.dof_e[k] * NED + A.dof_i[k]];
    }
    if (zero_fixed && V.fixed[g]) s = 0.0;
    out[g] = s;
    if (prod) prod[g] = v[g] * s;
  }
  fem_sync();
}

// Sums of t1[0..n) and t2[0..n) (t2 may be null) by the whole workgroup in a
// fixed tree order: per-thread strided partial sums, wave sums by lane
// shuffles, then the waves' sums in wave order, which every thread adds
// itself (a broadcast without a second barrier).  The forms whose results are
// held to a tolerance rather than to the reference's rounding (FP32-deviation
// blocks) use it for the conjugate-gradient dot products; the caller passes a
// block barrier before the next call.
__device__ inline void fem_tree_sums(const double *t1, const double *t2, int n, double &r1, double &r2) {
  __shared__ double part[2][16];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    a = a + t1[i];
    if (t2) b = b + t2[i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = a + __shfl_down(a, off, 64);
    if (t2) b = b + __shfl_down(b, off, 64);
  }
  const int nw = ((int)blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = a;
    part[1][threadIdx.x >> 6] = b;
  }
  fem_sync();
  double s1 = 0.0, s2 = 0.0;
  for (int w = 0; w < nw; ++w) {
    s1 = s1 + part[0][w];
    s2 = s2 + part[1][w];
  }
  r1 = s1;
  r2 = s2;
}

// sum of terms[0..n) in order by one thread (tree: fem_tree_sums), broadcast.
// The caller passes a block barrier before *shared is written again.
__device__ inline double fem_serial_sum(const double *terms, int n, double *shared, int tree = 0) {
  if (tree) {
    double r1, r2;
    fem_tree_sums(terms, nullptr, n, r1, r2);
    return r1;
  }
  if (threadIdx.x == 0) *shared = ordered_sum(0.0, terms, n);
  fem_sync();
  return *shared;
}

// z = M^-1 r (block Jacobi); optionally leaves the products r_g r_g in
// prod[g] and r_g z_g in prod[cap + g]
__device__ inline void fem_precond(const FemArrays &A, const FemBodyDesc &D, const FemVecs &V,
                                   double *prod = nullptr, int cap = 0) {
  const int g0 = D.goff;
  for (int g = threadIdx.x; g < 6 * D.nND; g += blockDim.x) {
    const int node = g / 6, i = g - node * 6;
    const double *B = V.blk + (size_t)node * 36;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) s = s + B[i * 6 + k] * V.r[6 * node + k];
    const double z = V.fixed[g] ? 0.0 : s;
    V.z[g] = z;
    if (prod) {
      const double r = V.r[g];
      prod[g] = r * r;
      prod[cap + g] = r * z;
    }
  }
  fem_sync();
}

// the two ordered sums of prod[0..n) and prod[cap..cap+n) on two waves,
// broadcast.  The caller passes a block barrier before shared is written again.
__device__ inline void fem_pair_sums(const double *prod, int cap, int n, double *shared, double &r1, double &r2,
                                     int tree = 0) {
  if (tree) {
    fem_tree_sums(prod, prod + cap, n, r1, r2);
    return;
  }
  if (threadIdx.x == 0) shared[0] = ordered_sum(0.0, prod, n);
  else if (threadIdx.x == 64) shared[1] = ordered_sum(0.0, prod + cap, n);
  fem_sync();
  r1 = shared[0];
  r2 = shared[1];
}

// Dot products in DOF order, broadcast through LDS.  The products a_i*b_i
// (each one rounding, as in the sequential loop) are formed by the whole
// workgroup into LDS; one thread then adds them in order, so the serial part
// is only the chain of additions.  fem_dot2 runs two such chains side by side
// on two waves.
__device__ inline void fem_dot_chunks(const double *a, const double *b, const double *c, const double *d, int n,
                                      double *shared, double *stage, int cap, double &r1, double &r2, int tree = 0) {
  if (tree) {
    double s1 = 0.0, s2 = 0.0;
    for (int base = 0; base < n; base += cap) {
      const int m = min(cap, n - base);
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        stage[i] = a[base + i] * b[base + i];
        if (c) stage[cap + i] = c[base + i] * d[base + i];
      }
      fem_sync();
      double t1, t2;
      fem_tree_sums(stage, c ? stage + cap : nullptr, m, t1, t2);
      s1 = s1 + t1;
      s2 = s2 + t2;
      fem_sync();
    }
    r1 = s1;
    r2 = s2;
    return;
  }
  double s1 = 0.0, s2 = 0.0;
  for (int base = 0; base < n; base += cap) {
    const int m = min(cap, n - base);
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      stage[i] = a[base + i] * b[base + i];
      if (c) stage[cap + i] = c[base + i] * d[base + i];
    }
    fem_sync();
    if (threadIdx.x == 0) s1 = ordered_sum(s1, stage, m);
    else if (c && threadIdx.x == 64) s2 = ordered_sum(s2, stage + cap, m);
    fem_sync();
  }
  if (threadIdx.x == 0) shared[0] = s1;
  if (threadIdx.x == 64) shared[1] = s2;
  fem_sync();
  r1 = shared[0];
  r2 = shared[1];
  fem_sync();
}

__device__ inline double fem_dot(const double *a, const double *b, int n, double *shared, double *stage, int cap) {
  double r1, r2;
  fem_dot_chunks(a, b, nullptr, nullptr, n, shared, stage, cap, r1, r2);
  return r1;
}

__device__ inline void fem_dot2(const double *a, const double *b, const double *c, const double *d, int n,
                                double *shared, double *stage, int cap, double &r1, double &r2, int tree = 0) {
  fem_dot_chunks(a, b, c, d, n, shared, stage, cap, r1, r2, tree);
}

// v from the lane the DPP control selects (0 where it selects none or the
// masks leave the lane unwritten), as two 32-bit moves
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, BANK_MASK, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, BANK_MASK, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Wave sum by DPP (lane 63 holds it): row_shr 1, 2, 4, 8 leave each row's
// sum of lanes 0..15 in its lane 15, row_bcast 15 and 31 then add the rows
// into lane 63.  A fixed order; no LDS round trips (the shuffle tree takes
// six ds_bpermute pairs).
__device__ inline double wave_sum_dpp(double v) {
  v = v + dpp_f64<0x111, 0xf, 0xf>(v);  // row_shr:1
  v = v + dpp_f64<0x112, 0xf, 0xf>(v);  // row_shr:2
  v = v + dpp_f64<0x114, 0xf, 0xf>(v);  // row_shr:4
  v = v + dpp_f64<0x118, 0xf, 0xf>(v);  // row_shr:8
  v = v + dpp_f64<0x142, 0xa, 0xf>(v);  // row_bcast:15 into rows 1, 3
  v = v + dpp_f64<0x143, 0xc, 0xf>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Row g of the element-by-element product A v (Beam_MatrixMultipy,
// SolidSolver.f90:2035-2049, then LocToGlobal): each element product in
// ascending k, the products added in the DOF's element order, as fem_matvec
// forms them.  vk(d) gives v at the body's DOF d.
template <class VK>
__device__ inline double fem_row(const FemVecs &V, const CgLds &C, int g, VK &&vk) {
  double y = 0.0;
  for (int q = C.ptr[g]; q < C.ptr[g + 1]; ++q) {
    const int ent = C.ent[q], e = ent / NED, i = ent - e * NED;
    const int n0 = 6 * V.en[2 * e], n1 = 6 * V.en[2 * e + 1];
    const double *cc = C.coef + (size_t)e * kCoefStride + i;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NED; ++k) s = s + cc[k * NED] * vk(k < 6 ? n0 + k : n1 + (k - 6));
    y = y + s;
  }
  return y;
}

// CG_Solve (SolidSolver.f90:1949-2020) for the forms whose dot products are
// summed in a tree (FP32-deviation blocks), with the CG state in LDS: three
// workgroup barriers per iteration (after A p and p.Ap, after the x, r, z
// updates and r.r, r.z, after the direction update).  DOFs are owned sixty
// (ten whole nodes) per wave, so that a node's block-Jacobi row reads r of its
// own wave; each owner forms its row of A p, its updates and its terms of the
// dot products, which are summed within each wave by DPP (wave_sum_dpp) and
// then over the waves in wave order.  (Forming p = z + beta p at every reader instead, which saves
// the third barrier, measured 29 us per sub-step for the A p phase against
// 17 us: the repeated reads cost more than the barrier.)  x holds xF on
// entry; returns 2 on a breakdown (zero p.Ap or r.z), else 0; adds the
// iterations to cg_total.
__device__ inline int fem_cg_fused(const FemArrays &A, const FemBodyDesc &D, const FemVecs &V, const CgLds &C,
                                   int &cg_total) {
  __shared__ double partA[16], partB[2][16];
  const int nG = 6 * D.nND, g0 = D.goff;
  const int nw = (int)blockDim.x >> 6, wv = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
  const int span = 60 * nw;
  const double err = 1.0e-6;
  auto zrow = [&](int g) {  // z_g = (M^-1 r)_g, block Jacobi
    const int node = g / 6, i = g - node * 6;
    const double *B = V.blk + (size_t)node * 36;
    double sz = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) sz = sz + B[i * 6 + k] * V.r[6 * node + k];
    return V.fixed[g] ? 0.0 : sz;
  };
  // r = lodEffe - A x (fixed DOFs 0), z = M^-1 r, r.r and r.z
  double a1 = 0.0, a2 = 0.0;
  for (int base = 0; base < nG; base += span) {
    const int g = base + wv * 60 + ln;
    const bool on = ln < 60 && g < nG;
    if (on) {
      const double y = fem_row(V, C, g, [&](int d) { return V.x[d]; });
      V.r[g] = V.fixed[g] ? 0.0 : A.lodEffe[g0 + g] - y;
    }
    wave_sync();
    if (on) {
      const double z = zrow(g), r = V.r[g];
      V.z[g] = z;
      a1 = a1 + r * r;
      a2 = a2 + r * z;
    }
  }
  a1 = wave_sum_dpp(a1);
  a2 = wave_sum_dpp(a2);
  if (ln == 63) { partB[0][wv] = a1; partB[1][wv] = a2; }
  fem_sync();
  double rr = 0.0, rsold = 0.0;
  for (int w = 0; w < nw; ++w) { rr = rr + partB[0][w]; rsold = rsold + partB[1][w]; }
  double rn = sqrt(rr);
  if (rn <= err) return 0;
  if (fabs(rsold) <= 1.0e-30) return 2;
  bool first = true;
  for (int it = 1; it <= 10000; ++it) {
    ++cg_total;
    // p (z on the first iteration), A p with the fixed DOFs zeroed, p.Ap
    auto pk = [&](int d) { return first ? V.z[d] : V.p[d]; };
    double a = 0.0;
    for (int base = 0; base < nG; base += span) {
      const int g = base + wv * 60 + ln;
      if (ln < 60 && g < nG) {
        double y = fem_row(V, C, g, pk);
        if (V.fixed[g]) y = 0.0;
        V.Ap[g] = y;
        const double pg = pk(g);
        if (first) V.p[g] = pg;  // p = z (read as z by every other reader)
        a = a + pg * y;
      }
    }
    a = wave_sum_dpp(a);
    if (ln == 63) partA[wv] = a;
    fem_sync();
    double pAp = 0.0;
    for (int w = 0; w < nw; ++w) pAp = pAp + partA[w];
    if (fabs(pAp) <= 1.0e-30) return 2;
    const double alpha = rsold / pAp;
    // x, r update, z = M^-1 r, r.r and r.z
    a1 = 0.0;
    a2 = 0.0;
    for (int base = 0; base < nG; base += span) {
      const int g = base + wv * 60 + ln;
      const bool on = ln < 60 && g < nG;
      if (on) {
        double xv = V.x[g] + alpha * V.p[g];
        if (V.fixed[g]) xv = V.xF[g];
        V.x[g] = xv;
        double rv = V.r[g] - alpha * V.Ap[g];
        if (V.fixed[g]) rv = 0.0;
        V.r[g] = rv;
      }
      wave_sync();
      if (on) {
        const double z = zrow(g), r = V.r[g];
        V.z[g] = z;
        a1 = a1 + r * r;
        a2 = a2 + r * z;
      }
    }
    a1 = wave_sum_dpp(a1);
    a2 = wave_sum_dpp(a2);
    if (ln == 63) { partB[0][wv] = a1; partB[1][wv] = a2; }
    fem_sync();
    double rsnew = 0.0;
    rr = 0.0;
    for (int w = 0; w < nw; ++w) { rr = rr + partB[0][w]; rsnew = rsnew + partB[1][w]; }
    rn = sqrt(rr);
    if (rn <= err) break;
    if (fabs(rsold) <= 1.0e-30) return 2;
    const double beta = rsnew / rsold;
    for (int base = 0; base < nG; base += span) {
      const int g = base + wv * 60 + ln;
      if (ln < 60 && g < nG) V.p[g] = V.fixed[g] ? 0.0 : V.z[g] + beta * V.p[g];
    }
    fem_sync();
    first = false;
    rsold = rsnew;
  }
  return 0;
}

// LDS of k_fem_substep, in doubles, behind one double of slack (fem_lds_base)
// and each area an even number of doubles: a work area of `work` doubles (the
// staging area of the dot products, 2 nG, and, when cg_lds, the CG's
// assembly lists and element products behind it; also the waves' maxima);
// then, when vec_lds, the five CG vectors, the inverse blocks, the lifted
// fixed values, the fixed flags and the element node pairs of the largest
// body; then, optionally, kCoefStride doubles per element of the largest body
// (the element matrices, at coef_off doubles); then, optionally, one BeamHot
// per element of the largest body (at hot_off doubles).
__host__ __device__ inline size_t fem_lds_bytes(int work, int vec_lds, int max_nd, int max_ne) {
  size_t d = 1 + (((size_t)work + 1) & ~(size_t)1);
  // five CG vectors, inverse blocks, xF; the fixed flags and element node pairs (ints)
  if (vec_lds)
    d += (6 * 6 * (size_t)max_nd + 36 * (size_t)max_nd + (6 * (size_t)max_nd + 2 * (size_t)max_ne + 1) / 2 + 1) &
         ~(size_t)1;
  return d * sizeof(double);
}

// work-area doubles the CG's assembly lists need for one body (cg_lds)
__host__ __device__ inline size_t fem_cg_work(int nND, int nEL) {
  const size_t nG = 6 * (size_t)nND, nent = NED * (size_t)nEL;
  return 2 * nG + nent + (nG + 1 + nent + 1) / 2;
}

__device__ constexpr HotMap kHotMapD{};

// The dynamic LDS of k_fem_substep from its first 16-byte boundary (the host
// reserves one double of slack; its offsets are even): measured on MI355X,
// the CG loop of the flexible plate takes 39 us per sub-step with its vectors
// on 16-byte boundaries against 48 us 8 bytes past them.
__device__ inline double *fem_lds_base() {
  extern __shared__ double lds_raw[];
  return lds_raw + (((size_t)lds_raw & 15) ? 1 : 0);
}

// Block BLK (block row BLK / 4, column BLK % 4) of an element's R^T K R,
// R^T G R and the Newmark tangent matrix (Segment_UpdateMatrix,
// SolidSolver.f90:131-157), each entry summed as el_rkr sums it.  The
// patterns of K and G are constants here: a product with a zero entry (and a
// product with an inner sum that has no terms) adds a signed zero to a sum
// that starts at +0 and is left out.  cl: the element's LDS matrices
// (transposed), or null for the element record's coef.
template <int BLK, bool HOT, class ET>
__device__ inline void fem_rkr_block(ET &E, double *cl, const double *c, const FemBodyDesc &D) {
  constexpr int j1 = (BLK >> 2) * 3, j2 = (BLK & 3) * 3;
  double r[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) r[q] = E.tee[(q % 3) * 3 + q / 3];  // R = tee^T (RotateMatrix)
  double ks[9], kg[9];
#pragma unroll
  for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
    for (int b2 = 0; b2 < 3; ++b2) {
      const int cs = kStiffPatternD.c[(j1 + a2) * NED + j2 + b2], cg = kGeoPatternD.c[(j1 + a2) * NED + j2 + b2];
      ks[a2 * 3 + b2] = cs == 0 ? 0.0 : (cs > 0 ? E.ksc[cs - 1] : -E.ksc[-cs - 1]);
      kg[a2 * 3 + b2] = cg == 0 ? 0.0 : (cg > 0 ? E.gsc[cg - 1] : -E.gsc[-cg - 1]);
    }
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int ii = 0; ii < 3; ++ii) {
      double sR = 0.0, gR = 0.0;
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        double kts = 0.0, ktg = 0.0;
        bool any_s = false, any_g = false;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          if (kStiffPatternD.c[(j1 + jj) * NED + j2 + l] != 0) {
            kts = kts + ks[jj * 3 + l] * r[l * 3 + ii];
            any_s = true;
          }
          if (kGeoPatternD.c[(j1 + jj) * NED + j2 + l] != 0) {
            ktg = ktg + kg[jj * 3 + l] * r[l * 3 + ii];
            any_g = true;
          }
        }
        if (any_s) sR = sR + r[jj * 3 + k] * kts;
        if (any_g) gR = gR + r[jj * 3 + k] * ktg;
      }
      const int idx = (j1 + k) * NED + j2 + ii;
      const double tanv = sR + D.GeoGamma * gR;
      // the rotated mass matrix is block diagonal: its off-diagonal 3x3
      // blocks hold exact +0 (the rotation of a zero block)
      const double m = j1 == j2 ? mas_entry(E, j1 + k, j2 + ii) : 0.0;
      double cv = tanv + c[0] * m + c[1] * D.dampM * m;
      if constexpr (!HOT) {  // the LDS form is taken without stiffness damping
        if (D.dampK > 0.0) cv = cv + c[1] * D.dampK * sR;
        if (D.dampK > 0.0) E.stf[idx] = sR;  // read only by the stiffness-damping load
      }
      if (cl) {
        cl[(j2 + ii) * NED + j1 + k] = cv;
      } else {
        if constexpr (!HOT) E.coef[idx] = cv;  // the LDS form always has cl
      }
    }
}

// One sub-step of one body by the whole workgroup.  HOT: the body's element
// fields of the Newton iterations are in LDS (H, BeamHot per element), else
// every element access goes to the global BeamElem array.
template <bool HOT>
__device__ inline void fem_substep_body(FemBodyDesc &D, const FemArrays &A, int work, int vec_lds, int cg_lds,
                                        int coef_off, BeamHot *H) {
  __shared__ double sh_val;
  __shared__ double sh_pair[2];
  __shared__ int sh_flag;
  double *lds = fem_lds_base();
  const int nE = D.nEL, nN = D.nND, nG = 6 * nN, g0 = D.goff;
  // element e of this body (body-local index)
  auto EL = [&](int e) -> auto & {
    if constexpr (HOT) return H[e];
    else return A.el[D.eoff + e];
  };
  // the element matrices (coef) in LDS, transposed per element ([k][i]), when
  // they fit behind the other areas
  double *coefL = coef_off ? lds + coef_off : nullptr;
  FemVecs V;
  if (vec_lds) {
    double *vb = lds + work;
    V.x = vb; V.r = vb + nG; V.p = vb + 2 * nG; V.z = vb + 3 * nG; V.Ap = vb + 4 * nG; V.blk = vb + 5 * nG;
    V.xF = V.blk + 36 * (size_t)nN;
    int *fx = reinterpret_cast<int *>(V.xF + nG);
    int *en = fx + nG;
    for (int g = threadIdx.x; g < nG; g += blockDim.x) fx[g] = A.fixed[g0 + g];
    for (int e = threadIdx.x; e < nE; e += blockDim.x) {
      en[2 * e] = A.el[D.eoff + e].node0;
      en[2 * e + 1] = A.el[D.eoff + e].node1;
    }
    V.fixed = fx;
    V.en = en;
    fem_sync();
  } else {
    V.x = A.x + g0; V.r = A.r + g0; V.p = A.p + g0; V.z = A.z + g0; V.Ap = A.Ap + g0;
    V.blk = A.blk + (size_t)D.noff * 36;
    V.xF = A.xF + g0;
    V.fixed = A.fixed + g0;
    V.en = nullptr;
  }
  if (D.iBodyModel == 1) {
    // rigid prescribed motion (SolidSolver.f90:1826-1858)
    for (int e = threadIdx.x; e < nE; e += blockDim.x) {
      BeamElem &E = A.el[D.eoff + e];
      el_map_pos(E, E.xnxt, D.T, D.XYZ, D.AoAd);
      double dn[3];
      bmv3(D.T, E.dirc00, dn);
      for (int k = 0; k < NED; ++k) E.x1[k] = E.xnxt[k];
      el_cptdxyz1(E);
      axis_dir_triad(E.xll1, E.xmm1, E.xnn1, dn, E.tn1);
      for (int k = 0; k < 9; ++k) { E.tn2[k] = E.tn1[k]; E.tee[k] = E.tn1[k]; }
    }
    fem_sync();
    for (int n = threadIdx.x; n < nN; n += blockDim.x) {
      const int last = A.node_ptr[D.noff + n + 1] - 1;
      const BeamElem &E = A.el[A.node_e[last]];
      const double *src = A.node_h[last] ? E.x1 + 6 : E.x1;
      for (int k = 0; k < 6; ++k) A.pos[g0 + 6 * n + k] = src[k];
      double rel[3];
      for (int k = 0; k < 3; ++k) rel[k] = A.pos[g0 + 6 * n + k] - D.XYZ[k];
      const double *W = D.W3;
      A.vel[g0 + 6 * n + 0] = (W[1] * rel[2] - W[2] * rel[1]) + D.UVW[0];
      A.vel[g0 + 6 * n + 1] = (W[2] * rel[0] - W[0] * rel[2]) + D.UVW[1];
      A.vel[g0 + 6 * n + 2] = (W[0] * rel[1] - W[1] * rel[0]) + D.UVW[2];
      for (int k = 0; k < 3; ++k) A.vel[g0 + 6 * n + 3 + k] = W[k];
    }
    if (threadIdx.x == 0) { D.newton_iters = 0; D.cg_iters = 0; D.status = 0; D.dnorm = 0.0; }
    return;
  }
  const double *c = D.coeffs;
  // prescribed positions of this sub-step and InitDspVelAccATTimeT (SolidSolver.f90:1916-1930)
  for (int e = threadIdx.x; e < nE; e += blockDim.x) {
    BeamElem &E = A.el[D.eoff + e];
    el_map_pos(E, E.xnxt, D.T, D.XYZ, D.AoAd);
  }
  fem_sync();
  for (int g = threadIdx.x; g < nG; g += blockDim.x) {
    const int last = A.dof_ptr[g0 + g + 1] - 1;
    const int li = A.dof_i[last];
    A.vBC[g0 + g] = A.el[A.dof_e[last]].xnxt[li] - EL(A.dof_e[last] - D.eoff).x1[li];
    A.lodExte[g0 + g] = (A.lod[g0 + g] + 0.0) + 0.0;   // lodFlow + lodGrav + lodRepl (gravity, repulsion zero)
    A.dspO[g0 + g] = A.dsp[g0 + g];
    A.velO[g0 + g] = A.vel[g0 + g];
    A.accO[g0 + g] = A.acc[g0 + g];
  }
  fem_sync();
  int iter;
  int cg_total = 0;
  int status = 0;
  double dnorm = 1.0;
  // the body's assembly lists for the CG in LDS (behind the dot-product
  // staging area [0, 2 nG) and the element products), once per sub-step
  CgLds C{};
  C.coef = coefL;
  if (cg_lds) {
    C.lbuf = lds + 2 * nG;
    int *ip = reinterpret_cast<int *>(lds + 2 * nG + NED * nE);
    const int p0 = A.dof_ptr[g0], nent = A.dof_ptr[g0 + nG] - p0;
    int *ep = ip + nG + 1;
    for (int g = threadIdx.x; g <= nG; g += blockDim.x) ip[g] = A.dof_ptr[g0 + g] - p0;
    for (int k = threadIdx.x; k < nent; k += blockDim.x) ep[k] = (A.dof_e[p0 + k] - D.eoff) * NED + A.dof_i[p0 + k];
    C.ptr = ip;
    C.ent = ep;
    C.on = true;
    fem_sync();
  }
  // UpdateMatrixANDLoad (SolidSolver.f90:1932-1947).  Per element: internal
  // force from the local stiffness, local geometric stiffness.
  auto stress_phase = [&](int t0, int step) {
    for (int e = t0; e < nE; e += step) {
      auto &E = EL(e);
      el_body_stress(E);
      el_geom_scalars(E);
    }
  };
  // Per 3x3 block of the element matrices: R^T K R, R^T G R and the Newmark
  // tangent matrix (Segment_UpdateMatrix, SolidSolver.f90:131-157); each
  // entry is summed as el_rkr sums it.
  auto matrices_phase = [&]() {
    // one 3x3 block per task, the block the same across a wave (64 element
    // slots per block), so that the block's stiffness and geometric patterns
    // are known where the products are formed (fem_rkr_block)
    const int epad = (nE + 63) & ~63;
    for (int t = threadIdx.x; t < 16 * epad; t += blockDim.x) {
      const int blk = t / epad, e = t - blk * epad;
      if (e >= nE) continue;
      auto &E = EL(e);
      double *cl = coefL ? coefL + (size_t)e * kCoefStride : nullptr;
      switch (blk) {
#define FEM_BLK(B) \
  case B: fem_rkr_block<B, HOT>(E, cl, c, D); break;
        FEM_BLK(0) FEM_BLK(1) FEM_BLK(2) FEM_BLK(3) FEM_BLK(4) FEM_BLK(5) FEM_BLK(6) FEM_BLK(7)
        FEM_BLK(8) FEM_BLK(9) FEM_BLK(10) FEM_BLK(11) FEM_BLK(12) FEM_BLK(13) FEM_BLK(14) FEM_BLK(15)
#undef FEM_BLK
        default: break;
      }
    }
  };
  auto load_phase = [&](int t0, int step) {
    for (int e = t0; e < nE; e += step)
      el_update_load(EL(e), c, D.dampM, D.dampK, A.dspO + g0, A.dsp + g0, A.vel + g0, A.acc + g0);
  };
  // internal, external and effective loads per DOF (LocToGlobal order)
  auto assembly_phase = [&](int t0, int step) {
    for (int g = t0; g < nG; g += step) {
      const double ext = A.lodExte[g0 + g];
      double li = 0.0, le;
      if (HOT && C.on) {
        // the same lists from LDS (the CG's copy, entry = element * 12 + index)
        const int q0 = C.ptr[g], q1 = C.ptr[g + 1];
        for (int q = q0; q < q1; ++q) {
          const int ent = C.ent[q], e = ent / NED;
          li = li + EL(e).fint[ent - e * NED];
        }
        le = ext - li;
        for (int q = q0; q < q1; ++q) {
          const int ent = C.ent[q], e = ent / NED;
          le = le + EL(e).fload[ent - e * NED];
        }
      } else {
        for (int k = A.dof_ptr[g0 + g]; k < A.dof_ptr[g0 + g + 1]; ++k)
          li = li + EL(A.dof_e[k] - D.eoff).fint[A.dof_i[k]];
        le = ext - li;
        for (int k = A.dof_ptr[g0 + g]; k < A.dof_ptr[g0 + g + 1]; ++k) {
          le = le + EL(A.dof_e[k] - D.eoff).fload[A.dof_i[k]];
          if constexpr (!HOT)
            if (D.dampK > 0.0) le = le + A.el[A.dof_e[k]].fdamp[A.dof_i[k]];
        }
      }
      A.lodInte[g0 + g] = li;
      A.lodEffe[g0 + g] = le;
    }
  };
  for (iter = 1; iter <= D.ntolFEM; ++iter) {
    if constexpr (HOT) {
      // Without stiffness damping the element loads do not read the rotated
      // stiffness: the internal forces (even waves) and the element loads
      // (odd waves, so on other SIMDs than the first forces) run side by
      // side, then the element matrices and the DOF loads (from the top
      // thread down).
      // The internal forces in two parts on the four SIMDs: the first end's
      // rotation (waves 0, 4), the element loads (1, 5), the second end's
      // rotation (2, 6) and the axial force with the geometric-stiffness
      // scalars (3, 7); then the forces themselves on waves 0, 4.  The end
      // rotations wait in the dot-product staging area (free here).
      const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63, nw4 = (int)blockDim.x >> 8;
      double *rot_ab = lds;  // [e][6]: a, b
      for (int e = (w >> 2) * 64 + lane; e < nE; e += nw4 * 64) {
        auto &E = EL(e);
        switch (w & 3) {
          case 0: el_stress_end(E, 0, rot_ab + 6 * e); break;
          case 1: el_update_load(E, c, D.dampM, D.dampK, A.dspO + g0, A.dsp + g0, A.vel + g0, A.acc + g0); break;
          case 2: el_stress_end(E, 1, rot_ab + 6 * e + 3); break;
          default:
            E.geoFRM = el_geo_frm(el_stress_dl(E), E.prop, E.len0);
            el_geom_scalars(E);
            break;
        }
      }
      fem_sync();
      if ((w & 3) == 0)
        for (int e = (w >> 2) * 64 + lane; e < nE; e += nw4 * 64) {
          auto &E = EL(e);
          el_stress_force(E, el_stress_dl(E), rot_ab + 6 * e, rot_ab + 6 * e + 3);
        }
      fem_sync();
      matrices_phase();
      assembly_phase((int)blockDim.x - 1 - (int)threadIdx.x, blockDim.x);
      fem_sync();
    } else {
      stress_phase(threadIdx.x, blockDim.x);
      fem_sync();
      matrices_phase();
      fem_sync();
      load_phase(threadIdx.x, blockDim.x);  // (stiffness damping reads what the matrices wrote)
      fem_sync();
      assembly_phase(threadIdx.x, blockDim.x);
      fem_sync();
    }
    // CG_Solve (SolidSolver.f90:1949-2020), block-Jacobi preconditioner
    for (int g = threadIdx.x; g < nG; g += blockDim.x) {
      const double xf = V.fixed[g] ? (iter == 1 ? A.vBC[g0 + g] : 0.0) : 0.0;
      V.xF[g] = xf;
      V.x[g] = xf;
    }
    // block-Jacobi preconditioner (SolidSolver.f90:1966-1975, Invert6x6
    // :2234-2279): Gauss-Jordan on each node's 6x12 block with one lane per
    // row holding its row in registers, ten nodes per wave (lanes 0..59).
    // Row exchanges are kept as a permutation (pr: logical row -> the node's
    // lane holding it, the same in the node's six lanes) instead of moving
    // rows; pivot columns and the pivot row move between the node's lanes by
    // shuffles.  Step i divides and eliminates only the columns later steps
    // read (left-half columns after i, the whole right half): a column's
    // values depend only on itself and the pivot column, and a left-half
    // column is not read after its own step, so the inverse (the right half)
    // and every pivot choice are the sequential routine's, operation for
    // operation.
    {
      const int lane = (int)threadIdx.x & 63;
      const int wv = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
      const int ln = lane / 6, row = lane - 6 * ln;
      const int lbase = 6 * ln;
      for (int base = wv * 10; base < nN; base += nw * 10) {
        const int n = base + ln;
        const bool act = lane < 60 && n < nN;
        double a[12];
#pragma unroll
        for (int j = 0; j < 12; ++j) a[j] = 0.0;
        int fxr = 0, fxc[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) fxc[j] = 0;
        if (act) {
          for (int k = A.node_ptr[D.noff + n]; k < A.node_ptr[D.noff + n + 1]; ++k) {
            const int o = A.node_h[k] ? 6 : 0;
            if (coefL) {
              const double *cl = coefL + (A.node_e[k] - D.eoff) * kCoefStride;
#pragma unroll
              for (int j = 0; j < 6; ++j) a[j] = a[j] + cl[(o + j) * NED + o + row];
            } else {
              const BeamElem &E = A.el[A.node_e[k]];
#pragma unroll
              for (int j = 0; j < 6; ++j) a[j] = a[j] + E.coef[(o + row) * NED + o + j];
            }
          }
#pragma unroll
          for (int j = 0; j < 6; ++j) fxc[j] = V.fixed[6 * n + j];
          fxr = V.fixed[6 * n + row];
        }
        // fixed DOFs: their rows and columns zeroed, unit diagonal
#pragma unroll
        for (int j = 0; j < 6; ++j)
          if (fxr || fxc[j]) a[j] = (j == row && fxr) ? 1.0 : 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) a[6 + j] = j == row ? 1.0 : 0.0;
        int bad = 0;
        int pr[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) pr[k] = k;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          // column i of the node's logical rows, pivot as the sequential search picks it
          double col[6];
#pragma unroll
          for (int k = i; k < 6; ++k) col[k] = __shfl(a[i], lbase + pr[k], 64);
          int p = i;
          double pivot = fabs(col[i]);
#pragma unroll
          for (int k = i + 1; k < 6; ++k)
            if (fabs(col[k]) > pivot) {
              pivot = fabs(col[k]);
              p = k;
            }
          if (!bad && pivot <= 1.0e-30) bad = i + 1;
          // logical rows i and p exchanged
          if (!bad) {
            const int pi = pr[i];
            int pp = pi;
#pragma unroll
            for (int k = i + 1; k < 6; ++k)
              if (k == p) pp = pr[k];
#pragma unroll
            for (int k = i + 1; k < 6; ++k)
              if (k == p) pr[k] = pi;
            pr[i] = pp;
          }
          const int P = pr[i];  // the node's lane holding the pivot row
          // the pivot row divided by the pivot: the 11 - i quotients are
          // shared out over the node's six lanes (two divisions per lane
          // instead of 11 - i by the pivot's lane, whose divisions the whole
          // wave would issue), then gathered by every lane
          const double pv = __shfl(a[i], lbase + P, 64);
          double raw[12];
#pragma unroll
          for (int j = i + 1; j < 12; ++j) raw[j] = __shfl(a[j], lbase + P, 64);
          double n0 = 0.0, n1 = 0.0;
#pragma unroll
          for (int kk = 0; kk < 6; ++kk) {
            if (row == kk && i + 1 + kk < 12) n0 = raw[i + 1 + kk];
            if (row == kk && i + 7 + kk < 12) n1 = raw[i + 7 + kk];
          }
          double q0 = 0.0, q1 = 0.0;
          if (!bad) {
            q0 = n0 / pv;
            q1 = n1 / pv;
          }
          double prow[12];
#pragma unroll
          for (int j = i + 1; j < 12; ++j) {
            const int kk = j - i - 1;
            prow[j] = kk < 6 ? __shfl(q0, lbase + kk, 64) : __shfl(q1, lbase + kk - 6, 64);
          }
          if (!bad && row == P) {
#pragma unroll
            for (int j = i + 1; j < 12; ++j) a[j] = prow[j];
          }
          if (!bad && row != P) {
            const double factor = a[i];
#pragma unroll
            for (int j = i + 1; j < 12; ++j) a[j] = a[j] - factor * prow[j];
          }
        }
        // this lane's logical row
        int L = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (pr[k] == row) L = k;
        if (act) {
          double *out = V.blk + (size_t)n * 36;
          // (a singular block leaves the identity; pr is then still the start)
#pragma unroll
          for (int j = 0; j < 6; ++j) out[L * 6 + j] = bad ? (L == j ? 1.0 : 0.0) : a[6 + j];
        }
      }
    }
    fem_sync();
    if (D.tree_sums && C.on && C.coef && V.en) {
      if (const int st = fem_cg_fused(A, D, V, C, cg_total)) { status = st; break; }
    } else {
    fem_matvec(A, D, V, C, V.x, V.Ap);
    for (int g = threadIdx.x; g < nG; g += blockDim.x)
      V.r[g] = V.fixed[g] ? 0.0 : A.lodEffe[g0 + g] - V.Ap[g];
    fem_sync();
    fem_precond(A, D, V);
    const double err = 1.0e-6;
    const int cap = nG;  // product staging: [0, nG) and [nG, 2 nG)
    double rr, rz;
    fem_dot2(V.r, V.r, V.r, V.z, nG, sh_pair, lds, cap, rr, rz, D.tree_sums);
    double rn = sqrt(rr);
    if (!(rn <= err)) {
      for (int g = threadIdx.x; g < nG; g += blockDim.x) V.p[g] = V.z[g];
      fem_sync();
      double rsold = rz;
      if (fabs(rsold) <= 1.0e-30) { status = 2; break; }
      for (int it = 1; it <= 10000; ++it) {
        ++cg_total;
        // A p with the fixed DOFs zeroed, and the products p_g (A p)_g
        fem_matvec(A, D, V, C, V.p, V.Ap, true, lds);
        const double pAp = fem_serial_sum(lds, nG, sh_pair, D.tree_sums);
        if (fabs(pAp) <= 1.0e-30) { status = 2; break; }
        const double alpha = rsold / pAp;
        // x, r update, then z = M^-1 r (formed before the convergence test,
        // unused when the loop stops there, so that r.r and r.z are summed
        // side by side), with sixty DOFs (ten whole nodes) per wave: a node's
        // block-Jacobi row reads only r values of its own wave
        {
          const int nw = (int)blockDim.x >> 6, wv = (int)threadIdx.x >> 6, ln = (int)threadIdx.x & 63;
          for (int base = 0; base < nG; base += 60 * nw) {
            const int g = base + wv * 60 + ln;
            const bool on = ln < 60 && g < nG;
            if (on) {
              double xv = V.x[g] + alpha * V.p[g];
              if (V.fixed[g]) xv = V.xF[g];
              V.x[g] = xv;
              double rv = V.r[g] - alpha * V.Ap[g];
              if (V.fixed[g]) rv = 0.0;
              V.r[g] = rv;
            }
            wave_sync();
            if (on) {
              const int node = g / 6, i = g - node * 6;
              const double *B = V.blk + (size_t)node * 36;
              double sz = 0.0;
#pragma unroll
              for (int k = 0; k < 6; ++k) sz = sz + B[i * 6 + k] * V.r[6 * node + k];
              const double z = V.fixed[g] ? 0.0 : sz;
              V.z[g] = z;
              const double r = V.r[g];
              lds[g] = r * r;
              lds[cap + g] = r * z;
            }
          }
        }
        fem_sync();
        double rsnew;
        fem_pair_sums(lds, cap, nG, sh_pair, rr, rsnew, D.tree_sums);
        rn = sqrt(rr);
        if (rn <= err) break;
        if (fabs(rsold) <= 1.0e-30) { status = 2; break; }
        const double beta = rsnew / rsold;
        for (int g = threadIdx.x; g < nG; g += blockDim.x) {
          double pv = V.z[g] + beta * V.p[g];
          if (V.fixed[g]) pv = 0.0;
          V.p[g] = pv;
        }
        fem_sync();
        rsold = rsnew;
      }
      if (status) break;
    }
    }
    // UpdateDspANDTride (SolidSolver.f90:2304-2338); the ramp factor is 1.
    // dnorm = |maxval(dsp_increment**2)| (SolidSolver.f90:2336): the maximum
    // does not depend on the order; each wave's maximum goes to lds[wave]
    {
      double mx = -INFINITY;
      for (int g = threadIdx.x; g < nG; g += blockDim.x) {
        A.dsp[g0 + g] = A.dsp[g0 + g] + V.x[g];
        const double v = V.x[g] * V.x[g];
        if (v > mx) mx = v;
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(mx, off, 64);
        if (o > mx) mx = o;
      }
      if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
    }
    fem_sync();
    // the two end triads on waves of different SIMDs (even waves: the first
    // end and the current geometry; odd waves: the second end), then the
    // element triad
    {
      const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63, nw2 = (int)blockDim.x >> 7;
      for (int e = (w >> 1) * 64 + lane; e < nE; e += nw2 * 64) {
        auto &E = EL(e);
        if ((w & 1) == 0) {
          const double *x0 = A.el[D.eoff + e].x0;
          for (int k = 0; k < 6; ++k) {
            E.x1[k] = x0[k] + A.dsp[g0 + 6 * E.node0 + k];
            E.x1[6 + k] = x0[6 + k] + A.dsp[g0 + 6 * E.node1 + k];
          }
          el_cptdxyz1(E);
          el_update_node_triad(E, V.x, 0);
        } else {
          el_update_node_triad(E, V.x, 1);
        }
      }
    }
    fem_sync();
    for (int e = threadIdx.x; e < nE; e += blockDim.x) el_make_triad_ee(EL(e));
    {
      double m = lds[0];
      for (int w = 1; w < (int)(blockDim.x + 63) / 64; ++w)
        if (lds[w] > m) m = lds[w];
      dnorm = fabs(m);
    }
    fem_sync();
    // (nothing before the next rewrite of lds[0..waves) reads pos)
    for (int n = threadIdx.x; n < nN; n += blockDim.x) {
      const int last = A.node_ptr[D.noff + n + 1] - 1;
      const auto &E = EL(A.node_e[last] - D.eoff);
      const double *src = A.node_h[last] ? E.x1 + 6 : E.x1;
      for (int k = 0; k < 6; ++k) A.pos[g0 + 6 * n + k] = src[k];
    }
    if (dnorm <= D.dtolFEM) break;
  }
  // UpdateVelAcc (SolidSolver.f90:2350-2359)
  for (int g = threadIdx.x; g < nG; g += blockDim.x) {
    const double a = c[0] * (A.dsp[g0 + g] - A.dspO[g0 + g]) - c[2] * A.velO[g0 + g] - c[3] * A.accO[g0 + g];
    A.acc[g0 + g] = a;
    A.vel[g0 + g] = A.velO[g0 + g] + c[6] * A.accO[g0 + g] + c[7] * a;
  }
  if (threadIdx.x == 0) {
    D.newton_iters = iter > D.ntolFEM ? D.ntolFEM : iter;
    D.cg_iters = cg_total;
    D.status = status;
    D.dnorm = dnorm;
  }
  (void)sh_flag;
}

// numsubstep sub-steps of every body in one launch: workgroup b runs body b's
// sub-steps in order, sub-step s from descs[s * nbodies + b].  The waves issue
// at raised priority (s_setprio): the collide kernel's waves that share the CU
// are HBM-bound and lose little, while this chain of dependent reductions
// would otherwise wait behind them for every instruction.
//
// HOT: a flexible body's element fields of the Newton iterations (BeamHot)
// are read from the element array into LDS at hot_off before the first
// sub-step and written back after the last; the host takes this form when
// the LDS holds it, no stiffness damping is set and every element's mass
// blocks repeat (BeamHot).
template <bool HOT>
__global__ __launch_bounds__(512) void k_fem_substep(FemBodyDesc *descs, int nbodies, int nsub, FemArrays A,
                                                   int work, int vec_lds, int cg_lds, int coef_off, int hot_off) {
  __builtin_amdgcn_s_setprio(3);
  double *lds = fem_lds_base();
  // descs is pinned host memory: each sub-step's descriptor is read into LDS
  // once, and its results are written back
  __shared__ FemBodyDesc sD;
  static_assert(sizeof(FemBodyDesc) % 8 == 0 && sizeof(FemBodyDesc) / 8 <= 512, "descriptor copy by 8-byte words");
  BeamHot *H = HOT ? reinterpret_cast<BeamHot *>(lds + hot_off) : nullptr;
  long long *hw = reinterpret_cast<long long *>(H);
  bool hot = false;
  for (int isub = 0; isub < nsub; ++isub) {
    FemBodyDesc *g = descs + (size_t)isub * nbodies + blockIdx.x;
    if (threadIdx.x < sizeof(FemBodyDesc) / 8)
      reinterpret_cast<long long *>(&sD)[threadIdx.x] = reinterpret_cast<const long long *>(g)[threadIdx.x];
    fem_sync();
    if constexpr (HOT) {
      if (isub == 0 && sD.iBodyModel != 1) {
        hot = true;
        const int nW = sD.nEL * kHotWords;
        for (int t = threadIdx.x; t < nW; t += blockDim.x) {
          const int e = t / kHotWords, w = t - e * kHotWords;
          hw[t] = reinterpret_cast<const long long *>(A.el + sD.eoff + e)[kHotMapD.src[w]];
        }
        fem_sync();
      }
    }
    // (a rigid body's sub-step does not reach the element fields)
    fem_substep_body<HOT>(sD, A, work, vec_lds, cg_lds, coef_off, H);
    fem_sync();
    if (threadIdx.x == 0) {
      g->newton_iters = sD.newton_iters;
      g->cg_iters = sD.cg_iters;
      g->status = sD.status;
      g->dnorm = sD.dnorm;
    }
  }
  if (hot) {
    const int nW = sD.nEL * kHotWords;
    for (int t = threadIdx.x; t < nW; t += blockDim.x) {
      const int e = t / kHotWords, w = t - e * kHotWords;
      reinterpret_cast<long long *>(A.el + sD.eoff + e)[kHotMapD.src[w]] = hw[t];
    }
  }
}

// PlateUpdatePosVelArea_ (Solidbody.f90:604-646): one thread per element,
// markers written in the body's marker order; elem_xc = beam axis midpoints.
struct FemMarkerArgs {
  const BeamElem *el;
  const double *pos, *vel;
  const int *body_of_elem, *elem_first_marker, *body_goff;
  const double *beta;   // per body: -(alpha*2*denIn)
  int nEL_total;
  double *Exyz, *Evel, *Ea, *elem_xc;
  double *Enrm;         // marker normals (chord x span direction), for the warm start; may be null
};

// one wave per element: every lane forms the element's mid-axis state, lane
// l writes span markers l, l+64, ...
__global__ __launch_bounds__(256) void k_fem_markers(FemMarkerArgs a) {
  const int lane = threadIdx.x & 63;
  const int e = blockIdx.x * 4 + ((int)threadIdx.x >> 6);
  if (e >= a.nEL_total) return;
  const BeamElem &E = a.el[e];
  const int body = a.body_of_elem[e];
  const int g0 = a.body_goff[body];
  const int i1 = E.node0, i2 = E.node1;
  double txyz[3], tvel[3], om[3], dirc[3];
  for (int k = 0; k < 3; ++k) {
    txyz[k] = 0.5 * (a.pos[g0 + 6 * i1 + k] + a.pos[g0 + 6 * i2 + k]);
    tvel[k] = 0.5 * (a.vel[g0 + 6 * i1 + k] + a.vel[g0 + 6 * i2 + k]);
    om[k] = 0.5 * (a.vel[g0 + 6 * i1 + 3 + k] + a.vel[g0 + 6 * i2 + 3 + k]);
  }
  const double dl = E.spanlen / (double)E.Nspan;
  const double area = dl * E.len1 * a.beta[body];
  for (int k = 0; k < 3; ++k) dirc[k] = 0.5 * (E.tn1[k * 3 + 1] + E.tn2[k * 3 + 1]);
  const double nrm = sqrt(bdot3(dirc, dirc));
  if (nrm > 1.0e-12) for (int k = 0; k < 3; ++k) dirc[k] = dirc[k] / nrm;
  else for (int k = 0; k < 3; ++k) dirc[k] = E.tee[k * 3 + 1];
  const int m0 = a.elem_first_marker[e];
  for (int s = lane; s < E.Nspan; s += 64) {
    const double ls = dl * (0.5 + (double)s) - E.Lspan;
    double r[3], w[3];
    for (int k = 0; k < 3; ++k) r[k] = dirc[k] * ls;
    w[0] = om[1] * r[2] - om[2] * r[1];
    w[1] = om[2] * r[0] - om[0] * r[2];
    w[2] = om[0] * r[1] - om[1] * r[0];
    for (int k = 0; k < 3; ++k) {
      a.Exyz[3LL * (m0 + s) + k] = txyz[k] + r[k];
      a.Evel[3LL * (m0 + s) + k] = tvel[k] + w[k];
    }
    a.Ea[m0 + s] = area;
  }
  if (a.Enrm) {
    double ax[3], nv[3];
    for (int k = 0; k < 3; ++k) ax[k] = E.x1[6 + k] - E.x1[k];
    nv[0] = ax[1] * dirc[2] - ax[2] * dirc[1];
    nv[1] = ax[2] * dirc[0] - ax[0] * dirc[2];
    nv[2] = ax[0] * dirc[1] - ax[1] * dirc[0];
    const double nn = sqrt(bdot3(nv, nv));
    for (int s = lane; s < E.Nspan; s += 64)
      for (int k = 0; k < 3; ++k) a.Enrm[3LL * (m0 + s) + k] = nn > 0.0 ? nv[k] / nn : 0.0;
  }
  if (lane == 0)
    for (int k = 0; k < 3; ++k) a.elem_xc[3LL * e + k] = 0.5 * (E.x1[k] + E.x1[6 + k]);
}

// ---------------------------------------------------------------------------
// Beam_write_solid_info (SolidSolver.f90:1729-1799): the values of one row of
// each DatInfo file of one body, every sum in the reference's order.
// refs = Lref, Uref, Aref, Fref, Pref, Eref (Solidbody.f90:276-280).
// rows (78): firstNode[15] lastNode[15] centerNode[15] nodeAverage[12]
// forces[6] power[7] energy[8].  One workgroup: the element energies in
// parallel (parts[2*nEL..]), then thread 0 adds everything in order.
struct SolidInfoArgs {
  const BeamElem *el;   // the body's elements
  const double *pos, *vel, *acc, *lod;  // the body's node vectors (6 per node) and lodFlow
  int nND, nEL;
  double XYZo[3], refs[6];
  double *parts;        // 3 * nEL scratch: strain (2) and kinetic energy per element
  double *rows;         // 78 outputs
};

__device__ inline void solid_node_row(const SolidInfoArgs &a, int n, double *r) {
  const double Lref = a.refs[0], Uref = a.refs[1], Aref = a.refs[2];
  for (int k = 0; k < 3; ++k) {
    r[k] = a.XYZo[k] / Lref;
    r[3 + k] = (a.pos[6 * n + k] - a.XYZo[k]) / Lref;
    r[6 + k] = a.pos[6 * n + 3 + k];
    r[9 + k] = a.vel[6 * n + k] / Uref;
    r[12 + k] = a.acc[6 * n + k] / Aref;
  }
}

__global__ __launch_bounds__(256) void k_solid_info(SolidInfoArgs a) {
  for (int i = threadIdx.x; i < a.nEL; i += blockDim.x) {
    const BeamElem &e = a.el[i];
    double es, eb;
    el_strain_energy(e, &es, &eb);
    double ve[NED];
    for (int k = 0; k < 6; ++k) {
      ve[k] = a.vel[6 * e.node0 + k];
      ve[6 + k] = a.vel[6 * e.node1 + k];
    }
    double d = 0.0;
    for (int j = 0; j < NED; ++j) {
      double mv = 0.0;
      for (int k = 0; k < NED; ++k) mv = mv + e.mas[j * NED + k] * ve[k];
      d = d + ve[j] * mv;
    }
    a.parts[3 * i] = es;
    a.parts[3 * i + 1] = eb;
    a.parts[3 * i + 2] = 0.5 * d;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double Lref = a.refs[0], Uref = a.refs[1], Aref = a.refs[2], Fref = a.refs[3], Pref = a.refs[4],
               Eref = a.refs[5];
  const int nND = a.nND;
  solid_node_row(a, 0, a.rows);
  solid_node_row(a, nND - 1, a.rows + 15);
  solid_node_row(a, (nND + 1) / 2 - 1, a.rows + 30);
  double *r = a.rows + 45;
  for (int k = 0; k < 3; ++k) {
    // lumped node masses mss(k, n): diagonal of the element mass matrices, element order
    double ms = 0.0, xs = 0.0, vs = 0.0, as = 0.0;
    for (int n = 0; n < nND; ++n) {
      double m = 0.0;
      for (int i = 0; i < a.nEL; ++i) {
        const BeamElem &e = a.el[i];
        if (e.node0 == n) m = m + e.mas[k * NED + k];
        if (e.node1 == n) m = m + e.mas[(k + 6) * NED + k + 6];
      }
      ms = ms + m;
      xs = xs + a.pos[6 * n + k] * m;
      vs = vs + a.vel[6 * n + k] * m;
      as = as + a.acc[6 * n + k] * m;
    }
    r[k] = a.XYZo[k] / Lref;
    r[3 + k] = (xs / ms - a.XYZo[k]) / Lref;
    r[6 + k] = (vs / ms) / Uref;
    r[9 + k] = (as / ms) / Aref;
  }
  r = a.rows + 57;
  double P[3];
  for (int k = 0; k < 3; ++k) {
    double f = 0.0, p = 0.0;
    for (int n = 0; n < nND; ++n) f = f + a.lod[6 * n + k];
    for (int n = 0; n < nND; ++n) p = p + a.lod[6 * n + k] * a.vel[6 * n + k];
    r[k] = a.XYZo[k] / Lref;
    r[3 + k] = f / Fref;
    P[k] = p / Pref;
  }
  double *pw = a.rows + 63;
  for (int k = 0; k < 3; ++k) pw[k] = a.XYZo[k] / Lref;
  pw[3] = (P[0] + P[1]) + P[2];
  pw[4] = P[0];
  pw[5] = P[1];
  pw[6] = P[2];
  double Es = 0.0, Eb = 0.0, Ev = 0.0;
  for (int i = 0; i < a.nEL; ++i) {
    Es = Es + a.parts[3 * i];
    Eb = Eb + a.parts[3 * i + 1];
  }
  Es = Es / Eref;
  Eb = Eb / Eref;
  const double Ep = Es + Eb;
  for (int i = 0; i < a.nEL; ++i) Ev = Ev + a.parts[3 * i + 2];
  Ev = Ev / Eref;
  double *en = a.rows + 70;
  for (int k = 0; k < 3; ++k) en[k] = a.XYZo[k] / Lref;
  en[3] = Ev + Ep;
  en[4] = Ev;
  en[5] = Ep;
  en[6] = Es;
  en[7] = Eb;
}

}  // namespace fsil
