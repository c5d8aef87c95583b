// Immersed-boundary multi-direct-forcing kernels for gfx950 (FP64).
//
// Data model.  Every marker m owns 64 stencil entries e = 64*m + s with
// s = 16*ix + 4*iy + iz in the reference's x-outer, y, z-inner loop order
// (Solidbody.f90:1009-1015).  The cells those entries touch form the "band".
// Entries are sorted by (cell, e), which is the order in which the reference's
// serial loops over markers (Solidbody.f90:1034-1048, 968-977) update a cell;
// a per-cell sequential reduction over its sorted entries therefore rounds
// exactly as the reference does, with no atomics.
//
// The whole structure is built on the device each coupling call: stencils
// (k_ibm_stencil), a radix sort of the 64-bit keys (cell << 26 | e), and the
// band assembly (k_ibm_band).  Band arrays have a fixed stride `cap` = 64*M
// (an upper bound of the band size) and the band size itself lives in device
// memory (ctl[kNband]), so nothing is read back between the kernels.
#pragma once
#include "fsil_common.h"
#include "fsil_lbm_kernels.h"

namespace fsil {

// control words of the coupling kept in device memory
enum IbmCtl {
  kBoxLo = 0,      // [0..2] lowest band cell coordinate per axis (atomic min)
  kBoxHi = 3,      // [3..5] highest band cell coordinate per axis (atomic max)
  kNband = 6,      // number of band cells
  kErr = 7,        // first failing marker (1-based) * 4 + reason, INT_MAX when none
  kDone = 8,       // multi-direct-forcing loop finished
  kIters = 9,      // iterations run
  kWarm = 10,      // learned warm start in use for this call
  kFallback = 11,  // 1: the warm start did not lower the residual, restart from zero; 2: restarted
  kDone0 = 12,     // the done word at the start of a loop (no iteration at all)
  kCtlWords = 16
};
constexpr int kKeyShift = 26;  // e = 64*m + s < 2^26

// 4-point Peskin kernel (Solidbody.f90:822-833), FP64; sqrt is correctly rounded
__host__ __device__ inline double peskin4(double x) {
  const double r = fabs(x);
  if (r < 1.0) return (3.0 - 2.0 * r + sqrt(1.0 + 4.0 * r * (1.0 - r))) * 0.125;
  if (r < 2.0) return (5.0 - 2.0 * r - sqrt(-7.0 + 4.0 * r * (3.0 - r))) * 0.125;
  return 0.0;
}

// trimedindex (Solidbody.f90:834-866): 1-based indices i-1..i+2 wrapped or
// mirrored by the face codes; returns 0, or 1 (below the min face) / 2
// (beyond the max face) for an index the face code cannot place, in which
// case the index is clamped into the block so later kernels stay in bounds.
__host__ __device__ inline int trim4(int i, int nd, int *out, int bclo, int bchi) {
  int err = 0;
  for (int k = -1; k <= 2; ++k) {
    int v = i + k;
    if (v < 1) {
      if (bclo == BC_PERIODIC) v += nd;
      else if ((bclo == BC_SYMMETRIC || bclo == BC_WALL) && v == 0) v = 2;
      else if (bclo == BC_WALL_HALFWAY && v == 0) v = 1;
      else { err = err ? err : 1; v = 1; }
    } else if (v > nd) {
      if (bchi == BC_PERIODIC) v -= nd;
      else if ((bchi == BC_SYMMETRIC || bchi == BC_WALL) && v == nd + 1) v = nd - 1;
      else if (bchi == BC_WALL_HALFWAY && v == nd + 1) v = nd;
      else { err = err ? err : 2; v = nd; }
    }
    if (v < 1 || v > nd) { err = err ? err : (v < 1 ? 1 : 2); v = v < 1 ? 1 : nd; }
    out[k + 1] = v;
  }
  return err;
}

// reset of the control words; clears the previous band from the cell map
__global__ __launch_bounds__(256) void k_ibm_reset(int *__restrict__ ctl, const long long *__restrict__ prev_cell,
                                                   int *__restrict__ cellmap, double *__restrict__ dstate,
                                                   int done0, int warm) {
  const int nprev = ctl[kNband];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nprev; i += gridDim.x * blockDim.x)
    cellmap[prev_cell[i]] = -1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl[kBoxLo] = ctl[kBoxLo + 1] = ctl[kBoxLo + 2] = 0x7fffffff;
    ctl[kBoxHi] = ctl[kBoxHi + 1] = ctl[kBoxHi + 2] = -1;
    ctl[kErr] = 0x7fffffff;
    ctl[kDone] = done0;
    ctl[kIters] = 0;
    ctl[kWarm] = warm;
    ctl[kFallback] = 0;
    ctl[kDone0] = done0;
    dstate[0] = 1e10;  // dmax before the first iteration (Solidbody.f90:893)
  }
}

struct StencilArgs {
  int M;
  const double *__restrict__ Exyz;       // 3*M
  const int *__restrict__ first_marker;  // per marker: first marker of its body
  double org[3];
  double dh, invdh;
  int nd[3];
  int bc[6];
  float *__restrict__ Ew;                // 12*M
  unsigned long long *__restrict__ keys; // 64*M
  int *__restrict__ ctl;
};

// UpdateElmtInterp_ (Solidbody.f90:760-821) per marker: anchor from the
// body's first marker, trimmed 4x3 indices, FP32 weights, and the 64 keys.
__global__ __launch_bounds__(256) void k_ibm_stencil(StencilArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.M) return;
  const int m0 = a.first_marker[m];
  int ix[12];
  int err = 0;
  int lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ii = (int)floor((a.Exyz[3LL * m0 + k] - a.org[k]) * a.invdh);
    const double x0 = a.org[k] + (double)ii * a.dh;
    double off = (a.Exyz[3LL * m + k] - x0) * a.invdh;
    int idx = (int)floor(off);
    off = off - (double)idx;
    idx = idx + ii + 1;
    const int e = trim4(idx, a.nd[k], &ix[4 * k], a.bc[2 * k], a.bc[2 * k + 1]);
    if (e && !err) err = e;
    lo[k] = hi[k] = ix[4 * k];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a.Ew[12LL * m + 4 * k + j] = (float)peskin4((double)(j - 1) - off);
      lo[k] = min(lo[k], ix[4 * k + j]);
      hi[k] = max(hi[k], ix[4 * k + j]);
    }
  }
  if (err) atomicMin(&a.ctl[kErr], (m + 1) * 4 + err);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    atomicMin(&a.ctl[kBoxLo + k], lo[k] - 1);
    atomicMax(&a.ctl[kBoxHi + k], hi[k] - 1);
  }
  unsigned long long *K = a.keys + 64LL * m;
  for (int x = 0; x < 4; ++x)
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int z = 0; z < 4; ++z) {
        const unsigned long long cell =
            ((unsigned long long)(ix[x] - 1) * a.nd[1] + (unsigned long long)(ix[4 + y] - 1)) * a.nd[2] +
            (unsigned long long)(ix[8 + z] - 1);
        const int e = 64 * m + 16 * x + 4 * y + z;
        K[16 * x + 4 * y + z] = (cell << kKeyShift) | (unsigned long long)e;
      }
}

// head flags of the sorted keys (first entry of each band cell)
__global__ __launch_bounds__(256) void k_ibm_heads(int E, const unsigned long long *__restrict__ keys,
                                                   int *__restrict__ flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E) return;
  flag[k] = (k == 0 || (keys[k] >> kKeyShift) != (keys[k - 1] >> kKeyShift)) ? 1 : 0;
}

// band assembly from the sorted keys and the inclusive scan of the heads
__global__ __launch_bounds__(256) void k_ibm_band(int E, const unsigned long long *__restrict__ keys,
                                                  const int *__restrict__ scan, int *__restrict__ entry,
                                                  int *__restrict__ bidx, long long *__restrict__ band_cell,
                                                  int *__restrict__ band_ptr, int *__restrict__ cellmap,
                                                  int *__restrict__ ctl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E) return;
  const unsigned long long key = keys[k];
  const long long cell = (long long)(key >> kKeyShift);
  const int e = (int)(key & ((1ULL << kKeyShift) - 1));
  const int b = scan[k] - 1;
  entry[k] = e;
  bidx[e] = b;
  if (k == 0 || (keys[k - 1] >> kKeyShift) != (key >> kKeyShift)) {
    band_cell[b] = cell;
    band_ptr[b] = k;
    cellmap[cell] = b;
  }
  if (k == E - 1) {
    band_ptr[b + 1] = E;
    ctl[kNband] = b + 1;
  }
}

// band macros: den/u of every band cell from the stored state, same
// expressions as the collision kernel (FluidDomain.f90:1136-1139)
__global__ __launch_bounds__(256) void k_band_macro(const double *__restrict__ f, long long N,
                                                    const long long *__restrict__ cell,
                                                    const int *__restrict__ ctl, int cap,
                                                    double hv0, double hv1, double hv2,
                                                    double *__restrict__ band_u, int only_on_fallback = 0) {
  if (only_on_fallback && ctl[kFallback] != 1) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ctl[kNband]) return;
  const long long n = cell[b];
  double g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) g[q] = f[q * N + n];
  double s, m1, m2, m3;
  moments(g, s, m1, m2, m3);
  band_u[b] = (m1 + hv0) / s;
  band_u[cap + b] = (m2 + hv1) / s;
  band_u[2LL * cap + b] = (m3 + hv2) / s;
}

// band macros of the FP32-deviation state (the immersed-boundary coupling
// runs in FP64): rho = 1 + sum g, j = sum e g, summed in FP64 from the stored
// deviations
__global__ __launch_bounds__(256) void k_band_macro32(const float *__restrict__ g, long long N,
                                                      const long long *__restrict__ cell,
                                                      const int *__restrict__ ctl, int cap, double hv0,
                                                      double hv1, double hv2, double *__restrict__ band_u,
                                                      int only_on_fallback = 0) {
  if (only_on_fallback && ctl[kFallback] != 1) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ctl[kNband]) return;
  const long long n = cell[b];
  double dr = 0.0, j1 = 0.0, j2 = 0.0, j3 = 0.0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double v = (double)g[q * N + n];
    dr = dr + v;
    if (ex(q) != 0) j1 = ex(q) > 0 ? j1 + v : j1 - v;
    if (ey(q) != 0) j2 = ey(q) > 0 ? j2 + v : j2 - v;
    if (ez(q) != 0) j3 = ez(q) > 0 ? j3 + v : j3 - v;
  }
  const double rho = 1.0 + dr;
  band_u[b] = (j1 + hv0) / rho;
  band_u[cap + b] = (j2 + hv1) / rho;
  band_u[2LL * cap + b] = (j3 + hv2) / rho;
}

__global__ void k_zero(double *__restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// plain copy (bandwidth calibration of the HBM counters and the roofline)
__global__ __launch_bounds__(256) void k_copy(const double *__restrict__ a, double *__restrict__ b,
                                              long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// wave64 sum through DPP/permute lanes (deterministic order)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = v + __shfl_down(v, off, 64);
  return v;
}

// One marker per wave: each lane forms one of the 64 interpolation terms
// ((u*rx)*ry)*rz of the marker's stencil (Solidbody.f90:1009-1015), lane 0
// adds them in the reference's x, y, z order, then the velocity defect,
// penalty force and force accumulation (PenaltyForce_, Solidbody.f90:1016-1026).
// The marker's |dV|_1 goes to tl[m] for the residual.
__global__ __launch_bounds__(256) void k_interp(int M, const int *__restrict__ bidx,
                                                const float *__restrict__ Ew,
                                                const double *__restrict__ band_u, int cap,
                                                const int *__restrict__ ctl,
                                                const double *__restrict__ Evel,
                                                const double *__restrict__ Ea, double invh3,
                                                double *__restrict__ Eforce,
                                                double *__restrict__ fT,
                                                double *__restrict__ tl) {
  if (ctl[kDone]) return;  // uniform over the grid
  __shared__ double t[4][3][64];
  const int w = threadIdx.x >> 6, s = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + w;
  if (m < M) {
    const int b = bidx[64LL * m + s];
    const float *W = Ew + 12LL * m;
    const double rx = (double)W[s >> 4], ry = (double)W[4 + ((s >> 2) & 3)], rz = (double)W[8 + (s & 3)];
    t[w][0][s] = ((band_u[b] * rx) * ry) * rz;
    t[w][1][s] = ((band_u[cap + b] * rx) * ry) * rz;
    t[w][2][s] = ((band_u[2LL * cap + b] * rx) * ry) * rz;
  }
  __syncthreads();
  if (m < M && s == 0) {
    double v1 = 0.0, v2 = 0.0, v3 = 0.0;
#pragma unroll 16
    for (int k = 0; k < 64; ++k) {
      v1 = v1 + t[w][0][k];
      v2 = v2 + t[w][1][k];
      v3 = v3 + t[w][2][k];
    }
    const double d1 = Evel[3LL * m] - v1, d2 = Evel[3LL * m + 1] - v2, d3 = Evel[3LL * m + 2] - v3;
    const double a = Ea[m];
    const double F1 = d1 * a, F2 = d2 * a, F3 = d3 * a;
    tl[m] = (fabs(d1) + fabs(d2)) + fabs(d3);
    Eforce[3LL * m] = Eforce[3LL * m] + F1;
    Eforce[3LL * m + 1] = Eforce[3LL * m + 1] + F2;
    Eforce[3LL * m + 2] = Eforce[3LL * m + 2] + F3;
    fT[3LL * m] = F1 * invh3;
    fT[3LL * m + 1] = F2 * invh3;
    fT[3LL * m + 2] = F3 * invh3;
  }
}

// velocity correction of every band cell: u -= fT*rx*ry*rz over its entries
// in (marker, stencil) order (Solidbody.f90:1034-1048)
__global__ __launch_bounds__(256) void k_correct(const int *__restrict__ ctl, int cap,
                                                 const int *__restrict__ ptr,
                                                 const int *__restrict__ entry,
                                                 const float *__restrict__ Ew,
                                                 const double *__restrict__ fT,
                                                 double *__restrict__ band_u) {
  if (ctl[kDone]) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ctl[kNband]) return;
  double u1 = band_u[b], u2 = band_u[cap + b], u3 = band_u[2LL * cap + b];
  for (int k = ptr[b]; k < ptr[b + 1]; ++k) {
    const int e = entry[k];
    const int m = e >> 6, s = e & 63;
    const float *W = Ew + 12LL * m;
    const double rx = (double)W[s >> 4], ry = (double)W[4 + ((s >> 2) & 3)],
                 rz = (double)W[8 + (s & 3)];
    u1 = u1 - ((fT[3LL * m] * rx) * ry) * rz;
    u2 = u2 - ((fT[3LL * m + 1] * rx) * ry) * rz;
    u3 = u3 - ((fT[3LL * m + 2] * rx) * ry) * rz;
  }
  band_u[b] = u1;
  band_u[cap + b] = u2;
  band_u[2LL * cap + b] = u3;
}

// stop test of the multi-direct-forcing loop (Solidbody.f90:895-906): the
// markers' |dV|_1 summed in a fixed order (one workgroup: strided partial sums,
// then a wave and a block tree), dmax = tol / (sum M_b * Uref), one iteration
// counted; sets the done word when the loop condition fails.
__global__ __launch_bounds__(256) void k_mdf_check(int *__restrict__ ctl, double *__restrict__ dstate,
                                                   const double *__restrict__ tl, int M, double dsum, double Uref,
                                                   int ntol, double dtol) {
  if (ctl[kDone]) return;
  double p = 0.0;
  for (int m = threadIdx.x; m < M; m += blockDim.x) p = p + tl[m];
  __shared__ double part[4];
  p = wave_sum(p);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = p;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tol = 0.0;
  for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) tol = tol + part[k];
  const int it = ctl[kIters] + 1;
  ctl[kIters] = it;
  if (!isfinite(tol)) {
    atomicMin(&ctl[kErr], 3);  // "Nan found in PenaltyForce" (Solidbody.f90:1028-1031)
    ctl[kDone] = 1;
    return;
  }
  const double dm = tol / (dsum * Uref);
  dstate[0] = dm;
  if (it == 1 && ctl[kWarm] == 1 && !(dm < dstate[1])) {
    // the guessed force did not lower the residual below the zero guess's:
    // restart the loop from zero (k_ws_fallback, k_band_macro on fallback)
    ctl[kWarm] = 0;
    ctl[kFallback] = 1;
    ctl[kIters] = 0;
    ctl[kDone] = ctl[kDone0];
    dstate[0] = 1e10;
    return;
  }
  if (!(it < ntol && dm > dtol)) ctl[kDone] = 1;
}

// Eulerian force of every band cell: (0 - sum F/h^3*rx*ry*rz) + VF
// (ResetVolumeForce_ FluidDomain.f90:1195-1206, FluidVolumeForce_
// Solidbody.f90:968-977, add_volume_force_ FluidDomain.f90:1182-1193)
__global__ __launch_bounds__(256) void k_spread(const int *__restrict__ ctl, int cap,
                                                const int *__restrict__ ptr,
                                                const int *__restrict__ entry,
                                                const float *__restrict__ Ew,
                                                const double *__restrict__ Eforce, double invh3,
                                                double VF1, double VF2, double VF3,
                                                double *__restrict__ band_F) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ctl[kNband]) return;
  double g1 = 0.0, g2 = 0.0, g3 = 0.0;
  for (int k = ptr[b]; k < ptr[b + 1]; ++k) {
    const int e = entry[k];
    const int m = e >> 6, s = e & 63;
    const float *W = Ew + 12LL * m;
    const double rx = (double)W[s >> 4], ry = (double)W[4 + ((s >> 2) & 3)],
                 rz = (double)W[8 + (s & 3)];
    const double h1 = Eforce[3LL * m] * invh3, h2 = Eforce[3LL * m + 1] * invh3,
                 h3 = Eforce[3LL * m + 2] * invh3;
    g1 = g1 + (-(((h1 * rx) * ry) * rz));
    g2 = g2 + (-(((h2 * rx) * ry) * rz));
    g3 = g3 + (-(((h3 * rx) * ry) * rz));
  }
  band_F[b] = g1 + VF1;
  band_F[cap + b] = g2 + VF2;
  band_F[2LL * cap + b] = g3 + VF3;
}

// Beam nodal loads lodFlow: per degree of freedom, the sum over its markers in
// marker order of 0.5*F (translational) or 0.5*(r x F) (rotational) with r the
// arm from the element axis midpoint (Solidbody.f90:944-967).
// dof_list entry: marker id; dof_kind[d] = component (0..2 force, 3..5 moment).
// k_load_terms forms every term (one thread per list entry), k_loads adds
// each DOF's terms in list order.
__global__ __launch_bounds__(256) void k_load_terms(int nlist, const int *__restrict__ dof_of_entry,
                                                    const int *__restrict__ dof_list,
                                                    const int *__restrict__ dof_kind,
                                                    const int *__restrict__ marker_elem,
                                                    const double *__restrict__ elem_xc,
                                                    const double *__restrict__ Exyz,
                                                    const double *__restrict__ Eforce,
                                                    double *__restrict__ term) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlist) return;
  const int k = dof_kind[dof_of_entry[j]];
  const int m = dof_list[j];
  double val;
  if (k < 3) {
    val = Eforce[3LL * m + k];
  } else {
    const int el = marker_elem[m];
    const double r1 = Exyz[3LL * m] - elem_xc[3LL * el];
    const double r2 = Exyz[3LL * m + 1] - elem_xc[3LL * el + 1];
    const double r3 = Exyz[3LL * m + 2] - elem_xc[3LL * el + 2];
    const double F1 = Eforce[3LL * m], F2 = Eforce[3LL * m + 1], F3 = Eforce[3LL * m + 2];
    if (k == 3) val = r2 * F3 - r3 * F2;
    else if (k == 4) val = r3 * F1 - r1 * F3;
    else val = r1 * F2 - r2 * F1;
  }
  term[j] = 0.5 * val;
}

__global__ __launch_bounds__(256) void k_loads(int ndof, const int *__restrict__ dof_ptr,
                                               const double *__restrict__ term, double *__restrict__ lod) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndof) return;
  double acc = 0.0;
#pragma unroll 8
  for (int j = dof_ptr[d]; j < dof_ptr[d + 1]; ++j) acc = acc + term[j];
  lod[d] = acc;
}

// The control words and the last residual of a coupling call, written into
// pinned host memory (read by the host after the stream's synchronisation).
struct IbmReport {
  int ctl[kCtlWords];
  double dmax;
};

__global__ __launch_bounds__(64) void k_ibm_report(const int *__restrict__ ctl, const double *__restrict__ dstate,
                                                   IbmReport *__restrict__ out) {
  const int t = threadIdx.x;
  if (t < kCtlWords) out->ctl[t] = ctl[t];
  if (t == 0) out->dmax = dstate[0];
}

// ---------------------------------------------------------------------------
// Learned warm start (no counterpart in the reference)

// Features of every marker for the network, one wave per marker: the
// interpolated velocity of the uncorrected band (as k_interp forms it), the
// marker velocity, the marker's force of the previous call and its normal:
//   X = [U/Uref, V/Uref, Fprev/(Uref |Ea|), n]
// and the zero guess's |V - U|_1 (the residual a plain loop starts from).
__global__ __launch_bounds__(256) void k_ws_features(int M, const int *__restrict__ bidx,
                                                     const float *__restrict__ Ew,
                                                     const double *__restrict__ band_u, int cap,
                                                     const double *__restrict__ Evel,
                                                     const double *__restrict__ Ea,
                                                     const double *__restrict__ Eprev,
                                                     const double *__restrict__ Enrm, double Uref,
                                                     float *__restrict__ X, double *__restrict__ tl0) {
  __shared__ double t[4][3][64];
  const int w = threadIdx.x >> 6, s = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + w;
  if (m < M) {
    const int b = bidx[64LL * m + s];
    const float *W = Ew + 12LL * m;
    const double rx = (double)W[s >> 4], ry = (double)W[4 + ((s >> 2) & 3)], rz = (double)W[8 + (s & 3)];
    t[w][0][s] = ((band_u[b] * rx) * ry) * rz;
    t[w][1][s] = ((band_u[cap + b] * rx) * ry) * rz;
    t[w][2][s] = ((band_u[2LL * cap + b] * rx) * ry) * rz;
  }
  __syncthreads();
  if (m < M && s == 0) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 64; ++k) {
      v[0] = v[0] + t[w][0][k];
      v[1] = v[1] + t[w][1][k];
      v[2] = v[2] + t[w][2][k];
    }
    const double fs = Uref * fabs(Ea[m]);
    float *x = X + 12LL * m;
    double r = 0.0;
    for (int c = 0; c < 3; ++c) {
      x[c] = (float)(v[c] / Uref);
      x[3 + c] = (float)(Evel[3LL * m + c] / Uref);
      x[6 + c] = (float)(fs > 0.0 ? Eprev[3LL * m + c] / fs : 0.0);
      x[9 + c] = Enrm ? (float)Enrm[3LL * m + c] : 0.f;
      r = r + fabs(Evel[3LL * m + c] - v[c]);
    }
    tl0[m] = r;
  }
}

// dstate[1] = the zero guess's residual, reduced as k_mdf_check reduces
__global__ __launch_bounds__(256) void k_ws_dm0(const double *__restrict__ tl0, int M, double dsum, double Uref,
                                                double *__restrict__ dstate) {
  double p = 0.0;
  for (int m = threadIdx.x; m < M; m += blockDim.x) p = p + tl0[m];
  __shared__ double part[4];
  p = wave_sum(p);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = p;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tol = 0.0;
  for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) tol = tol + part[k];
  dstate[1] = tol / (dsum * Uref);
}

// the network's guess as the starting force: Eforce = F0, fT = F0 * invh3
// (applied to the band by k_correct as one forcing round)
__global__ __launch_bounds__(256) void k_ws_apply(int M, const float *__restrict__ Y, const double *__restrict__ Ea,
                                                  double Uref, double invh3, double *__restrict__ Eforce,
                                                  double *__restrict__ fT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * M) return;
  const double F0 = (double)Y[i] * (Uref * fabs(Ea[i / 3]));
  Eforce[i] = F0;
  fT[i] = F0 * invh3;
}

// restart from the zero guess when k_mdf_check asked for it
__global__ __launch_bounds__(256) void k_ws_fallback(int M, int *__restrict__ ctl, double *__restrict__ Eforce) {
  if (ctl[kFallback] != 1) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * M) Eforce[i] = 0.0;
}
__global__ void k_ws_fallback_end(int *__restrict__ ctl) {
  if (threadIdx.x == 0 && ctl[kFallback] == 1) ctl[kFallback] = 2;
}

// training targets (the converged force over Uref |Ea|) and the previous
// force of the next call
__global__ __launch_bounds__(256) void k_ws_targets(int M, const double *__restrict__ Eforce,
                                                    const double *__restrict__ Ea, double Uref,
                                                    float *__restrict__ Y, double *__restrict__ Eprev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * M) return;
  const double fs = Uref * fabs(Ea[i / 3]);
  Y[i] = (float)(fs > 0.0 ? Eforce[i] / fs : 0.0);
  Eprev[i] = Eforce[i];
}

}  // namespace fsil
