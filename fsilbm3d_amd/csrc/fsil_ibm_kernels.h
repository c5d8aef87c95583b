// Immersed-boundary multi-direct-forcing kernels for gfx950 (FP64).
//
// Data model.  Every marker m owns 64 stencil entries e = 64*m + s with
// s = 16*ix + 4*iy + iz in the reference's x-outer, y, z-inner loop order
// (Solidbody.f90:1009-1015).  The cells those entries touch form the "band".
// Each band cell keeps its entries in ascending e, which is the order in
// which the reference's serial loops over markers (Solidbody.f90:1034-1048,
// 968-977) update a cell; a per-cell sequential reduction over its entries
// therefore rounds exactly as the reference does, with no atomics on values.
//
// The whole structure is built on the device each coupling call, with no
// global sort: k_ibm_stencil (one wave per marker) forms the stencil and
// every entry bids for its cell in the cell map; k_ibm_claim gives each
// cell's winning entry a band number (from an atomic counter, so the band's
// order varies from call to call while every per-cell result does not);
// k_ibm_count gives every entry its band cell and a slot; an inclusive scan
// of the per-cell counts makes the CSR pointers (and the band's box);
// k_ibm_fill places the
// entries in arrival order and k_ibm_rank moves each to its rank by e among
// its cell's entries.  Band arrays
// have a fixed stride `cap` = 64*M (an upper bound of the band size) and the
// band size itself lives in device memory (ctl[kNband]), so nothing is read
// back between the kernels.
#pragma once
#include "fsil_common.h"
#include "fsil_lbm_kernels.h"

namespace fsil {

// control words of the coupling kept in device memory
enum IbmCtl {
  kBoxLo = 0,      // [0..2] lowest band cell coordinate per axis (the fluid kernels look
  kBoxHi = 3,      // [3..5] highest       up the cell map only inside this box)
  kNband = 6,      // number of band cells
  kErr = 7,        // first failing marker (1-based) * 4 + reason, INT_MAX when none
  kDone = 8,       // multi-direct-forcing loop finished
  kIters = 9,      // iterations run
  kWarm = 10,      // learned warm start in use for this call
  kFallback = 11,  // 1: the warm start did not lower the residual, restart from zero; 2: restarted
  kDone0 = 12,     // the done word at the start of a loop (no iteration at all)
  kNew = 13,       // band cells claimed so far in this call (becomes kNband)
  kRan = 14,       // the last forcing round whose interpolation ran (-1: none)
  kGuess = 15,     // 1: the learned guess was applied in this call (kept on fallback)
  kSkip = 16,      // forcing rounds the applied guess stands for (0 without a guess)
  kCtlWords = 20
};

static_assert(kErr == kBandErrWord, "the fused kernels read the coupling error word at kBandErrWord");

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

// reset of the control words; clears the previous band from the cell map,
// the per-cell entry counts (E of them) of this call and the marker forces
__global__ __launch_bounds__(256) void k_ibm_reset(int *__restrict__ ctl, const long long *__restrict__ prev_cell,
                                                   int *__restrict__ cellmap, double *__restrict__ dstate,
                                                   int done0, int warm, int skip, int E,
                                                   int *__restrict__ count, int *__restrict__ band_ptr,
                                                   double *__restrict__ Eforce, int n3M) {
  const int nprev = ctl[kNband];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nprev; i += gridDim.x * blockDim.x)
    cellmap[prev_cell[i]] = -1;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) count[i] = 0;
  // v_Eforce = 0 (Solidbody.f90:889)
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n3M; i += gridDim.x * blockDim.x) Eforce[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    band_ptr[0] = 0;
    ctl[kNew] = 0;
    ctl[kRan] = -1;
    ctl[kErr] = 0x7fffffff;
    ctl[kDone] = done0;
    // a learned guess stands for the first `skip` forcing rounds: the loop
    // count continues from there
    ctl[kIters] = warm ? skip : 0;
    ctl[kWarm] = warm;
    ctl[kGuess] = warm;
    ctl[kSkip] = warm ? skip : 0;
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
  unsigned *__restrict__ keys;           // 64*M: cell of entry e = 64*m + s
  int *__restrict__ ctl;
  int *__restrict__ cellmap;             // the entries' bids
};

// UpdateElmtInterp_ (Solidbody.f90:760-821), one wave per marker: every lane
// forms the marker's anchor (from its body's first marker) and trimmed 4x3
// indices, lanes 0..11 write the FP32 weights, lane s the cell of entry s and
// that entry's bid in the cell map: every entry bids -(e+2) for its cell, the
// map keeps the lowest bid, so the cell's highest entry wins (cells outside
// the band hold -1).
__global__ __launch_bounds__(256) void k_ibm_stencil(StencilArgs a) {
  const int s = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + ((int)threadIdx.x >> 6);
  if (m >= a.M) return;
  const int m0 = a.first_marker[m];
  const int sel[3] = {s >> 4, (s >> 2) & 3, s & 3};
  int cidx[3];
  int err = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ii = (int)floor((a.Exyz[3LL * m0 + k] - a.org[k]) * a.invdh);
    const double x0 = a.org[k] + (double)ii * a.dh;
    double off = (a.Exyz[3LL * m + k] - x0) * a.invdh;
    int idx = (int)floor(off);
    off = off - (double)idx;
    idx = idx + ii + 1;
    int ix[4];
    const int e = trim4(idx, a.nd[k], ix, a.bc[2 * k], a.bc[2 * k + 1]);
    if (e && !err) err = e;
    cidx[k] = ix[sel[k]];
    if ((s >> 2) == k) a.Ew[12LL * m + s] = (float)peskin4((double)((s & 3) - 1) - off);
  }
  if (s == 0 && err) atomicMin(&a.ctl[kErr], (m + 1) * 4 + err);
  const unsigned key =
      ((unsigned)(cidx[0] - 1) * (unsigned)a.nd[1] + (unsigned)(cidx[1] - 1)) * (unsigned)a.nd[2] +
      (unsigned)(cidx[2] - 1);
  a.keys[64LL * m + s] = key;
  // the entry's bid for its cell
  atomicMin(&a.cellmap[key], -(64 * m + s + 2));
}

// Cell-map accesses of the band build go through device-scope atomics (the
// workgroups that bid for, claim and look up one cell run on different XCDs).
__device__ __forceinline__ int cellmap_load(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the winning entry of each cell takes a band number and writes it into the
// cell map.  One 1024-thread workgroup makes one atomic add on the band
// counter (atomics on one line from many workgroups serialize), and leaves
// its part of the band's bounding box in blkbox[6 * workgroup ..] for
// k_band_scan.
constexpr int kClaimThreads = 1024;
__global__ __launch_bounds__(kClaimThreads) void k_ibm_claim(int E, const unsigned *__restrict__ keys,
                                                             int *__restrict__ cellmap,
                                                             long long *__restrict__ band_cell, int ny, int nz,
                                                             int *__restrict__ blkbox, int *__restrict__ ctl) {
  constexpr int NW = kClaimThreads / 64;
  __shared__ int wave_won[NW], wave_off[NW], blk_base, wbox[NW][6];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const unsigned cell = e < E ? keys[e] : 0u;
  const bool won = e < E && cellmap_load(&cellmap[cell]) == -(e + 2);
  const unsigned long long mask = __ballot(won);
  int v[3];
  {
    const unsigned plane = (unsigned)ny * (unsigned)nz;
    const unsigned x = cell / plane, r = cell - x * plane, y = r / (unsigned)nz;
    v[0] = (int)x;
    v[1] = (int)y;
    v[2] = (int)(r - y * (unsigned)nz);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int lo = won ? v[k] : 0x7fffffff, hi = won ? v[k] : -1;
    for (int off = 32; off > 0; off >>= 1) {
      lo = min(lo, __shfl_xor(lo, off, 64));
      hi = max(hi, __shfl_xor(hi, off, 64));
    }
    if (s == 0) {
      wbox[w][k] = lo;
      wbox[w][3 + k] = hi;
    }
  }
  if (s == 0) wave_won[w] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < NW; ++k) {
      wave_off[k] = t;
      t += wave_won[k];
    }
    blk_base = t ? atomicAdd(&ctl[kNew], t) : 0;
  }
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    int r = wbox[0][k];
    for (int q = 1; q < NW; ++q) r = k < 3 ? min(r, wbox[q][k]) : max(r, wbox[q][k]);
    blkbox[6 * blockIdx.x + k] = r;
  }
  __syncthreads();
  if (won) {
    const int b = blk_base + wave_off[w] + __popcll(mask & ((1ULL << s) - 1ULL));
    band_cell[b] = (long long)cell;
    atomicExch(&cellmap[cell], b);
  }
}

__device__ __forceinline__ void band_macro_cell(const double *__restrict__ f, long long N, long long n, int b,
                                                int cap, double hv0, double hv1, double hv2,
                                                double *__restrict__ band_u);
__device__ __forceinline__ void band_macro32_cell(const float *__restrict__ g, long long N, long long n, int b,
                                                  int cap, double hv0, double hv1, double hv2,
                                                  double *__restrict__ band_u);

// the band velocity of the stored state (FP64 field f, or FP32 deviations g)
struct BandMacroArgs {
  const double *f;
  const float *g;
  long long N;
  const long long *cell;
  int cap;
  double hv0, hv1, hv2;
  double *band_u;
};

// every entry: its band cell and a slot among the cell's entries (the slots
// are in arrival order; k_ibm_rank restores entry order).  A cell without a
// band number would be an inconsistent build: reported in ctl[kErr] (code 4),
// the entry skipped.  Thread b < nband also forms band cell b's velocity
// from the stored state (k_band_macro's value).
__global__ __launch_bounds__(256) void k_ibm_count(int E, const unsigned *__restrict__ keys,
                                                   const int *__restrict__ cellmap, int *__restrict__ bidx,
                                                   int *__restrict__ count, int *__restrict__ slot,
                                                   int *__restrict__ ctl, BandMacroArgs ma) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  if (e < ctl[kNew]) {
    if (ma.f) band_macro_cell(ma.f, ma.N, ma.cell[e], e, ma.cap, ma.hv0, ma.hv1, ma.hv2, ma.band_u);
    else band_macro32_cell(ma.g, ma.N, ma.cell[e], e, ma.cap, ma.hv0, ma.hv1, ma.hv2, ma.band_u);
  }
  const int b = cellmap_load(&cellmap[keys[e]]);
  if (b < 0 || b >= E) {
    atomicMin(&ctl[kErr], 4);
    bidx[e] = 0;
    slot[e] = 0;
    return;
  }
  bidx[e] = b;
  slot[e] = atomicAdd(&count[b], 1);
}

// entries into their cells' CSR ranges in slot order (band_ptr: inclusive
// scan of the counts, shifted by one)
__global__ __launch_bounds__(256) void k_ibm_fill(int E, const int *__restrict__ bidx,
                                                  const int *__restrict__ slot, const int *__restrict__ band_ptr,
                                                  int *__restrict__ tmp) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  tmp[band_ptr[bidx[e]] + slot[e]] = e;
}

// every entry at its rank in its cell's range (the number of the cell's
// entries below it, its range read sixteen entries at a time), so each cell
// lists its entries in ascending e, its weights (rx, ry, rz) beside it in
// went; the band size is published
__global__ __launch_bounds__(256) void k_ibm_rank(int E, int *__restrict__ ctl, const int *__restrict__ bidx,
                                                  const int *__restrict__ band_ptr, const int *__restrict__ tmp,
                                                  int *__restrict__ entry, const float *__restrict__ Ew,
                                                  float4 *__restrict__ went) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e == 0) ctl[kNband] = ctl[kNew];
  if (e >= E) return;
  // the entry's weights are read with its cell (they need nothing else)
  const int m = e >> 6, sidx = e & 63;
  const float *W = Ew + 12LL * m;
  const float wx = W[sidx >> 4], wy = W[4 + ((sidx >> 2) & 3)], wz = W[8 + (sidx & 3)];
  const int b = bidx[e];
  const int p0 = band_ptr[b], n = band_ptr[b + 1] - p0;
  int r = 0;
  for (int j0 = 0; j0 < n; j0 += 16) {
    int v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = j0 + j < n ? tmp[p0 + j0 + j] : 0x7fffffff;
#pragma unroll
    for (int j = 0; j < 16; ++j) r += v[j] < e ? 1 : 0;
  }
  entry[p0 + r] = e;
  // the entry's three stencil weights beside it, in the cells' order
  went[p0 + r] = make_float4(wx, wy, wz, 0.0f);
}

// band macros: den/u of every band cell from the stored state, same
// expressions as the collision kernel (FluidDomain.f90:1136-1139)
__device__ __forceinline__ void band_macro_cell(const double *__restrict__ f, long long N, long long n, int b,
                                                int cap, double hv0, double hv1, double hv2,
                                                double *__restrict__ band_u) {
  double g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) g[q] = f[q * N + n];
  double s, m1, m2, m3;
  moments(g, s, m1, m2, m3);
  band_u[b] = (m1 + hv0) / s;
  band_u[cap + b] = (m2 + hv1) / s;
  band_u[2LL * cap + b] = (m3 + hv2) / s;
}

__global__ __launch_bounds__(256) void k_band_macro(const double *__restrict__ f, long long N,
                                                    const long long *__restrict__ cell,
                                                    const int *__restrict__ ctl, int cap,
                                                    double hv0, double hv1, double hv2,
                                                    double *__restrict__ band_u, int only_on_fallback = 0) {
  if (only_on_fallback && ctl[kFallback] != 1) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ctl[kNband]) return;
  band_macro_cell(f, N, cell[b], b, cap, hv0, hv1, hv2, band_u);
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
  band_macro32_cell(g, N, cell[b], b, cap, hv0, hv1, hv2, band_u);
}

__device__ __forceinline__ void band_macro32_cell(const float *__restrict__ g, long long N, long long n, int b,
                                                  int cap, double hv0, double hv1, double hv2,
                                                  double *__restrict__ band_u) {
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

// every cell set to the same 19 populations (a uniform initial state)
struct Fill19 {
  double v[Q];
};
struct Fill19f {
  float v[Q];
};
__global__ __launch_bounds__(256) void k_fill19(double *__restrict__ f, long long N, Fill19 a) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q * N + n] = a.v[q];
}
__global__ __launch_bounds__(256) void k_fill19f(float *__restrict__ g, long long N, Fill19f a) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
#pragma unroll
  for (int q = 0; q < Q; ++q) g[q * N + n] = a.v[q];
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

// sum of tl[0..M) in a fixed order by one 256-thread workgroup: strided
// partial sums, a wave tree, then the four wave sums in order.  Each thread
// reads its terms sixteen at a time before adding them (in the same order),
// so the reads do not wait for one another; the first sixteen are read by
// the caller (mdf_tl_first), ahead of its other reads.
constexpr int kTlBatch = 16;
__device__ inline void mdf_tl_first(const double *tl, int M, double *v) {
  const int step = blockDim.x;
#pragma unroll
  for (int j = 0; j < kTlBatch; ++j) {
    const int m = (int)threadIdx.x + j * step;
    v[j] = m < M ? tl[m] : 0.0;
  }
}
__device__ inline double mdf_tl_sum(const double *tl, int M, const double *first) {
  constexpr int B = kTlBatch;
  double p = 0.0;
  const int step = blockDim.x;
  for (int base = threadIdx.x; base < M; base += B * step) {
    double v[B];
#pragma unroll
    for (int j = 0; j < B; ++j)
      v[j] = base == (int)threadIdx.x ? first[j] : (base + j * step < M ? tl[base + j * step] : 0.0);
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (base + j * step < M) p = p + v[j];
  }
  __shared__ double part[4];
  p = wave_sum(p);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = p;
  __syncthreads();
  double tol = 0.0;
  for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) tol = tol + part[k];
  return tol;
}

// stop test of the multi-direct-forcing loop (Solidbody.f90:895-906) from the
// summed |dV|_1 of the markers: dmax = tol / (sum M_b * Uref), one iteration
// counted; sets the done word when the loop condition fails.
// dstate[2] keeps the residual of loop iteration skip+1 (the first one after
// a learned guess standing for `skip` rounds) for the next call: a guessed
// call whose first residual exceeds twice that of the previous call restarts
// from the zero guess (k_ws_fallback, k_band_macro on fallback), which then
// runs the plain loop.  A negative dstate[2] (no previous call) always
// falls back.
// The words it reads, read up front (with the residual terms) so that the
// decision waits for no further reads.
struct MdfWords {
  int iters, warm, done0;
  double ref;
};
__device__ inline MdfWords mdf_words(const int *ctl, const double *dstate) {
  return MdfWords{ctl[kIters], ctl[kWarm], ctl[kDone0], dstate[2]};
}
__device__ inline void mdf_decide(int *ctl, double *dstate, double tol, double dsum, double Uref, int ntol,
                                  double dtol, int skip, const MdfWords &wd) {
  const int it = wd.iters + 1;
  ctl[kIters] = it;
  if (!isfinite(tol)) {
    atomicMin(&ctl[kErr], 3);  // "Nan found in PenaltyForce" (Solidbody.f90:1028-1031)
    ctl[kDone] = 1;
    return;
  }
  const double dm = tol / (dsum * Uref);
  dstate[0] = dm;
  if (it == skip + 1) {
    const double ref = wd.ref;
    dstate[2] = dm;
    if (wd.warm == 1 && !(ref >= 0.0 && dm <= 2.0 * ref)) {
      ctl[kWarm] = 0;
      ctl[kFallback] = 1;
      ctl[kIters] = 0;
      ctl[kDone] = wd.done0;
      dstate[0] = 1e10;
      return;
    }
  }
  if (!(it < ntol && dm > dtol)) ctl[kDone] = 1;
}

// One marker per wave: each lane forms one of the 64 interpolation terms
// ((u*rx)*ry)*rz of the marker's stencil (Solidbody.f90:1009-1015), lane 0
// adds them in the reference's x, y, z order, then the velocity defect,
// penalty force and force accumulation (PenaltyForce_, Solidbody.f90:1016-1026).
// The marker's |dV|_1 goes to tl[m] for the residual; ctl[kRan] = round
// records that this forcing round ran.
__global__ __launch_bounds__(256) void k_interp(int M, int round, const int *__restrict__ bidx,
                                                const float *__restrict__ Ew,
                                                const double *__restrict__ band_u, int cap,
                                                int *__restrict__ ctl,
                                                const double *__restrict__ Evel,
                                                const double *__restrict__ Ea, double invh3,
                                                double *__restrict__ Eforce,
                                                double *__restrict__ fT,
                                                double *__restrict__ tl, int snap = -1,
                                                double *__restrict__ Esnap = nullptr) {
  __shared__ double t[4][3][64];
  const int w = threadIdx.x >> 6, s = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + w;
  // the stencil's band cells and weights and the marker's own values are read
  // with the control word (the reads are the kernel's latency)
  int b = 0;
  float wx = 0.0f, wy = 0.0f, wz = 0.0f;
  if (m < M) {
    b = bidx[64LL * m + s];
    const float *W = Ew + 12LL * m;
    wx = W[s >> 4];
    wy = W[4 + ((s >> 2) & 3)];
    wz = W[8 + (s & 3)];
  }
  double ev1 = 0.0, ev2 = 0.0, ev3 = 0.0, am = 0.0, ef1 = 0.0, ef2 = 0.0, ef3 = 0.0;
  if (m < M && s == 0) {
    ev1 = Evel[3LL * m]; ev2 = Evel[3LL * m + 1]; ev3 = Evel[3LL * m + 2];
    am = Ea[m];
    ef1 = Eforce[3LL * m]; ef2 = Eforce[3LL * m + 1]; ef3 = Eforce[3LL * m + 2];
  }
  if (ctl[kDone]) return;  // uniform over the grid
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl[kRan] = round;
  if (m < M) {
    const double rx = (double)wx, ry = (double)wy, rz = (double)wz;
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
    const double d1 = ev1 - v1, d2 = ev2 - v2, d3 = ev3 - v3;
    const double F1 = d1 * am, F2 = d2 * am, F3 = d3 * am;
    tl[m] = (fabs(d1) + fabs(d2)) + fabs(d3);
    Eforce[3LL * m] = ef1 + F1;
    Eforce[3LL * m + 1] = ef2 + F2;
    Eforce[3LL * m + 2] = ef3 + F3;
    if (round == snap) {  // the force after round snap+1 (a training target of the warm start)
      Esnap[3LL * m] = ef1 + F1;
      Esnap[3LL * m + 1] = ef2 + F2;
      Esnap[3LL * m + 2] = ef3 + F3;
    }
    fT[3LL * m] = F1 * invh3;
    fT[3LL * m + 1] = F2 * invh3;
    fT[3LL * m + 2] = F3 * invh3;
  }
}

// The per-cell sums over a band cell's entries (the velocity correction and
// the spreading), by workgroups of 256 threads taking kCellsPerPass
// consecutive cells at a time: the cells' entries are one contiguous range of
// the CSR arrays, so each thread forms one entry's term from coalesced reads
// (entry, weights) and one gather (the marker's value) into LDS, and then one
// thread per cell adds its terms in entry order, as the sequential loop does.
// A pass whose range exceeds the LDS area (pass_cap <= kPassEntries entries)
// falls back to one thread per cell reading its entries itself.  SPREAD: term = -(((v h^-3 rx)
// ry) rz), g = sum + VF (out = band_F); else u -= ((v rx) ry) rz (out =
// band_u, read first).
constexpr int kCellsPerPass = 32;
constexpr int kPassEntries = 1024;  // 24 KiB of LDS: fits beside the bulk collide of a split step
template <bool SPREAD>
__device__ inline void band_cells_reduce(int blk, int nb, int nblk, int cap, const int *__restrict__ ptr,
                                         const int *__restrict__ entry, const float4 *__restrict__ went,
                                         const double *__restrict__ val, double invh3, double *__restrict__ out,
                                         double VF1, double VF2, double VF3, int pass_cap) {
  __shared__ int sptr[kCellsPerPass + 1];
  __shared__ double term[3][kPassEntries];
  const int t = threadIdx.x;
  for (int c0 = blk * kCellsPerPass; c0 < nb; c0 += nblk * kCellsPerPass) {
    const int nc = min(kCellsPerPass, nb - c0);
    double o1 = 0.0, o2 = 0.0, o3 = 0.0;
    if (t <= nc) sptr[t] = ptr[c0 + t];
    if (!SPREAD && t < nc) {
      o1 = out[c0 + t];
      o2 = out[cap + c0 + t];
      o3 = out[2LL * cap + c0 + t];
    }
    __syncthreads();
    const int e0 = sptr[0], ne = sptr[nc] - e0;
    const bool fits = ne <= pass_cap;
    if (fits) {
      for (int k = t; k < ne; k += blockDim.x) {
        const int m = entry[e0 + k] >> 6;
        const float4 w = went[e0 + k];
        const double rx = (double)w.x, ry = (double)w.y, rz = (double)w.z;
        double v1 = val[3LL * m], v2 = val[3LL * m + 1], v3 = val[3LL * m + 2];
        if (SPREAD) {
          v1 = v1 * invh3; v2 = v2 * invh3; v3 = v3 * invh3;
          term[0][k] = -(((v1 * rx) * ry) * rz);
          term[1][k] = -(((v2 * rx) * ry) * rz);
          term[2][k] = -(((v3 * rx) * ry) * rz);
        } else {
          term[0][k] = ((v1 * rx) * ry) * rz;
          term[1][k] = ((v2 * rx) * ry) * rz;
          term[2][k] = ((v3 * rx) * ry) * rz;
        }
      }
    }
    __syncthreads();
    if (t < nc) {
      const int k0 = sptr[t] - e0, k1 = sptr[t + 1] - e0;
      for (int k = k0; k < k1; ++k) {
        double x1, x2, x3;
        if (fits) {
          x1 = term[0][k]; x2 = term[1][k]; x3 = term[2][k];
        } else {
          const int m = entry[e0 + k] >> 6;
          const float4 w = went[e0 + k];
          const double rx = (double)w.x, ry = (double)w.y, rz = (double)w.z;
          double v1 = val[3LL * m], v2 = val[3LL * m + 1], v3 = val[3LL * m + 2];
          if (SPREAD) {
            v1 = v1 * invh3; v2 = v2 * invh3; v3 = v3 * invh3;
            x1 = -(((v1 * rx) * ry) * rz); x2 = -(((v2 * rx) * ry) * rz); x3 = -(((v3 * rx) * ry) * rz);
          } else {
            x1 = ((v1 * rx) * ry) * rz; x2 = ((v2 * rx) * ry) * rz; x3 = ((v3 * rx) * ry) * rz;
          }
        }
        if (SPREAD) {
          o1 = o1 + x1; o2 = o2 + x2; o3 = o3 + x3;
        } else {
          o1 = o1 - x1; o2 = o2 - x2; o3 = o3 - x3;
        }
      }
      if (SPREAD) {
        out[c0 + t] = o1 + VF1;
        out[cap + c0 + t] = o2 + VF2;
        out[2LL * cap + c0 + t] = o3 + VF3;
      } else {
        out[c0 + t] = o1;
        out[cap + c0 + t] = o2;
        out[2LL * cap + c0 + t] = o3;
      }
    }
    __syncthreads();  // (the next pass rewrites sptr and term)
  }
}

struct MdfTest {
  const double *tl;
  double *dstate;
  int M, ntol;
  double dsum, Uref, dtol;
  int skip;  // rounds a learned guess stands for (the fallback test is at iteration skip+1)
};

// velocity correction of every band cell: u -= fT*rx*ry*rz over its entries
// in (marker, stencil) order (Solidbody.f90:1034-1048), by the first
// cell_blocks workgroups (band_cells_reduce; the host sizes the grid to the
// last call's band).
// round >= 0: the correction of that forcing round (runs when its k_interp
// ran), and workgroup cell_blocks makes the round's stop test (the markers'
// |dV|_1 summed in a fixed order, then mdf_decide); round < 0 (the
// warm-start guess): runs unless the loop is done, no stop test.
__global__ __launch_bounds__(256) void k_correct(int *__restrict__ ctl, int round, MdfTest mt, int cap,
                                                 const int *__restrict__ ptr,
                                                 const int *__restrict__ entry,
                                                 const float4 *__restrict__ went,
                                                 const double *__restrict__ fT,
                                                 double *__restrict__ band_u, int cell_blocks, int pass_cap) {
  if ((int)blockIdx.x >= cell_blocks) {
    // the stop test: its reads issued together, then the sums
    double first[kTlBatch];
    mdf_tl_first(mt.tl, mt.M, first);
    const MdfWords wd = mdf_words(ctl, mt.dstate);
    if (round >= 0 ? ctl[kRan] != round : ctl[kDone] != 0) return;  // uniform over the grid
    if (round >= 0) {
      const double tol = mdf_tl_sum(mt.tl, mt.M, first);
      if (threadIdx.x == 0) mdf_decide(ctl, mt.dstate, tol, mt.dsum, mt.Uref, mt.ntol, mt.dtol, mt.skip, wd);
    }
    return;
  }
  if (round >= 0 ? ctl[kRan] != round : ctl[kDone] != 0) return;  // uniform over the grid
  band_cells_reduce<false>(blockIdx.x, ctl[kNband], cell_blocks, cap, ptr, entry, went, fT, 1.0, band_u, 0.0, 0.0,
                           0.0, pass_cap);
}

// Inclusive scan of the per-cell entry counts of the band's cells (ctl[kNew],
// final after k_ibm_claim) into the CSR pointers ptr[1..] (ptr[0] = 0), one
// 1024-thread workgroup: each thread reads a run of kScanRun counts into
// registers at once (the reads are the scan's latency), the run totals are
// scanned within each wave by lane shuffles and across the sixteen waves by
// one wave, then the runs are written with their offsets; bands above 16384
// cells take further passes.  Integer sums: the result does not depend on the
// order.  The workgroup also reduces the claim workgroups' parts of the
// band's bounding box (nbox of them) into ctl[kBoxLo..kBoxHi+2].
constexpr int kScanRun = 16;  // (a multiple of 4: the runs are read as int4)
__global__ __launch_bounds__(1024) void k_band_scan(int *__restrict__ ctl, const int *__restrict__ count,
                                                     int *__restrict__ ptr, const int *__restrict__ blkbox,
                                                     int nbox) {
  __shared__ int wtot[16];
  __shared__ int wbox[16][6];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  {
    int r[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, -1, -1, -1};
    for (int q = t; q < nbox; q += blockDim.x)
#pragma unroll
      for (int k = 0; k < 6; ++k) r[k] = k < 3 ? min(r[k], blkbox[6 * q + k]) : max(r[k], blkbox[6 * q + k]);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      int v = r[k];
      for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = k < 3 ? min(v, o) : max(v, o);
      }
      if (lane == 0) wbox[w][k] = v;
    }
  }
  // the cells claimed in this call (k_ibm_rank copies this count to kNband only later)
  const int n = ctl[kNew];
  int carry = 0;
  for (int base = 0; base < n; base += 1024 * kScanRun) {
    const int lo = base + t * kScanRun;
    int c[kScanRun];
    if (lo < n) {
      // the run in 16-byte reads (count holds kScanRun spare words past the band)
      const int4 *q = reinterpret_cast<const int4 *>(count + lo);
#pragma unroll
      for (int j = 0; j < kScanRun / 4; ++j) {
        const int4 v = q[j];
        c[4 * j] = v.x; c[4 * j + 1] = v.y; c[4 * j + 2] = v.z; c[4 * j + 3] = v.w;
      }
    }
#pragma unroll
    for (int j = 0; j < kScanRun; ++j)
      if (!(lo + j < n)) c[j] = 0;
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanRun; ++j) s += c[j];
    int v = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(v, off, 64);
      if (lane >= off) v += o;
    }
    __syncthreads();  // (the previous pass is done with wtot)
    if (lane == 63) wtot[w] = v;
    __syncthreads();
    if (w == 0) {
      int x = lane < 16 ? wtot[lane] : 0;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        const int o = __shfl_up(x, off, 64);
        if (lane >= off) x += o;
      }
      if (lane < 16) wtot[lane] = x;
    }
    __syncthreads();
    int run = carry + (v - s) + (w ? wtot[w - 1] : 0);
#pragma unroll
    for (int j = 0; j < kScanRun; ++j) {
      run += c[j];
      if (lo + j < n) ptr[lo + j + 1] = run;
    }
    carry += wtot[15];
  }
  __syncthreads();
  if (t < 6) {
    int r = wbox[0][t];
    for (int q = 1; q < 16; ++q) r = t < 3 ? min(r, wbox[q][t]) : max(r, wbox[q][t]);
    ctl[kBoxLo + t] = r;
  }
}

// The control words and the last residual of a coupling call, written into
// pinned host memory (read by the host after the stream's synchronisation).
struct IbmReport {
  int ctl[kCtlWords];
  double dmax;
};

// The end of a coupling call in one launch: the beam loads, the spreading and
// the report.
// Beam nodal loads lodFlow (Solidbody.f90:944-967): per degree of freedom, the
// sum over its markers in marker order of 0.5*F (translational) or
// 0.5*(r x F) (rotational) with r the arm from the element axis midpoint;
// dof_list entry: marker id, dof_kind[d] = component (0..2 force, 3..5
// moment).  The leading dof_blocks workgroups form them with one wave per
// degree of freedom: the lanes form the DOF's terms into LDS, then lane 0 adds
// them in list order.
// Eulerian force of every band cell (ResetVolumeForce_ FluidDomain.f90:
// 1195-1206, FluidVolumeForce_ Solidbody.f90:968-977, add_volume_force_
// FluidDomain.f90:1182-1193): (0 - sum F/h^3*rx*ry*rz) + VF, by the other
// band_blocks workgroups (band_cells_reduce, workgroup blk of them), the first of which copies the
// control words into the report (the last k_correct has made the stop test,
// so they are final).
constexpr int kLoadChunk = 256;

struct IbmFinishArgs {
  const int *ctl;
  const double *dstate;
  IbmReport *rep;
  int cap;
  const int *ptr, *entry;
  const float4 *went;
  const double *Eforce;
  double invh3, VF1, VF2, VF3;
  double *band_F;
  int ndof, dof_blocks, band_blocks, pass_cap;
  const int *dof_ptr, *dof_list, *dof_elem, *dof_kind, *marker_elem;
  const double *elem_xc, *Exyz;
  double *lod;
};

__device__ __forceinline__ double load_term(const IbmFinishArgs &a, int kind, int m, int el) {
  if (kind < 3) return 0.5 * a.Eforce[3LL * m + kind];
  const double r1 = a.Exyz[3LL * m] - a.elem_xc[3LL * el];
  const double r2 = a.Exyz[3LL * m + 1] - a.elem_xc[3LL * el + 1];
  const double r3 = a.Exyz[3LL * m + 2] - a.elem_xc[3LL * el + 2];
  const double F1 = a.Eforce[3LL * m], F2 = a.Eforce[3LL * m + 1], F3 = a.Eforce[3LL * m + 2];
  double val;
  if (kind == 3) val = r2 * F3 - r3 * F2;
  else if (kind == 4) val = r3 * F1 - r1 * F3;
  else val = r1 * F2 - r2 * F1;
  return 0.5 * val;
}

__global__ __launch_bounds__(256) void k_ibm_finish(IbmFinishArgs a) {
  if ((int)blockIdx.x < a.dof_blocks) {
    __shared__ double term[4][kLoadChunk];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int d = blockIdx.x * 4 + w;
    if (d >= a.ndof) return;  // whole waves; no workgroup barrier follows
    const int kind = a.dof_kind[d];
    const int j0 = a.dof_ptr[d], j1 = a.dof_ptr[d + 1];
    double acc = 0.0;
    for (int base = j0; base < j1; base += kLoadChunk) {
      const int n = min(kLoadChunk, j1 - base);
      for (int q = lane; q < n; q += 64) term[w][q] = load_term(a, kind, a.dof_list[base + q], a.dof_elem[base + q]);
      wave_sync();
      if (lane == 0)
        for (int q = 0; q < n; ++q) acc = acc + term[w][q];
      wave_sync();
    }
    if (lane == 0) a.lod[d] = acc;
    return;
  }
  const int blk = (int)blockIdx.x - a.dof_blocks;
  if (blk == 0 && threadIdx.x < kCtlWords) a.rep->ctl[threadIdx.x] = a.ctl[threadIdx.x];
  if (blk == 0 && threadIdx.x == 0) a.rep->dmax = a.dstate[0];
  band_cells_reduce<true>(blk, a.ctl[kNband], a.band_blocks, a.cap, a.ptr, a.entry, a.went, a.Eforce, a.invh3,
                          a.band_F, a.VF1, a.VF2, a.VF3, a.pass_cap);
}

// ---------------------------------------------------------------------------
// Learned warm start (no counterpart in the reference)
//
// The forcing loop is linear in the accumulated marker force: after j rounds
// from the zero guess, Eforce = E_j and the band velocity is
// u0 - S(E_j h^-3 / (2 rho)) with S the spreading of the stencil weights.  A
// guess of E_k applied as one correction of the uncorrected band velocity u0
// therefore puts the loop in the state of round k, and the loop goes on from
// iteration k + 1 to the same cap and tolerance.  The network predicts E_k
// as a correction to the marker's force of the previous call:
//   E_k ~ Eprev + Uref |Ea| net(X)
// and is trained on E_k recorded from plain calls.

// Features of every marker for the network, one wave per marker: the
// interpolated velocity of the uncorrected band (as k_interp forms it), the
// marker velocity, the marker's force of the previous call and its normal:
//   X = [U/Uref, V/Uref, Fprev/(Uref |Ea|), n]
__global__ __launch_bounds__(256) void k_ws_features(int M, const int *__restrict__ bidx,
                                                     const float *__restrict__ Ew,
                                                     const double *__restrict__ band_u, int cap,
                                                     const double *__restrict__ Evel,
                                                     const double *__restrict__ Ea,
                                                     const double *__restrict__ Eprev,
                                                     const double *__restrict__ Enrm, double Uref,
                                                     float *__restrict__ X) {
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
    for (int c = 0; c < 3; ++c) {
      x[c] = (float)(v[c] / Uref);
      x[3 + c] = (float)(Evel[3LL * m + c] / Uref);
      x[6 + c] = (float)(fs > 0.0 ? Eprev[3LL * m + c] / fs : 0.0);
      x[9 + c] = Enrm ? (float)Enrm[3LL * m + c] : 0.f;
    }
  }
}

// the network's guess of E_k as the starting force: Eforce = Eprev + Y Uref
// |Ea|, fT = Eforce * invh3 (applied to the band by k_correct as one
// correction of the uncorrected velocity)
__global__ __launch_bounds__(256) void k_ws_apply(int M, const float *__restrict__ Y, const double *__restrict__ Ea,
                                                  const double *__restrict__ Eprev, double Uref, double invh3,
                                                  double *__restrict__ Eforce, double *__restrict__ fT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * M) return;
  const double F0 = Eprev[i] + (double)Y[i] * (Uref * fabs(Ea[i / 3]));
  Eforce[i] = F0;
  fT[i] = F0 * invh3;
}

// restart from the zero guess when the stop test asked for it
__global__ __launch_bounds__(256) void k_ws_fallback(int M, int *__restrict__ ctl, double *__restrict__ Eforce) {
  if (ctl[kFallback] != 1) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * M) Eforce[i] = 0.0;
}
__global__ void k_ws_fallback_end(int *__restrict__ ctl) {
  if (threadIdx.x == 0 && ctl[kFallback] == 1) ctl[kFallback] = 2;
}

// training target of a plain call, (E_k - Eprev) / (Uref |Ea|) with E_k the
// force after `skip` rounds (the final force when the loop stopped earlier),
// then Eprev = the call's final force, the previous force of the next call
__global__ __launch_bounds__(256) void k_ws_targets(int M, const int *__restrict__ ctl, int skip,
                                                    const double *__restrict__ Eforce,
                                                    const double *__restrict__ Esnap,
                                                    const double *__restrict__ Ea, double Uref, int record,
                                                    float *__restrict__ Y, double *__restrict__ Eprev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * M) return;
  if (record) {
    const double fs = Uref * fabs(Ea[i / 3]);
    const double Ek = (skip > 0 && ctl[kIters] >= skip) ? Esnap[i] : Eforce[i];
    Y[i] = (float)(fs > 0.0 ? (Ek - Eprev[i]) / fs : 0.0);
  }
  Eprev[i] = Eforce[i];
}

}  // namespace fsil
