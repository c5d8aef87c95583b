// Immersed-boundary multi-direct-forcing kernels for gfx950 (FP64).
//
// Data model.  Every marker m owns 64 stencil entries e = 64*m + s with
// s = 16*ix + 4*iy + iz in the reference's x-outer, y, z-inner loop order
// (Solidbody.f90:1009-1015).  The cells those entries touch form the "band".
// Entries are sorted by (cell, e), which is the order in which the reference's
// serial loops over markers (Solidbody.f90:1034-1048, 968-977) update a cell;
// a per-cell sequential reduction over its sorted entries therefore rounds
// exactly as the reference does, with no atomics.
#pragma once
#include "fsil_common.h"
#include "fsil_lbm_kernels.h"

namespace fsil {

// band macros: den/u of every band cell from the stored state, same
// expressions as the collision kernel (FluidDomain.f90:1136-1139)
__global__ __launch_bounds__(256) void k_band_macro(const double *__restrict__ f, long long N,
                                                    const long long *__restrict__ cell, int nband,
                                                    double hv0, double hv1, double hv2,
                                                    double *__restrict__ band_u) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nband) return;
  const long long n = cell[b];
  double g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) g[q] = f[q * N + n];
  double s, m1, m2, m3;
  moments(g, s, m1, m2, m3);
  band_u[b] = (m1 + hv0) / s;
  band_u[nband + b] = (m2 + hv1) / s;
  band_u[2 * nband + b] = (m3 + hv2) / s;
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

// One marker per lane: 64-point interpolation of the band velocity, velocity
// defect, penalty force, force accumulation (PenaltyForce_, Solidbody.f90:1000-1026).
// Per-block partial sums of |dV|_1 go to tolpart[blockIdx.x].
__global__ __launch_bounds__(256) void k_interp(int M, const int *__restrict__ bidx,
                                                const float *__restrict__ Ew,
                                                const double *__restrict__ band_u, int nband,
                                                const double *__restrict__ Evel,
                                                const double *__restrict__ Ea, double invh3,
                                                double *__restrict__ Eforce,
                                                double *__restrict__ fT,
                                                double *__restrict__ tolpart) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  double tl = 0.0;
  if (m < M) {
    double v1 = 0.0, v2 = 0.0, v3 = 0.0;
    const int *I = bidx + 64LL * m;
    const float *W = Ew + 12LL * m;
    for (int x = 0; x < 4; ++x) {
      const double rx = (double)W[x];
      for (int y = 0; y < 4; ++y) {
        const double ry = (double)W[4 + y];
#pragma unroll
        for (int z = 0; z < 4; ++z) {
          const double rz = (double)W[8 + z];
          const int b = I[16 * x + 4 * y + z];
          v1 = v1 + ((band_u[b] * rx) * ry) * rz;
          v2 = v2 + ((band_u[nband + b] * rx) * ry) * rz;
          v3 = v3 + ((band_u[2 * nband + b] * rx) * ry) * rz;
        }
      }
    }
    const double d1 = Evel[3LL * m] - v1, d2 = Evel[3LL * m + 1] - v2, d3 = Evel[3LL * m + 2] - v3;
    const double a = Ea[m];
    const double F1 = d1 * a, F2 = d2 * a, F3 = d3 * a;
    tl = (fabs(d1) + fabs(d2)) + fabs(d3);
    Eforce[3LL * m] = Eforce[3LL * m] + F1;
    Eforce[3LL * m + 1] = Eforce[3LL * m + 1] + F2;
    Eforce[3LL * m + 2] = Eforce[3LL * m + 2] + F3;
    fT[3LL * m] = F1 * invh3;
    fT[3LL * m + 1] = F2 * invh3;
    fT[3LL * m + 2] = F3 * invh3;
  }
  __shared__ double part[4];
  tl = wave_sum(tl);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = tl;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) s = s + part[k];
    tolpart[blockIdx.x] = s;
  }
}

// velocity correction of every band cell: u -= fT*rx*ry*rz over its entries
// in (marker, stencil) order (Solidbody.f90:1034-1048)
__global__ __launch_bounds__(256) void k_correct(int nband, const int *__restrict__ ptr,
                                                 const int *__restrict__ entry,
                                                 const float *__restrict__ Ew,
                                                 const double *__restrict__ fT,
                                                 double *__restrict__ band_u) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nband) return;
  double u1 = band_u[b], u2 = band_u[nband + b], u3 = band_u[2 * nband + b];
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
  band_u[nband + b] = u2;
  band_u[2 * nband + b] = u3;
}

// Eulerian force of every band cell: (0 - sum F/h^3*rx*ry*rz) + VF
// (ResetVolumeForce_ FluidDomain.f90:1195-1206, FluidVolumeForce_
// Solidbody.f90:968-977, add_volume_force_ FluidDomain.f90:1182-1193)
__global__ __launch_bounds__(256) void k_spread(int nband, const int *__restrict__ ptr,
                                                const int *__restrict__ entry,
                                                const float *__restrict__ Ew,
                                                const double *__restrict__ Eforce, double invh3,
                                                double VF1, double VF2, double VF3,
                                                double *__restrict__ band_F) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nband) return;
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
  band_F[nband + b] = g2 + VF2;
  band_F[2 * nband + b] = g3 + VF3;
}

// Beam nodal loads lodFlow: per degree of freedom, the sum over its markers in
// marker order of 0.5*F (translational) or 0.5*(r x F) (rotational) with r the
// arm from the element axis midpoint (Solidbody.f90:944-967).
// dof_list entry: marker id; dof_kind[d] = component (0..2 force, 3..5 moment).
__global__ __launch_bounds__(256) void k_loads(int ndof, const int *__restrict__ dof_ptr,
                                               const int *__restrict__ dof_list,
                                               const int *__restrict__ dof_kind,
                                               const int *__restrict__ marker_elem,
                                               const double *__restrict__ elem_xc,
                                               const double *__restrict__ Exyz,
                                               const double *__restrict__ Eforce,
                                               double *__restrict__ lod) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndof) return;
  const int k = dof_kind[d];
  double acc = 0.0;
  for (int j = dof_ptr[d]; j < dof_ptr[d + 1]; ++j) {
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
    acc = acc + 0.5 * val;
  }
  lod[d] = acc;
}

}  // namespace fsil
