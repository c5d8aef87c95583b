// FP32-deviation storage form of the fused collide + push-stream kernel.
//
// Storage: g_q = f_q - w_q in FP32 (half the bytes of the FP64 form).  The
// deviations are O(1e-2) where f is O(1e-1), so FP32 keeps ~3 more significant
// bits of the physical signal than storing f itself would.  The density
// deviation and momentum are accumulated with compensated (Kahan) sums; the
// collision works on deviations:
//   feq_q - w_q = w_q * (drho + rho*(3 u.e + 4.5 (u.e)^2 - 1.5 u.u))
//   g'_q = g_q + Omega*((feq_q - w_q) - g_q) + (1 - Omega/2) * Flb_q
// which is the reference's SRT + Guo update (FluidDomain.f90:1217-1227)
// rearranged around w_q.  Cells of an immersed-boundary band take the
// coupling's corrected velocity and force (BAND), as in the FP64 form.
#pragma once
#include "fsil_common.h"
#include "fsil_lbm_kernels.h"

namespace fsil {

struct Collide32Args {
  const float *__restrict__ gin;
  float *__restrict__ gout;
  int xd, yd, zd;
  long long N;
  float Omega, c1, dt3;
  float hO, hO2, cT1, cT2;                  // TRT: 0.5 Omega, 0.5 Omega2, 0.5 - 0.25 Omega, 0.5 - 0.25 Omega2
  const float *__restrict__ Mc, *__restrict__ Mf;  // MRT: M_COLLID, M_FORCE rounded to FP32, column-major
  float hv[3];
  float VF[3];
  // immersed-boundary band overlay (corrected velocity and force, FP64)
  const int *__restrict__ ctl;        // band bounding box in ctl[0..5]
  const int *__restrict__ cellmap;    // cell -> band index, -1 outside
  const double *__restrict__ band_u;  // [3][cap]
  const double *__restrict__ band_F;  // [3][cap]
  int cap;
  int nt;    // nontemporal population loads and stores (selects the kernel instantiation)
  int zrow;  // workgroups are whole z-rows covering the block exactly: 1 shift_rows_z,
             // 2 shift_rows_z_waves (rows of whole waves)
  int region;   // CollideRegion
  int hole[4];  // kRegionBulk / kRegionHull: x0, x1, y0, y1
  // auxiliary planes of boundary codes 102 (den/u of the first interior
  // plane) and 203/204 (post-collision boundary plane as f = g + w), FP64
  int aux_face_mask102, aux_face_mask_hw;
  double *nb_den[6];
  double *nb_u[6];  // [3][plane]
  double *hw[6];    // [q][plane]
};

__device__ __forceinline__ void kahan_add(float &s, float &c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <int q>
__device__ __forceinline__ float dotu32(float u1, float u2, float u3) {
  return (float)ex(q) * u1 + (float)ey(q) * u2 + (float)ez(q) * u3;
}

// the nine opposite pairs (q, oppo(q)) by their lower member
constexpr int kPairQ[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};

// e_q . v for a D3Q19 direction, written out so that zero components cost nothing
template <typename T>
__device__ __forceinline__ T edot(int q, T v1, T v2, T v3) {
  T s = T(-0.0);  // -0 + x folds to x
  if (ex(q) > 0) s = s + v1;
  if (ex(q) < 0) s = s - v1;
  if (ey(q) > 0) s = s + v2;
  if (ey(q) < 0) s = s - v2;
  if (ez(q) > 0) s = s + v3;
  if (ez(q) < 0) s = s - v3;
  return s;
}

// Density deviation and momentum of the stored deviations, summed in FP64 (the
// FP32 values convert exactly; the pair sums g_q + g_qbar and differences
// g_q - g_qbar are formed first, so the momentum uses nine terms).
__device__ __forceinline__ void moments32(const float *g, double &dr, double &j1, double &j2, double &j3) {
  dr = (double)g[0];
  j1 = -0.0; j2 = -0.0; j3 = -0.0;
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int q = kPairQ[p];
    const double a = (double)g[q], b = (double)g[oppo(q)];
    dr = dr + (a + b);
    const double d = a - b;
    if (ex(q) > 0) j1 = j1 + d;
    if (ex(q) < 0) j1 = j1 - d;
    if (ey(q) > 0) j2 = j2 + d;
    if (ey(q) < 0) j2 = j2 - d;
    if (ez(q) > 0) j3 = j3 + d;
    if (ez(q) < 0) j3 = j3 - d;
  }
}

// keeps a uniform base address in scalar registers and out of the compiler's
// address arithmetic, so a load takes it with a 32-bit per-lane offset
// Collision of one cell in the deviation form, per opposite pair (q, qbar):
//   A = drho + rho*(4.5 (u.e_q)^2 - 1.5 u.u),  B = 3 rho (u.e_q)
//   g'_q = (1-Omega) g_q + Omega w (A + B),  g'_qbar = (1-Omega) g_qbar + Omega w (A - B)
// and, in forced cells, the Guo term c1*3dh*w*(e.F (1 + 3 u.e) - u.F) added
// (FluidDomain.f90:1217-1227 rearranged around w_q).
// TRT and MRT in the deviation form: with fe_q = (feq_q - w_q) - g_q and the
// Guo term fl_q = 3 dh w_q (e_q.F (1 + 3 u.e_q) - u.F), the reference's
// updates (FluidDomain.f90:1228-1238) read g'_q = g_q + (its f_q update);
// fe of an opposite pair is w (A +- B) - g with the pair's A and B of the SRT
// form.  MRT's matrices (rounded to FP32) are staged in LDS (ms, 2 x 361).
template <int MODEL, bool FORCE, bool BAND, bool NT>
__device__ __forceinline__ void collide_push32_cell(const Collide32Args &a, int x, int y, int z, float *zx,
                                                    const float *ms) {
  const unsigned n = ((unsigned)x * (unsigned)a.yd + (unsigned)y) * (unsigned)a.zd + (unsigned)z;
  const unsigned nb = n << 2;  // blocks of this form hold at most 2^30 cells
  const size_t plane = (size_t)a.N * sizeof(float);
  float g[Q];
  load_cell(g, a.gin, plane, nb, NT);
  double dr64, j1, j2, j3;
  moments32(g, dr64, j1, j2, j3);
  const float dr = (float)dr64;
  const float rho = (float)(1.0 + dr64);
  const float inv = __builtin_amdgcn_rcpf(rho);
  float u1 = ((float)j1 + a.hv[0]) * inv, u2 = ((float)j2 + a.hv[1]) * inv, u3 = ((float)j3 + a.hv[2]) * inv;
  bool forced = FORCE;
  float F1 = a.VF[0], F2 = a.VF[1], F3 = a.VF[2];
  if constexpr (BAND) {
    if (a.ctl[kBandErrWord] == kBandNoErr && x >= a.ctl[0] && x <= a.ctl[3] && y >= a.ctl[1] && y <= a.ctl[4] && z >= a.ctl[2] && z <= a.ctl[5]) {
      const int b = a.cellmap[n];
      if (b >= 0) {
        u1 = (float)a.band_u[b];
        u2 = (float)a.band_u[a.cap + b];
        u3 = (float)a.band_u[2 * a.cap + b];
        F1 = (float)a.band_F[b];
        F2 = (float)a.band_F[a.cap + b];
        F3 = (float)a.band_F[2 * a.cap + b];
        forced = true;
      }
    }
  }
  if (a.aux_face_mask102) {
    // the first interior plane of a code-102 face (as the FP64 form: density
    // of the stored state, velocity after the band overlay)
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      if (!((a.aux_face_mask102 >> face) & 1)) continue;
      const int axis = face >> 1, hi = face & 1;
      const int c = axis == 0 ? x : (axis == 1 ? y : z);
      const int nd = axis == 0 ? a.xd : (axis == 1 ? a.yd : a.zd);
      if (c == (hi ? nd - 2 : 1)) {
        const int p = face_plane_index(face, x, y, z, a.yd, a.zd);
        const int np = axis == 0 ? a.yd * a.zd : (axis == 1 ? a.xd * a.zd : a.xd * a.yd);
        a.nb_den[face][p] = 1.0 + dr64;
        a.nb_u[face][p] = (double)u1;
        a.nb_u[face][np + p] = (double)u2;
        a.nb_u[face][2 * np + p] = (double)u3;
      }
    }
  }
  const float omc = 1.0f - a.Omega;
  const float m15 = -1.5f * (u1 * u1 + u2 * u2 + u3 * u3);
  const float rho3 = 3.0f * rho;
  float post[Q];
  if constexpr (MODEL == 2 || MODEL == 3) {
    float fe[Q], fl[Q];
    fe[0] = (float)wt(0) * __builtin_fmaf(rho, m15, dr) - g[0];
    const float uF = u1 * F1 + u2 * F2 + u3 * F3;
    fl[0] = a.dt3 * (float)wt(0) * -uF;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
      const int q = kPairQ[p], o = oppo(q);
      const float w = (float)wt(q);
      const float ux = edot(q, u1, u2, u3);
      const float A = __builtin_fmaf(rho, __builtin_fmaf(4.5f * ux, ux, m15), dr);
      const float B = rho3 * ux;
      fe[q] = w * (A + B) - g[q];
      fe[o] = w * (A - B) - g[o];
      const float eF = edot(q, F1, F2, F3);
      const float c = __builtin_fmaf(3.0f * ux, eF, -uF);
      fl[q] = a.dt3 * w * (c + eF);
      fl[o] = a.dt3 * w * (c - eF);
    }
    if constexpr (MODEL == 2) {
      // TRT (FluidDomain.f90:1230-1235): rest direction as SRT, pairs split
      // into their symmetric (Omega) and antisymmetric (Omega2) parts
      post[0] = g[0] + a.Omega * fe[0] + (forced ? a.c1 * fl[0] : 0.0f);
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        const int q = kPairQ[p], o = oppo(q);
        float tp = a.hO * (fe[q] + fe[o]);
        float tm = a.hO2 * (fe[q] - fe[o]);
        if (forced) {
          tp = __builtin_fmaf(a.cT1, fl[q] + fl[o], tp);
          tm = __builtin_fmaf(a.cT2, fl[q] - fl[o], tm);
        }
        post[q] = g[q] + (tp + tm);
        post[o] = g[o] + (tp - tm);
      }
    } else {
      // MRT (FluidDomain.f90:1238): g + M_COLLID^T fe + M_FORCE^T fl
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < Q; ++j) s = __builtin_fmaf(ms[j * Q + i], fe[j], s);
        if (forced) {
#pragma unroll
          for (int j = 0; j < Q; ++j) s = __builtin_fmaf(ms[Q * Q + j * Q + i], fl[j], s);
        }
        post[i] = g[i] + s;
      }
    }
  } else {
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int q = kPairQ[p], o = oppo(q);
    const float Ow = a.Omega * (float)wt(q);
    const float ux = edot(q, u1, u2, u3);
    const float A = __builtin_fmaf(rho, __builtin_fmaf(4.5f * ux, ux, m15), dr);
    const float B = rho3 * ux;
    const float OwA = Ow * A, OwB = Ow * B;
    post[q] = __builtin_fmaf(omc, g[q], OwA + OwB);
    post[o] = __builtin_fmaf(omc, g[o], OwA - OwB);
  }
  post[0] = __builtin_fmaf(omc, g[0], (a.Omega * (float)wt(0)) * __builtin_fmaf(rho, m15, dr));
  if (FORCE || BAND) {
    if (forced) {
      const float cdt = a.c1 * a.dt3;
      const float uF = u1 * F1 + u2 * F2 + u3 * F3;
      post[0] = __builtin_fmaf(cdt * (float)wt(0), -uF, post[0]);
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        const int q = kPairQ[p], o = oppo(q);
        const float k = cdt * (float)wt(q);
        const float ux = edot(q, u1, u2, u3);
        const float eF = edot(q, F1, F2, F3);
        const float c = __builtin_fmaf(3.0f * ux, eF, -uF);
        post[q] = __builtin_fmaf(k, c + eF, post[q]);
        post[o] = __builtin_fmaf(k, c - eF, post[o]);
      }
    }
  }
  }
  const unsigned yz = (unsigned)a.yd * (unsigned)a.zd, zd = (unsigned)a.zd;
  const unsigned sxp = x + 1 == a.xd ? 0u - (unsigned)(a.xd - 1) * yz : yz;
  const unsigned sxm = x == 0 ? (unsigned)(a.xd - 1) * yz : 0u - yz;
  const unsigned syp = y + 1 == a.yd ? 0u - (unsigned)(a.yd - 1) * zd : zd;
  const unsigned sym = y == 0 ? (unsigned)(a.yd - 1) * zd : 0u - zd;
  const unsigned szp = z + 1 == a.zd ? 0u - (unsigned)(a.zd - 1) : 1u;
  const unsigned szm = z == 0 ? (unsigned)(a.zd - 1) : 0u - 1u;
  if (a.aux_face_mask_hw) {
    // the post-collision boundary plane of a code-203/204 face, before the
    // z-row exchange moves the z-going values between threads
#pragma unroll
    for (int face = 0; face < 6; ++face) {
      if (!((a.aux_face_mask_hw >> face) & 1)) continue;
      const int axis = face >> 1, hi = face & 1;
      const int c = axis == 0 ? x : (axis == 1 ? y : z);
      const int nd = axis == 0 ? a.xd : (axis == 1 ? a.yd : a.zd);
      if (c == (hi ? nd - 1 : 0)) {
        const int p = face_plane_index(face, x, y, z, a.yd, a.zd);
        const int np = axis == 0 ? a.yd * a.zd : (axis == 1 ? a.xd * a.zd : a.xd * a.yd);
        for (int q = 0; q < Q; ++q) a.hw[face][(long long)q * np + p] = (double)post[q] + wt(q);
      }
    }
  }
  if (a.zrow == 2) shift_rows_z_waves(post, zx);
  else if (a.zrow) shift_rows_z(post, zx);
  const unsigned st[6] = {sxp, sxm, syp, sym, a.zrow ? 0u : szp, a.zrow ? 0u : szm};
  store_cell(post, a.gout, plane, n, st, NT);
}

// The z-row exchange has workgroup barriers: a workgroup of whole z-rows is
// skipped or covered as a whole (the host aligns the hole's y-range to the
// workgroup's rows, and a hull tile is a workgroup's shape inside the block).
template <int MODEL, bool FORCE, bool BAND, bool NT>
__global__ __launch_bounds__(512) void k_collide_push32(Collide32Args a) {
  // LDS of the exchange, sized at launch (zrow32_lds_bytes)
  extern __shared__ float zx[];
  const float *ms = nullptr;
  if constexpr (MODEL == 3) {
    // the MRT matrices into LDS (every thread reads them in the same order: broadcasts)
    __shared__ float mlds[2 * Q * Q];
    const int nt = (int)(blockDim.x * blockDim.y);
    for (int i = (int)(threadIdx.y * blockDim.x + threadIdx.x); i < 2 * Q * Q; i += nt)
      mlds[i] = i < Q * Q ? a.Mc[i] : a.Mf[i - Q * Q];
    __syncthreads();
    ms = mlds;
  }
  if (a.region == kRegionHull) {
    for_hull_cells(a.ctl, a.hole, a.xd, a.yd, a.zd, [&](int x, int y, int z) {
      if (z < a.zd && y < a.yd) collide_push32_cell<MODEL, FORCE, BAND, NT>(a, x, y, z, zx, ms);
    });
    return;
  }
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int x = blockIdx.z;
  if (z >= a.zd || y >= a.yd) return;
  if (a.region == kRegionBulk && in_hole(a.hole, x, y)) return;
  collide_push32_cell<MODEL, FORCE, BAND, NT>(a, x, y, z, zx, ms);
}

// dynamic LDS of k_collide_push32 for its exchange mode and workgroup size
inline size_t zrow32_lds_bytes(int zrow, unsigned threads) {
  return zrow == 2 ? (threads / 64) * 2 * 10 * sizeof(float) : (zrow == 1 ? 10 * (size_t)threads * sizeof(float) : 0);
}

// den/u of the FP32-deviation state, in FP64 output arrays
__global__ __launch_bounds__(256) void k_macro32(const float *__restrict__ g, long long N, float hv0,
                                                 float hv1, float hv2, double *__restrict__ den,
                                                 double *__restrict__ u) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float dr = 0.f, cr = 0.f, j1 = 0.f, c1 = 0.f, j2 = 0.f, c2 = 0.f, j3 = 0.f, c3 = 0.f;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const float v = g[q * N + n];
    kahan_add(dr, cr, v);
    if (ex(q) != 0) kahan_add(j1, c1, ex(q) > 0 ? v : -v);
    if (ey(q) != 0) kahan_add(j2, c2, ey(q) > 0 ? v : -v);
    if (ez(q) != 0) kahan_add(j3, c3, ez(q) > 0 ? v : -v);
  }
  const double rho = 1.0 + (double)dr;
  den[n] = rho;
  u[n] = ((double)j1 + (double)hv0) / rho;
  u[N + n] = ((double)j2 + (double)hv1) / rho;
  u[2 * N + n] = ((double)j3 + (double)hv2) / rho;
}

// ---------------------------------------------------------------------------
// The instantaneous DatFlow fields of write_flow_ (FluidDomain.f90:1652-1664,
// outputtype 1): over the output window (offsetOutput cells left out on every
// face), real(4) p = Cs2 (den - denIn), u, v, w = uuu / Uref, with den/uuu as
// calculate_macro_quantities_ leaves them (FP64 form: k_macro's arithmetic;
// FP32-deviation form: k_macro32's).  out: [4][nx*ny*nz], z fastest.
template <bool F32>
__global__ __launch_bounds__(256) void k_flow_pack(const double *__restrict__ f, const float *__restrict__ g,
                                                   long long N, int xd, int yd, int zd, int off, double hv0,
                                                   double hv1, double hv2, double denIn, double invUref,
                                                   float *__restrict__ out) {
  const int nx = xd - 2 * off, ny = yd - 2 * off, nz = zd - 2 * off;
  const long long nw = (long long)nx * ny * nz;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const int z = (int)(i % nz) + off, y = (int)((i / nz) % ny) + off, x = (int)(i / ((long long)nz * ny)) + off;
  const long long n = ((long long)x * yd + y) * zd + z;
  double rho, u1, u2, u3;
  if constexpr (F32) {
    float dr = 0.f, cr = 0.f, j1 = 0.f, c1 = 0.f, j2 = 0.f, c2 = 0.f, j3 = 0.f, c3 = 0.f;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const float v = g[q * N + n];
      kahan_add(dr, cr, v);
      if (ex(q) != 0) kahan_add(j1, c1, ex(q) > 0 ? v : -v);
      if (ey(q) != 0) kahan_add(j2, c2, ey(q) > 0 ? v : -v);
      if (ez(q) != 0) kahan_add(j3, c3, ez(q) > 0 ? v : -v);
    }
    rho = 1.0 + (double)dr;
    u1 = ((double)j1 + (double)(float)hv0) / rho;
    u2 = ((double)j2 + (double)(float)hv1) / rho;
    u3 = ((double)j3 + (double)(float)hv2) / rho;
  } else {
    double h[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) h[q] = f[q * N + n];
    double m1, m2, m3;
    moments(h, rho, m1, m2, m3);
    u1 = (m1 + hv0) / rho;
    u2 = (m2 + hv1) / rho;
    u3 = (m3 + hv2) / rho;
  }
  out[i] = (float)((1.0 / 3.0) * (rho - denIn));
  out[nw + i] = (float)(u1 * invUref);
  out[2 * nw + i] = (float)(u2 * invUref);
  out[3 * nw + i] = (float)(u3 * invUref);
}

}  // namespace fsil
