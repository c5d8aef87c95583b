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

namespace fsil {

struct Collide32Args {
  const float *__restrict__ gin;
  float *__restrict__ gout;
  int xd, yd, zd;
  long long N;
  float Omega, c1, dt3;
  float hv[3];
  float VF[3];
  // immersed-boundary band overlay (corrected velocity and force, FP64)
  const int *__restrict__ ctl;        // band bounding box in ctl[0..5]
  const int *__restrict__ cellmap;    // cell -> band index, -1 outside
  const double *__restrict__ band_u;  // [3][cap]
  const double *__restrict__ band_F;  // [3][cap]
  int cap;
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

template <bool FORCE, bool BAND>
__global__ __launch_bounds__(256) void k_collide_push32(Collide32Args a) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int x = blockIdx.z;
  if (z >= a.zd || y >= a.yd) return;
  const unsigned n = ((unsigned)x * (unsigned)a.yd + (unsigned)y) * (unsigned)a.zd + (unsigned)z;
  const size_t plane = (size_t)a.N * sizeof(float);
  const char *gin = reinterpret_cast<const char *>(a.gin);
  const unsigned nb = n << 2;
  float g[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q)
    g[q] = __builtin_nontemporal_load(reinterpret_cast<const float *>(gin + q * plane + nb));
  float dr = 0.f, cr = 0.f, j1 = 0.f, c1 = 0.f, j2 = 0.f, c2 = 0.f, j3 = 0.f, c3 = 0.f;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    kahan_add(dr, cr, g[q]);
    if (ex(q) != 0) kahan_add(j1, c1, ex(q) > 0 ? g[q] : -g[q]);
    if (ey(q) != 0) kahan_add(j2, c2, ey(q) > 0 ? g[q] : -g[q]);
    if (ez(q) != 0) kahan_add(j3, c3, ez(q) > 0 ? g[q] : -g[q]);
  }
  const float rho = 1.0f + dr;
  const float inv = 1.0f / rho;
  float u1 = (j1 + a.hv[0]) * inv, u2 = (j2 + a.hv[1]) * inv, u3 = (j3 + a.hv[2]) * inv;
  bool forced = FORCE;
  float F1 = a.VF[0], F2 = a.VF[1], F3 = a.VF[2];
  if constexpr (BAND) {
    if (x >= a.ctl[0] && x <= a.ctl[3] && y >= a.ctl[1] && y <= a.ctl[4] && z >= a.ctl[2] && z <= a.ctl[5]) {
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
  const float uSqr = u1 * u1 + u2 * u2 + u3 * u3;
  float post[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    float ux;
    switch (q) {
#define C(k) case k: ux = dotu32<k>(u1, u2, u3); break;
      C(0) C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16) C(17) C(18)
#undef C
    }
    const float w = (float)wt(q);
    const float feqd = w * (dr + rho * (ux * (3.0f + 4.5f * ux) - 1.5f * uSqr));
    float t = g[q] + a.Omega * (feqd - g[q]);
    if (FORCE || BAND) {
      if (forced) {
        const float e1 = (float)ex(q), e2 = (float)ey(q), e3 = (float)ez(q);
        const float fl = a.dt3 * w *
                         ((e1 - u1 + 3.0f * ux * e1) * F1 + (e2 - u2 + 3.0f * ux * e2) * F2 +
                          (e3 - u3 + 3.0f * ux * e3) * F3);
        t = t + a.c1 * fl;
      }
    }
    post[q] = t;
  }
  const unsigned yz = (unsigned)a.yd * (unsigned)a.zd, zd = (unsigned)a.zd;
  const unsigned sxp = x + 1 == a.xd ? 0u - (unsigned)(a.xd - 1) * yz : yz;
  const unsigned sxm = x == 0 ? (unsigned)(a.xd - 1) * yz : 0u - yz;
  const unsigned syp = y + 1 == a.yd ? 0u - (unsigned)(a.yd - 1) * zd : zd;
  const unsigned sym = y == 0 ? (unsigned)(a.yd - 1) * zd : 0u - zd;
  const unsigned szp = z + 1 == a.zd ? 0u - (unsigned)(a.zd - 1) : 1u;
  const unsigned szm = z == 0 ? (unsigned)(a.zd - 1) : 0u - 1u;
  char *gout = reinterpret_cast<char *>(a.gout);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const unsigned sx = ex(q) > 0 ? sxp : (ex(q) < 0 ? sxm : 0u);
    const unsigned sy = ey(q) > 0 ? syp : (ey(q) < 0 ? sym : 0u);
    const unsigned sz = ez(q) > 0 ? szp : (ez(q) < 0 ? szm : 0u);
    const unsigned d = n + sx + sy + sz;
    __builtin_nontemporal_store(post[q], reinterpret_cast<float *>(gout + q * plane + (d << 2)));
  }
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

}  // namespace fsil
