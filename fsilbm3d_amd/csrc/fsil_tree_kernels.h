// Father/son (refinement ratio 2) halo exchange of the block tree, FP64
// plain form (gaoak/FSILBM3D src/LBMBlockComm.f90).
//
// A pair keeps, for every son face whose boundary code is BCfluid (0), two
// copies of the father plane under it: level 1 (taken before the father's
// collision) and level 2 (after its boundary faces).  Planes are stored
// [q][a][b] with b the fast in-plane index (z for x and y faces, y for z
// faces), as the reference's fIn_F*t* arrays (LBMBlockComm.f90:201-245).
//
// The reference builds each fine plane in passes over a scratch array
// (interpolate_fIn, LBMBlockComm.f90:808-914); here each son node evaluates
// the same pass formulas on demand from the coarse plane, in the same
// association order, so every value is the one the passes produce.
#pragma once
#include "fsil_lbm_kernels.h"

namespace fsil {

struct TreeFace {
  int active;       // sds != 0
  int fixedF;       // father coordinate (0-based) of the plane along the face axis
  int fixedS;       // son coordinate (0-based) of the son boundary plane
  int fixedFi;      // father coordinate of the first inner father plane (son-to-father)
  int fixedSi;      // son coordinate of the son layer delivered to the father
  int aF, bF;       // coarse plane extent
  int aS, bS;       // son plane extent
  long long off;    // offset of the level-1 plane in the pair buffer; level 2 follows
};

struct TreeArgs {
  double *__restrict__ ff;  // father state (current buffer)
  double *__restrict__ fs;  // son state (current buffer)
  double *__restrict__ planes;
  int fxd, fyd, fzd;
  long long fN;
  int sxd, syd, szd;
  long long sN;
  int f0[3];        // father start (0-based) of the plane window, per axis
  int si0[3];       // son inner-layer start (0-based), per axis (son-to-father)
  int sicnt[3];     // number of delivered son nodes per axis (step 2)
  int fi0[3];       // father inner start (0-based), per axis
  TreeFace face[6];
  double coeff;     // grid-transform coefficient of this launch
  double hv[3];     // 0.5*VF*dh of the block whose VF the transform uses
};

// fIn_GridTransform (LBMBlockComm.f90:958-979): Dupuis-Chopard rescaling of
// the non-equilibrium part by coeff, equilibrium from the node's own moments.
__device__ __forceinline__ void grid_transform(double (&f)[Q], double coeff, const double (&hv)[3]) {
  double rho, m1, m2, m3;
  moments(f, rho, m1, m2, m3);
  const double u1 = (m1 + hv[0]) / rho;
  const double u2 = (m2 + hv[1]) / rho;
  const double u3 = (m3 + hv[2]) / rho;
  const double uSqr = (u1 * u1 + u2 * u2) + u3 * u3;
  const double base = 1.0 - 1.5 * uSqr;
#define FSIL_GT(q)                                                          \
  {                                                                         \
    const double ux = dotu<q>(u1, u2, u3);                                  \
    const double feq = (wt(q) * rho) * (base + ux * (3.0 + 4.5 * ux));    \
    f[q] = feq + coeff * (f[q] - feq);                                      \
  }
  FSIL_GT(0) FSIL_GT(1) FSIL_GT(2) FSIL_GT(3) FSIL_GT(4) FSIL_GT(5) FSIL_GT(6)
  FSIL_GT(7) FSIL_GT(8) FSIL_GT(9) FSIL_GT(10) FSIL_GT(11) FSIL_GT(12) FSIL_GT(13)
  FSIL_GT(14) FSIL_GT(15) FSIL_GT(16) FSIL_GT(17) FSIL_GT(18)
#undef FSIL_GT
}

// one coarse plane of one direction q; (b, a) are 1-based as in the reference
struct PlaneQ {
  const double *p;
  int bF;
  __device__ __forceinline__ double F(int b, int a) const { return p[(a - 1) * bF + (b - 1)]; }
};

// first pass (odd a, b <= bT): copies and b-interpolation (LBMBlockComm.f90:831-846, 885-893)
template <int SCHEME>
__device__ __forceinline__ double interp_row(const PlaneQ &P, int b, int a, int bT) {
  const int a1 = a / 2 + 1;
  if (b & 1) return P.F(b / 2 + 1, a1);
  const int b0 = b - 1, b1 = b0 / 2 + 1;
  if constexpr (SCHEME == 2) {
    if (b0 == 1) return (0.375 * P.F(b1, a1) + 0.75 * P.F(b1 + 1, a1)) - 0.125 * P.F(b1 + 2, a1);
    if (b0 == bT - 2) return (0.375 * P.F(b1 + 1, a1) + 0.75 * P.F(b1, a1)) - 0.125 * P.F(b1 - 1, a1);
    return ((-(0.0625 * P.F(b1 - 1, a1)) + 0.5625 * P.F(b1, a1)) + 0.5625 * P.F(b1 + 1, a1)) -
           0.0625 * P.F(b1 + 2, a1);
  } else {
    return (P.F(b1, a1) + P.F(b1 + 1, a1)) * 0.5;
  }
}

// first and second pass (a <= aT, b <= bT) (LBMBlockComm.f90:848-861, 895-901)
template <int SCHEME>
__device__ __forceinline__ double interp_inner(const PlaneQ &P, int b, int a, int aT, int bT) {
  if (a & 1) return interp_row<SCHEME>(P, b, a, bT);
  if constexpr (SCHEME == 2) {
    if (a == 2)
      return (0.375 * interp_row<2>(P, b, 1, bT) + 0.75 * interp_row<2>(P, b, 3, bT)) -
             0.125 * interp_row<2>(P, b, 5, bT);
    if (a == aT - 1)
      return (0.375 * interp_row<2>(P, b, a + 1, bT) + 0.75 * interp_row<2>(P, b, a - 1, bT)) -
             0.125 * interp_row<2>(P, b, a - 3, bT);
    return ((-(0.0625 * interp_row<2>(P, b, a - 3, bT)) + 0.5625 * interp_row<2>(P, b, a - 1, bT)) +
            0.5625 * interp_row<2>(P, b, a + 1, bT)) -
           0.0625 * interp_row<2>(P, b, a + 3, bT);
  } else {
    return (interp_row<1>(P, b, a - 1, bT) + interp_row<1>(P, b, a + 1, bT)) * 0.5;
  }
}

// last column b = bT+1 of an even-b plane (wraps to b = 1, 3) (LBMBlockComm.f90:862-868, 902-908)
template <int SCHEME>
__device__ __forceinline__ double interp_colR(const PlaneQ &P, int a, int aT, int bT) {
  if constexpr (SCHEME == 2)
    return ((-(0.0625 * interp_inner<2>(P, bT - 2, a, aT, bT)) + 0.5625 * interp_inner<2>(P, bT, a, aT, bT)) +
            0.5625 * interp_inner<2>(P, 1, a, aT, bT)) -
           0.0625 * interp_inner<2>(P, 3, a, aT, bT);
  else
    return (interp_inner<1>(P, bT, a, aT, bT) + interp_inner<1>(P, 1, a, aT, bT)) * 0.5;
}

// any son plane node (b <= bS, a <= aS)
template <int SCHEME>
__device__ double interp_node(const PlaneQ &P, int b, int a, int aT, int bT) {
  if (a <= aT) {
    if (b <= bT) return interp_inner<SCHEME>(P, b, a, aT, bT);
    return interp_colR<SCHEME>(P, a, aT, bT);
  }
  // last row a = aT+1 of an even-a plane (LBMBlockComm.f90:869-877, 909-913)
  if (b <= bT) {
    if constexpr (SCHEME == 2)
      return ((-(0.0625 * interp_inner<2>(P, b, aT - 2, aT, bT)) + 0.5625 * interp_inner<2>(P, b, aT, aT, bT)) +
              0.5625 * interp_inner<2>(P, b, 1, aT, bT)) -
             0.0625 * interp_inner<2>(P, b, 3, aT, bT);
    else
      return (interp_inner<1>(P, b, aT, aT, bT) + interp_inner<1>(P, b, 1, aT, bT)) * 0.5;
  }
  if constexpr (SCHEME == 2)
    return ((-(0.0625 * interp_colR<2>(P, aT - 2, aT, bT)) + 0.5625 * interp_colR<2>(P, aT, aT, bT)) +
            0.5625 * interp_colR<2>(P, 1, aT, bT)) -
           0.0625 * interp_colR<2>(P, 3, aT, bT);
  else
    return (interp_colR<1>(P, aT, aT, bT) + interp_colR<1>(P, 1, aT, bT)) * 0.5;
}

// cell coordinates of in-plane (a, b) on a face of `axis` at `fixed`
__device__ __forceinline__ void plane_cell(int axis, int fixed, int a, int b, int &x, int &y, int &z) {
  if (axis == 0) { x = fixed; y = a; z = b; }
  else if (axis == 1) { x = a; y = fixed; z = b; }
  else { x = a; y = b; z = fixed; }
}

// extract_interpolate_layer (LBMBlockComm.f90:340-505): copy the father planes
// under the son's fluid faces; at level 2 also level1 = 0.5*(level1 + level2).
// grid: x over q*aF*bF of the largest face, y = face.
__global__ void k_tree_extract(TreeArgs t, int level) {
  const int face = blockIdx.y;
  const TreeFace &F = t.face[face];
  if (!F.active) return;
  const int np = F.aF * F.bF;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Q * np) return;
  const int q = i / np, r = i - q * np;
  const int a = r / F.bF, b = r - a * F.bF;
  const int axis = face >> 1;
  const int oa = axis == 0 ? t.f0[1] : t.f0[0];
  const int ob = axis == 2 ? t.f0[1] : t.f0[2];
  int x, y, z;
  plane_cell(axis, F.fixedF, oa + a, ob + b, x, y, z);
  const double v = t.ff[(size_t)q * t.fN + ((size_t)x * t.fyd + y) * t.fzd + z];
  double *l1 = t.planes + F.off + i;
  if (level == 1) {
    *l1 = v;
  } else {
    l1[(size_t)Q * np] = v;
    *l1 = 0.5 * (*l1 + v);
  }
}

// is the son node (x, y, z) on an active son face after `face` in the
// reference's face order?  Later faces overwrite earlier ones there.
__device__ __forceinline__ bool later_face_owns(const TreeArgs &t, int face, int x, int y, int z, bool son) {
  for (int g = face + 1; g < 6; ++g) {
    if (!t.face[g].active) continue;
    const int c = (g >> 1) == 0 ? x : ((g >> 1) == 1 ? y : z);
    if (c == (son ? t.face[g].fixedS : t.face[g].fixedFi)) return true;
  }
  return false;
}

// interpolation_father_to_son (LBMBlockComm.f90:655-806): interpolate the
// coarse plane (level 1 at n_timeStep 0, level 2 at 1) onto the son boundary
// plane and rescale it with the father's volume force.
template <int SCHEME>
__global__ void k_tree_father_to_son(TreeArgs t, int level) {
  const int face = blockIdx.y;
  const TreeFace &F = t.face[face];
  if (!F.active) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F.aS * F.bS) return;
  const int a0 = i / F.bS, b0 = i - a0 * F.bS;
  const int axis = face >> 1;
  int x, y, z;
  plane_cell(axis, F.fixedS, a0, b0, x, y, z);
  if (later_face_owns(t, face, x, y, z, true)) return;
  const int bT = (F.bS & 1) ? F.bS : F.bS - 1;
  const int aT = (F.aS & 1) ? F.aS : F.aS - 1;
  const size_t np = (size_t)F.aF * F.bF;
  const double *pl = t.planes + F.off + (level == 2 ? (size_t)Q * np : 0);
  double f[Q];
  for (int q = 0; q < Q; ++q) {
    PlaneQ P{pl + q * np, F.bF};
    f[q] = interp_node<SCHEME>(P, b0 + 1, a0 + 1, aT, bT);
  }
  grid_transform(f, t.coeff, t.hv);
  const size_t n = ((size_t)x * t.syd + y) * t.szd + z;
#pragma unroll
  for (int q = 0; q < Q; ++q) t.fs[(size_t)q * t.sN + n] = f[q];
}

// deliver_son_to_father (LBMBlockComm.f90:546-653): every second son node of
// the son's first inner layer, rescaled with the son's volume force, replaces
// the father node it coincides with.
__global__ void k_tree_son_to_father(TreeArgs t) {
  const int face = blockIdx.y;
  const TreeFace &F = t.face[face];
  if (!F.active) return;
  const int axis = face >> 1;
  const int ea = axis == 0 ? 1 : 0, eb = axis == 2 ? 1 : 2;  // in-plane axes
  const int na = t.sicnt[ea], nb = t.sicnt[eb];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na * nb) return;
  const int ia = i / nb, ib = i - ia * nb;
  int s[3], fc[3];
  s[axis] = F.fixedSi;
  fc[axis] = F.fixedFi;
  s[ea] = t.si0[ea] + 2 * ia;
  s[eb] = t.si0[eb] + 2 * ib;
  fc[ea] = t.fi0[ea] + ia;
  fc[eb] = t.fi0[eb] + ib;
  if (later_face_owns(t, face, fc[0], fc[1], fc[2], false)) return;
  const size_t ns = ((size_t)s[0] * t.syd + s[1]) * t.szd + s[2];
  double f[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = t.fs[(size_t)q * t.sN + ns];
  grid_transform(f, t.coeff, t.hv);
  const size_t nf = ((size_t)fc[0] * t.fyd + fc[1]) * t.fzd + fc[2];
#pragma unroll
  for (int q = 0; q < Q; ++q) t.ff[(size_t)q * t.fN + nf] = f[q];
}

}  // namespace fsil
