// Shared definitions of the FSILBM3D MI355X hot-path library (device + host).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace fsil {

constexpr int Q = 19;

// barrier of the lanes of one wave (they run in lockstep; the fences keep the
// compiler from moving LDS accesses across it)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// D3Q19 lattice, in the reference's direction order (ConstParams.f90:11-25).
__host__ __device__ constexpr int ex(int q) {
  constexpr int e[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return e[q];
}
__host__ __device__ constexpr int ey(int q) {
  constexpr int e[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  return e[q];
}
__host__ __device__ constexpr int ez(int q) {
  constexpr int e[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  return e[q];
}
__host__ __device__ constexpr int oppo(int q) {
  constexpr int o[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  return o[q];
}
// weights 1/3, 1/18 (axis), 1/36 (diagonal), each the correctly rounded double
__host__ __device__ constexpr double wt(int q) {
  return q == 0 ? 1.0 / 3.0 : (q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
}

// Population planes in global memory.  A uniform plane base is kept in scalar
// registers (the empty asm stops the compiler from folding it back into
// per-lane 64-bit address arithmetic), so every access is one scalar base plus
// a 32-bit per-lane byte offset, and the base advances by one plane per
// direction with two scalar adds.
template <typename T> using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> sbase(gptr<T> g) {
  asm volatile("" : "+s"(g));
  return g;
}

// a 32-bit per-lane offset defined afresh where it is used, so the address
// selection there sees its zero extension (scalar base + 32-bit offset form)
__device__ __forceinline__ unsigned voff(unsigned o) {
  asm volatile("" : "+v"(o));
  return o;
}

// The 19 values of one cell, v[q] = plane q at byte offset off.  nt streams
// them past the caches (nontemporal); the choice is uniform over a launch.
template <typename T>
__device__ __forceinline__ void load_cell(T (&v)[Q], const T *planes, size_t plane_bytes, unsigned off, bool nt) {
  gptr<const char> base = sbase((gptr<const char>)planes);
  if (nt) {
    const unsigned o = voff(off);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      v[q] = __builtin_nontemporal_load((gptr<const T>)(base + o));
      base = sbase(base + plane_bytes);
    }
  } else {
    const unsigned o = voff(off);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      v[q] = *((gptr<const T>)(base + o));
      base = sbase(base + plane_bytes);
    }
  }
}

// Push of one cell's post-collision values: v[q] to plane q at cell
// n + step_x + step_y + step_z, where the steps are st[0..5] = (+x, -x, +y,
// -y, +z, -z) for the axes e_q moves along (periodic wrap folded into them).
// The destinations are formed inside each branch, next to their stores.
template <typename T>
__device__ __forceinline__ unsigned push_offset(int q, unsigned n, const unsigned (&st)[6]) {
  const unsigned sx = ex(q) > 0 ? st[0] : (ex(q) < 0 ? st[1] : 0u);
  const unsigned sy = ey(q) > 0 ? st[2] : (ey(q) < 0 ? st[3] : 0u);
  const unsigned sz = ez(q) > 0 ? st[4] : (ez(q) < 0 ? st[5] : 0u);
  return (n + sx + sy + sz) * (unsigned)sizeof(T);
}

template <typename T>
__device__ __forceinline__ void store_cell(const T (&v)[Q], T *planes, size_t plane_bytes, unsigned n,
                                           const unsigned (&st)[6], bool nt) {
  gptr<char> base = sbase((gptr<char>)planes);
  if (nt) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      __builtin_nontemporal_store(v[q], (gptr<T>)(base + push_offset<T>(q, n, st)));
      base = sbase(base + plane_bytes);
    }
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      *((gptr<T>)(base + push_offset<T>(q, n, st))) = v[q];
      base = sbase(base + plane_bytes);
    }
  }
}

// Whole z-rows per workgroup (blockDim.x == zDim and every thread a cell of
// the block): the ten values that move along z (e_z = +-1) are exchanged
// through LDS so that each thread stores, at its own z, the value arriving
// there from z - e_z (periodic within the row).  A wave's stores then cover
// whole cache lines instead of lines shifted by one value.  lds holds
// 10 * blockDim.x * blockDim.y values.
template <typename T>
__device__ __forceinline__ void shift_rows_z(T (&post)[Q], T *lds) {
  const int nt = (int)(blockDim.x * blockDim.y);
  const int zr = (int)threadIdx.x, zd = (int)blockDim.x, row = (int)threadIdx.y * zd;
  const int t = row + zr;
  int k = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q)
    if (ez(q) != 0) lds[(k++) * nt + t] = post[q];
  __syncthreads();
  const int from_m = row + (zr == 0 ? zd - 1 : zr - 1);  // sender of e_z = +1
  const int from_p = row + (zr == zd - 1 ? 0 : zr + 1);  // sender of e_z = -1
  k = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q)
    if (ez(q) != 0) post[q] = lds[(k++) * nt + (ez(q) > 0 ? from_m : from_p)];
}

// The same exchange for rows of whole waves (blockDim.x a multiple of 64):
// inside a wave the values move by lane permutes (no LDS allocation); only
// each wave's first and last lanes go through edge[], 20 values per wave
// (80 bytes per wave for FP32).  Leaves the LDS free for a large workgroup
// that shares the CU.
__device__ __forceinline__ void shift_rows_z_waves(float (&post)[Q], float *edge) {
  const int zd = (int)blockDim.x, wpr = zd >> 6;
  const int lane = (int)threadIdx.x & 63, wr = (int)threadIdx.x >> 6;
  const int w = (int)threadIdx.y * wpr + wr, rw0 = w - wr;
  if (lane == 0 || lane == 63) {
    int k = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q)
      if (ez(q) != 0) edge[(w * 2 + (lane == 63 ? 1 : 0)) * 10 + (k++)] = post[q];
  }
  __syncthreads();
  const int prev_w = rw0 + (wr == 0 ? wpr - 1 : wr - 1), next_w = rw0 + (wr == wpr - 1 ? 0 : wr + 1);
  const int src_m = (lane == 0 ? 0 : lane - 1) << 2, src_p = (lane == 63 ? 63 : lane + 1) << 2;
  int k = 0;
#pragma unroll
  for (int q = 0; q < Q; ++q)
    if (ez(q) != 0) {
      const int v = __builtin_amdgcn_ds_bpermute(ez(q) > 0 ? src_m : src_p, __float_as_int(post[q]));
      float f = __int_as_float(v);
      if (ez(q) > 0 && lane == 0) f = edge[(prev_w * 2 + 1) * 10 + k];
      if (ez(q) < 0 && lane == 63) f = edge[(next_w * 2 + 0) * 10 + k];
      post[q] = f;
      ++k;
    }
}

// Boundary codes (ConstParams.f90:31-34)
enum BcCode {
  BC_FLUID = 0,
  BC_FLUID_FATHER = 1,
  BC_EQ_DIRICHLET = 101,
  BC_NEQ_DIRICHLET = 102,
  BC_EXTRAP1 = 103,
  BC_EXTRAP2 = 104,
  BC_WALL = 201,
  BC_MOVING_WALL = 202,
  BC_WALL_HALFWAY = 203,
  BC_MOVING_WALL_HALFWAY = 204,
  BC_PERIODIC = 301,
  BC_SYMMETRIC = 302,
};

}  // namespace fsil
