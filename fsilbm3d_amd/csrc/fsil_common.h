// Shared definitions of the FSILBM3D MI355X hot-path library (device + host).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace fsil {

constexpr int Q = 19;

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
