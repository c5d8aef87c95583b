// Learned warm start of the multi-direct-forcing loop: a 12 -> 64 -> 64 -> 64
// -> 3 perceptron (tanh hidden layers) per immersed-boundary marker.
//
// Training runs in FP32 with plain dense kernels (forward, backward, Adam),
// every sum in a fixed order so a run is reproducible.  Inference runs on the
// matrix cores: BF16 operands, FP32 accumulation (v_mfma_f32_16x16x16_bf16),
// one wave per 16 markers, the weights of all layers held in LDS as BF16.
//
// Parameter layout (floats): W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64]
// b3[64] W4[3][64] b4[3]; W[o][i] multiplies input i into output o.
#pragma once
#include "fsil_common.h"

namespace fsil {

constexpr int kMlpIn = 12, kMlpH = 64, kMlpOut = 3;
constexpr int kMlpOffW1 = 0, kMlpOffB1 = kMlpOffW1 + kMlpH * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH, kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH, kMlpOffB3 = kMlpOffW3 + kMlpH * kMlpH;
constexpr int kMlpOffW4 = kMlpOffB3 + kMlpH, kMlpOffB4 = kMlpOffW4 + kMlpOut * kMlpH;
constexpr int kMlpParams = kMlpOffB4 + kMlpOut;

// ---------------------------------------------------------------------------
// FP32 training kernels

// Y[s][o] = act(b[o] + sum_k W[o][k] X[s][k]), k ascending; act: 1 tanh, 0 none
__global__ __launch_bounds__(256) void k_dense_fwd(const float *__restrict__ X, int n, int nin,
                                                   const float *__restrict__ W, const float *__restrict__ b,
                                                   int nout, int act, float *__restrict__ Y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nout) return;
  const int s = t / nout, o = t - s * nout;
  const float *x = X + (size_t)s * nin;
  const float *w = W + (size_t)o * nin;
  float z = b[o];
  for (int k = 0; k < nin; ++k) z = z + w[k] * x[k];
  Y[t] = act ? tanhf(z) : z;
}

// output gradient of the mean squared error over n x 3 values, and the
// squared errors summed per block (block order) for the loss
__global__ __launch_bounds__(256) void k_mse_grad(const float *__restrict__ Yp, const float *__restrict__ Yt,
                                                  int count, float *__restrict__ dY,
                                                  double *__restrict__ part) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (t < count) {
    const float d = Yp[t] - Yt[t];
    dY[t] = 2.0f * d / (float)count;
    e2 = (double)d * (double)d;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dH[s][k] = (sum_o dZ[s][o] W[o][k]) * (1 - H[s][k]^2)   (tanh derivative)
__global__ __launch_bounds__(256) void k_dense_bwd_in(const float *__restrict__ dZ, int n, int nout,
                                                      const float *__restrict__ W, int nin,
                                                      const float *__restrict__ H, float *__restrict__ dH) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nin) return;
  const int s = t / nin, k = t - s * nin;
  float g = 0.0f;
  for (int o = 0; o < nout; ++o) g = g + dZ[(size_t)s * nout + o] * W[(size_t)o * nin + k];
  const float h = H[t];
  dH[t] = g * (1.0f - h * h);
}

// gW[o][k] = sum_s dZ[s][o] Hin[s][k] (s ascending); gb[o] = sum_s dZ[s][o]
__global__ __launch_bounds__(256) void k_dense_bwd_w(const float *__restrict__ dZ, const float *__restrict__ Hin,
                                                     int n, int nout, int nin, float *__restrict__ gW,
                                                     float *__restrict__ gb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nout * nin) {
    const int o = t / nin, k = t - o * nin;
    float g = 0.0f;
    for (int s = 0; s < n; ++s) g = g + dZ[(size_t)s * nout + o] * Hin[(size_t)s * nin + k];
    gW[t] = g;
  } else if (t < nout * nin + nout) {
    const int o = t - nout * nin;
    float g = 0.0f;
    for (int s = 0; s < n; ++s) g = g + dZ[(size_t)s * nout + o];
    gb[o] = g;
  }
}

// Adam (Kingma & Ba 2015, algorithm 1); bc1 = 1 - beta1^t, bc2 = 1 - beta2^t
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ th, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, int P, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] = th[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---------------------------------------------------------------------------
// BF16 matrix-core inference

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bfloat16 bits (finite inputs)
__device__ __forceinline__ short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u = u + 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// LDS image of the network: BF16 weights padded to 16-multiples
// (W1: 64 x 16, W2, W3: 64 x 64, W4: 16 x 64), FP32 biases (b4 padded to 16)
struct MlpLds {
  short w1[64 * 16];
  short w2[64 * 64];
  short w3[64 * 64];
  short w4[16 * 64];
  float b1[64], b2[64], b3[64], b4[16];
  short act[4][16 * 64];  // per wave: 16 markers x 64 activations
};

// One wave: C(16 x 16 tile nt) = A(16 x K) * W^T, A from LDS rows (K = 16*kt),
// W rows nt*16..nt*16+15 from LDS, both BF16.  Lane l holds A[l%16][4*(l/16)+j]
// and B[4*(l/16)+j][l%16] = W[l%16][4*(l/16)+j]; the result C[4*(l/16)+j][l%16].
__device__ __forceinline__ f32x4 mfma_tile(const short *A, int lda, const short *W, int ldw, int ksteps, int nt) {
  const int l = threadIdx.x & 63;
  const int r = l & 15, kq = (l >> 4) * 4;
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < ksteps; ++ks) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = A[r * lda + ks * 16 + kq + j];
      b[j] = W[(nt * 16 + r) * ldw + ks * 16 + kq + j];
    }
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  return c;
}

// Y[m][0..2] for n markers, X[m][0..11] inputs; one wave per 16 markers
__global__ __launch_bounds__(256) void k_mlp_infer_bf16(const float *__restrict__ X, int n,
                                                        const float *__restrict__ th, float *__restrict__ Y) {
  __shared__ MlpLds L;
  const int tid = threadIdx.x;
  // weights -> LDS (BF16), zero padding
  for (int i = tid; i < 64 * 16; i += blockDim.x) {
    const int o = i / 16, k = i % 16;
    L.w1[i] = k < kMlpIn ? f2bf(th[kMlpOffW1 + o * kMlpIn + k]) : (short)0;
  }
  for (int i = tid; i < 64 * 64; i += blockDim.x) {
    L.w2[i] = f2bf(th[kMlpOffW2 + i]);
    L.w3[i] = f2bf(th[kMlpOffW3 + i]);
  }
  for (int i = tid; i < 16 * 64; i += blockDim.x) {
    const int o = i / 64;
    L.w4[i] = o < kMlpOut ? f2bf(th[kMlpOffW4 + i]) : (short)0;
  }
  for (int i = tid; i < 64; i += blockDim.x) {
    L.b1[i] = th[kMlpOffB1 + i];
    L.b2[i] = th[kMlpOffB2 + i];
    L.b3[i] = th[kMlpOffB3 + i];
  }
  for (int i = tid; i < 16; i += blockDim.x) L.b4[i] = i < kMlpOut ? th[kMlpOffB4 + i] : 0.f;
  const int w = tid >> 6, l = tid & 63;
  const int m0 = (blockIdx.x * 4 + w) * 16;
  short *act = L.act[w];
  // inputs -> LDS rows (16 x 16, K padded)
  for (int i = l; i < 16 * 16; i += 64) {
    const int r = i / 16, k = i % 16;
    const int m = m0 + r;
    act[r * 64 + k] = (m < n && k < kMlpIn) ? f2bf(X[(size_t)m * kMlpIn + k]) : (short)0;
  }
  __syncthreads();
  const int col = l & 15, rq = (l >> 4) * 4;
  // layer 1: K = 16
  f32x4 c1[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) c1[nt] = mfma_tile(act, 64, L.w1, 16, 1, nt);
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int j = 0; j < 4; ++j) act[(rq + j) * 64 + nt * 16 + col] = f2bf(tanhf(c1[nt][j] + L.b1[nt * 16 + col]));
  __syncthreads();
  // layers 2 and 3: K = 64
  for (int layer = 0; layer < 2; ++layer) {
    const short *Wl = layer == 0 ? L.w2 : L.w3;
    const float *bl = layer == 0 ? L.b2 : L.b3;
    f32x4 c[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) c[nt] = mfma_tile(act, 64, Wl, 64, 4, nt);
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) act[(rq + j) * 64 + nt * 16 + col] = f2bf(tanhf(c[nt][j] + bl[nt * 16 + col]));
    __syncthreads();
  }
  // output layer: N = 16 (3 used)
  const f32x4 c4 = mfma_tile(act, 64, L.w4, 64, 4, 0);
  if (col < kMlpOut)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + rq + j;
      if (m < n) Y[(size_t)m * kMlpOut + col] = c4[j] + L.b4[col];
    }
}

}  // namespace fsil
