// Spinor storage policies and the per-rhs block reduction shared by the
// operator and vector kernels.
#pragma once

#include "cvchip_internal.h"

namespace cvchip {

// Plain storage: cplx<R>[12][V][N]; idx = site*N + rhs; planes VN apart.
template <typename R>
struct IOPlain {
  using real = R;
  using C = cplx<R>;
  C *__restrict__ p;
  uint32_t VN;
  __device__ __forceinline__ void load(C v[12], uint32_t idx) const {
#pragma unroll
    for (int c = 0; c < 12; c++) v[c] = p[(size_t)c * VN + idx];
  }
  __device__ __forceinline__ void store(const C v[12], uint32_t idx) const {
#pragma unroll
    for (int c = 0; c < 12; c++) p[(size_t)c * VN + idx] = v[c];
  }
};

// Half storage with one float scale per (site, rhs): value = half * norm[idx].
// Arithmetic in FP32.  The scale is the largest |component| of the site.
struct IOHalf {
  using real = float;
  using C = float2;
  __half2 *__restrict__ p;
  float *__restrict__ nrm;
  uint32_t VN;
  __device__ __forceinline__ void load(C v[12], uint32_t idx) const {
    const float n = nrm[idx];
#pragma unroll
    for (int c = 0; c < 12; c++) {
      float2 f = __half22float2(p[(size_t)c * VN + idx]);
      v[c] = make_float2(f.x * n, f.y * n);
    }
  }
  __device__ __forceinline__ void store(const C v[12], uint32_t idx) const {
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 12; c++) m = fmaxf(m, fmaxf(fabsf(v[c].x), fabsf(v[c].y)));
    const float inv = m > 0.f ? 1.0f / m : 0.f;
    nrm[idx] = m;
#pragma unroll
    for (int c = 0; c < 12; c++)
      p[(size_t)c * VN + idx] = __floats2half2_rn(v[c].x * inv, v[c].y * inv);
  }
};

// Sums Q per-thread values over the threads of a block that share a
// right-hand side (thread t holds rhs t % N; blockDim.x is a multiple of N)
// with a fixed-order tree, and writes part[(block*Q + q)*N + rhs].
template <int Q>
__device__ __forceinline__ void block_reduce_store(const double v[Q], double *__restrict__ part, int N) {
  __shared__ double sh[Q][256];
  const int tid = threadIdx.x;
  const int B = blockDim.x;
#pragma unroll
  for (int q = 0; q < Q; q++) sh[q][tid] = v[q];
  __syncthreads();
  const int S = B / N;
  const int row = tid / N;
  for (int active = S; active > 1;) {
    const int half = (active + 1) >> 1;
    if (row < active - half) {
#pragma unroll
      for (int q = 0; q < Q; q++) sh[q][tid] += sh[q][tid + half * N];
    }
    __syncthreads();
    active = half;
  }
  if (tid < N) {
#pragma unroll
    for (int q = 0; q < Q; q++) part[((size_t)blockIdx.x * Q + q) * N + tid] = sh[q][tid];
  }
}

// Reduces part[nblk][Q][N] over blocks into res[q*N + rhs]; launched with one
// block of 256 threads per (q, rhs); fixed order, so results are reproducible.
__global__ void k_reduce_partials(const double *__restrict__ part, int nblk, int Q, int N,
                                  double *__restrict__ res);

int reduce_partials(cvchip_ctx *c, const double *part, int nblk, int Q, int N, double *res);

}  // namespace cvchip
