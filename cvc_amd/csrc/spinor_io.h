// Spinor storage policies and the per-rhs block reduction shared by the
// operator and vector kernels.
#pragma once

#include "cvchip_internal.h"

namespace cvchip {

// Every load and store of a (site, rhs) moves 16 bytes per lane (the widest
// vector memory access), because the texture addresser's cost is per lane
// and per instruction rather than per byte.  Components are therefore grouped
// K = 16 B / sizeof(complex) at a time:
//   FP64: K = 1, 12 planes of double2       element (c, idx) at  c*VN + idx
//   FP32: K = 2,  6 planes of float4        element (c, idx) at (c/2)*VN + idx, half c%2
//   FP16: K = 4,  3 planes of 4 x half2     element (c, idx) at (c/4)*VN + idx, quarter c%4
// with idx = site*N + rhs and VN = V*N.  cplx_index() gives the position in
// units of one complex number.
__host__ __device__ __forceinline__ size_t cplx_index(int c, uint32_t idx, uint32_t VN, int K) {
  return ((size_t)(c / K) * VN + idx) * K + (c % K);
}

template <typename R> struct IOPlain;

template <>
struct IOPlain<double> {
  using real = double;
  using C = double2;
  static constexpr int K = 1;
  C *__restrict__ p;
  uint32_t VN;
  __device__ __forceinline__ void load(C v[12], uint32_t idx) const {
#pragma unroll
    for (int c = 0; c < 12; c++) v[c] = p[(size_t)c * VN + idx];
  }
  __device__ __forceinline__ C load_c(int c, uint32_t idx) const { return p[(size_t)c * VN + idx]; }
  // streaming load (no allocation in the caches): for a field read once per launch
  __device__ __forceinline__ void load_nt(C v[12], uint32_t idx) const {
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v *q = reinterpret_cast<const d2v *>(p);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const d2v w = __builtin_nontemporal_load(q + (size_t)c * VN + idx);
      v[c] = make_double2(w.x, w.y);
    }
  }
  __device__ __forceinline__ void store(const C v[12], uint32_t idx) const {
#pragma unroll
    for (int c = 0; c < 12; c++) p[(size_t)c * VN + idx] = v[c];
  }
  // streaming store (no allocation in the caches)
  __device__ __forceinline__ void store_nt(const C v[12], uint32_t idx) const {
    typedef double d2v __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const d2v w = {v[c].x, v[c].y};
      __builtin_nontemporal_store(w, reinterpret_cast<d2v *>(p + (size_t)c * VN + idx));
    }
  }
};

template <>
struct IOPlain<float> {
  using real = float;
  using C = float2;
  static constexpr int K = 2;
  C *__restrict__ p;   // viewed as float4 planes
  uint32_t VN;
  __device__ __forceinline__ void load(C v[12], uint32_t idx) const {
    const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const float4 f = q[(size_t)k * VN + idx];
      v[2 * k] = make_float2(f.x, f.y);
      v[2 * k + 1] = make_float2(f.z, f.w);
    }
  }
  __device__ __forceinline__ void load_nt(C v[12], uint32_t idx) const {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v *q = reinterpret_cast<const f4v *>(p);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const f4v f = __builtin_nontemporal_load(q + (size_t)k * VN + idx);
      v[2 * k] = make_float2(f.x, f.y);
      v[2 * k + 1] = make_float2(f.z, f.w);
    }
  }
  __device__ __forceinline__ void store(const C v[12], uint32_t idx) const {
    float4 *q = reinterpret_cast<float4 *>(p);
#pragma unroll
    for (int k = 0; k < 6; k++) q[(size_t)k * VN + idx] = make_float4(v[2 * k].x, v[2 * k].y, v[2 * k + 1].x, v[2 * k + 1].y);
  }
  __device__ __forceinline__ void store_nt(const C v[12], uint32_t idx) const {
    typedef float f4v __attribute__((ext_vector_type(4)));
    f4v *q = reinterpret_cast<f4v *>(p);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const f4v w = {v[2 * k].x, v[2 * k].y, v[2 * k + 1].x, v[2 * k + 1].y};
      __builtin_nontemporal_store(w, q + (size_t)k * VN + idx);
    }
  }
};

// Half storage with one float scale per (site, rhs): value = half * norm[idx].
// Arithmetic in FP32.  The scale is the largest |component| of the site.
struct IOHalf {
  using real = float;
  using C = float2;
  static constexpr int K = 4;
  __half2 *__restrict__ p;   // viewed as uint4 planes (4 x half2)
  float *__restrict__ nrm;
  uint32_t VN;
  __device__ __forceinline__ void load(C v[12], uint32_t idx) const {
    const float n = nrm[idx];
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const uint4 w = q[(size_t)k * VN + idx];
      const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        __half2 h;
        __builtin_memcpy(&h, &u[j], 4);
        const float2 f = __half22float2(h);
        v[4 * k + j] = make_float2(f.x * n, f.y * n);
      }
    }
  }
  __device__ __forceinline__ void load_nt(C v[12], uint32_t idx) const {
    typedef unsigned u4v __attribute__((ext_vector_type(4)));
    const float n = __builtin_nontemporal_load(nrm + idx);
    const u4v *q = reinterpret_cast<const u4v *>(p);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const u4v w = __builtin_nontemporal_load(q + (size_t)k * VN + idx);
      const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        __half2 h;
        __builtin_memcpy(&h, &u[j], 4);
        const float2 f = __half22float2(h);
        v[4 * k + j] = make_float2(f.x * n, f.y * n);
      }
    }
  }
  __device__ __forceinline__ void store(const C v[12], uint32_t idx) const { store_t<false>(v, idx); }
  __device__ __forceinline__ void store_nt(const C v[12], uint32_t idx) const { store_t<true>(v, idx); }
  template <bool NT>
  __device__ __forceinline__ void store_t(const C v[12], uint32_t idx) const {
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 12; c++) m = fmaxf(m, fmaxf(fabsf(v[c].x), fabsf(v[c].y)));
    // a site whose largest component is below kTiny is stored as zero: its
    // reciprocal scale would overflow FP32 below ~3e-39 (the far tail of a
    // point-source Krylov vector reaches the denormal range on large lattices)
    constexpr float kTiny = 1e-37f;
    const bool keep = m >= kTiny;
    const float inv = keep ? 1.0f / m : 0.f;
    typedef unsigned u4v __attribute__((ext_vector_type(4)));
    if constexpr (NT) __builtin_nontemporal_store(keep ? m : 0.f, nrm + idx);
    else nrm[idx] = keep ? m : 0.f;
    u4v *q = reinterpret_cast<u4v *>(p);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      uint32_t u[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const __half2 h = __floats2half2_rn(v[4 * k + j].x * inv, v[4 * k + j].y * inv);
        __builtin_memcpy(&u[j], &h, 4);
      }
      const u4v w = {u[0], u[1], u[2], u[3]};
      if constexpr (NT) __builtin_nontemporal_store(w, q + (size_t)k * VN + idx);
      else q[(size_t)k * VN + idx] = w;
    }
  }
};

// Sums Q per-thread values over the threads of a block that share a
// right-hand side (thread t holds rhs t % N; blockDim.x is a multiple of N and
// at most 256) in a fixed order, and writes part[(q*N + rhs)*gridDim.x + block],
// so that the second stage reads each (q, rhs) row contiguously.
// N <= 64: lanes l and l + N of a wave hold the same rhs, so a cross-lane tree
// with offsets N, 2N, 4N, ... (ds_bpermute shuffles, no LDS traffic and no
// barrier) leaves lane l < N of each wave with the wave's sum for its rhs; the
// per-wave sums meet in LDS once and one thread per rhs adds them in wave order.
// N > 64: a tree over the block's rows in LDS.
template <int Q>
__device__ __forceinline__ void block_reduce_store(const double v[Q], double *__restrict__ part, int N) {
  const uint32_t slot = blockIdx.x, nb = gridDim.x;
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  if (N <= 64) {
    __shared__ double wsum[Q][4][64];   // [q][wave][rhs]
    const int lane = tid & 63, w = tid >> 6, nw = (B + 63) >> 6;
    const int nl = min(64, B - 64 * w);    // lanes of this wave
    double s[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) s[q] = v[q];
    for (int off = N; off < 64; off <<= 1) {
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const double o = __shfl_down(s[q], (unsigned)off, 64);
        if (lane + off < nl) s[q] += o;
      }
    }
    // zero the table (a partial last wave may hold fewer than N rhs), then fill
    for (int e = tid; e < Q * nw * 64; e += B) (&wsum[0][0][0])[(e / (nw * 64)) * 256 + e % (nw * 64)] = 0.;
    __syncthreads();
    if (lane < N && lane < nl) {
      const int r = (64 * w + lane) % N;
#pragma unroll
      for (int q = 0; q < Q; q++) wsum[q][w][r] = s[q];
    }
    __syncthreads();
    if (tid < N) {
#pragma unroll
      for (int q = 0; q < Q; q++) {
        double a = 0.;
        for (int ww = 0; ww < nw; ww++) a += wsum[q][ww][tid];
        part[((size_t)q * N + tid) * nb + slot] = a;
      }
    }
    return;
  }
  __shared__ double sh[Q][256];
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
    for (int q = 0; q < Q; q++) part[((size_t)q * N + tid) * nb + slot] = sh[q][tid];
  }
}

// Reduces part[Q][N][nblk] over blocks into res[q*N + rhs] (reduce_partials:
// segment sums, then one block of 256 threads per (q, rhs)); fixed order, so
// results are reproducible.
__global__ void k_reduce_partials(const double *__restrict__ part, int nblk, int Q, int N,
                                  double *__restrict__ res);

int reduce_partials(cvchip_ctx *c, const double *part, int nblk, int Q, int N, double *res);

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD b % 8).
constexpr unsigned kXcds = 8;

}  // namespace cvchip
