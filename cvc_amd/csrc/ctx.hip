// Context, device memory and layout conversion of libcvchip.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "spinor_io.h"

namespace cvchip {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  if (e == hipErrorOutOfMemory) return CVCHIP_ERR_OOM;
  return CVCHIP_ERR_HIP;
}

int ensure_stage(cvchip_ctx *c, size_t bytes) {
  if (c->stage_bytes >= bytes) return 0;
  if (c->stage) CVC_HIP(hipFree(c->stage));
  c->stage = nullptr;
  c->stage_bytes = 0;
  CVC_HIP(hipMalloc(&c->stage, bytes));
  c->stage_bytes = bytes;
  return 0;
}

// Device-to-host copies of 16 MB and more go through two pinned host buffers
// in turn: the copy of chunk i out of its pinned buffer (over up to 8 host
// threads) overlaps the DMA of chunk i + 1.  A direct copy into pageable memory
// moves 403 MB in about 23 ms (17 GB/s).
constexpr size_t kPinChunk = (size_t)64 << 20;
constexpr size_t kPinMin = (size_t)16 << 20;

static void par_memcpy(void *dst, const void *src, size_t bytes) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(8u, nt));
  nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, bytes / ((size_t)4 << 20)));
  const size_t per = (bytes + nt - 1) / nt;
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; t++) {
    const size_t lo = (size_t)t * per;
    if (lo >= bytes) break;
    const size_t len = std::min(per, bytes - lo);
    try {
      th.emplace_back([=] { memcpy((char *)dst + lo, (const char *)src + lo, len); });
    } catch (...) {   // no thread to be had: this slice on the calling thread
      memcpy((char *)dst + lo, (const char *)src + lo, len);
    }
  }
  memcpy(dst, src, std::min(per, bytes));
  for (auto &x : th) x.join();
}

static int ensure_pinned(cvchip_ctx *c) {
  for (int k = 0; k < 2; k++) {
    if (!c->pin[k]) CVC_HIP(hipHostMalloc(&c->pin[k], kPinChunk, hipHostMallocDefault));
    if (!c->pin_ev[k]) CVC_HIP(hipEventCreateWithFlags(&c->pin_ev[k], hipEventDisableTiming));
  }
  return 0;
}

int copy_to_host(cvchip_ctx *c, void *host, const void *dev, size_t bytes) {
  if (bytes < kPinMin) {
    CVC_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    CVC_HIP(hipStreamSynchronize(c->stream));
    return 0;
  }
  // a DMA into a pinned buffer follows, in stream order, the one that last read it
  CVC_TRY(ensure_pinned(c));
  const size_t n = (bytes + kPinChunk - 1) / kPinChunk;
  auto issue = [&](size_t i) -> int {
    const size_t off = i * kPinChunk, len = std::min(kPinChunk, bytes - off);
    CVC_HIP(hipMemcpyAsync(c->pin[i % 2], (const char *)dev + off, len, hipMemcpyDeviceToHost, c->stream));
    CVC_HIP(hipEventRecord(c->pin_ev[i % 2], c->stream));
    return 0;
  };
  for (size_t i = 0; i < n && i < 2; i++) CVC_TRY(issue(i));
  for (size_t i = 0; i < n; i++) {
    CVC_HIP(hipEventSynchronize(c->pin_ev[i % 2]));
    const size_t off = i * kPinChunk, len = std::min(kPinChunk, bytes - off);
    par_memcpy((char *)host + off, c->pin[i % 2], len);
    if (i + 2 < n) CVC_TRY(issue(i + 2));   // this buffer is free again
  }
  return 0;
}

// Host-to-device copies of 16 MB and more: each chunk is copied into a pinned
// buffer (up to 8 host threads) while the DMA of the previous chunk runs (a
// direct copy from pageable memory moves 403 MB in about 13.5 ms, 30 GB/s).
// The copies are enqueued on the context's stream; a large host buffer has
// been read when the call returns, a small one (a plain asynchronous copy) only
// once the stream has reached it, so callers synchronize the stream before the
// host buffer may change.
int copy_to_device(cvchip_ctx *c, void *dev, const void *host, size_t bytes) {
  if (bytes < kPinMin) {
    CVC_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream));
    return 0;
  }
  CVC_TRY(ensure_pinned(c));
  const size_t n = (bytes + kPinChunk - 1) / kPinChunk;
  for (size_t i = 0; i < n; i++) {
    const int k = (int)(i % 2);
    const size_t off = i * kPinChunk, len = std::min(kPinChunk, bytes - off);
    CVC_HIP(hipEventSynchronize(c->pin_ev[k]));   // the last DMA from or into this buffer has ended
    par_memcpy(c->pin[k], (const char *)host + off, len);
    CVC_HIP(hipMemcpyAsync((char *)dev + off, c->pin[k], len, hipMemcpyHostToDevice, c->stream));
    CVC_HIP(hipEventRecord(c->pin_ev[k], c->stream));
  }
  return 0;
}

int ensure_red(cvchip_ctx *c, size_t bytes) {
  if (c->red.bytes >= bytes) return 0;
  if (c->red.ptr) CVC_HIP(hipFree(c->red.ptr));
  c->red.ptr = nullptr;
  c->red.bytes = 0;
  CVC_HIP(hipMalloc(&c->red.ptr, bytes));
  c->red.bytes = bytes;
  return 0;
}

int ensure_red2(cvchip_ctx *c, size_t bytes) {
  if (c->red2.bytes >= bytes) return 0;
  if (c->red2.ptr) CVC_HIP(hipFree(c->red2.ptr));
  c->red2.ptr = nullptr;
  c->red2.bytes = 0;
  CVC_HIP(hipMalloc(&c->red2.ptr, bytes));
  c->red2.bytes = bytes;
  return 0;
}

static size_t elem_bytes(int prec) {
  return prec == CVCHIP_PREC_F64 ? sizeof(double2) : prec == CVCHIP_PREC_F32 ? sizeof(float2) : sizeof(__half2);
}

// hipMalloc that, when the device is full, gives the context's idle buffers
// back, oldest first, as many bytes at a time as the request, and tries again
// after each round: the buffers a call has just returned (the ones its next
// call of the same size takes again) stay while older ones can make room.
// (Releasing the whole pool instead made a 60-column even/odd solve, repeated
// after solves of other widths near the device's capacity, allocate its
// workspaces anew on every call: 4.4 s against 1.9 s.)
static int device_malloc(cvchip_ctx *c, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  while (e != hipSuccess && !c->idle.empty()) {
    (void)hipGetLastError();   // clear the failed allocation's error, then retry with fewer idle buffers
    size_t freed = 0;
    while (!c->idle.empty() && freed < bytes) {
      (void)hipFree(c->idle.front().first);
      freed += c->idle.front().second;
      c->idle_bytes -= c->idle.front().second;
      c->idle.erase(c->idle.begin());
    }
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return CVCHIP_ERR_OOM;
  }
  return 0;
}

int pool_get(cvchip_ctx *c, void **p, size_t bytes) {
  for (size_t k = c->idle.size(); k-- > 0;)
    if (c->idle[k].second == bytes) {   // the most recently freed buffer of this size
      *p = c->idle[k].first;
      c->idle_bytes -= bytes;
      c->idle.erase(c->idle.begin() + (std::ptrdiff_t)k);
      return 0;
    }
  return device_malloc(c, p, bytes);
}

// Buffers return to the pool in stream order: every use of a context's
// temporaries is enqueued on the context's one stream, so a later reuse
// cannot overtake an earlier kernel.
void pool_put(cvchip_ctx *c, void *p, size_t bytes) {
  if (!p) return;
  c->idle.emplace_back(p, bytes);
  c->idle_bytes += bytes;
  while (c->idle_bytes > kPoolIdleMax && !c->idle.empty()) {   // oldest first
    (void)hipFree(c->idle.front().first);
    c->idle_bytes -= c->idle.front().second;
    c->idle.erase(c->idle.begin());
  }
}

void pool_release(cvchip_ctx *c) {
  for (auto &b : c->idle) (void)hipFree(b.first);
  c->idle.clear();
  c->idle_bytes = 0;
}

int field_alloc(cvchip_ctx *c, Field &f, int nrhs, int prec, int64_t sites, bool pooled) {
  if (nrhs <= 0 || nrhs > 256) {
    set_error("nrhs %d outside 1..256", nrhs);
    return CVCHIP_ERR_ARG;
  }
  if (prec < CVCHIP_PREC_F64 || prec > CVCHIP_PREC_F16) {
    set_error("unknown precision %d", prec);
    return CVCHIP_ERR_ARG;
  }
  if ((uint64_t)c->lat.V * (uint64_t)nrhs >= ((uint64_t)1 << 32) - 1024) {
    set_error("volume x nrhs = %lld exceeds the 32-bit site-rhs index", (long long)(c->lat.V * nrhs));
    return CVCHIP_ERR_ARG;
  }
  f.nrhs = nrhs;
  f.prec = prec;
  f.V = sites > 0 ? sites : c->lat.V;
  f.bytes = (size_t)12 * f.V * nrhs * elem_bytes(prec);
  if (!pooled) {
    f.pool = nullptr;
    CVC_TRY(device_malloc(c, &f.data, f.bytes));
    if (prec == CVCHIP_PREC_F16) {
      const int rc = device_malloc(c, (void **)&f.norm, sizeof(float) * f.V * nrhs);
      if (rc) {
        (void)hipFree(f.data);
        f.data = nullptr;
        return rc;
      }
    }
    return 0;
  }
  f.pool = c;
  CVC_TRY(pool_get(c, &f.data, f.bytes));
  if (prec == CVCHIP_PREC_F16) CVC_TRY(pool_get(c, (void **)&f.norm, sizeof(float) * f.V * nrhs));
  return 0;
}

void field_free(Field &f) {
  if (f.pool) {
    pool_put(f.pool, f.data, f.bytes);
    if (f.norm) pool_put(f.pool, f.norm, sizeof(float) * f.V * f.nrhs);
  } else {
    if (f.data) (void)hipFree(f.data);
    if (f.norm) (void)hipFree(f.norm);
  }
  f.data = nullptr;
  f.norm = nullptr;
  f.pool = nullptr;
}

// ------------------------------------------------------------- kernels

// AoS staging double[n][V][24] -> field rhs rhs0..rhs0+n-1
template <typename R>
__global__ void k_aos_to_field(cplx<R> *__restrict__ f, const double *__restrict__ aos, int64_t V,
                               int N, int rhs0, int n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // over n*V*12, comp fastest
  if (i >= (int64_t)n * V * 12) return;
  int c = (int)(i % 12);
  int64_t sv = i / 12;
  int64_t s = sv % V;
  int r = (int)(sv / V);
  const double *p = aos + ((int64_t)r * V + s) * 24 + 2 * c;
  f[cplx_index(c, (uint32_t)(s * N + rhs0 + r), (uint32_t)(V * N), 16 / (int)sizeof(cplx<R>))] =
      mk<R>((R)p[0], (R)p[1]);
}

template <typename R>
__global__ void k_field_to_aos(double *__restrict__ aos, const cplx<R> *__restrict__ f, int64_t V,
                               int N, int rhs0, int n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * V * 12) return;
  int c = (int)(i % 12);
  int64_t sv = i / 12;
  int64_t s = sv % V;
  int r = (int)(sv / V);
  cplx<R> v = f[cplx_index(c, (uint32_t)(s * N + rhs0 + r), (uint32_t)(V * N), 16 / (int)sizeof(cplx<R>))];
  double *p = aos + ((int64_t)r * V + s) * 24 + 2 * c;
  p[0] = (double)v.x;
  p[1] = (double)v.y;
}

// gauge AoS double[V][4][18] -> recon-18 FP64 [4][9][V], recon-12 FP64 [4][6][V],
// recon-12 FP32 [4][6][V]
__global__ void k_gauge_convert(double2 *__restrict__ g18, double2 *__restrict__ g12,
                                float2 *__restrict__ g12f, const double *__restrict__ aos, int64_t V) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // over V*4*9
  if (i >= V * 36) return;
  int k = (int)(i % 9);
  int mu = (int)((i / 9) % 4);
  int64_t s = i / 36;
  const double *p = aos + s * 72 + 18 * mu + 2 * k;
  double2 v = make_double2(p[0], p[1]);
  g18[((int64_t)mu * 9 + k) * V + s] = v;
  if (k < 6) {
    g12[((int64_t)mu * 6 + k) * V + s] = v;
    g12f[((int64_t)mu * 6 + k) * V + s] = make_float2((float)v.x, (float)v.y);
  }
}

// deterministic fill: each real = uniform(-1, 1) from a 64-bit mix of (seed, index)
__device__ __forceinline__ double mix_uniform(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__global__ void k_fill_random(double2 *__restrict__ f, uint64_t n, uint64_t seed) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f[i] = make_double2(mix_uniform(seed * 0x100000001B3ull + 2 * i), mix_uniform(seed * 0x100000001B3ull + 2 * i + 1));
}

// Point sources of several source points (the reference's
// source[_GSI(ix) + 2 ia] = 1, cvc_exact2_xspace.c:521-524): column 12 k + a of
// the N = 12 nsrc columns holds the unit vector of spin-colour a at site[k].
constexpr int kMaxPointSources = 21;   // 12 * 21 = 252 columns (at most 256)
struct PointSites {
  uint32_t site[kMaxPointSources];
};

__global__ void k_point_sources_n(double2 *__restrict__ b, uint32_t V, uint32_t N, PointSites ps) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;   // (c V + site) N + r
  if (i >= (uint64_t)12 * V * N) return;
  const uint32_t r = (uint32_t)(i % N);
  const uint64_t cs = i / N;
  const uint32_t site = (uint32_t)(cs % V), c = (uint32_t)(cs / V);
  b[i] = make_double2((site == ps.site[r / 12] && c == r % 12) ? 1.0 : 0.0, 0.0);
}

__global__ void k_zero16(uint4 *__restrict__ p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// dst[.., d0 + r] = src[.., s0 + r] for r < n (FP64 sets of the same volume)
__global__ void k_copy_columns(double2 *__restrict__ dst, const double2 *__restrict__ src, uint32_t V, uint32_t ND,
                               uint32_t NS, uint32_t d0, uint32_t s0, uint32_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)12 * V * n) return;
  const uint32_t r = (uint32_t)(i % n);
  const uint64_t cs = i / n;   // c V + site
  dst[cs * ND + d0 + r] = src[cs * NS + s0 + r];
}

}  // namespace cvchip

using namespace cvchip;

// =================================================================== C ABI

extern "C" int cvchip_field_zero(cvchip_ctx *c, cvchip_field *fh) {
  return cvchip::guarded("cvchip_field_zero", [&]() -> int {
    if (!c || !fh) {
      set_error("NULL argument to cvchip_field_zero");
      return CVCHIP_ERR_NULL;
    }
    Field &f = fh->f;
    const uint64_t n = f.bytes / 16;   // every set is a whole number of 16-byte planes
    if (n) hipLaunchKernelGGL(k_zero16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (uint4 *)f.data, n);
    if (f.norm) {
      const uint64_t m = (uint64_t)f.V * f.nrhs;   // one float per (site, rhs): zero scale
      hipLaunchKernelGGL(k_zero16, dim3((unsigned)((m / 4 + 255) / 256)), dim3(256), 0, c->stream, (uint4 *)f.norm, m / 4);
      if (m % 4) CVC_HIP(hipMemsetAsync(f.norm + (m / 4) * 4, 0, sizeof(float) * (m % 4), c->stream));
    }
    CVC_HIP(hipGetLastError());
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_point_sources(cvchip_ctx *c, cvchip_field *fh, int nsrc, const int *coords) {
  return cvchip::guarded("cvchip_field_point_sources", [&]() -> int {
    if (!c || !fh || !coords) {
      set_error("NULL argument to cvchip_field_point_sources");
      return CVCHIP_ERR_NULL;
    }
    Field &f = fh->f;
    if (f.prec != CVCHIP_PREC_F64 || f.V != c->lat.V || nsrc < 1 || nsrc > kMaxPointSources || f.nrhs != 12 * nsrc) {
      set_error("point sources need a full-lattice FP64 set of 12 x nsrc columns, 1 <= nsrc <= %d (nsrc %d, %d columns)",
                kMaxPointSources, nsrc, f.nrhs);
      return CVCHIP_ERR_ARG;
    }
    const int L[4] = {c->lat.T, c->lat.LX, c->lat.LY, c->lat.LZ};
    PointSites ps{};
    for (int k = 0; k < nsrc; k++) {
      const int *x = coords + 4 * k;
      for (int m = 0; m < 4; m++)
        if (x[m] < 0 || x[m] >= L[m]) {
          set_error("source %d coordinate %d = %d outside 0..%d", k, m, x[m], L[m] - 1);
          return CVCHIP_ERR_ARG;
        }
      ps.site[k] = (uint32_t)((((int64_t)x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3]);
    }
    const uint64_t n = (uint64_t)12 * f.V * f.nrhs;
    hipLaunchKernelGGL(k_point_sources_n, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       (double2 *)f.data, (uint32_t)f.V, (uint32_t)f.nrhs, ps);
    CVC_HIP(hipGetLastError());
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_copy_columns(cvchip_ctx *c, cvchip_field *dst, int dst_col0, const cvchip_field *src,
                                         int src_col0, int n) {
  return cvchip::guarded("cvchip_field_copy_columns", [&]() -> int {
    if (!c || !dst || !src) {
      set_error("NULL argument to cvchip_field_copy_columns");
      return CVCHIP_ERR_NULL;
    }
    const Field &d = dst->f, &s = src->f;
    if (d.prec != CVCHIP_PREC_F64 || s.prec != CVCHIP_PREC_F64 || d.V != s.V) {
      set_error("column copies need FP64 sets of the same volume");
      return CVCHIP_ERR_ARG;
    }
    if (n < 0 || dst_col0 < 0 || src_col0 < 0 || dst_col0 + n > d.nrhs || src_col0 + n > s.nrhs ||
        (d.data == s.data && dst_col0 < src_col0 + n && src_col0 < dst_col0 + n)) {
      set_error("column ranges %d+%d of %d and %d+%d of %d are invalid or overlap", dst_col0, n, d.nrhs, src_col0, n,
                s.nrhs);
      return CVCHIP_ERR_ARG;
    }
    if (n == 0) return CVCHIP_OK;
    const uint64_t tot = (uint64_t)12 * d.V * n;
    hipLaunchKernelGGL(k_copy_columns, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, (double2 *)d.data,
                       (const double2 *)s.data, (uint32_t)d.V, (uint32_t)d.nrhs, (uint32_t)s.nrhs, (uint32_t)dst_col0,
                       (uint32_t)src_col0, (uint32_t)n);
    CVC_HIP(hipGetLastError());
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_fill_random(cvchip_ctx *c, cvchip_field *fh, uint64_t seed) {
  return cvchip::guarded("cvchip_field_fill_random", [&]() -> int {
    if (!c || !fh) return CVCHIP_ERR_NULL;
    Field &f = fh->f;
    if (f.prec != CVCHIP_PREC_F64) {
      set_error("random fill is defined for FP64 sets (convert with cvchip_field_copy)");
      return CVCHIP_ERR_ARG;
    }
    const uint64_t n = (uint64_t)12 * f.V * f.nrhs;
    hipLaunchKernelGGL(k_fill_random, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (double2 *)f.data,
                       n, seed);
    CVC_HIP(hipGetLastError());
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}

extern "C" const char *cvchip_last_error(void) { return g_err; }

extern "C" int cvchip_ctx_create(cvchip_ctx **ctx, const int dims[4], const double bc_angle[4],
                                 double kappa, double mu) {
  return cvchip::guarded("cvchip_ctx_create", [&]() -> int {
    if (!ctx || !dims || !bc_angle) {
      set_error("NULL argument to cvchip_ctx_create");
      return CVCHIP_ERR_NULL;
    }
    for (int m = 0; m < 4; m++)
      if (dims[m] < 2) {
        set_error("lattice extent %d in direction %d is below 2", dims[m], m);
        return CVCHIP_ERR_ARG;
      }
    int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
    if (V >= ((int64_t)1 << 31)) {
      set_error("lattice volume %lld too large", (long long)V);
      return CVCHIP_ERR_ARG;
    }
    if (!(kappa > 0.0)) {
      set_error("kappa must be positive");
      return CVCHIP_ERR_ARG;
    }
    cvchip_ctx *c = new cvchip_ctx();
    c->lat.T = dims[0];
    c->lat.LX = dims[1];
    c->lat.LY = dims[2];
    c->lat.LZ = dims[3];
    c->lat.V = V;
    c->kappa = kappa;
    c->mu = mu;
    for (int m = 0; m < 4; m++) {
      c->bc[m] = bc_angle[m];
      // cvc_geometry.c:452-459: co_phase_up[mu] = exp(i pi BCangle[mu] / L_mu)
      c->ph.re[m] = cos(bc_angle[m] * M_PI / (double)dims[m]);
      c->ph.im[m] = sin(bc_angle[m] * M_PI / (double)dims[m]);
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) {
      const int rs = twopoint_scratch_init(c->device);
      if (rs) {
        cvchip_ctx_destroy(c);
        return rs;
      }
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
      int rc = hip_fail(e, "context setup", __FILE__, __LINE__);
      cvchip_ctx_destroy(c);
      return rc;
    }
    *ctx = c;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_ctx_destroy(cvchip_ctx *c) {
  return cvchip::guarded("cvchip_ctx_destroy", [&]() -> int {
    if (!c) return CVCHIP_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pool_release(c);
    if (c->gauge64) (void)hipFree(c->gauge64);
    if (c->gauge64r12) (void)hipFree(c->gauge64r12);
    if (c->gauge32r12) (void)hipFree(c->gauge32r12);
    if (c->stage) (void)hipFree(c->stage);
    for (int k = 0; k < 2; k++) {
      if (c->pin[k]) (void)hipHostFree(c->pin[k]);
      if (c->pin_ev[k]) (void)hipEventDestroy(c->pin_ev[k]);
    }
    if (c->red.ptr) (void)hipFree(c->red.ptr);
    if (c->red2.ptr) (void)hipFree(c->red2.ptr);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_ctx_set_params(cvchip_ctx *c, double kappa, double mu) {
  return cvchip::guarded("cvchip_ctx_set_params", [&]() -> int {
    if (!c) return CVCHIP_ERR_NULL;
    if (!(kappa > 0.0)) {
      set_error("kappa must be positive");
      return CVCHIP_ERR_ARG;
    }
    c->kappa = kappa;
    c->mu = mu;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_ctx_set_link_recon(cvchip_ctx *c, int recon) {
  return cvchip::guarded("cvchip_ctx_set_link_recon", [&]() -> int {
    if (!c) return CVCHIP_ERR_NULL;
    if (recon != 18 && recon != 12) {
      set_error("link reconstruction must be 18 or 12, got %d", recon);
      return CVCHIP_ERR_ARG;
    }
    c->link_recon = recon;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_ctx_set_solver_flags(cvchip_ctx *c, unsigned flags) {
  return cvchip::guarded("cvchip_ctx_set_solver_flags", [&]() -> int {
    if (!c) return CVCHIP_ERR_NULL;
    if (flags & ~(CVCHIP_SOLVE_LAGGED | CVCHIP_SOLVE_VERBOSE | CVCHIP_SOLVE_EO)) {
      set_error("unknown solver flags 0x%x", flags);
      return CVCHIP_ERR_ARG;
    }
    if ((flags & CVCHIP_SOLVE_EO) && c->lat.LZ % 2) {
      set_error("even/odd preconditioning needs an even LZ (got %d)", c->lat.LZ);
      return CVCHIP_ERR_ARG;
    }
    c->solve_flags = flags;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_ctx_set_dslash_policy(cvchip_ctx *c, int policy) {
  return cvchip::guarded("cvchip_ctx_set_dslash_policy", [&]() -> int {
    if (!c) return CVCHIP_ERR_NULL;
    const int mode = (policy >> 4) & 15, yb = (policy >> 8) & 255;
    const int bt = (policy >> 16) & 15, bx = (policy >> 20) & 15;
    if (policy < 0 || policy >= (1 << 28) || mode > 6) {
      set_error("operator policy 0x%x: bits 0-3 staging/streaming, 4-7 traversal 0..6, 8-15 y-block "
                "(traversal 6: march length), 16-19 brick t (traversal 5: log2 of the column's z extent; 6: patch x), "
                "20-23 brick x (traversal 6: patch y), 24-27 LDS padding",
                policy);
      return CVCHIP_ERR_ARG;
    }
    if (mode == 6) {
      // t march: patch extents divide LX and LY
      if ((bt && c->lat.LX % bt) || (bx && c->lat.LY % bx)) {
        set_error("operator policy 0x%x: t-march patch extents x %d, y %d must divide LX=%d, LY=%d", policy, bt, bx,
                  c->lat.LX, c->lat.LY);
        return CVCHIP_ERR_ARG;
      }
    } else if (mode == 5) {
      // column extents: powers of two dividing the lattice
      const int zb = bt ? 1 << bt : 8;
      const int cy = yb ? yb : 4, cx = bx ? bx : 4;
      auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
      if (bt > 5 || !pow2(cy) || !pow2(cx) || c->lat.LY % cy || c->lat.LX % cx || c->lat.LZ % zb) {
        set_error("operator policy 0x%x: column extents x %d, y %d, z %d must be powers of two dividing LX=%d, "
                  "LY=%d, LZ=%d", policy, cx, cy, zb, c->lat.LX, c->lat.LY, c->lat.LZ);
        return CVCHIP_ERR_ARG;
      }
    } else if ((yb && c->lat.LY % yb) || (bt && c->lat.T % bt) || (bx && c->lat.LX % bx)) {
      set_error("operator policy 0x%x: block extents must divide the lattice (T=%d LX=%d LY=%d)", policy,
                c->lat.T, c->lat.LX, c->lat.LY);
      return CVCHIP_ERR_ARG;
    }
    c->dslash_policy = policy;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_set_gauge(cvchip_ctx *c, const double *gauge) {
  return cvchip::guarded("cvchip_set_gauge", [&]() -> int {
    if (!c || !gauge) {
      set_error("NULL argument to cvchip_set_gauge");
      return CVCHIP_ERR_NULL;
    }
    const int64_t V = c->lat.V;
    if (!c->gauge64) CVC_HIP(hipMalloc(&c->gauge64, sizeof(double2) * 36 * V));
    if (!c->gauge64r12) CVC_HIP(hipMalloc(&c->gauge64r12, sizeof(double2) * 24 * V));
    if (!c->gauge32r12) CVC_HIP(hipMalloc(&c->gauge32r12, sizeof(float2) * 24 * V));
    CVC_TRY(ensure_stage(c, sizeof(double) * 72 * V));
    CVC_TRY(copy_to_device(c, c->stage, gauge, sizeof(double) * 72 * V));
    const int64_t n = 36 * V;
    hipLaunchKernelGGL(k_gauge_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->gauge64, c->gauge64r12, c->gauge32r12, (const double *)c->stage, V);
    CVC_HIP(hipGetLastError());
    CVC_HIP(hipStreamSynchronize(c->stream));
    c->gauge_set = 1;
    return CVCHIP_OK;
  });
}

namespace cvchip {

int upload_aos(cvchip_ctx *c, Field &f, const double *host, int rhs0, int n) {
  if (rhs0 < 0 || n < 0 || rhs0 + n > f.nrhs) {
    set_error("rhs range %d..%d outside field of %d", rhs0, rhs0 + n, f.nrhs);
    return CVCHIP_ERR_ARG;
  }
  if (n == 0) return 0;
  const int64_t V = f.V;
  const size_t bytes = sizeof(double) * 24 * V * n;
  CVC_TRY(ensure_stage(c, bytes));
  CVC_TRY(copy_to_device(c, c->stage, host, bytes));
  const int64_t tot = (int64_t)n * V * 12;
  dim3 g((unsigned)((tot + 255) / 256)), b(256);
  if (f.prec == CVCHIP_PREC_F64)
    hipLaunchKernelGGL(k_aos_to_field<double>, g, b, 0, c->stream, (double2 *)f.data,
                       (const double *)c->stage, V, f.nrhs, rhs0, n);
  else if (f.prec == CVCHIP_PREC_F32)
    hipLaunchKernelGGL(k_aos_to_field<float>, g, b, 0, c->stream, (float2 *)f.data,
                       (const double *)c->stage, V, f.nrhs, rhs0, n);
  else {
    set_error("host upload into FP16 fields is not supported; upload FP32 and convert");
    return CVCHIP_ERR_ARG;
  }
  CVC_HIP(hipGetLastError());
  CVC_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int download_aos(cvchip_ctx *c, const Field &f, double *host, int rhs0, int n) {
  if (rhs0 < 0 || n < 0 || rhs0 + n > f.nrhs) {
    set_error("rhs range %d..%d outside field of %d", rhs0, rhs0 + n, f.nrhs);
    return CVCHIP_ERR_ARG;
  }
  if (n == 0) return 0;
  const int64_t V = f.V;
  const size_t bytes = sizeof(double) * 24 * V * n;
  CVC_TRY(ensure_stage(c, bytes));
  const int64_t tot = (int64_t)n * V * 12;
  dim3 g((unsigned)((tot + 255) / 256)), b(256);
  if (f.prec == CVCHIP_PREC_F64)
    hipLaunchKernelGGL(k_field_to_aos<double>, g, b, 0, c->stream, (double *)c->stage,
                       (const double2 *)f.data, V, f.nrhs, rhs0, n);
  else if (f.prec == CVCHIP_PREC_F32)
    hipLaunchKernelGGL(k_field_to_aos<float>, g, b, 0, c->stream, (double *)c->stage,
                       (const float2 *)f.data, V, f.nrhs, rhs0, n);
  else {
    set_error("host download from FP16 fields is not supported");
    return CVCHIP_ERR_ARG;
  }
  CVC_HIP(hipGetLastError());
  return copy_to_host(c, host, c->stage, bytes);
}

}  // namespace cvchip

extern "C" int cvchip_field_create(cvchip_ctx *c, int nrhs, int prec, cvchip_field **out) {
  return cvchip::guarded("cvchip_field_create", [&]() -> int {
    if (!c || !out) return CVCHIP_ERR_NULL;
    cvchip_field *f = new cvchip_field();
    int rc = field_alloc(c, f->f, nrhs, prec, 0, false);
    if (rc) {
      field_free(f->f);
      delete f;
      return rc;
    }
    *out = f;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_destroy(cvchip_field *f) {
  return cvchip::guarded("cvchip_field_destroy", [&]() -> int {
    if (!f) return CVCHIP_OK;
    field_free(f->f);
    delete f;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_upload(cvchip_ctx *c, cvchip_field *f, const double *host, int rhs0, int n) {
  return cvchip::guarded("cvchip_field_upload", [&]() -> int {
    if (!c || !f || !host) return CVCHIP_ERR_NULL;
    return upload_aos(c, f->f, host, rhs0, n);
  });
}

extern "C" int cvchip_field_download(cvchip_ctx *c, const cvchip_field *f, double *host, int rhs0, int n) {
  return cvchip::guarded("cvchip_field_download", [&]() -> int {
    if (!c || !f || !host) return CVCHIP_ERR_NULL;
    return download_aos(c, f->f, host, rhs0, n);
  });
}
