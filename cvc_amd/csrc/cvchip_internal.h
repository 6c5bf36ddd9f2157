// Internal definitions of libcvchip: context, device field layout, complex
// and SU(3) helpers, the Dirac-matrix structure used by the kernels.
//
// Device layout (all precisions):
//   spinor field set of N right-hand sides:  cplx[12][V][N]
//     element (component c = 3*spin + colour, site s, rhs r) at (c*V + s)*N + r,
//     so that the flat thread index i = s*N + r walks memory contiguously.
//   gauge (recon-18):                        cplx[4][9][V]  (link mu, entry 3*i+j, site)
//   gauge (recon-12):                        cplx[4][6][V]  (first two rows; the third
//                                                             row is rebuilt in registers)
// Sites are lexicographic with z fastest: s = ((t*LX + x)*LY + y)*LZ + z, as in
// the reference (cvc_geometry.c:17-26).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "../../include/cvchip.h"
#include "entry_guard.h"

namespace cvchip {

// ------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define CVC_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) return ::cvchip::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define CVC_TRY(call)                                                        \
  do {                                                                       \
    int _rc = (call);                                                        \
    if (_rc != 0) return _rc;                                                \
  } while (0)

// ------------------------------------------------------------- complex
template <typename R> struct cx;
template <> struct cx<double> { using t = double2; };
template <> struct cx<float> { using t = float2; };
template <typename R> using cplx = typename cx<R>::t;

template <typename R> __host__ __device__ __forceinline__ cplx<R> mk(R a, R b) {
  cplx<R> c; c.x = a; c.y = b; return c;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { a.x += b.x; a.y += b.y; return a; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { a.x -= b.x; a.y -= b.y; return a; }
// Products and sums of products are written as explicit FMAs wherever rounding
// matters: left to the compiler, the contraction of a b - c d may pick either
// product to fuse, differently in different kernels, and the traversal and
// staging variants of one operator would no longer give the same bits.
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  C c; c.x = fmar(a.x, b.x, -(a.y * b.y)); c.y = fmar(a.x, b.y, a.y * b.x); return c;
}
// conj(a) * b
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  C c; c.x = fmar(a.x, b.x, a.y * b.y); c.y = fmar(a.x, b.y, -(a.y * b.x)); return c;
}
template <typename C, typename R> __device__ __forceinline__ C cscale(C a, R s) { a.x *= s; a.y *= s; return a; }
// a + b * c as two chained FMAs per real part (written as a + (b c - b c)
// it contracts to mul, fma and add: three instructions instead of two)
template <typename C> __device__ __forceinline__ C cfma(C a, C b, C c) {
  a.x = fmar(b.x, c.x, fmar(-b.y, c.y, a.x));
  a.y = fmar(b.x, c.y, fmar(b.y, c.x, a.y));
  return a;
}
// a + conj(b) * c
template <typename C> __device__ __forceinline__ C cfmac(C a, C b, C c) {
  a.x = fmar(b.x, c.x, fmar(b.y, c.y, a.x));
  a.y = fmar(b.x, c.y, fmar(-b.y, c.x, a.y));
  return a;
}
// multiplication by a unit u in {+1, -1, +i, -i} (code 0,1,2,3), resolved at compile time
template <int U, typename C> __device__ __forceinline__ C cunit(C a) {
  if constexpr (U == 0) return a;
  else if constexpr (U == 1) { a.x = -a.x; a.y = -a.y; return a; }
  else if constexpr (U == 2) { C c; c.x = -a.y; c.y = a.x; return c; }
  else { C c; c.x = a.y; c.y = -a.x; return c; }
}

// ------------------------------------------------------- Dirac matrices
// gamma_mu of the reference basis (cvc_utils.c:709-745): one unit entry per
// row.  GCOL[mu][a] is the column of the non-zero in row a, GVAL[mu][a] its
// value as a unit code (0:+1 1:-1 2:+i 3:-i).  gamma5 = diag(1,1,-1,-1).
constexpr int GCOL[4][4] = {{2, 3, 0, 1}, {3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}};
constexpr int GVAL[4][4] = {{1, 1, 1, 1}, {3, 3, 2, 2}, {1, 0, 0, 1}, {3, 2, 2, 3}};
constexpr int uneg(int u) { return u ^ 1; }   // -u

// --------------------------------------------------------------- context
struct Lattice {
  int T, LX, LY, LZ;
  int64_t V;
};

struct Phases { double re[4], im[4]; };

struct Field {
  int nrhs = 0;
  int prec = CVCHIP_PREC_F64;
  int64_t V = 0;
  void *data = nullptr;      // cplx<R>[12][V][N]
  float *norm = nullptr;     // FP16 storage: per (site, rhs) scale
  size_t bytes = 0;
  cvchip_ctx *pool = nullptr;  // context whose idle-buffer pool takes the memory back (temporaries)
};

struct Scratch {
  void *ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace cvchip

struct cvchip_field {
  cvchip::Field f;
};

struct cvchip_ctx {
  cvchip::Lattice lat;
  double kappa, mu;
  double bc[4];
  cvchip::Phases ph;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // gauge copies
  double2 *gauge64 = nullptr;    // recon-18, FP64
  double2 *gauge64r12 = nullptr; // recon-12, FP64
  float2 *gauge32r12 = nullptr;  // recon-12, FP32
  int gauge_set = 0;
  int link_recon = 18;           // 18 or 12 for the FP64 Dslash
  // multi-rhs operator staging: bit 0 links in LDS, bit 1 own + z-neighbour spinors in LDS
  int dslash_policy = 5;   // links staged in LDS, streaming stores, automatic traversal
  unsigned solve_flags = 0;   // OR-ed into every solve
  // while a solver runs: per-rhs flags (op_mask[r * op_mask_stride] != 0:
  // right-hand side r has converged); the operator skips those rhs
  const int *op_mask = nullptr;
  int op_mask_stride = 0;
  // staging for host <-> device transfers (AoS reference layout)
  double *stage = nullptr;
  size_t stage_bytes = 0;
  // pinned host buffers (used in turn) and their events for device-to-host copies
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  // reduction scratch
  cvchip::Scratch red;
  cvchip::Scratch red2;   // first-stage sums of the two-stage partial reduction
  // idle device buffers of freed temporaries (solver, propagator and
  // contraction workspaces), reused by exact size: repeated calls allocate
  // nothing; at most kPoolIdleMax bytes stay idle
  std::vector<std::pair<void *, size_t>> idle;
  size_t idle_bytes = 0;
};

namespace cvchip {

int ensure_stage(cvchip_ctx *c, size_t bytes);
// device -> host copy on the context's stream (after the work enqueued before
// it), complete on return; large copies stream through pinned buffers
int copy_to_host(cvchip_ctx *c, void *host, const void *dev, size_t bytes);
// host -> device copy enqueued on the context's stream; the caller
// synchronizes the stream before the host buffer may change
int copy_to_device(cvchip_ctx *c, void *dev, const void *host, size_t bytes);
// device address of the allocation-free two-point entry point's scratch on
// `device` (the current device), taken once per device
int twopoint_scratch_init(int device);
// row a of the gamma matrix id (0-3 gamma_mu, 4 identity, 5 gamma5, 6-9
// gamma_mu gamma5, 10-15 sigma-type products): entry re[a] + i im[a] at col[a]
void gamma_matrix_rows(int id, int col[4], int re[4], int im[4]);
int ensure_red(cvchip_ctx *c, size_t bytes);
int ensure_red2(cvchip_ctx *c, size_t bytes);
// sites: 0 for the whole lattice, or the half lattice V/2 of the even/odd solver
// Temporaries of the library (pooled); user sets of cvchip_field_create pass
// pooled = false, so they own their memory outright.
int field_alloc(cvchip_ctx *c, Field &f, int nrhs, int prec, int64_t sites = 0, bool pooled = true);
void field_free(Field &f);
// device memory from the context's idle pool (or hipMalloc; on failure idle
// buffers are released, oldest first, until the allocation fits, also for the
// sets that own their memory); back into the pool
constexpr size_t kPoolIdleMax = (size_t)160 << 30;   // of 288 GB: the workspaces of a 48-rhs solve stay
int pool_get(cvchip_ctx *c, void **p, size_t bytes);
void pool_put(cvchip_ctx *c, void *p, size_t bytes);
void pool_release(cvchip_ctx *c);

// host AoS [V][24] (nrhs contiguous fields) <-> device field rhs range
int upload_aos(cvchip_ctx *c, Field &f, const double *host, int rhs0, int n);
int download_aos(cvchip_ctx *c, const Field &f, double *host, int rhs0, int n);

// Dslash launcher: out = D(mutm) in (gamma5_out: out = gamma5 D in).
// Optional fused reductions, written as per-block partials into c->red:
//   DOT_NONE, DOT_TS (<out,in> complex and <out,out>; <out,r2> when r2 is given),
//   DOT_R2 (<r2,out> complex), DOT_NN (<out,out>)
enum DotMode { DOT_NONE = 0, DOT_TS = 1, DOT_R2 = 2, DOT_NN = 3 };
int dslash(cvchip_ctx *c, Field &out, const Field &in, double mutm, int gamma5_out,
           DotMode dm = DOT_NONE, const Field *r2 = nullptr, int *nblocks_out = nullptr);

// Even/odd pieces on half-lattice fields (output parity par):
//   dmode 0: A own + hopc H in     1: A^-1 (hopc H in)
//   dmode 2: own + hopc H in       3: A^-1 (own + hopc H in)
// with A = 1/(2 kappa) + i mutm gamma5, H the hopping sum (D = A - H/2).
int dslash_eo(cvchip_ctx *c, Field &out, const Field &in, const Field *own, double mutm, int par, int dmode,
              double hopc, int gamma5_out, DotMode dm = DOT_NONE, const Field *r2 = nullptr,
              int *nblocks_out = nullptr);

// number of blocks (= reduction partial slots) of an operator launch on N rhs
// (the larger of the whole-lattice and the half-lattice traversal)
unsigned dslash_blocks(const cvchip_ctx *c, int N);

// block-size convention shared by all per-rhs reductions
inline int block_for(int nrhs) {
  int s = 256 / nrhs;
  if (s < 1) s = 1;
  return s * nrhs;
}

}  // namespace cvchip
