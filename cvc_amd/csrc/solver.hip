// BiCGStab for D(mu) x = b on N right-hand sides at once (reference
// invert_Qtm.c:86-270, invert_Qtm_her.c:276-478), device resident.
//
// Per iteration (r0rn, alpha, omega, beta per rhs):
//   alpha = r0rn / <r2,p2>                      (scalar kernel)
//   s     = r1 - alpha p2                       (vector kernel)
//   t     = D s,   <t,s>, <t,t>                 (operator with fused dot epilogue)
//   omega = <t,s> / <t,t>                       (scalar kernel)
//   r1    = s - omega t,  <r1,r1>, <r2,r1>      (vector kernel with dots)
//   test  |r1|^2 <= sp |b|^2; beta = (<r2,r1>/r0rn)(alpha/omega); r0rn = <r2,r1>
//   x    += omega s + alpha p ;  p = r1 + beta (p - omega p2)   (one pass)
//   p2    = D p,  <r2,p2>                       (operator with fused dot epilogue)
// Dot products are per-thread sums of a site's 12 components, block trees per
// rhs and a fixed-order tree over blocks: results are reproducible run to run.
// A rhs that has converged is frozen (its vectors are no longer written).
// The reference returns the iterate from before the last update
// (invert_Qtm.c:199-211); CVCHIP_SOLVE_LAGGED reproduces that, the default
// applies the update so that x matches the tested residual.
//
// The mixed-precision solver runs the same recurrence on FP32 (or FP16 +
// norm) vectors and, for each rhs, replaces the iterated residual by the true
// FP64 residual b - D x whenever |r| has dropped by the factor delta since the
// previous such update (reliable updates), accumulating the low-precision
// solution into the FP64 x at each update.  Convergence is always decided on
// the FP64 true residual.
#include <type_traits>
#include <stdio.h>
#include <algorithm>
#include <chrono>
#include <vector>

#include "spinor_io.h"

namespace cvchip {

struct RhsState {
  double alpha[2], omega[2], beta[2], r0rn[2], r0rn_prev[2];
  double normb, norm, tol2;
  double rmax;        // mixed: |r|^2 at the last reliable update
  double true_res2;
  double best;        // mixed: the smallest true |r|^2 at a reliable update so far
  int done, final_upd, iter, upd, last_upd;
  int stall;          // mixed: own reliable updates in a row that did not halve best
  int stalled;        // mixed: stopped at the floor of the true residual, not converged
};

// ------------------------------------------------------------ vector kernels

// Streaming (non-temporal) stores for the vector kernels' outputs when a
// vector set is 256 MB or more: it does not stay in the caches between kernels
// anyway, and the stores then leave them to the inputs (measured at 64x32^3,
// 12 rhs: FP64 solve 4.01 -> 3.67 s)
template <typename IO>
__device__ __forceinline__ void vstore(const IO &io, const typename IO::C v[12], uint32_t i, int nt) {
  if (nt) io.store_nt(v, i);
  else io.store(v, i);
}

template <typename IO> static IO mkio(const Field &f);
template <> IOPlain<double> mkio<IOPlain<double>>(const Field &f) {
  return IOPlain<double>{(double2 *)f.data, (uint32_t)(f.V * f.nrhs)};
}
template <> IOPlain<float> mkio<IOPlain<float>>(const Field &f) {
  return IOPlain<float>{(float2 *)f.data, (uint32_t)(f.V * f.nrhs)};
}
template <> IOHalf mkio<IOHalf>(const Field &f) {
  return IOHalf{(__half2 *)f.data, f.norm, (uint32_t)(f.V * f.nrhs)};
}

#define CVC_ELEMWISE_PROLOGUE                                  \
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;    \
  const int rhs = (int)(threadIdx.x % (unsigned)N);            \
  const bool act = i < VN;                                     \
  (void)rhs;

// s = r1 - alpha p2
template <typename IO>
__global__ void __launch_bounds__(256) k_s_update(IO s, IO r1, IO p2, const RhsState *__restrict__ st,
                                                  uint32_t VN, int N, int nt) {
  CVC_ELEMWISE_PROLOGUE
  if (!act || st[rhs].done) return;
  using R = typename IO::real;
  using C = cplx<R>;
  const C a = mk<R>((R)st[rhs].alpha[0], (R)st[rhs].alpha[1]);
  C vr[12], vp[12];
  r1.load(vr, i);
  p2.load(vp, i);
#pragma unroll
  for (int c = 0; c < 12; c++) vr[c] = csub(vr[c], cmul(a, vp[c]));
  vstore(s, vr, i, nt);
}

// r1 = s - omega t ; partial dots <r1,r1>, <r2,r1>  (r2 stored as IOR)
template <typename IO, typename IOR = IO>
__global__ void __launch_bounds__(256) k_r1_update(IO r1, IO s, IO t, IOR r2, const RhsState *__restrict__ st,
                                                   uint32_t VN, int N, double *__restrict__ part, int nt) {
  CVC_ELEMWISE_PROLOGUE
  using R = typename IO::real;
  using C = cplx<R>;
  double dv[3] = {0., 0., 0.};
  if (act && !st[rhs].done) {
    const C w = mk<R>((R)st[rhs].omega[0], (R)st[rhs].omega[1]);
    C vs[12], vt[12], v2[12];
    s.load(vs, i);
    t.load(vt, i);
    r2.load(v2, i);
    // products in FP64: squares of a small-scale rhs underflow FP32
    double n = 0, a = 0, b = 0;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      C o = csub(vs[c], cmul(w, vt[c]));
      vs[c] = o;
      const double ox = o.x, oy = o.y, qx = v2[c].x, qy = v2[c].y;
      n += ox * ox + oy * oy;
      a += qx * ox + qy * oy;
      b += qx * oy - qy * ox;
    }
    vstore(r1, vs, i, nt);
    dv[0] = n; dv[1] = a; dv[2] = b;
  }
  block_reduce_store<3>(dv, part, N);
}

// x += omega s + alpha p  (while running, or once more on convergence when
// not lagged); p = r1 + beta (p - omega p2) (while running)
template <typename IOX, typename IO>
__global__ void __launch_bounds__(256) k_xp_update(IOX x, IO p, IO s, IO r1, IO p2,
                                                   const RhsState *__restrict__ st, uint32_t VN, int N, int nt) {
  CVC_ELEMWISE_PROLOGUE
  if (!act) return;
  const RhsState &S = st[rhs];
  const bool upx = !S.done;               // includes the final (non-lagged) update
  const bool upp = !S.done && !S.final_upd && !S.upd;
  if (!upx) return;
  using R = typename IO::real;
  using C = cplx<R>;
  using RX = typename IOX::real;
  using CX = cplx<RX>;
  const C a = mk<R>((R)S.alpha[0], (R)S.alpha[1]);
  const C w = mk<R>((R)S.omega[0], (R)S.omega[1]);
  C vp[12], vs[12], vr[12], v2[12];
  CX vx[12];
  p.load(vp, i);
  s.load(vs, i);
  x.load(vx, i);
  // FP16 fields are a quarter of the FP64 bytes per lane: the loads of all
  // five fields go out together (the FP16-inner kernel 1140 -> 1047 us at
  // 32^3x64 even/odd; the FP32 one was 2 % slower, the FP64 one would halve
  // its waves)
  constexpr bool kAll = IO::K == 4;
  if (kAll && upp) {
    r1.load(vr, i);
    p2.load(v2, i);
  }
#pragma unroll
  for (int c = 0; c < 12; c++) {
    C d = cadd(cmul(w, vs[c]), cmul(a, vp[c]));
    vx[c].x += (RX)d.x;
    vx[c].y += (RX)d.y;
  }
  vstore(x, vx, i, nt);
  if (upp) {
    const C be = mk<R>((R)S.beta[0], (R)S.beta[1]);
    if (!kAll) {
      r1.load(vr, i);
      p2.load(v2, i);
    }
#pragma unroll
    for (int c = 0; c < 12; c++) vp[c] = cadd(vr[c], cmul(be, csub(vp[c], cmul(w, v2[c]))));
    vstore(p, vp, i, nt);
  }
}

// r = b - t  (t = D x), optionally copied to r2 (stored as IOR) and p; partial dots <r,r>, <b,b>
template <typename IOB, typename IO, typename IOR = IO>
__global__ void __launch_bounds__(256) k_residual(IO r, IOR r2, IO p, IOB b, IOB t, int copies,
                                                  const RhsState *__restrict__ st, int only_upd,
                                                  uint32_t VN, int N, double *__restrict__ part) {
  CVC_ELEMWISE_PROLOGUE
  using RB = typename IOB::real;
  using CB = cplx<RB>;
  using R = typename IO::real;
  using C = cplx<R>;
  double dv[2] = {0., 0.};
  if (act && (!only_upd || st[rhs].upd)) {
    CB vb[12], vt[12];
    b.load(vb, i);
    t.load(vt, i);
    C o[12];
    double n = 0., nb = 0.;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      CB d = csub(vb[c], vt[c]);
      n += (double)d.x * d.x + (double)d.y * d.y;
      nb += (double)vb[c].x * vb[c].x + (double)vb[c].y * vb[c].y;
      o[c] = mk<R>((R)d.x, (R)d.y);
    }
    r.store(o, i);
    if (copies) {
      r2.store(o, i);
      p.store(o, i);
    }
    dv[0] = n; dv[1] = nb;
  }
  block_reduce_store<2>(dv, part, N);
}

// dst = src (precision conversion), optionally only rhs flagged in st[].upd
template <typename IOD, typename IOS>
__global__ void __launch_bounds__(256) k_copy(IOD d, IOS s, const RhsState *__restrict__ st, int only_upd,
                                              uint32_t VN, int N) {
  CVC_ELEMWISE_PROLOGUE
  if (!act || (only_upd && !st[rhs].upd)) return;
  using RS = typename IOS::real;
  using RD = typename IOD::real;
  cplx<RS> v[12];
  cplx<RD> o[12];
  s.load(v, i);
#pragma unroll
  for (int c = 0; c < 12; c++) o[c] = mk<RD>((RD)v[c].x, (RD)v[c].y);
  d.store(o, i);
}

// x64 += xl (low precision accumulator), xl = 0, for rhs flagged in st[].upd
template <typename IOX, typename IOL>
__global__ void __launch_bounds__(256) k_accum_zero(IOX x, IOL xl, const RhsState *__restrict__ st,
                                                    uint32_t VN, int N) {
  CVC_ELEMWISE_PROLOGUE
  if (!act || !st[rhs].upd) return;
  cplx<double> vx[12];
  cplx<typename IOL::real> vl[12];
  x.load(vx, i);
  xl.load(vl, i);
#pragma unroll
  for (int c = 0; c < 12; c++) {
    vx[c].x += (double)vl[c].x;
    vx[c].y += (double)vl[c].y;
    vl[c] = mk<typename IOL::real>(0, 0);
  }
  x.store(vx, i);
  xl.store(vl, i);
}

// <a,b> complex and <a,a>: partials for Q=3
template <typename IOA, typename IOB = IOA>
__global__ void __launch_bounds__(256) k_dot3(IOA a, IOB b, uint32_t VN, int N, double *__restrict__ part) {
  CVC_ELEMWISE_PROLOGUE
  double dv[3] = {0., 0., 0.};
  if (act) {
    cplx<typename IOA::real> va[12];
    cplx<typename IOB::real> vb[12];
    a.load(va, i);
    b.load(vb, i);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      dv[0] += (double)va[c].x * vb[c].x + (double)va[c].y * vb[c].y;
      dv[1] += (double)va[c].x * vb[c].y - (double)va[c].y * vb[c].x;
      dv[2] += (double)va[c].x * va[c].x + (double)va[c].y * va[c].y;
    }
  }
  block_reduce_store<3>(dv, part, N);
}

// ------------------------------------------------------------ scalar kernels

// res = reduced dots of the preceding kernel (layout [q][N]).
__global__ void k_sc_init(RhsState *st, const double *res, int N, double sp, int mixed) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  S.norm = res[r];
  S.normb = res[N + r];
  S.tol2 = sp * S.normb;
  S.r0rn[0] = S.norm;
  S.r0rn[1] = 0.;
  S.rmax = S.norm;
  S.best = S.norm;
  S.stall = 0;
  S.stalled = 0;
  S.final_upd = 0;
  S.upd = 0;
  S.last_upd = 0;
  S.iter = -1;
  S.done = (S.norm <= S.tol2) ? 1 : 0;       // invert_Qtm.c:142-147: return 0
  if (S.done) S.iter = 0;
  S.alpha[0] = S.alpha[1] = S.omega[0] = S.omega[1] = S.beta[0] = S.beta[1] = 0.;
  (void)mixed;
}

__device__ __forceinline__ void cdiv(double o[2], const double a[2], const double b[2]) {
  const double d = b[0] * b[0] + b[1] * b[1];
  const double re = (a[0] * b[0] + a[1] * b[1]) / d;
  const double im = (a[1] * b[0] - a[0] * b[1]) / d;
  o[0] = re;
  o[1] = im;
}

// alpha = r0rn / <r2,p2>
__global__ void k_sc_alpha(RhsState *st, const double *res, int N) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  if (S.final_upd) {
    S.done = 1;
    S.final_upd = 0;
  }
  if (S.done) return;
  const double w[2] = {res[r], res[N + r]};
  cdiv(S.alpha, S.r0rn, w);
}

// omega = <t,s>/<t,t>
__global__ void k_sc_omega(RhsState *st, const double *res, int N) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  if (S.done) return;
  const double inv = 1. / res[2 * N + r];
  S.omega[0] = res[r] * inv;
  S.omega[1] = res[N + r] * inv;
}

constexpr int kMaxGap = 48;

// after r1: convergence test (or reliable-update trigger), beta, r0rn
__global__ void k_sc_beta(RhsState *st, const double *res, int N, int it, int lagged, int mixed,
                          double delta2) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  if (S.done) return;
  S.norm = res[r];
  const double w[2] = {res[N + r], res[2 * N + r]};
  if (!mixed) {
    if (S.norm <= S.tol2) {
      S.iter = it;
      if (lagged) S.done = 1;
      else S.final_upd = 1;
      if (S.done) return;
    }
  } else {
    // reliable update when the iterated residual fell by delta since the last
    // update, reached the tolerance (then decided in FP64), or has not done so
    // for kMaxGap iterations (a low-precision recurrence can stall above delta)
    // (the tolerance trigger waits for a factor 10 margin in |r|: the iterated
    // low-precision residual runs slightly ahead of the true one)
    S.upd = (S.norm <= delta2 * S.rmax || S.norm <= 0.01 * S.tol2 || it - S.last_upd >= kMaxGap) ? 1 : 0;
  }
  double w2[2], w3[2];
  cdiv(w2, w, S.r0rn);
  cdiv(w3, S.alpha, S.omega);
  S.beta[0] = w2[0] * w3[0] - w2[1] * w3[1];
  S.beta[1] = w2[0] * w3[1] + w2[1] * w3[0];
  S.r0rn_prev[0] = S.r0rn[0];
  S.r0rn_prev[1] = S.r0rn[1];
  S.r0rn[0] = w[0];
  S.r0rn[1] = w[1];
}

// A column whose own reliable-update trigger fired kMaxStall times in a row
// without the true |r|^2 falling below half its best value has reached the
// rounding floor of the FP64 residual b - D x (a tolerance below that floor is
// never met; at the floor the true residual only fluctuates): it stops, not
// converged, with x as folded in at that update.  Updates forced on a column by
// another column's trigger do not count.
constexpr int kMaxStall = 6;

// mixed: after the FP64 residual recompute (res: <r,r>, unused) decide convergence
__global__ void k_sc_reliable(RhsState *st, const double *res, int N, int it) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  if (!S.upd) return;
  S.norm = res[r];
  S.rmax = S.norm;
  S.last_upd = it;
  if (S.norm <= S.tol2) {
    S.done = 1;
    S.iter = it;
    return;
  }
  if (S.norm < 0.5 * S.best) {
    S.best = S.norm;
    S.stall = 0;
  } else if (S.upd == 1 && ++S.stall >= kMaxStall) {
    S.done = 1;
    S.stalled = 1;
    S.iter = it;
  }
}

// mixed: after a reliable update, beta = (<r2,r>/r0rn_prev)(alpha/omega) and
// r0rn = <r2,r> with the replaced residual r (res: <r2,r> complex)
__global__ void k_sc_r0rn(RhsState *st, const double *res, int N) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  if (!S.upd || S.done) return;
  const double w[2] = {res[r], res[N + r]};
  double w2[2], w3[2];
  cdiv(w2, w, S.r0rn_prev);
  cdiv(w3, S.alpha, S.omega);
  S.beta[0] = w2[0] * w3[0] - w2[1] * w3[1];
  S.beta[1] = w2[0] * w3[1] + w2[1] * w3[0];
  S.r0rn[0] = w[0];
  S.r0rn[1] = w[1];
}

// r0rn = <r2, r> (res: complex) for the running rhs, at the start of a solve
// whose shadow residual r2 is stored in lower precision than r
__global__ void k_sc_set_r0rn(RhsState *st, const double *res, int N) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  if (S.done) return;
  S.r0rn[0] = res[r];
  S.r0rn[1] = res[N + r];
}

// mixed with restart (FP16 storage): after a reliable update the recurrence
// starts afresh from the replaced residual: r2 = p = r, r0rn = |r|^2
__global__ void k_sc_restart(RhsState *st, int N) {
  const int r = threadIdx.x;
  if (r >= N) return;
  RhsState &S = st[r];
  if (S.upd && !S.done) {
    S.r0rn[0] = S.norm;
    S.r0rn[1] = 0.;
  }
  S.upd = 0;
}

// the FP64 residual recompute runs over all rhs at once: give every active rhs
// its reliable update in the same pass.  An rhs whose own trigger did not fire
// (upd 0) has already advanced its search direction p in k_xp_update; it is
// marked 2 so that k_p_update leaves that p alone (only its r and r0rn are
// replaced).
__global__ void k_sc_upd_all(RhsState *st, int N) {
  const int r = threadIdx.x;
  if (r < N && !st[r].done && !st[r].upd) st[r].upd = 2;
}

__global__ void k_sc_clear_upd(RhsState *st, int N) {
  const int r = threadIdx.x;
  if (r < N) st[r].upd = 0;
}

// mixed, rhs whose own trigger fired (upd 1): p = r + beta (p - omega p2)
// with the replaced residual r (k_xp_update skipped their p)
template <typename IO>
__global__ void __launch_bounds__(256) k_p_update(IO p, IO r1, IO p2, const RhsState *__restrict__ st,
                                                  uint32_t VN, int N) {
  CVC_ELEMWISE_PROLOGUE
  if (!act || st[rhs].upd != 1 || st[rhs].done) return;
  using R = typename IO::real;
  using C = cplx<R>;
  const C w = mk<R>((R)st[rhs].omega[0], (R)st[rhs].omega[1]);
  const C be = mk<R>((R)st[rhs].beta[0], (R)st[rhs].beta[1]);
  C vp[12], vr[12], v2[12];
  p.load(vp, i);
  r1.load(vr, i);
  p2.load(v2, i);
#pragma unroll
  for (int c = 0; c < 12; c++) vp[c] = cadd(vr[c], cmul(be, csub(vp[c], cmul(w, v2[c]))));
  p.store(vp, i);
}

// ------------------------------------------------------------------ driver

struct Solve {
  cvchip_ctx *c;
  int N;
  uint32_t VN;
  int B;
  unsigned nblk;
  RhsState *st = nullptr;     // device
  double *res = nullptr;      // device [4][N]
  RhsState *hst = nullptr;    // pinned host mirror
  int nt = 0;                 // streaming stores in the vector kernels
};

// The operator applications of a running solve skip converged right-hand
// sides (their results are not read again); cleared before the applications
// that need every rhs (true residual, solution assembly) and on exit.
struct OpMask {
  cvchip_ctx *c;
  OpMask(cvchip_ctx *cc, const RhsState *st) : c(cc) {
    c->op_mask = &st->done;
    c->op_mask_stride = (int)(sizeof(RhsState) / sizeof(int));
  }
  void clear() {
    c->op_mask = nullptr;
    c->op_mask_stride = 0;
  }
  ~OpMask() { clear(); }
};

// V: sites of the solver's vectors (the lattice, or one parity of it)
static int solve_setup(cvchip_ctx *c, Solve &S, int N, int64_t V = 0) {
  S.c = c;
  S.N = N;
  S.VN = (uint32_t)((V > 0 ? V : c->lat.V) * N);
  S.B = block_for(N);
  S.nblk = (S.VN + S.B - 1) / S.B;
  // an FP64 set of this size is 192 B per site and rhs
  S.nt = (size_t)S.VN * 192 >= ((size_t)256 << 20) ? 1 : 0;
  // room for the operator's partials too (its traversal has its own block count)
  const size_t nb_max = std::max<size_t>(S.nblk, dslash_blocks(c, N));
  CVC_TRY(ensure_red(c, sizeof(double) * (nb_max * 3 * N + 64 * 256)));
  CVC_HIP(hipMalloc(&S.st, sizeof(RhsState) * N));
  CVC_HIP(hipMalloc(&S.res, sizeof(double) * 4 * N));
  CVC_HIP(hipHostMalloc(&S.hst, sizeof(RhsState) * N, hipHostMallocDefault));
  CVC_HIP(hipMemsetAsync(S.st, 0, sizeof(RhsState) * N, c->stream));
  return 0;
}

static void solve_free(Solve &S) {
  if (S.st) (void)hipFree(S.st);
  if (S.res) (void)hipFree(S.res);
  if (S.hst) (void)hipHostFree(S.hst);
}

static int reduce(Solve &S, int nblk, int Q) {
  return reduce_partials(S.c, (const double *)S.c->red.ptr, nblk, Q, S.N, S.res);
}

static int pull_state(Solve &S) {
  CVC_HIP(hipMemcpyAsync(S.hst, S.st, sizeof(RhsState) * S.N, hipMemcpyDeviceToHost, S.c->stream));
  CVC_HIP(hipStreamSynchronize(S.c->stream));
  return 0;
}

#define LAUNCH_EW(kern, ...)                                                                     \
  do {                                                                                           \
    hipLaunchKernelGGL(kern, dim3(S.nblk), dim3(S.B), 0, c->stream, __VA_ARGS__);                \
    CVC_HIP(hipGetLastError());                                                                  \
  } while (0)

// true residual |D x - b|^2 / |b|^2 per rhs into stats / state
static int true_residual(Solve &S, Field &x, const Field &b, Field &tmp, double mutm, double *maxres,
                         std::vector<double> *per_rhs) {
  cvchip_ctx *c = S.c;
  CVC_TRY(dslash(c, tmp, x, mutm, 0));
  IOPlain<double> ib = mkio<IOPlain<double>>(b), it = mkio<IOPlain<double>>(tmp);
  // tmp <- b - D x (sign irrelevant for the norm)
  LAUNCH_EW((k_residual<IOPlain<double>, IOPlain<double>>), it, it, it, ib, it, 0, S.st, 0, S.VN, S.N,
            (double *)c->red.ptr);
  CVC_TRY(reduce(S, (int)S.nblk, 2));
  std::vector<double> h(2 * S.N);
  CVC_HIP(hipMemcpyAsync(h.data(), S.res, sizeof(double) * 2 * S.N, hipMemcpyDeviceToHost, c->stream));
  CVC_HIP(hipStreamSynchronize(c->stream));
  double m = 0.;
  if (per_rhs) per_rhs->resize(S.N);
  for (int r = 0; r < S.N; r++) {
    double v = h[S.N + r] > 0 ? h[r] / h[S.N + r] : 0.;
    if (v > m) m = v;
    if (per_rhs) (*per_rhs)[r] = v;
  }
  *maxres = m;
  return 0;
}

__global__ void k_zero_doubles(double *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.;
}

// The operator of the recurrence: D(mu), or for the gamma5-hermitian system
// A = gamma5 D(mu) gamma5 D(-mu) = Qf5(mu) Qf5(-mu) (invert_Qtm.c:325-327).
struct Op {
  cvchip_ctx *c;
  int her;
  Field *aux;
  int eo = 0;
  Field *tmp = nullptr;   // even/odd: the even-parity intermediate
  // Schur complement on the odd sites, M = A_oo - D_oe A_ee^-1 D_eo
  //   = A_oo - H_oe A_ee^-1 H_eo / 4  (gamma5 on the output when g5); the fused
  // <out,in> of DOT_TS pairs the output with `partner` (default: in)
  int applyM(Field &out, const Field &in, double mutm, int g5, DotMode dm, const Field *partner, int *nb) {
    CVC_TRY(dslash_eo(c, *tmp, in, nullptr, mutm, 0, 1, 0.5, 0));
    return dslash_eo(c, out, *tmp, &in, mutm, 1, 0, -0.5, g5, dm,
                     dm == DOT_TS ? (partner ? partner : &in) : partner, nb);
  }
  // t = A in with <t,in> (complex) in res[0..N) / res[N..2N) and <t,t> in
  // res[2N..3N).  For the gamma5-hermitian operator A = B^dagger B with
  // B = gamma5 D(-mu) (gamma5 M(-mu) on the even/odd system), so
  // <t,in> = <B in, B in> is the norm of the first application's output: both
  // applications fuse a norm, and neither reads the input again.
  int apply_ts(Field &out, const Field &in, cvchip_ctx *cc, int N, double *res) {
    int nb1 = 0, nb2 = 0;
    if (!her) {
      CVC_TRY(apply(out, in, DOT_TS, nullptr, &nb1));
      return reduce_partials(cc, (const double *)cc->red.ptr, nb1, 3, N, res);
    }
    if (eo) CVC_TRY(applyM(*aux, in, -c->mu, 1, DOT_NN, nullptr, &nb1));
    else CVC_TRY(dslash(c, *aux, in, -c->mu, 1, DOT_NN, nullptr, &nb1));
    CVC_TRY(reduce_partials(cc, (const double *)cc->red.ptr, nb1, 1, N, res));
    hipLaunchKernelGGL(k_zero_doubles, dim3(1), dim3(256), 0, cc->stream, res + N, N);
    if (eo) CVC_TRY(applyM(out, *aux, c->mu, 1, DOT_NN, nullptr, &nb2));
    else CVC_TRY(dslash(c, out, *aux, c->mu, 1, DOT_NN, nullptr, &nb2));
    return reduce_partials(cc, (const double *)cc->red.ptr, nb2, 1, N, res + 2 * N);
  }
  int apply(Field &out, const Field &in, DotMode dm, const Field *r2, int *nb) {
    if (eo) {
      if (!her) return applyM(out, in, c->mu, 0, dm, r2, nb);
      CVC_TRY(applyM(*aux, in, -c->mu, 1, DOT_NONE, nullptr, nullptr));
      return applyM(out, *aux, c->mu, 1, dm, dm == DOT_TS ? &in : r2, nb);
    }
    if (!her) return dslash(c, out, in, c->mu, 0, dm, r2, nb);
    CVC_TRY(dslash(c, *aux, in, -c->mu, 1));
    // the fused <t,s> must pair the output with the operator's input, not with aux
    return dslash(c, out, *aux, c->mu, 1, dm, dm == DOT_TS ? &in : r2, nb);
  }
};

// gamma5 on all rhs of a FP64 set (in place)
__global__ void k_g5(double2 *__restrict__ f, uint32_t VN) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= VN) return;
#pragma unroll
  for (int c = 6; c < 12; c++) {
    double2 v = f[(size_t)c * VN + i];
    v.x = -v.x;
    v.y = -v.y;
    f[(size_t)c * VN + i] = v;
  }
}

// FP64 BiCGStab (on the gamma5-hermitian system when her != 0; then b is
// expected as gamma5 b and x receives y of A y = gamma5 b)
static int bicgstab64(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, unsigned flags,
                      cvchip_stats *stats, int her = 0, int eo = 0) {
  const int N = b.nrhs;
  Solve S;
  int rc = solve_setup(c, S, N, x.V);
  Field r1, r2, s, t, p, p2, aux, tmp;
  Field *all[8] = {&r1, &r2, &s, &t, &p, &p2, &aux, &tmp};
  for (int k = 0; k < 6 + (her ? 1 : 0) && !rc; k++) rc = field_alloc(c, *all[k], N, CVCHIP_PREC_F64, x.V);
  if (!rc && eo) rc = field_alloc(c, tmp, N, CVCHIP_PREC_F64, x.V);
  auto cleanup = [&]() {
    for (int k = 0; k < 8; k++) field_free(*all[k]);
    solve_free(S);
  };
  Op op{c, her, &aux, eo, &tmp};
  if (rc) {
    cleanup();
    return rc;
  }
  using IO = IOPlain<double>;
  IO ir1 = mkio<IO>(r1), ir2 = mkio<IO>(r2), is = mkio<IO>(s), iT = mkio<IO>(t), ip = mkio<IO>(p),
     ip2 = mkio<IO>(p2), ix = mkio<IO>(x), ib = mkio<IO>(b);
  const int lagged = (flags & CVCHIP_SOLVE_LAGGED) ? 1 : 0;
  const int verbose = (flags & CVCHIP_SOLVE_VERBOSE) ? 1 : 0;
  int dcalls = 0, nb = 0;
  int iters = 0, all_done = 0;

#define TRY(x)                                                                                   \
  do {                                                                                           \
    rc = (x);                                                                                    \
    if (rc) {                                                                                    \
      cleanup();                                                                                 \
      return rc;                                                                                 \
    }                                                                                            \
  } while (0)

  // r = b - D x ; r2 = p = r ; <r,r>, <b,b>
  TRY(op.apply(t, x, DOT_NONE, nullptr, nullptr));
  dcalls++;
  hipLaunchKernelGGL((k_residual<IO, IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, ir1, ir2, ip, ib, iT, 1,
                     S.st, 0, S.VN, N, (double *)c->red.ptr);
  TRY(reduce(S, (int)S.nblk, 2));
  hipLaunchKernelGGL(k_sc_init, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N, sp, 0);
  OpMask mask(c, S.st);
  TRY(pull_state(S));
  all_done = 1;
  for (int r = 0; r < N; r++) all_done &= S.hst[r].done;
  if (!all_done) {
    TRY(op.apply(p2, p, DOT_R2, &r2, &nb));
    dcalls++;
    TRY(reduce(S, nb, 2));
    for (int it = 0; it <= niter_max; it++) {
      hipLaunchKernelGGL(k_sc_alpha, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N);
      hipLaunchKernelGGL((k_s_update<IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, is, ir1, ip2, S.st, S.VN, N, S.nt);
      TRY(op.apply_ts(t, s, c, N, S.res));
      dcalls++;
      hipLaunchKernelGGL(k_sc_omega, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N);
      hipLaunchKernelGGL((k_r1_update<IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, ir1, is, iT, ir2, S.st,
                         S.VN, N, (double *)c->red.ptr, S.nt);
      TRY(reduce(S, (int)S.nblk, 3));
      hipLaunchKernelGGL(k_sc_beta, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N, it, lagged, 0, 0.0);
      hipLaunchKernelGGL((k_xp_update<IO, IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, ix, ip, is, ir1, ip2,
                         S.st, S.VN, N, S.nt);
      CVC_HIP(hipGetLastError());
      iters = it + 1;
      TRY(pull_state(S));
      all_done = 1;
      for (int r = 0; r < N; r++) all_done &= (S.hst[r].done | S.hst[r].final_upd);
      if (verbose)
        for (int r = 0; r < N; r++)
          printf("# [%d] residuum after iteration %d: %25.16e\n", r, it, S.hst[r].norm);
      if (all_done) break;
      TRY(op.apply(p2, p, DOT_R2, &r2, &nb));
      dcalls++;
      TRY(reduce(S, nb, 2));
    }
  }
  // final state (final_upd counts as converged)
  TRY(pull_state(S));
  int nconv = 0, maxit = 0;
  for (int r = 0; r < N; r++) {
    if (S.hst[r].done || S.hst[r].final_upd) {
      nconv++;
      if (S.hst[r].iter > maxit) maxit = S.hst[r].iter;
    }
  }
  double tr = 0.;
  mask.clear();
  if (!her && !eo) {
    TRY(true_residual(S, x, b, t, c->mu, &tr, nullptr));
    dcalls++;
  }
  if (stats) {
    stats->iterations = iters;
    stats->converged = nconv;
    stats->reliable_updates = 0;
    stats->dslash_calls = dcalls;
    stats->true_res2_max = tr;
  }
  cleanup();
#undef TRY
  if (nconv < N) {
    set_error("BiCGStab: %d of %d right-hand sides not converged after %d iterations", N - nconv, N,
              niter_max + 1);
    return CVCHIP_ERR_NOCONV;
  }
  return maxit;
}

// Mixed precision: FP32 or FP16 recurrence, FP64 reliable updates (on the
// gamma5-hermitian system when her != 0, b then being gamma5 b).
template <typename IOL>
static int bicgstab_mixed(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, int prec,
                          double delta, cvchip_stats *stats, int her = 0, int verbose = 0, int eo = 0) {
  const int N = b.nrhs;
  Solve S;
  int rc = solve_setup(c, S, N, x.V);
  Field r1, r2, s, t, p, p2, xl, t64, auxl, aux64, tmpl, tmp64;
  Field *lo[7] = {&r1, &r2, &s, &t, &p, &p2, &xl};
  const int64_t V = x.V;
  // The shadow residual r2 (r^ of BiCGStab) stays fixed through an FP32 solve
  // (no restarts), and any fixed shadow vector serves when every <r2, .> uses
  // the same one: it is kept in FP16 with a per-site scale, which halves the
  // bytes of the two <r2, .> products per iteration.
  using IOR = typename std::conditional<std::is_same<IOL, IOPlain<float>>::value, IOHalf, IOL>::type;
  const int prec_r2 = std::is_same<IOR, IOHalf>::value ? CVCHIP_PREC_F16 : prec;
  for (int k = 0; k < 6 && !rc; k++) rc = field_alloc(c, *lo[k], N, k == 1 ? prec_r2 : prec, V);
  // the solution accumulator since the last reliable update stays FP32
  if (!rc) rc = field_alloc(c, xl, N, CVCHIP_PREC_F32, V);
  if (!rc) rc = field_alloc(c, t64, N, CVCHIP_PREC_F64, V);
  if (!rc && her) rc = field_alloc(c, auxl, N, prec, V);
  if (!rc && her) rc = field_alloc(c, aux64, N, CVCHIP_PREC_F64, V);
  if (!rc && eo) rc = field_alloc(c, tmpl, N, prec, V);
  if (!rc && eo) rc = field_alloc(c, tmp64, N, CVCHIP_PREC_F64, V);
  auto cleanup = [&]() {
    for (int k = 0; k < 7; k++) field_free(*lo[k]);
    field_free(t64);
    field_free(auxl);
    field_free(aux64);
    field_free(tmpl);
    field_free(tmp64);
    solve_free(S);
  };
  Op opl{c, her, &auxl, eo, &tmpl}, op64{c, her, &aux64, eo, &tmp64};
  // FP16 storage keeps too few digits for the Krylov space to survive a
  // residual replacement: restart the recurrence at every reliable update
  const bool restart = prec == CVCHIP_PREC_F16;
  if (rc) {
    cleanup();
    return rc;
  }
  using IO = IOL;
  using I64 = IOPlain<double>;
  IO ir1 = mkio<IO>(r1), is = mkio<IO>(s), iT = mkio<IO>(t), ip = mkio<IO>(p), ip2 = mkio<IO>(p2);
  IOR ir2 = mkio<IOR>(r2);
  IOPlain<float> ixl = mkio<IOPlain<float>>(xl);
  I64 ix = mkio<I64>(x), ib = mkio<I64>(b), it64 = mkio<I64>(t64);
  const double delta2 = delta * delta;
  int dcalls = 0, nb = 0, nrel = 0, iters = 0, all_done = 0;

#define TRY(x)                                                                                   \
  do {                                                                                           \
    rc = (x);                                                                                    \
    if (rc) {                                                                                    \
      cleanup();                                                                                 \
      return rc;                                                                                 \
    }                                                                                            \
  } while (0)

  // r = b - D x in FP64, converted; r2 = p = r; xl = 0
  TRY(op64.apply(t64, x, DOT_NONE, nullptr, nullptr));
  dcalls++;
  hipLaunchKernelGGL((k_residual<I64, IO, IOR>), dim3(S.nblk), dim3(S.B), 0, c->stream, ir1, ir2, ip, ib, it64, 1,
                     S.st, 0, S.VN, N, (double *)c->red.ptr);
  TRY(reduce(S, (int)S.nblk, 2));
  hipLaunchKernelGGL(k_sc_init, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N, sp, 1);
  if constexpr (!std::is_same<IOR, IO>::value) {
    // r0rn = <r2, r> with the stored (FP16) shadow residual, not |r|^2
    hipLaunchKernelGGL((k_dot3<IOR, IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, ir2, ir1, S.VN, N,
                       (double *)c->red.ptr);
    TRY(reduce(S, (int)S.nblk, 3));
    hipLaunchKernelGGL(k_sc_set_r0rn, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N);
  }
  CVC_HIP(hipMemsetAsync(xl.data, 0, xl.bytes, c->stream));
  OpMask mask(c, S.st);
  TRY(pull_state(S));
  all_done = 1;
  for (int r = 0; r < N; r++) all_done &= S.hst[r].done;
  if (!all_done) {
    TRY(opl.apply(p2, p, DOT_R2, &r2, &nb));
    dcalls++;
    TRY(reduce(S, nb, 2));
    for (int it = 0; it <= niter_max; it++) {
      hipLaunchKernelGGL(k_sc_alpha, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N);
      hipLaunchKernelGGL((k_s_update<IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, is, ir1, ip2, S.st, S.VN, N, S.nt);
      TRY(opl.apply_ts(t, s, c, N, S.res));
      dcalls++;
      hipLaunchKernelGGL(k_sc_omega, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N);
      hipLaunchKernelGGL((k_r1_update<IO, IOR>), dim3(S.nblk), dim3(S.B), 0, c->stream, ir1, is, iT, ir2, S.st,
                         S.VN, N, (double *)c->red.ptr, S.nt);
      TRY(reduce(S, (int)S.nblk, 3));
      hipLaunchKernelGGL(k_sc_beta, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N, it, 0, 1, delta2);
      // xl += omega s + alpha p ; p = r1 + beta (p - omega p2)
      hipLaunchKernelGGL((k_xp_update<IOPlain<float>, IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, ixl, ip, is, ir1, ip2,
                         S.st, S.VN, N, S.nt);
      CVC_HIP(hipGetLastError());
      iters = it + 1;
      TRY(pull_state(S));
      int any_upd = 0;
      for (int r = 0; r < N; r++) any_upd |= (!S.hst[r].done && S.hst[r].upd);
      if (verbose && (any_upd || it % 16 == 0)) {
        double mx = 0.;
        for (int r = 0; r < N; r++) {
          if (S.hst[r].done) continue;
          const double q = S.hst[r].norm / S.hst[r].normb;
          if (!(q <= mx)) mx = q;   // a NaN propagates
        }
        printf("# [mixed] iteration %d: max iterated |r|^2/|b|^2 = %e\n", it, mx);
      }
      if (any_upd) {
        // reliable update for all active rhs: x += xl; xl = 0; r = b - D64 x
        nrel++;
        hipLaunchKernelGGL(k_sc_upd_all, dim3(1), dim3(256), 0, c->stream, S.st, N);
        hipLaunchKernelGGL((k_accum_zero<I64, IOPlain<float>>), dim3(S.nblk), dim3(S.B), 0, c->stream, ix, ixl, S.st,
                           S.VN, N);
        TRY(op64.apply(t64, x, DOT_NONE, nullptr, nullptr));
        dcalls++;
        hipLaunchKernelGGL((k_residual<I64, IO, IOR>), dim3(S.nblk), dim3(S.B), 0, c->stream, ir1, ir2, ip, ib,
                           it64, restart ? 1 : 0, S.st, 1, S.VN, N, (double *)c->red.ptr);
        TRY(reduce(S, (int)S.nblk, 2));
        hipLaunchKernelGGL(k_sc_reliable, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N, it);
        if (verbose) {
          TRY(pull_state(S));
          double mx = 0.;
          for (int r = 0; r < N; r++) {
            const double q = S.hst[r].norm / S.hst[r].normb;
            if (!(q <= mx)) mx = q;   // a NaN propagates
          }
          printf("# [mixed] reliable update %d at iteration %d: max true |r|^2/|b|^2 = %e\n", nrel, it, mx);
          fflush(stdout);
        }
        if (restart) {
          hipLaunchKernelGGL(k_sc_restart, dim3(1), dim3(256), 0, c->stream, S.st, N);
        } else {
          // r0rn and beta with the replaced residual: <r2, r>
          hipLaunchKernelGGL((k_dot3<IOR, IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, ir2, ir1, S.VN, N,
                             (double *)c->red.ptr);
          TRY(reduce(S, (int)S.nblk, 3));
          hipLaunchKernelGGL(k_sc_r0rn, dim3(1), dim3(256), 0, c->stream, S.st, S.res, N);
          hipLaunchKernelGGL((k_p_update<IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, ip, ir1, ip2, S.st, S.VN, N);
          hipLaunchKernelGGL(k_sc_clear_upd, dim3(1), dim3(256), 0, c->stream, S.st, N);
        }
        CVC_HIP(hipGetLastError());
        TRY(pull_state(S));
      }
      all_done = 1;
      for (int r = 0; r < N; r++) all_done &= S.hst[r].done;
      if (all_done) break;
      TRY(opl.apply(p2, p, DOT_R2, &r2, &nb));
      dcalls++;
      TRY(reduce(S, nb, 2));
    }
  }
  TRY(pull_state(S));
  int nconv = 0, maxit = 0, nstall = 0;
  for (int r = 0; r < N; r++)
    if (S.hst[r].done && !S.hst[r].stalled) {
      nconv++;
      if (S.hst[r].iter > maxit) maxit = S.hst[r].iter;
    } else if (S.hst[r].stalled) {
      nstall++;
    }
  // fold in what the low-precision accumulator still holds (non-converged rhs;
  // converged ones were folded at their last reliable update)
  if (nconv < N) {
    hipLaunchKernelGGL(k_sc_upd_all, dim3(1), dim3(256), 0, c->stream, S.st, N);
    hipLaunchKernelGGL((k_accum_zero<I64, IOPlain<float>>), dim3(S.nblk), dim3(S.B), 0, c->stream, ix, ixl, S.st,
                       S.VN, N);
    CVC_HIP(hipGetLastError());
  }
  double tr = 0.;
  mask.clear();
  if (!her && !eo) {
    TRY(true_residual(S, x, b, t64, c->mu, &tr, nullptr));
    dcalls++;
  }
  if (stats) {
    stats->iterations = iters;
    stats->converged = nconv;
    stats->reliable_updates = nrel;
    stats->dslash_calls = dcalls;
    stats->true_res2_max = tr;
  }
  cleanup();
#undef TRY
  if (nconv < N) {
    if (nstall)
      set_error("mixed BiCGStab: %d of %d right-hand sides not converged (%d stopped at the floor of the FP64 true "
                "residual, %d reliable updates without progress) after %d iterations", N - nconv, N, nstall,
                kMaxStall, iters);
    else
      set_error("mixed BiCGStab: %d of %d right-hand sides not converged after %d iterations", N - nconv, N,
                niter_max + 1);
    return CVCHIP_ERR_NOCONV;
  }
  return maxit;
}

// ------------------------------------------------------------ even/odd

// FP64 set <-> its two parity halves (half-site h of parity p is the
// lexicographic site 2h or 2h+1 of that parity: z = 2 (h mod LZ/2) + ...)
__global__ void __launch_bounds__(256) k_eo_split(double2 *__restrict__ full, double2 *__restrict__ ev,
                                                  double2 *__restrict__ od, uint32_t V, int N, int LX, int LY,
                                                  int LZ, int to_full) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t VN = V * (uint32_t)N;
  if (i >= VN) return;
  const uint32_t site = i / (uint32_t)N, r = i - site * (uint32_t)N;
  const uint32_t z = site % (uint32_t)LZ;
  uint32_t q = site / (uint32_t)LZ;
  const uint32_t y = q % (uint32_t)LY;
  q /= (uint32_t)LY;
  const uint32_t x = q % (uint32_t)LX, t = q / (uint32_t)LX;
  double2 *h = ((t + x + y + z) & 1u) ? od : ev;
  const uint32_t j = (site >> 1) * (uint32_t)N + r, VhN = VN / 2;
#pragma unroll
  for (int c = 0; c < 12; c++) {
    if (to_full) full[(size_t)c * VN + i] = h[(size_t)c * VhN + j];
    else h[(size_t)c * VhN + j] = full[(size_t)c * VN + i];
  }
}

static int eo_split(cvchip_ctx *c, Field &full, Field &ev, Field &od, int to_full) {
  const uint32_t VN = (uint32_t)(full.V * full.nrhs);
  hipLaunchKernelGGL(k_eo_split, dim3((VN + 255) / 256), dim3(256), 0, c->stream, (double2 *)full.data,
                     (double2 *)ev.data, (double2 *)od.data, (uint32_t)full.V, full.nrhs, c->lat.LX, c->lat.LY,
                     c->lat.LZ, to_full);
  CVC_HIP(hipGetLastError());
  return 0;
}

// out = A^-1 in on one parity, A = 1/(2 kappa) + i mutm gamma5 (FP64)
__global__ void __launch_bounds__(256) k_diag_inv(double2 *__restrict__ out, const double2 *__restrict__ in,
                                                  uint32_t VN, double k2, double mutm, double dinv) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= VN) return;
#pragma unroll
  for (int c = 0; c < 12; c++) {
    const double m = c < 6 ? mutm : -mutm;
    const double2 v = in[(size_t)c * VN + i];
    out[(size_t)c * VN + i] = make_double2((k2 * v.x + m * v.y) * dinv, (k2 * v.y - m * v.x) * dinv);
  }
}

// D(mu) x = b through the Schur complement on the odd sites:
//   M x_o = b_o - D_oe A_ee^-1 b_e = b_o + H_oe A_ee^-1 b_e / 2,
//   x_e = A_ee^-1 (b_e - D_eo x_o) = A_ee^-1 (b_e + H_eo x_o / 2),
// M = A_oo - H_oe A_ee^-1 H_eo / 4.  her: M is solved like the full operator,
// as gamma5 M(mu) gamma5 M(-mu) y = gamma5 b'_o, x_o = gamma5 M(-mu) y
// (M(mu)^dagger = gamma5 M(-mu) gamma5).  The convergence test is the
// reference's, on the odd system; the returned statistics carry the true
// residual |D x - b|^2 / |b|^2 of the whole lattice.
static int invert_eo(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, int prec, double delta,
                     unsigned flags, cvchip_stats *stats, int her) {
  if (c->lat.LZ % 2) {
    set_error("even/odd preconditioning needs an even LZ (got %d)", c->lat.LZ);
    return CVCHIP_ERR_ARG;
  }
  if (prec != CVCHIP_PREC_F64 && !(delta > 0. && delta < 1.)) {
    set_error("reliable-update factor delta must lie in (0,1), got %g", delta);
    return CVCHIP_ERR_ARG;
  }
  if (prec < CVCHIP_PREC_F64 || prec > CVCHIP_PREC_F16) {
    set_error("unknown precision %d", prec);
    return CVCHIP_ERR_ARG;
  }
  auto t0 = std::chrono::steady_clock::now();
  const int N = b.nrhs;
  const int64_t Vh = c->lat.V / 2;
  const double mu = c->mu;
  Field be, bo, xe, xo, bp, y;
  int rc = 0;
  Field *hf[6] = {&be, &bo, &xe, &xo, &bp, &y};
  for (int k = 0; k < 5 + (her ? 1 : 0) && !rc; k++) rc = field_alloc(c, *hf[k], N, CVCHIP_PREC_F64, Vh);
  auto cleanup = [&]() {
    for (int k = 0; k < 6; k++) field_free(*hf[k]);
  };
#define TRY(x)                                                                                   \
  do {                                                                                           \
    rc = (x);                                                                                    \
    if (rc) {                                                                                    \
      cleanup();                                                                                 \
      return rc;                                                                                 \
    }                                                                                            \
  } while (0)
  if (rc) {
    cleanup();
    return rc;
  }
  TRY(eo_split(c, const_cast<Field &>(b), be, bo, 0));
  TRY(eo_split(c, x, xe, xo, 0));
  const uint32_t VhN = (uint32_t)(Vh * N);
  const double k2 = 0.5 / c->kappa, dinv = 1.0 / (k2 * k2 + mu * mu);
  hipLaunchKernelGGL(k_diag_inv, dim3((VhN + 255) / 256), dim3(256), 0, c->stream, (double2 *)xe.data,
                     (const double2 *)be.data, VhN, k2, mu, dinv);
  TRY(dslash_eo(c, bp, xe, &bo, mu, 1, 2, 0.5, 0));      // b'_o = b_o + H_oe A_ee^-1 b_e / 2
  int ret;
  if (her) {
    hipLaunchKernelGGL(k_g5, dim3((VhN + 255) / 256), dim3(256), 0, c->stream, (double2 *)bp.data, VhN);
    CVC_HIP(hipMemcpyAsync(y.data, xo.data, xo.bytes, hipMemcpyDeviceToDevice, c->stream));
    if (prec == CVCHIP_PREC_F64) ret = bicgstab64(c, y, bp, sp, niter_max, flags, stats, 1, 1);
    else if (prec == CVCHIP_PREC_F32)
      ret = bicgstab_mixed<IOPlain<float>>(c, y, bp, sp, niter_max, prec, delta, stats, 1,
                                           (flags & CVCHIP_SOLVE_VERBOSE) ? 1 : 0, 1);
    else
      ret = bicgstab_mixed<IOHalf>(c, y, bp, sp, niter_max, prec, delta, stats, 1,
                                   (flags & CVCHIP_SOLVE_VERBOSE) ? 1 : 0, 1);
    if (ret >= 0 || ret == CVCHIP_ERR_NOCONV) {
      Op m{c, 0, nullptr, 1, &xe};
      TRY(m.applyM(xo, y, -mu, 1, DOT_NONE, nullptr, nullptr));    // x_o = gamma5 M(-mu) y
    }
  } else {
    if (prec == CVCHIP_PREC_F64) ret = bicgstab64(c, xo, bp, sp, niter_max, flags, stats, 0, 1);
    else if (prec == CVCHIP_PREC_F32)
      ret = bicgstab_mixed<IOPlain<float>>(c, xo, bp, sp, niter_max, prec, delta, stats, 0,
                                           (flags & CVCHIP_SOLVE_VERBOSE) ? 1 : 0, 1);
    else
      ret = bicgstab_mixed<IOHalf>(c, xo, bp, sp, niter_max, prec, delta, stats, 0,
                                   (flags & CVCHIP_SOLVE_VERBOSE) ? 1 : 0, 1);
  }
  if (ret >= 0 || ret == CVCHIP_ERR_NOCONV) {
    TRY(dslash_eo(c, xe, xo, &be, mu, 0, 3, 0.5, 0));    // x_e = A_ee^-1 (b_e + H_eo x_o / 2)
    TRY(eo_split(c, x, xe, xo, 1));
    if (stats) {
      Solve S;
      Field tmp;
      double tr = 0.;
      int rc3 = solve_setup(c, S, N);
      if (!rc3) rc3 = field_alloc(c, tmp, N, CVCHIP_PREC_F64);
      if (!rc3) rc3 = true_residual(S, x, b, tmp, mu, &tr, nullptr);
      field_free(tmp);
      solve_free(S);
      if (rc3) ret = rc3;
      stats->true_res2_max = tr;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  }
  cleanup();
#undef TRY
  return ret;
}

// invert_Qtm_her (invert_Qtm.c:276-478): y from A y = gamma5 b, then x = gamma5 D(-mu) y.
int invert_her_fields(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, int prec,
                      double delta, unsigned flags, cvchip_stats *stats) {
  if (!c->gauge_set) {
    set_error("gauge field not set (call cvchip_set_gauge first)");
    return CVCHIP_ERR_STATE;
  }
  flags |= c->solve_flags;
  if (flags & CVCHIP_SOLVE_EO) {
    if (niter_max < 0 || !(sp >= 0.)) {
      set_error("invalid niter_max %d or tolerance %g", niter_max, sp);
      return CVCHIP_ERR_ARG;
    }
    return invert_eo(c, x, b, sp, niter_max, prec, delta, flags, stats, 1);
  }
  auto t0 = std::chrono::steady_clock::now();
  Field g5b, y;
  int rc = field_alloc(c, g5b, b.nrhs, CVCHIP_PREC_F64);
  if (!rc) rc = field_alloc(c, y, b.nrhs, CVCHIP_PREC_F64);
  if (!rc) {
    const uint32_t VN = (uint32_t)(b.V * b.nrhs);
    rc = hipMemcpyAsync(g5b.data, b.data, b.bytes, hipMemcpyDeviceToDevice, c->stream) == hipSuccess ? 0 : CVCHIP_ERR_HIP;
    if (!rc && hipMemcpyAsync(y.data, x.data, x.bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) rc = CVCHIP_ERR_HIP;
    if (!rc) {
      hipLaunchKernelGGL(k_g5, dim3((VN + 255) / 256), dim3(256), 0, c->stream, (double2 *)g5b.data, VN);
      if (prec == CVCHIP_PREC_F64) rc = bicgstab64(c, y, g5b, sp, niter_max, flags, stats, 1);
      else if (!(delta > 0. && delta < 1.)) {
        set_error("reliable-update factor delta must lie in (0,1), got %g", delta);
        rc = CVCHIP_ERR_ARG;
      } else if (prec == CVCHIP_PREC_F32)
        rc = bicgstab_mixed<IOPlain<float>>(c, y, g5b, sp, niter_max, prec, delta, stats, 1,
                                            (flags & CVCHIP_SOLVE_VERBOSE) ? 1 : 0);
      else if (prec == CVCHIP_PREC_F16)
        rc = bicgstab_mixed<IOHalf>(c, y, g5b, sp, niter_max, prec, delta, stats, 1,
                                    (flags & CVCHIP_SOLVE_VERBOSE) ? 1 : 0);
      else {
        set_error("unknown precision %d", prec);
        rc = CVCHIP_ERR_ARG;
      }
    }
    // x = gamma5 D(-mu) y, done whether or not the solve converged (invert_Qtm.c:441-446)
    if (rc >= 0 || rc == CVCHIP_ERR_NOCONV) {
      const int rc2 = dslash(c, x, y, -c->mu, 1);
      if (rc2) rc = rc2;
    }
    if (stats && (rc >= 0 || rc == CVCHIP_ERR_NOCONV)) {
      Solve S;
      double tr = 0.;
      Field tmp;
      int rc3 = solve_setup(c, S, b.nrhs);
      if (!rc3) rc3 = field_alloc(c, tmp, b.nrhs, CVCHIP_PREC_F64);
      if (!rc3) rc3 = true_residual(S, x, b, tmp, c->mu, &tr, nullptr);
      field_free(tmp);
      solve_free(S);
      if (rc3) rc = rc3;
      stats->true_res2_max = tr;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  }
  field_free(g5b);
  field_free(y);
  return rc;
}

int invert_fields(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, int prec, double delta,
                  unsigned flags, cvchip_stats *stats) {
  if (!c->gauge_set) {
    set_error("gauge field not set (call cvchip_set_gauge first)");
    return CVCHIP_ERR_STATE;
  }
  if (x.prec != CVCHIP_PREC_F64 || b.prec != CVCHIP_PREC_F64 || x.nrhs != b.nrhs) {
    set_error("solution and source must be FP64 sets of equal size");
    return CVCHIP_ERR_ARG;
  }
  if (niter_max < 0 || !(sp >= 0.)) {
    set_error("invalid niter_max %d or tolerance %g", niter_max, sp);
    return CVCHIP_ERR_ARG;
  }
  flags |= c->solve_flags;
  if (flags & CVCHIP_SOLVE_EO) return invert_eo(c, x, b, sp, niter_max, prec, delta, flags, stats, 0);
  auto t0 = std::chrono::steady_clock::now();
  int rc;
  if (prec == CVCHIP_PREC_F64) rc = bicgstab64(c, x, b, sp, niter_max, flags, stats);
  else {
    if (!(delta > 0. && delta < 1.)) {
      set_error("reliable-update factor delta must lie in (0,1), got %g", delta);
      return CVCHIP_ERR_ARG;
    }
    if (prec == CVCHIP_PREC_F32) rc = bicgstab_mixed<IOPlain<float>>(c, x, b, sp, niter_max, prec, delta, stats);
    else if (prec == CVCHIP_PREC_F16) rc = bicgstab_mixed<IOHalf>(c, x, b, sp, niter_max, prec, delta, stats);
    else {
      set_error("unknown precision %d", prec);
      return CVCHIP_ERR_ARG;
    }
  }
  if (stats)
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // namespace cvchip

using namespace cvchip;

extern "C" int cvchip_field_invert(cvchip_ctx *c, cvchip_field *x, const cvchip_field *b, double sp,
                                   int niter_max, int prec, double delta, unsigned flags,
                                   cvchip_stats *stats) {
  return cvchip::guarded("cvchip_field_invert", [&]() -> int {
    if (!c || !x || !b) {
      set_error("NULL argument to cvchip_field_invert");
      return CVCHIP_ERR_NULL;
    }
    return invert_fields(c, x->f, b->f, sp, niter_max, prec, delta, flags, stats);
  });
}

extern "C" int cvchip_invert_Qtm(cvchip_ctx *c, double *x, const double *b, int nrhs, double sp,
                                 int niter_max, int prec, double delta, unsigned flags, cvchip_stats *stats) {
  return cvchip::guarded("cvchip_invert_Qtm", [&]() -> int {
    if (!c || !x || !b) {
      set_error("NULL argument to cvchip_invert_Qtm");
      return CVCHIP_ERR_NULL;
    }
    Field fx, fb;
    int rc = field_alloc(c, fx, nrhs, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, fb, nrhs, CVCHIP_PREC_F64);
    if (!rc) rc = upload_aos(c, fx, x, 0, nrhs);
    if (!rc) rc = upload_aos(c, fb, b, 0, nrhs);
    int ret = rc;
    if (!rc) {
      ret = invert_fields(c, fx, fb, sp, niter_max, prec, delta, flags, stats);
      // the reference leaves x as computed also when not converged
      if (ret >= 0 || ret == CVCHIP_ERR_NOCONV) {
        int rc2 = download_aos(c, fx, x, 0, nrhs);
        if (rc2) ret = rc2;
      }
    }
    field_free(fx);
    field_free(fb);
    return ret;
  });
}

extern "C" int cvchip_invert_Qtm_her(cvchip_ctx *c, double *x, const double *b, int nrhs, double sp,
                                     int niter_max, int prec, double delta, unsigned flags,
                                     cvchip_stats *stats) {
  return cvchip::guarded("cvchip_invert_Qtm_her", [&]() -> int {
    if (!c || !x || !b) {
      set_error("NULL argument to cvchip_invert_Qtm_her");
      return CVCHIP_ERR_NULL;
    }
    Field fx, fb;
    int rc = field_alloc(c, fx, nrhs, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, fb, nrhs, CVCHIP_PREC_F64);
    if (!rc) rc = upload_aos(c, fx, x, 0, nrhs);
    if (!rc) rc = upload_aos(c, fb, b, 0, nrhs);
    int ret = rc;
    if (!rc) {
      ret = invert_her_fields(c, fx, fb, sp, niter_max, prec, delta, flags, stats);
      if (ret >= 0 || ret == CVCHIP_ERR_NOCONV) {
        int rc2 = download_aos(c, fx, x, 0, nrhs);
        if (rc2) ret = rc2;
      }
    }
    field_free(fx);
    field_free(fb);
    return ret;
  });
}

// spinor_scalar_product_co / _re (invert_Qtm.c:28-84) on host arrays
static int scalar_product(cvchip_ctx *c, double w[3], const double *xi, const double *phi) {
  Field a, b;
  Solve S;
  int rc = field_alloc(c, a, 1, CVCHIP_PREC_F64);
  if (!rc) rc = field_alloc(c, b, 1, CVCHIP_PREC_F64);
  if (!rc) rc = upload_aos(c, a, xi, 0, 1);
  if (!rc) rc = upload_aos(c, b, phi, 0, 1);
  if (!rc) rc = solve_setup(c, S, 1);
  if (!rc) {
    using IO = IOPlain<double>;
    hipLaunchKernelGGL((k_dot3<IO>), dim3(S.nblk), dim3(S.B), 0, c->stream, mkio<IO>(a), mkio<IO>(b), S.VN, 1,
                       (double *)c->red.ptr);
    rc = reduce(S, (int)S.nblk, 3);
    if (!rc) {
      hipError_t e = hipMemcpyAsync(w, S.res, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) rc = hip_fail(e, "scalar product", __FILE__, __LINE__);
    }
  }
  solve_free(S);
  field_free(a);
  field_free(b);
  return rc;
}

extern "C" int cvchip_scalar_product_co(cvchip_ctx *c, double w[2], const double *xi, const double *phi) {
  return cvchip::guarded("cvchip_scalar_product_co", [&]() -> int {
    if (!c || !w || !xi || !phi) return CVCHIP_ERR_NULL;
    double r[3];
    CVC_TRY(scalar_product(c, r, xi, phi));
    w[0] = r[0];
    w[1] = r[1];
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_scalar_product_re(cvchip_ctx *c, double *out, const double *xi, const double *phi) {
  return cvchip::guarded("cvchip_scalar_product_re", [&]() -> int {
    if (!c || !out || !xi || !phi) return CVCHIP_ERR_NULL;
    double r[3];
    CVC_TRY(scalar_product(c, r, xi, phi));
    *out = r[0];
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_invert_her(cvchip_ctx *c, cvchip_field *x, const cvchip_field *b, double sp,
                                       int niter_max, int prec, double delta, unsigned flags,
                                       cvchip_stats *stats) {
  return cvchip::guarded("cvchip_field_invert_her", [&]() -> int {
    if (!c || !x || !b) {
      set_error("NULL argument to cvchip_field_invert_her");
      return CVCHIP_ERR_NULL;
    }
    if (x->f.prec != CVCHIP_PREC_F64 || b->f.prec != CVCHIP_PREC_F64 || x->f.nrhs != b->f.nrhs) {
      set_error("solution and source must be FP64 sets of equal size");
      return CVCHIP_ERR_ARG;
    }
    return invert_her_fields(c, x->f, b->f, sp, niter_max, prec, delta, flags, stats);
  });
}

extern "C" int cvchip_field_copy(cvchip_ctx *c, cvchip_field *dst, const cvchip_field *src) {
  return cvchip::guarded("cvchip_field_copy", [&]() -> int {
    if (!c || !dst || !src) return CVCHIP_ERR_NULL;
    Field &d = dst->f;
    const Field &s = src->f;
    if (d.nrhs != s.nrhs || d.V != s.V) {
      set_error("field copy between sets of different size");
      return CVCHIP_ERR_ARG;
    }
    const int N = s.nrhs;
    const uint32_t VN = (uint32_t)(s.V * N);
    const int B = block_for(N);
    dim3 g((VN + B - 1) / B), b(B);
    using I64 = IOPlain<double>;
    using I32 = IOPlain<float>;
  #define CP(TD, TS) hipLaunchKernelGGL((k_copy<TD, TS>), g, b, 0, c->stream, mkio<TD>(d), mkio<TS>(s), \
                                        (const RhsState *)nullptr, 0, VN, N)
    if (d.prec == CVCHIP_PREC_F64 && s.prec == CVCHIP_PREC_F64) CP(I64, I64);
    else if (d.prec == CVCHIP_PREC_F64 && s.prec == CVCHIP_PREC_F32) CP(I64, I32);
    else if (d.prec == CVCHIP_PREC_F64 && s.prec == CVCHIP_PREC_F16) CP(I64, IOHalf);
    else if (d.prec == CVCHIP_PREC_F32 && s.prec == CVCHIP_PREC_F64) CP(I32, I64);
    else if (d.prec == CVCHIP_PREC_F32 && s.prec == CVCHIP_PREC_F32) CP(I32, I32);
    else if (d.prec == CVCHIP_PREC_F32 && s.prec == CVCHIP_PREC_F16) CP(I32, IOHalf);
    else if (d.prec == CVCHIP_PREC_F16 && s.prec == CVCHIP_PREC_F64) CP(IOHalf, I64);
    else if (d.prec == CVCHIP_PREC_F16 && s.prec == CVCHIP_PREC_F32) CP(IOHalf, I32);
    else CP(IOHalf, IOHalf);
  #undef CP
    CVC_HIP(hipGetLastError());
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}
