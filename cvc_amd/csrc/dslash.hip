// Twisted-mass Wilson operator on gfx950:
//   xi = D(mutm) phi = 1/(2 kappa) phi + i mutm gamma5 phi
//        - 1/2 sum_mu [ (1 - gamma_mu) U~_mu(x) phi(x+mu) + (1 + gamma_mu) U~_mu(x-mu)^dagger phi(x-mu) ]
// with U~_mu = U_mu * exp(i pi bc_mu / L_mu) (reference Q_phi.c:25-36, 38-188;
// phases cvc_geometry.c:452-459).
//
// One thread per (site, rhs); the rhs index is the fastest-moving index of the
// field layout, so the 64 lanes of a wave read 64 consecutive complex numbers
// of each colour-spin plane and lanes of the same site share link loads.
// Each hop projects the neighbour spinor to two spin components with
// (1 -+ gamma_mu) (rank 2 in the chiral basis), multiplies the half spinor by
// the link, and rebuilds the lower spin components from the upper ones.
// Periodic wrap and the boundary phase are resolved in the kernel: there is no
// halo.  Optional epilogues fuse the solver's per-rhs dot products.
#include <algorithm>

#include <type_traits>

#include "spinor_io.h"

namespace cvchip {

template <typename R>
struct DParams {
  uint32_t V, VN, sX, sT;
  int N, T, LX, LY, LZ;
  // log2 of LZ, LY, LX, sX, sT and of the traversal's LZt when a power of two,
  // else -1 (coordinates then come from shifts and masks instead of divisions);
  // 1/N, 1/S and 1/(S KL) for the small quotients of thread and staging indices
  int shLZ, shLY, shLX, shSX, shST, shLZt;
  float invN, invS, invKS;
  int partner;   // DOT_TS: dot with the separate field r2 instead of the input
  // traversal (mode 1): the lattice is cut into chunks of csz consecutive
  // sites (one eighth of a time slice when possible), bpc blocks per chunk;
  // the 8 XCDs take the 8 chunks of a time slice side by side, so the whole
  // GPU advances through the slices together.  Mode 0: the same chunks with
  // no XCD interleave.  Mode 2: XCD j owns the time slices t = j mod 8 and
  // walks each slice in runs of yb consecutive y rows (yb*LZ consecutive
  // sites), x within y block: the x and y neighbours of a run were touched a
  // few runs earlier and are still in the XCD's L2, and the t neighbours are
  // the slices the adjacent XCDs are working on at the same time.
  // Mode 4: bricks of bt time slices x bx x-planes x yb y-rows (all z); the
  // bricks of one (x, y) column follow each other in t on one XCD, so the
  // sites in flight on an XCD form a compact 4D region.
  // Mode 5: columns of bx x-planes x yb y-rows x zb z-sites (powers of two)
  // swept along the whole t extent, one column at a time per XCD, with the
  // column's sites ordered z, y, x within a time slice: the t neighbour of a
  // site lies one cross-section (bx yb zb sites, a few blocks) ahead in the
  // sweep, inside the XCD's L2 window even for 12 rhs in FP64.
  uint32_t csz, bpc;
  int mode;
  uint32_t yb, runlen, bpr, nrun;
  uint32_t bt, bx, ncx, bpb, bpcol;
  uint32_t zb, ncy, ncz, lzb, lyz, lA, colsz, sTt, sXt;
  // Mode 6 (k_dslash_march): mS columns per workgroup marched through mM
  // slices; patches of mXP x mYP x LZ columns (mNPX along x), mCP chunks of mS
  // columns per patch, mNR = T / mM slice ranges; mXcd: patch p on XCD p mod 8.
  uint32_t mS, mM, mXP, mYP, mCP, mNR, mNPX;
  int mXcd;
  int nt;   // streaming output stores
  // Even/odd operator (EO kernels): the traversal runs over the half lattice
  // (T, LX, LY, LZ/2); half-site h of parity par is the lexicographic site
  // with z = 2 (h mod LZ/2) + ((par + t + x + y) mod 2), i.e. lex = 2h or 2h+1.
  // The epilogue forms, with A = 1/(2 kappa) + i mutm gamma5 on the output
  // parity and hop the hopping sum:
  //   dmode 0: A own + hopc hop     1: A^-1 (hopc hop)
  //   dmode 2: own + hopc hop       3: A^-1 (own + hopc hop)
  uint32_t LZt;   // z extent of the traversal (LZ, or LZ/2 for EO)
  int par, dmode;
  R hopc, dinv;   // dinv = 1 / ((1/(2 kappa))^2 + mutm^2)
  R k2, mutm;
  cplx<R> ph[4];
  // converged right-hand sides of a running solve (done[r * dstride] != 0):
  // their lanes compute and store nothing
  const int *done;
  int dstride;
};

// n / d and n % d for a divisor fixed for the launch: a shift and a mask when
// d = 2^sh (sh >= 0, a branch uniform across the grid)
__device__ __forceinline__ uint32_t qdiv(uint32_t n, uint32_t d, int sh) { return sh >= 0 ? n >> sh : n / d; }
__device__ __forceinline__ uint32_t qmod(uint32_t n, uint32_t d, int sh) { return sh >= 0 ? n & (d - 1u) : n % d; }
// floor(n / d) for n < 2^21 from inv, the float nearest 1/d (or within two
// roundings of it): (n + 1/2) / d stays at least 1/(2d) away from an integer,
// and the float product errs by less than (n + 1/2) 2^-22 / d
__device__ __forceinline__ uint32_t sdiv(uint32_t n, float inv) { return (uint32_t)(((float)n + 0.5f) * inv); }

// Spin projection of a neighbour spinor: h[a] = psi_a -+ g_a psi_{GCOL[a]}
// for the upper spins a = 0, 1 ((1 -+ gamma_mu) has rank 2 in the chiral
// basis).  get(c) returns component c = 3*spin + colour; FP64 loads each
// component where it is used, so no more than the half spinor and one pair of
// spin rows are live at a time.
template <int MU, bool FWD, typename C, typename Get>
__device__ __forceinline__ void project(C h[2][3], Get get) {
  constexpr int p0 = GCOL[MU][0], p1 = GCOL[MU][1];
  constexpr int g0 = FWD ? uneg(GVAL[MU][0]) : GVAL[MU][0];
  constexpr int g1 = FWD ? uneg(GVAL[MU][1]) : GVAL[MU][1];
#pragma unroll
  for (int c = 0; c < 3; c++) h[0][c] = cadd(get(c), cunit<g0>(get(3 * p0 + c)));
#pragma unroll
  for (int c = 0; c < 3; c++) h[1][c] = cadd(get(3 + c), cunit<g1>(get(3 * p1 + c)));
}

// FP16 storage: the projection straight from the stored halves with the site's
// scale n, each output real one pair of mixed-precision FMAs (half operand,
// FP32 scale and accumulator): h = n A + u n B for the unit u of cunit.
__device__ __forceinline__ float half_lo(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
}
__device__ __forceinline__ float half_hi(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16));
}
template <int U>
__device__ __forceinline__ float2 half_pair(uint32_t A, uint32_t B, float n) {
  const float ax = __builtin_fmaf(half_lo(A), n, 0.0f), ay = __builtin_fmaf(half_hi(A), n, 0.0f);
  if constexpr (U == 0) return make_float2(__builtin_fmaf(half_lo(B), n, ax), __builtin_fmaf(half_hi(B), n, ay));
  else if constexpr (U == 1) return make_float2(__builtin_fmaf(half_lo(B), -n, ax), __builtin_fmaf(half_hi(B), -n, ay));
  else if constexpr (U == 2) return make_float2(__builtin_fmaf(half_hi(B), -n, ax), __builtin_fmaf(half_lo(B), n, ay));
  else return make_float2(__builtin_fmaf(half_hi(B), n, ax), __builtin_fmaf(half_lo(B), -n, ay));
}
template <int MU, bool FWD>
__device__ __forceinline__ void project_half(float2 h[2][3], const uint32_t w[12], float n) {
  constexpr int p0 = GCOL[MU][0], p1 = GCOL[MU][1];
  constexpr int g0 = FWD ? uneg(GVAL[MU][0]) : GVAL[MU][0];
  constexpr int g1 = FWD ? uneg(GVAL[MU][1]) : GVAL[MU][1];
#pragma unroll
  for (int c = 0; c < 3; c++) h[0][c] = half_pair<g0>(w[c], w[3 * p0 + c], n);
#pragma unroll
  for (int c = 0; c < 3; c++) h[1][c] = half_pair<g1>(w[3 + c], w[3 * p1 + c], n);
}

// acc += (1 - gamma_MU) U ph psi (FWD) or (1 + gamma_MU) U^dagger conj(ph) psi
// (!FWD), from the projected h: one colour row of the link at a time,
// w_a = row . h_a, then the upper spins take w_a and the lower spins
// -+g_b w_{GCOL[b]}.  base[k * stride] is the stored link entry k (9 entries,
// or the first two rows, the third rebuilt as conj(row0 x row1)).
// third row of an SU(3) link from its first two: conj(row0 x row1)
template <typename C>
__device__ __forceinline__ void rebuild_row3(C u[9]) {
  const C a = csub(cmul(u[1], u[5]), cmul(u[2], u[4]));
  const C b = csub(cmul(u[2], u[3]), cmul(u[0], u[5]));
  const C d = csub(cmul(u[0], u[4]), cmul(u[1], u[3]));
  u[6] = mk<decltype(a.x)>(a.x, -a.y);
  u[7] = mk<decltype(a.x)>(b.x, -b.y);
  u[8] = mk<decltype(a.x)>(d.x, -d.y);
}

// PHASE false: the entries at base already carry the boundary phase
template <typename R, int RECON, int MU, bool FWD, bool PHASE = true>
__device__ __forceinline__ void link_apply(cplx<R> acc[12], const cplx<R> h[2][3],
                                           const cplx<R> *__restrict__ base, uint32_t stride, cplx<R> ph) {
  using C = cplx<R>;
  constexpr int q2 = GCOL[MU][2], q3 = GCOL[MU][3];
  constexpr int l2 = FWD ? uneg(GVAL[MU][2]) : GVAL[MU][2];
  constexpr int l3 = FWD ? uneg(GVAL[MU][3]) : GVAL[MU][3];
  C u[9];
  if constexpr (RECON == 12) {
#pragma unroll
    for (int k = 0; k < 6; k++) u[k] = base[(size_t)k * stride];
    rebuild_row3(u);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    C w0 = mk<R>(0, 0), w1 = mk<R>(0, 0);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int k = FWD ? 3 * i + j : 3 * j + i;
      const C e = RECON == 18 ? base[(size_t)k * stride] : u[k];
      if constexpr (FWD) {
        w0 = cfma(w0, e, h[0][j]);
        w1 = cfma(w1, e, h[1][j]);
      } else {
        w0 = cfmac(w0, e, h[0][j]);
        w1 = cfmac(w1, e, h[1][j]);
      }
    }
    if constexpr (!PHASE) {
    } else if constexpr (FWD) {
      w0 = cmul(ph, w0);
      w1 = cmul(ph, w1);
    } else {
      w0 = cmulc(ph, w0);
      w1 = cmulc(ph, w1);
    }
    acc[i] = cadd(acc[i], w0);
    acc[3 + i] = cadd(acc[3 + i], w1);
    acc[6 + i] = cadd(acc[6 + i], cunit<l2>(q2 == 0 ? w0 : w1));
    acc[9 + i] = cadd(acc[9 + i], cunit<l3>(q3 == 0 ? w0 : w1));
  }
}

// Links of a block staged in LDS: [8 hops][KL entries][LDS_SITES], hop 2*mu is
// U_mu(x - mu) (backward), 2*mu + 1 is U_mu(x) (forward).
constexpr int LDS_SITES = 32;

// Projected neighbour spinors.  With LDSS the block's own sites and the two
// sites before and after them are staged in LDS, so the z neighbours
// (adjacent sites) and the own spinor come from LDS instead of the texture
// path.
template <typename IO, bool LDSS>
struct NbLoader {
  using R = typename IO::real;
  using C = cplx<R>;
  const IO &in;
  const C *Sp;
  uint32_t N, r, s0, S;
  // the staged slot of site, or -1
  __device__ __forceinline__ int slot(uint32_t site, bool from_lds) const {
    if constexpr (LDSS) {
      if (from_lds) {
        const int64_t d = (int64_t)site - (int64_t)s0 + 1;
        if (d >= 0 && d <= (int64_t)S + 1) return (int)d;
      }
    }
    return -1;
  }
  template <int MU, bool FWD>
  __device__ __forceinline__ void proj(C h[2][3], uint32_t site, bool from_lds) const {
    const int d = slot(site, from_lds);
    if (d >= 0) {
      const C *q = Sp + (uint32_t)d * 12 * N + r;
      project<MU, FWD>(h, [&](int c) { return q[c * N]; });
      return;
    }
    const uint32_t idx = site * N + r;
    if constexpr (IO::K == 1) {
      project<MU, FWD>(h, [&](int c) { return in.load_c(c, idx); });
    } else if constexpr (IO::K == 4) {
      const uint4 *q = reinterpret_cast<const uint4 *>(in.p);
      uint32_t w[12];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const uint4 v = q[(size_t)k * in.VN + idx];
        w[4 * k] = v.x;
        w[4 * k + 1] = v.y;
        w[4 * k + 2] = v.z;
        w[4 * k + 3] = v.w;
      }
      project_half<MU, FWD>(h, w, in.nrm[idx]);
    } else {
      C v[12];
      in.load(v, idx);
      project<MU, FWD>(h, [&](int c) { return v[c]; });
    }
  }
  __device__ __forceinline__ void full(C v[12], uint32_t site, bool from_lds) const {
    const int d = slot(site, from_lds);
    if (d >= 0) {
#pragma unroll
      for (int c = 0; c < 12; c++) v[c] = Sp[((uint32_t)d * 12 + c) * N + r];
      return;
    }
    in.load(v, site * N + r);
  }
};

// Multi-rhs operator (links from LDS): the compiler would otherwise issue the
// loads of all eight hops up front and defer the arithmetic to the epilogue
// (the accumulator is not needed before), keeping every loaded value live,
// spilling, and running one wave per SIMD.  An empty asm that reads and writes
// the accumulator (and, at PIN 2, the projected half spinor) and clobbers
// memory bounds each hop: its arithmetic is done before, and the next hop's
// loads cannot move above it, so three to four waves fit per SIMD and cover
// each other's latency.  The single-rhs operator keeps the compiler's
// schedule (all loads in flight, one wave per SIMD): it is faster there.
#ifndef CVC_HOP_PIN
#define CVC_HOP_PIN 2
#endif
template <typename C>
__device__ __forceinline__ void pin_regs(C *v, int n) {
#pragma unroll
  for (int k = 0; k < n; k++) asm volatile("" : "+v"(v[k].x), "+v"(v[k].y) : : "memory");
}

template <typename IO, int RECON, int MU, bool LDSL, bool LDSS, bool EO>
__device__ __forceinline__ void hop_pair(cplx<typename IO::real> acc[12], const NbLoader<IO, LDSS> &ld,
                                         const cplx<typename IO::real> *__restrict__ U,
                                         const cplx<typename IO::real> *__restrict__ Ul,
                                         const DParams<typename IO::real> &P, uint32_t s, uint32_t sl,
                                         int coord, int L, uint32_t stride) {
  using R = typename IO::real;
  constexpr int KL = RECON == 18 ? 9 : 6;
  // staged whole with the phase applied (see k_dslash): read as full links, no phase multiply
  constexpr bool FOLD = LDSL && RECON == 12 && sizeof(R) == 4;
  constexpr int KS = FOLD ? 9 : KL;
  constexpr int RL = FOLD ? 18 : RECON;
  cplx<R> h[2][3];
  const uint32_t sm = coord == 0 ? s + (uint32_t)(L - 1) * stride : s - stride;
  const uint32_t sp = coord == L - 1 ? s - (uint32_t)(L - 1) * stride : s + stride;
  // backward: (1 + gamma_mu) U_mu(x-mu)^dagger phi(x-mu)
  ld.template proj<MU, false>(h, EO ? sm >> 1 : sm, MU == 3);
  if constexpr (LDSL && CVC_HOP_PIN == 2) pin_regs(&h[0][0], 6);
  if constexpr (LDSL) link_apply<R, RL, MU, false, !FOLD>(acc, h, Ul + (2 * MU) * KS * LDS_SITES + sl, LDS_SITES, P.ph[MU]);
  else link_apply<R, RECON, MU, false>(acc, h, U + (size_t)MU * KL * P.V + sm, P.V, P.ph[MU]);
  if constexpr (LDSL && (CVC_HOP_PIN == 1 || CVC_HOP_PIN == 2)) pin_regs(acc, 12);
  // forward: (1 - gamma_mu) U_mu(x) phi(x+mu)
  ld.template proj<MU, true>(h, EO ? sp >> 1 : sp, MU == 3);
  if constexpr (LDSL && CVC_HOP_PIN == 2) pin_regs(&h[0][0], 6);
  if constexpr (LDSL) link_apply<R, RL, MU, true, !FOLD>(acc, h, Ul + (2 * MU + 1) * KS * LDS_SITES + sl, LDS_SITES, P.ph[MU]);
  else link_apply<R, RECON, MU, true>(acc, h, U + (size_t)MU * KL * P.V + s, P.V, P.ph[MU]);
  if constexpr (LDSL && CVC_HOP_PIN > 0) pin_regs(acc, 12);
}

// Minimum resident waves per SIMD the register allocation must allow (the
// compiler otherwise hoists every neighbour load and ends at one wave).
#ifndef CVC_DSLASH_WAVES64
#define CVC_DSLASH_WAVES64 1
#endif
#ifndef CVC_DSLASH_WAVES32
#define CVC_DSLASH_WAVES32 3
#endif
template <typename R> struct DslashWaves {
  static constexpr int value = sizeof(R) == 8 ? CVC_DSLASH_WAVES64 : CVC_DSLASH_WAVES32;
};

// LDSL: the block's sites share their links through LDS (N >= 8, at most
// LDS_SITES sites per block): every link entry crosses the texture path once
// per block instead of once per rhs lane.
// IOR: the storage of r2, the shadow residual of the <r2, out> epilogue (the
// FP32 mixed solver keeps it in FP16 with a per-site scale)
template <typename IOI, typename IOO, int RECON, bool G5OUT, int DM, bool LDSL, bool LDSS, bool EO, typename IOR = IOI>
__global__ void __launch_bounds__(256, DslashWaves<typename IOI::real>::value)
k_dslash(IOO out, IOI in, IOI own_f, const cplx<typename IOI::real> *__restrict__ U,
         DParams<typename IOI::real> P, IOR r2, double *__restrict__ part) {
  using R = typename IOI::real;
  using C = cplx<R>;
  constexpr int KL = RECON == 18 ? 9 : 6;
  static_assert(!(EO && LDSS), "the even/odd operator stages no spinors");
  // lexicographic site of traversal index v
  auto lex_of = [&](uint32_t v) -> uint32_t {
    if constexpr (!EO) {
      return v;
    } else {
      const uint32_t zl = qmod(v, P.LZt, P.shLZt);
      uint32_t q = qdiv(v, P.LZt, P.shLZt);
      const uint32_t y = qmod(q, (uint32_t)P.LY, P.shLY);
      q = qdiv(q, (uint32_t)P.LY, P.shLY);
      const uint32_t x = qmod(q, (uint32_t)P.LX, P.shLX), t = qdiv(q, (uint32_t)P.LX, P.shLX);
      return ((q * (uint32_t)P.LY + y) * (uint32_t)P.LZ) + 2u * zl + (((uint32_t)P.par + t + x + y) & 1u);
    }
  };
  const uint32_t S = blockDim.x / (uint32_t)P.N;
  uint32_t s0 = 0, s_end = 0;   // modes 0-4: the block's contiguous traversal range
  uint32_t jb = 0, cbase = 0;   // mode 5: first column-local index, column origin
  if (P.mode == 5) {
    const uint32_t k = blockIdx.x / kXcds;
    const uint32_t col = (k / P.bpcol) * kXcds + blockIdx.x % kXcds;
    jb = (k - (k / P.bpcol) * P.bpcol) * S;
    const uint32_t cz = col % P.ncz, q = col / P.ncz;
    const uint32_t cy = q % P.ncy, cx = q / P.ncy;
    cbase = ((cx * P.bx) * (uint32_t)P.LY + cy * P.yb) * P.LZt + cz * P.zb;
  } else if (P.mode == 4) {
    const uint32_t k = blockIdx.x / kXcds;
    const uint32_t col = k / P.bpcol, rem = k - col * P.bpcol;
    const uint32_t brick = rem / P.bpb, rem2 = rem - brick * P.bpb;
    const uint32_t run = rem2 / P.bpr, blk = rem2 - run * P.bpr;
    const uint32_t c = col * kXcds + blockIdx.x % kXcds;
    const uint32_t cy = c / P.ncx, cx = c - cy * P.ncx;
    const uint32_t tin = run / P.bx, xin = run - tin * P.bx;
    const uint32_t t = brick * P.bt + tin, x = cx * P.bx + xin;
    const uint32_t start = ((t * (uint32_t)P.LX + x) * (uint32_t)P.LY + cy * P.yb) * P.LZt;
    s0 = start + blk * S;
    s_end = min(s0 + S, start + P.runlen);
  } else if (P.mode == 2) {
    const uint32_t k = blockIdx.x / kXcds;
    const uint32_t per_slice = P.nrun * P.bpr;
    const uint32_t tl = k / per_slice, rem = k - tl * per_slice;
    const uint32_t run = rem / P.bpr, blk = rem - run * P.bpr;
    const uint32_t t = tl * kXcds + blockIdx.x % kXcds;
    const uint32_t ybl = run / (uint32_t)P.LX, x = run - ybl * (uint32_t)P.LX;
    const uint32_t start = ((t * (uint32_t)P.LX + x) * (uint32_t)P.LY + ybl * P.yb) * P.LZt;
    s0 = start + blk * S;
    s_end = min(s0 + S, start + P.runlen);
  } else {
    uint32_t chunk, blk;
    if (P.mode == 1) {
      const uint32_t j = blockIdx.x / kXcds;
      chunk = (j / P.bpc) * kXcds + blockIdx.x % kXcds;
      blk = j % P.bpc;
    } else {
      chunk = blockIdx.x / P.bpc;
      blk = blockIdx.x % P.bpc;
    }
    s0 = chunk * P.csz + blk * S;
    s_end = min(s0 + S, (chunk + 1) * P.csz);
  }
  // traversal index (site, or half-site for EO) of block slot sl; false past
  // the end of the block's range
  auto slot_v = [&](uint32_t sl, uint32_t &v) -> bool {
    if (P.mode == 5) {
      const uint32_t j = jb + sl;
      v = 0;
      if (j >= P.colsz) return false;
      const uint32_t rem = j & ((1u << P.lA) - 1u);
      v = cbase + (j >> P.lA) * P.sTt + (rem >> P.lyz) * P.sXt + ((rem >> P.lzb) & (P.yb - 1u)) * P.LZt +
          (rem & (P.zb - 1u));
      return true;
    }
    v = s0 + sl;
    return v < s_end;
  };
  const uint32_t sl = sdiv(threadIdx.x, P.invN);
  const uint32_t r = threadIdx.x - sl * (uint32_t)P.N;
  uint32_t vs;
  const bool active = slot_v(sl, vs) && !(P.done && P.done[(size_t)r * P.dstride]);
  const uint32_t i = vs * (uint32_t)P.N + r;
  double dv[3] = {0., 0., 0.};
  // FP32 / FP16 with 2-row links: the staged links are whole (third row rebuilt)
  // and carry the boundary phase, both done once per site here instead of once
  // per rhs lane in every hop (the multi-rhs kernels in these precisions are
  // bound by their arithmetic)
  constexpr bool FOLD = LDSL && RECON == 12 && sizeof(R) == 4;
  constexpr int KS = FOLD ? 9 : KL;
  __shared__ C Ul[LDSL ? 8 * KS * LDS_SITES : 1];
  // the link of hop hopi (2 mu: backward, U_mu(x - mu); 2 mu + 1: forward) of lexicographic site
  auto link_site = [&](uint32_t hopi, uint32_t site) -> uint32_t {
    if (hopi & 1u) return site;
    const int mu = (int)(hopi >> 1);
    const uint32_t str = mu == 0 ? P.sT : mu == 1 ? P.sX : mu == 2 ? (uint32_t)P.LZ : 1u;
    const int shs = mu == 0 ? P.shST : mu == 1 ? P.shSX : mu == 2 ? P.shLZ : 0;
    const uint32_t L = (uint32_t)(mu == 0 ? P.T : mu == 1 ? P.LX : mu == 2 ? P.LY : P.LZ);
    const int shl = mu == 0 ? -1 : mu == 1 ? P.shLX : mu == 2 ? P.shLY : P.shLZ;
    const uint32_t crd = qmod(qdiv(site, str, shs), L, shl);
    return crd == 0 ? site + (L - 1) * str : site - str;
  };
  if constexpr (FOLD) {
    for (uint32_t e = threadIdx.x; e < 8u * S; e += blockDim.x) {
      const uint32_t hopi = sdiv(e, P.invS), sle = e - hopi * S;
      uint32_t ve;
      if (!slot_v(sle, ve)) continue;
      const uint32_t ls = link_site(hopi, lex_of(ve));
      const int mu = (int)(hopi >> 1);
      C u[9];
#pragma unroll
      for (int k = 0; k < 6; k++) u[k] = U[((size_t)mu * 6 + k) * P.V + ls];
      rebuild_row3(u);
#pragma unroll
      for (int k = 0; k < 9; k++) Ul[(hopi * 9 + k) * LDS_SITES + sle] = cmul(u[k], P.ph[mu]);
    }
    __syncthreads();
  } else if constexpr (LDSL) {
    const uint32_t tot = 8u * KL * S;
    for (uint32_t e = threadIdx.x; e < tot; e += blockDim.x) {
      const uint32_t hopi = sdiv(e, P.invKS), rem = e - hopi * (KL * S);
      const uint32_t k = sdiv(rem, P.invS), sle = rem - k * S;
      uint32_t ve;
      if (!slot_v(sle, ve)) continue;
      const uint32_t site = lex_of(ve);
      const int mu = (int)(hopi >> 1);
      uint32_t ls = site;
      if ((hopi & 1u) == 0) {          // backward hop: link at site - mu
        const uint32_t str = mu == 0 ? P.sT : mu == 1 ? P.sX : mu == 2 ? (uint32_t)P.LZ : 1u;
        const int shs = mu == 0 ? P.shST : mu == 1 ? P.shSX : mu == 2 ? P.shLZ : 0;
        const uint32_t L = (uint32_t)(mu == 0 ? P.T : mu == 1 ? P.LX : mu == 2 ? P.LY : P.LZ);
        const int shl = mu == 0 ? -1 : mu == 1 ? P.shLX : mu == 2 ? P.shLY : P.shLZ;
        const uint32_t crd = qmod(qdiv(site, str, shs), L, shl);
        ls = crd == 0 ? site + (L - 1) * str : site - str;
      }
      Ul[(hopi * KL + k) * LDS_SITES + sle] = U[((size_t)mu * KL + k) * P.V + ls];
    }
    __syncthreads();
  }
  extern __shared__ unsigned char dyn_lds[];
  C *Sp = reinterpret_cast<C *>(dyn_lds);
  if constexpr (LDSS) {
    // stage sites s0-1 .. s0+S (all rhs); (s0 - 1 + k) mod V, every index is a valid site
    const uint32_t nst = (S + 2) * (uint32_t)P.N;
    for (uint32_t u = threadIdx.x; u < nst; u += blockDim.x) {
      const uint32_t k = u / (uint32_t)P.N, rr = u - k * (uint32_t)P.N;
      const uint32_t site = (s0 + P.V - 1 + k) % P.V;
      C v[12];
      in.load(v, site * (uint32_t)P.N + rr);
#pragma unroll
      for (int c = 0; c < 12; c++) Sp[(k * 12 + c) * (uint32_t)P.N + rr] = v[c];
    }
    __syncthreads();
  }
  const NbLoader<IOI, LDSS> ld{in, Sp, (uint32_t)P.N, r, s0, S};
  if (active) {
    const uint32_t s = lex_of(vs);
    const uint32_t t = qdiv(s, P.sT, P.shST);
    uint32_t rem = s - t * P.sT;
    const uint32_t x = qdiv(rem, P.sX, P.shSX);
    rem -= x * P.sX;
    const uint32_t y = qdiv(rem, (uint32_t)P.LZ, P.shLZ);
    const uint32_t z = rem - y * (uint32_t)P.LZ;

    C acc[12];
#pragma unroll
    for (int c = 0; c < 12; c++) acc[c] = mk<R>(0, 0);
#ifndef CVC_HOP_ORDER
#define CVC_HOP_ORDER 0
#endif
#define CVC_HOP(MU, CRD, L, STR) hop_pair<IOI, RECON, MU, LDSL, LDSS, EO>(acc, ld, U, Ul, P, s, sl, (int)(CRD), L, STR)
#if CVC_HOP_ORDER == 0
    CVC_HOP(0, t, P.T, P.sT); CVC_HOP(1, x, P.LX, P.sX); CVC_HOP(2, y, P.LY, (uint32_t)P.LZ); CVC_HOP(3, z, P.LZ, 1u);
#elif CVC_HOP_ORDER == 1
    CVC_HOP(1, x, P.LX, P.sX); CVC_HOP(2, y, P.LY, (uint32_t)P.LZ); CVC_HOP(3, z, P.LZ, 1u); CVC_HOP(0, t, P.T, P.sT);
#elif CVC_HOP_ORDER == 2
    CVC_HOP(0, t, P.T, P.sT); CVC_HOP(3, z, P.LZ, 1u); CVC_HOP(2, y, P.LY, (uint32_t)P.LZ); CVC_HOP(1, x, P.LX, P.sX);
#else
    CVC_HOP(0, t, P.T, P.sT); CVC_HOP(1, x, P.LX, P.sX); CVC_HOP(3, z, P.LZ, 1u); CVC_HOP(2, y, P.LY, (uint32_t)P.LZ);
#endif
#undef CVC_HOP

    // the shadow residual of the <r2, out> epilogue is read once per launch:
    // streaming loads, issued beside the own spinor's (the latencies overlap)
    C w2[12];
    if constexpr (DM == DOT_R2) r2.load_nt(w2, i);
    C own[12];
    if constexpr (!EO) {
      ld.full(own, s, true);
      const R mh = (R)-0.5;
#pragma unroll
      for (int c = 0; c < 12; c++) {
        // -1/2 hop + phi/(2 kappa) + i mutm gamma5 phi  (gamma5 = +1 on spins 0,1, -1 on 2,3)
        const R m = c < 6 ? P.mutm : -P.mutm;
        C o;
        o.x = fmar(mh, acc[c].x, fmar(P.k2, own[c].x, -(m * own[c].y)));
        o.y = fmar(mh, acc[c].y, fmar(P.k2, own[c].y, m * own[c].x));
        if (G5OUT && c >= 6) { o.x = -o.x; o.y = -o.y; }
        acc[c] = o;
      }
    } else {
      if (P.dmode != 1) own_f.load(own, i);
#pragma unroll
      for (int c = 0; c < 12; c++) {
        const R m = c < 6 ? P.mutm : -P.mutm;
        C o;
        if (P.dmode == 0) {
          o.x = fmar(P.hopc, acc[c].x, fmar(P.k2, own[c].x, -(m * own[c].y)));
          o.y = fmar(P.hopc, acc[c].y, fmar(P.k2, own[c].y, m * own[c].x));
        } else {
          C h = mk<R>(P.hopc * acc[c].x, P.hopc * acc[c].y);
          if (P.dmode >= 2) h = mk<R>(fmar(P.hopc, acc[c].x, own[c].x), fmar(P.hopc, acc[c].y, own[c].y));
          if (P.dmode == 2) {
            o = h;
          } else {
            // A^-1 h = (1/(2 kappa) - i mutm gamma5) h / ((1/(2 kappa))^2 + mutm^2)
            o.x = fmar(P.k2, h.x, m * h.y) * P.dinv;
            o.y = fmar(P.k2, h.y, -(m * h.x)) * P.dinv;
          }
        }
        if (G5OUT && c >= 6) { o.x = -o.x; o.y = -o.y; }
        acc[c] = o;
      }
    }
    if (P.nt) out.store_nt(acc, i);
    else out.store(acc, i);
    if constexpr (DM == DOT_TS) {
      // <out, in> = sum conj(out) in ; <out, out>   (or <out, r2> when a partner is given)
      if (P.partner) r2.load(own, i);
      // products in FP64: squares of a small-scale rhs underflow FP32
      double a = 0, b = 0, n = 0;
#pragma unroll
      for (int c = 0; c < 12; c++) {
        const double ox = acc[c].x, oy = acc[c].y, qx = own[c].x, qy = own[c].y;
        a += ox * qx + oy * qy;
        b += ox * qy - oy * qx;
        n += ox * ox + oy * oy;
      }
      dv[0] = a; dv[1] = b; dv[2] = n;
    } else if constexpr (DM == DOT_R2) {
      double a = 0, b = 0;
#pragma unroll
      for (int c = 0; c < 12; c++) {
        const double ox = acc[c].x, oy = acc[c].y, qx = w2[c].x, qy = w2[c].y;
        a += qx * ox + qy * oy;
        b += qx * oy - qy * ox;
      }
      dv[0] = a; dv[1] = b;
    } else if constexpr (DM == DOT_NN) {
      double n = 0;
#pragma unroll
      for (int c = 0; c < 12; c++) n += (double)acc[c].x * acc[c].x + (double)acc[c].y * acc[c].y;
      dv[0] = n;
    }
  }
  if constexpr (DM == DOT_TS) block_reduce_store<3>(dv, part, P.N);
  else if constexpr (DM == DOT_R2) block_reduce_store<2>(dv, part, P.N);
  else if constexpr (DM == DOT_NN) block_reduce_store<1>(dv, part, P.N);
}

// ------------------------------------------------------- t-marching operator
//
// Multi-rhs operator on the whole lattice in which each workgroup owns S
// columns (sites of one (x, y, z) cross-section, all N rhs) and walks them
// through M consecutive time slices.  The backward t hop of slice t+1,
// (1 + gamma_0) U_0(t)^dagger phi(t), is formed from the own spinor phi(t)
// while slice t is processed and carried to the next step in registers as the
// six complex numbers it depends on, so the t neighbours are not read again
// from memory: per step a thread reads phi(t) (its own, which the same thread
// read one step before as the forward neighbour, so it comes from L2), the six
// spatial neighbours and phi(t+1).  Columns are grouped in patches of
// XP x YP x LZ sites; XCD j takes the patches p = j mod 8, and the workgroups
// of one XCD run the chunks of one patch over one range of M slices side by
// side, so the spatial neighbours of a chunk are read by neighbouring
// workgroups of the same XCD at about the same step.  The links of a step
// (U_0(t), U_k(t), U_k(t - k) for k = x, y, z) are staged in LDS once per
// workgroup and shared by the N rhs lanes of a column.
//
// Accumulation: acc starts at -2 (1/(2 kappa) + i mutm gamma5) phi(t), the
// eight hops are added, and the output is -acc/2 (the factors -2 and -1/2 are
// exact), i.e. phi/(2 kappa) + i mutm gamma5 phi - hop/2 as in Q_phi.c:58-174
// with a different rounding order.

// w_a = link row . h_a for the three rows (w[a][i]), the boundary phase
// applied (PHASE) as in link_apply; expand_w adds the hop's four spin rows.
template <typename R, int RECON, bool FWD, bool PHASE>
__device__ __forceinline__ void link_mul(cplx<R> w[2][3], const cplx<R> h[2][3], const cplx<R> *__restrict__ base,
                                         uint32_t stride, cplx<R> ph) {
  using C = cplx<R>;
  C u[9];
  if constexpr (RECON == 12) {
#pragma unroll
    for (int k = 0; k < 6; k++) u[k] = base[(size_t)k * stride];
    rebuild_row3(u);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    C w0 = mk<R>(0, 0), w1 = mk<R>(0, 0);
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int k = FWD ? 3 * i + j : 3 * j + i;
      const C e = RECON == 18 ? base[(size_t)k * stride] : u[k];
      if constexpr (FWD) {
        w0 = cfma(w0, e, h[0][j]);
        w1 = cfma(w1, e, h[1][j]);
      } else {
        w0 = cfmac(w0, e, h[0][j]);
        w1 = cfmac(w1, e, h[1][j]);
      }
    }
    if constexpr (!PHASE) {
    } else if constexpr (FWD) {
      w0 = cmul(ph, w0);
      w1 = cmul(ph, w1);
    } else {
      w0 = cmulc(ph, w0);
      w1 = cmulc(ph, w1);
    }
    w[0][i] = w0;
    w[1][i] = w1;
  }
}

template <typename C, int MU, bool FWD>
__device__ __forceinline__ void expand_w(C acc[12], const C w[2][3]) {
  constexpr int q2 = GCOL[MU][2], q3 = GCOL[MU][3];
  constexpr int l2 = FWD ? uneg(GVAL[MU][2]) : GVAL[MU][2];
  constexpr int l3 = FWD ? uneg(GVAL[MU][3]) : GVAL[MU][3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    acc[i] = cadd(acc[i], w[0][i]);
    acc[3 + i] = cadd(acc[3 + i], w[1][i]);
    acc[6 + i] = cadd(acc[6 + i], cunit<l2>(q2 == 0 ? w[0][i] : w[1][i]));
    acc[9 + i] = cadd(acc[9 + i], cunit<l3>(q3 == 0 ? w[0][i] : w[1][i]));
  }
}

// the links of one step, staged per column: slot 0 U_0(t), slots 1-3 U_k(t),
// slots 4-6 U_k(t - k)
constexpr int kMarchSlots = 7;

#ifndef CVC_MARCH_WAVES64
#define CVC_MARCH_WAVES64 3
#endif
#ifndef CVC_MARCH_WAVES32
#define CVC_MARCH_WAVES32 4
#endif
template <typename R> struct MarchWaves {
  static constexpr int value = sizeof(R) == 8 ? CVC_MARCH_WAVES64 : CVC_MARCH_WAVES32;
};

template <typename IOI, typename IOO, int RECON, bool G5OUT, int DM, typename IOR = IOI>
__global__ void __launch_bounds__(256, MarchWaves<typename IOI::real>::value)
k_dslash_march(IOO out, IOI in, const cplx<typename IOI::real> *__restrict__ U, DParams<typename IOI::real> P,
               IOR r2, double *__restrict__ part) {
  using R = typename IOI::real;
  using C = cplx<R>;
  constexpr int KL = RECON == 18 ? 9 : 6;
  // FP32 with 2-row links: staged whole with the phase applied (as in k_dslash)
  constexpr bool FOLD = RECON == 12 && sizeof(R) == 4;
  constexpr int KS = FOLD ? 9 : KL;
  constexpr int RL = FOLD ? 18 : RECON;
  const uint32_t S = P.mS;
  // workgroup -> (patch, range of slices, chunk of the patch's columns)
  uint32_t k = blockIdx.x, xcd = 0;
  if (P.mXcd) {
    xcd = blockIdx.x % kXcds;
    k = blockIdx.x / kXcds;
  }
  const uint32_t ch = k % P.mCP, q = k / P.mCP;
  const uint32_t rng = q % P.mNR, pl = q / P.mNR;
  const uint32_t patch = P.mXcd ? pl * kXcds + xcd : pl;
  const uint32_t px = patch % P.mNPX, py = patch / P.mNPX;
  const uint32_t r = threadIdx.x % (uint32_t)P.N, cl = threadIdx.x / (uint32_t)P.N;
  // column cl of the chunk: z fastest, then y, then x inside the patch
  auto column = [&](uint32_t c, uint32_t &sxyz) -> bool {
    const uint32_t j = ch * S + c;
    if (c >= S || j >= P.mXP * P.mYP * (uint32_t)P.LZ) return false;
    const uint32_t z = j % (uint32_t)P.LZ, j2 = j / (uint32_t)P.LZ;
    const uint32_t y = py * P.mYP + j2 % P.mYP, x = px * P.mXP + j2 / P.mYP;
    sxyz = (x * (uint32_t)P.LY + y) * (uint32_t)P.LZ + z;
    return true;
  };
  uint32_t sxyz = 0;
  const bool active = column(cl, sxyz) && !(P.done && P.done[(size_t)r * P.dstride]);
  uint32_t cx = 0, cy = 0, cz = 0;
  if (active) {
    cz = sxyz % (uint32_t)P.LZ;
    cy = (sxyz / (uint32_t)P.LZ) % (uint32_t)P.LY;
    cx = sxyz / P.sX;
  }
  const uint32_t t0 = rng * P.mM;
  extern __shared__ unsigned char dyn_lds[];
  C *Ul = reinterpret_cast<C *>(dyn_lds);   // [kMarchSlots][KS][S]
  const NbLoader<IOI, false> ld{in, nullptr, (uint32_t)P.N, r, 0, 0};
  double dv[3] = {0., 0., 0.};

  // carry: (U_0^dagger (1 + gamma_0) phi)(t - 1) as the six complex numbers w
  C carry[2][3];
  if (active) {
    const uint32_t tm = t0 == 0 ? (uint32_t)P.T - 1 : t0 - 1;
    const uint32_t sm = tm * P.sT + sxyz;
    C h[2][3];
    ld.template proj<0, false>(h, sm, false);
    link_mul<R, RECON, false, true>(carry, h, U + sm, P.V, P.ph[0]);
  }

  for (uint32_t step = 0; step < P.mM; step++) {
    const uint32_t t = t0 + step;
    __syncthreads();   // the previous step's links are no longer read
    for (uint32_t e = threadIdx.x; e < (uint32_t)kMarchSlots * S; e += blockDim.x) {
      const uint32_t slot = e / S, c = e - slot * S;
      uint32_t cs;
      if (!column(c, cs)) continue;
      const int mu = slot < 4 ? (int)slot : (int)slot - 3;
      uint32_t ls = t * P.sT + cs;
      if (slot >= 4) {   // U_k(t - k)
        const uint32_t str = mu == 1 ? P.sX : mu == 2 ? (uint32_t)P.LZ : 1u;
        const uint32_t L = (uint32_t)(mu == 1 ? P.LX : mu == 2 ? P.LY : P.LZ);
        const uint32_t crd = (cs / str) % L;
        ls = crd == 0 ? ls + (L - 1) * str : ls - str;
      }
      if constexpr (FOLD) {
        C u[9];
#pragma unroll
        for (int kk = 0; kk < 6; kk++) u[kk] = U[((size_t)mu * 6 + kk) * P.V + ls];
        rebuild_row3(u);
#pragma unroll
        for (int kk = 0; kk < 9; kk++) Ul[(slot * 9 + kk) * S + c] = cmul(u[kk], P.ph[mu]);
      } else {
#pragma unroll
        for (int kk = 0; kk < KL; kk++) Ul[(slot * KL + kk) * S + c] = U[((size_t)mu * KL + kk) * P.V + ls];
      }
    }
    __syncthreads();
    if (!active) continue;
    const uint32_t s = t * P.sT + sxyz;
    const uint32_t i = s * (uint32_t)P.N + r;
    // the shadow residual of the <r2, out> epilogue is read once per launch:
    // streaming loads, issued with the own spinor's
    C w2[12];
    if constexpr (DM == DOT_R2) r2.load_nt(w2, i);
    C acc[12];
    {
      C own[12];
      ld.full(own, s, false);
      // acc = -2 (1/(2 kappa) + i mutm gamma5) phi   (gamma5 = +1 on spins 0,1, -1 on 2,3)
#pragma unroll
      for (int c = 0; c < 12; c++) {
        const R m = c < 6 ? P.mutm : -P.mutm;
        acc[c].x = (R)-2 * fmar(P.k2, own[c].x, -(m * own[c].y));
        acc[c].y = (R)-2 * fmar(P.k2, own[c].y, m * own[c].x);
      }
      // backward t hop from the carry, then the carry for slice t+1 from phi(t)
      expand_w<C, 0, false>(acc, carry);
      C h[2][3];
      project<0, false>(h, [&](int c) { return own[c]; });
      link_mul<R, RL, false, !FOLD>(carry, h, Ul + cl, S, P.ph[0]);
    }
    pin_regs(acc, 12);
    pin_regs(&carry[0][0], 6);
    // spatial hops
#define CVC_MHOP(MU, CRD, L, STR)                                                                              \
    {                                                                                                          \
      const uint32_t sm_ = (CRD) == 0 ? s + (uint32_t)((L) - 1) * (STR) : s - (STR);                          \
      const uint32_t sp_ = (CRD) == (uint32_t)(L) - 1 ? s - (uint32_t)((L) - 1) * (STR) : s + (STR);           \
      C h[2][3], w[2][3];                                                                                      \
      ld.template proj<MU, false>(h, sm_, false);                                                              \
      pin_regs(&h[0][0], 6);                                                                                   \
      link_mul<R, RL, false, !FOLD>(w, h, Ul + (3 + MU) * KS * S + cl, S, P.ph[MU]);                           \
      expand_w<C, MU, false>(acc, w);                                                                          \
      pin_regs(acc, 12);                                                                                       \
      ld.template proj<MU, true>(h, sp_, false);                                                               \
      pin_regs(&h[0][0], 6);                                                                                   \
      link_mul<R, RL, true, !FOLD>(w, h, Ul + (MU)*KS * S + cl, S, P.ph[MU]);                                  \
      expand_w<C, MU, true>(acc, w);                                                                           \
      pin_regs(acc, 12);                                                                                       \
    }
    CVC_MHOP(1, cx, P.LX, P.sX);
    CVC_MHOP(2, cy, P.LY, (uint32_t)P.LZ);
    CVC_MHOP(3, cz, P.LZ, 1u);
#undef CVC_MHOP
    // forward t hop: (1 - gamma_0) U_0(t) phi(t+1), last, so that phi(t+1) is
    // still in L2 when the next step reads it as its own spinor
    {
      const uint32_t sp = t == (uint32_t)P.T - 1 ? sxyz : s + P.sT;
      C h[2][3], w[2][3];
      ld.template proj<0, true>(h, sp, false);
      pin_regs(&h[0][0], 6);
      link_mul<R, RL, true, !FOLD>(w, h, Ul + cl, S, P.ph[0]);
      expand_w<C, 0, true>(acc, w);
    }
    const R mh = (R)-0.5;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      acc[c].x *= mh;
      acc[c].y *= mh;
      if (G5OUT && c >= 6) { acc[c].x = -acc[c].x; acc[c].y = -acc[c].y; }
    }
    if (P.nt) out.store_nt(acc, i);
    else out.store(acc, i);
    if constexpr (DM == DOT_TS) {
      C own[12];
      if (P.partner) r2.load(own, i);
      else in.load(own, i);
      double a = 0, b = 0, n = 0;
#pragma unroll
      for (int c = 0; c < 12; c++) {
        const double ox = acc[c].x, oy = acc[c].y, qx = own[c].x, qy = own[c].y;
        a += ox * qx + oy * qy;
        b += ox * qy - oy * qx;
        n += ox * ox + oy * oy;
      }
      dv[0] += a; dv[1] += b; dv[2] += n;
    } else if constexpr (DM == DOT_R2) {
      double a = 0, b = 0;
#pragma unroll
      for (int c = 0; c < 12; c++) {
        const double ox = acc[c].x, oy = acc[c].y, qx = w2[c].x, qy = w2[c].y;
        a += qx * ox + qy * oy;
        b += qx * oy - qy * ox;
      }
      dv[0] += a; dv[1] += b;
    } else if constexpr (DM == DOT_NN) {
      double n = 0;
#pragma unroll
      for (int c = 0; c < 12; c++) n += (double)acc[c].x * acc[c].x + (double)acc[c].y * acc[c].y;
      dv[0] += n;
    }
  }
  if constexpr (DM == DOT_TS) block_reduce_store<3>(dv, part, P.N);
  else if constexpr (DM == DOT_R2) block_reduce_store<2>(dv, part, P.N);
  else if constexpr (DM == DOT_NN) block_reduce_store<1>(dv, part, P.N);
}

// ------------------------------------------------------------- launcher

static int log2_exact(uint32_t v) {   // -1 unless v is a power of two
  if (v == 0 || (v & (v - 1))) return -1;
  int l = 0;
  while ((1u << l) < v) l++;
  return l;
}

template <typename R>
static DParams<R> make_params(const cvchip_ctx *c, int N, double mutm) {
  DParams<R> P;
  P.V = (uint32_t)c->lat.V;
  P.N = N;
  P.VN = (uint32_t)(c->lat.V * N);
  P.T = c->lat.T;
  P.LX = c->lat.LX;
  P.LY = c->lat.LY;
  P.LZ = c->lat.LZ;
  P.sX = (uint32_t)(c->lat.LY * c->lat.LZ);
  P.sT = (uint32_t)(c->lat.LX * c->lat.LY * c->lat.LZ);
  P.shLZ = log2_exact((uint32_t)c->lat.LZ);
  P.shLY = log2_exact((uint32_t)c->lat.LY);
  P.shLX = log2_exact((uint32_t)c->lat.LX);
  P.shSX = log2_exact(P.sX);
  P.shST = log2_exact(P.sT);
  P.shLZt = P.shLZ;
  P.invN = 1.0f / (float)N;
  P.invS = P.invKS = 0.f;
  P.k2 = (R)(0.5 / c->kappa);
  P.mutm = (R)mutm;
  for (int m = 0; m < 4; m++) P.ph[m] = mk<R>((R)c->ph.re[m], (R)c->ph.im[m]);
  return P;
}

template <typename IO> static IO make_io(const Field &f);
template <> IOPlain<double> make_io<IOPlain<double>>(const Field &f) {
  return IOPlain<double>{(double2 *)f.data, (uint32_t)(f.V * f.nrhs)};
}
template <> IOPlain<float> make_io<IOPlain<float>>(const Field &f) {
  return IOPlain<float>{(float2 *)f.data, (uint32_t)(f.V * f.nrhs)};
}
template <> IOHalf make_io<IOHalf>(const Field &f) {
  return IOHalf{(__half2 *)f.data, f.norm, (uint32_t)(f.V * f.nrhs)};
}

struct Traversal {
  uint32_t csz, bpc, nblk;
  int mode;
  uint32_t yb, runlen, bpr, nrun;
  uint32_t bt, bx, ncx, bpb, bpcol;
  uint32_t zb, ncy, ncz, lzb, lyz, lA, colsz;
  uint32_t mS, mM, mXP, mYP, mCP, mNR, mNPX;
  int mXcd;
  unsigned threads;   // workgroup size
};

// t-march geometry (mode 6) for N rhs on the whole lattice L; bits 8-15 of the
// policy give the march length M (0: 16; the largest divisor of T not above
// it is used), bits 16-19 and 20-23 the patch's x and y extents (0: 8; a
// value that does not divide LX or LY falls back to 1).
static bool march_geometry(const cvchip_ctx *c, int N, Traversal &t, bool explicit_policy, uint32_t M_auto) {
  const Lattice &L = c->lat;
  if (N < 1 || N > 192) return false;
  uint32_t S = 192u / (uint32_t)N;
  if (S > 48) S = 48;
  const int pol = explicit_policy ? c->dslash_policy : 0;
  uint32_t M = (uint32_t)((pol >> 8) & 255);
  if (M == 0) M = explicit_policy ? 16 : M_auto;
  while (M > 1 && L.T % (int)M) M--;
  uint32_t XP = (uint32_t)((pol >> 16) & 15), YP = (uint32_t)((pol >> 20) & 15);
  if (XP == 0) XP = 8;
  if (YP == 0) YP = 8;
  if (L.LX % (int)XP) XP = 1;
  if (L.LY % (int)YP) YP = 1;
  t.mode = 6;
  t.mS = S;
  t.mM = M;
  t.mXP = XP;
  t.mYP = YP;
  t.mNPX = (uint32_t)L.LX / XP;
  const uint32_t npatch = t.mNPX * ((uint32_t)L.LY / YP);
  t.mCP = (XP * YP * (uint32_t)L.LZ + S - 1) / S;
  t.mNR = (uint32_t)L.T / M;
  t.mXcd = npatch % kXcds == 0 ? 1 : 0;
  t.nblk = npatch * t.mNR * t.mCP;
  t.threads = S * (uint32_t)N;
  return true;
}


// Block-to-site mapping; the policy's bits 4-7 pick the mode (0: automatic),
// bits 8-15 the y-block height of modes 2, 4 and 5 (0: automatic), bits 16-19
// and 20-23 the brick's t and x extents of mode 4 (0: 4); for mode 5, bits
// 16-19 are log2 of the column's z extent (0: 8) and bits 20-23 its x extent
// (0: 4).
// K: complexes per 16 bytes of the operator's spinors (1 FP64, 2 FP32, 4
// FP16+norm).  The FP32 operator's automatic traversal is the t march with
// marches of 4 slices for 8 <= N <= 20 (at 64x32^3, per 12 rhs: 1.72-1.80 ms
// against 1.92-2.06 ms for the best bricks with 12 rhs, 1.74 against 1.91 ms
// with 16) and bricks beyond (24 rhs: 1.83 ms against 1.93 ms for the march);
// the FP64 and FP16 operators keep the bricks
static Traversal traversal(const cvchip_ctx *c, int N, bool eo = false, int K = 1) {
  Lattice L = c->lat;
  if (eo) {   // the half lattice of one parity: LZ/2 half-sites per z line
    L.LZ /= 2;
    L.V /= 2;
  }
  const uint32_t V = (uint32_t)L.V, slice = (uint32_t)(L.LX * L.LY * L.LZ);
  const uint32_t S = (uint32_t)(block_for(N) / N);
  int want = (c->dslash_policy >> 4) & 15;
  if (want == 6) {
    Traversal t{};
    if (!eo && march_geometry(c, N, t, true, 16)) return t;
    want = 4;   // the even/odd operator does not march
  }
  if (want == 0 && K == 2 && !eo && N >= 8 && N <= 20) {
    Traversal t{};
    if (march_geometry(c, N, t, false, 4)) return t;
  }
  uint32_t yb = (uint32_t)((c->dslash_policy >> 8) & 255);
  // automatic: bricks of 4 slices x 4 x-planes x yb y-rows for the multi-rhs
  // operator, else one slice per XCD in y blocks (both measured at 64x32^3
  // with 12 rhs: 3.98 and 4.19 ms against 4.75 ms for XCD chunks of a
  // slice); XCD chunks of a slice for a single rhs (0.48 ms against 0.54).
  // The brick's y extent shrinks as N grows so that the spinor bytes of a
  // brick's x-y cross-section stay about constant, yb * N <= 96 K: measured at
  // 64x32^3, FP64 12 rhs best at yb 8, 24 rhs at 4, 48 rhs at 2; FP32 24 rhs
  // at 8, 48 rhs at 4 (1.86 ms per 12 rhs against 1.96-1.99 ms at yb 2); FP16
  // 12 rhs at 16 (1.18 ms against 1.20 ms at yb 8).
  const bool autom = want == 0;
  if (autom) want = N >= 8 ? 4 : 1;
  if (yb == 0) {
    yb = 1;
    if (autom && want == 4) {
      const uint32_t budget = 96u * (uint32_t)K;
      for (uint32_t y : {16u, 8u, 4u, 2u})
        if (y * (uint32_t)N <= budget && L.LY % (int)y == 0) { yb = y; break; }
    } else {
      for (uint32_t y : {16u, 8u, 4u, 2u})
        if (L.LY % (int)y == 0) { yb = y; break; }
    }
  }
  Traversal t{};
  t.threads = (unsigned)block_for(N);
  if (want == 5) {
    const uint32_t cyb = (c->dslash_policy >> 8) & 255 ? yb : 4u;
    const uint32_t lz = (c->dslash_policy >> 16) & 15;
    const uint32_t czb = lz ? (1u << lz) : 8u;
    const uint32_t cbx = (c->dslash_policy >> 20) & 15 ? (c->dslash_policy >> 20) & 15 : 4u;
    const int l1 = log2_exact(cbx), l2 = log2_exact(cyb), l3 = log2_exact(czb);
    if (l1 >= 0 && l2 >= 0 && l3 >= 0 && L.LX % (int)cbx == 0 && L.LY % (int)cyb == 0 && L.LZ % (int)czb == 0) {
      const uint32_t ncol = (uint32_t)(L.LX / (int)cbx) * (uint32_t)(L.LY / (int)cyb) * (uint32_t)(L.LZ / (int)czb);
      if (ncol % kXcds == 0) {
        t.mode = 5;
        t.bx = cbx;
        t.yb = cyb;
        t.zb = czb;
        t.ncy = (uint32_t)L.LY / cyb;
        t.ncz = (uint32_t)L.LZ / czb;
        t.lzb = (uint32_t)l3;
        t.lyz = (uint32_t)(l2 + l3);
        t.lA = (uint32_t)(l1 + l2 + l3);
        t.colsz = (uint32_t)L.T << t.lA;
        t.bpcol = (t.colsz + S - 1) / S;
        t.nblk = ncol * t.bpcol;
        return t;
      }
    }
    want = 4;   // a column request that does not tile the lattice falls back to bricks
  }
  uint32_t bt = (uint32_t)((c->dslash_policy >> 16) & 15), bx = (uint32_t)((c->dslash_policy >> 20) & 15);
  if (bt == 0) bt = 4;
  if (bx == 0) bx = 4;
  if (want == 4 && L.T % (int)bt == 0 && L.LX % (int)bx == 0 && L.LY % (int)yb == 0 &&
      ((uint32_t)(L.LX / (int)bx) * (uint32_t)(L.LY / (int)yb)) % kXcds == 0) {
    t.mode = 4;
    t.yb = yb;
    t.bt = bt;
    t.bx = bx;
    t.ncx = (uint32_t)L.LX / bx;
    t.runlen = yb * (uint32_t)L.LZ;
    t.bpr = (t.runlen + S - 1) / S;
    t.bpb = bt * bx * t.bpr;
    t.bpcol = (uint32_t)L.T / bt * t.bpb;
    const uint32_t ncol = t.ncx * ((uint32_t)L.LY / yb);
    t.nblk = ncol * t.bpcol;
    return t;
  }
  // (a brick request that does not tile the lattice falls back to mode 2)
  if ((want == 2 || want == 4) && L.T % (int)kXcds == 0 && L.LY % (int)yb == 0) {
    t.mode = 2;
    t.yb = yb;
    t.runlen = yb * (uint32_t)L.LZ;
    t.bpr = (t.runlen + S - 1) / S;
    t.nrun = (uint32_t)(L.LY / (int)yb) * (uint32_t)L.LX;
    t.nblk = (uint32_t)L.T * t.nrun * t.bpr;
    return t;
  }
  if (slice % kXcds == 0) { t.csz = slice / kXcds; t.mode = 1; }
  else if (V % kXcds == 0) { t.csz = V / kXcds; t.mode = 1; }
  else { t.csz = V; t.mode = 0; }
  if (want == 3) t.mode = 0;
  t.bpc = (t.csz + S - 1) / S;
  t.nblk = (V / t.csz) * t.bpc;
  return t;
}

unsigned dslash_blocks(const cvchip_ctx *c, int N) {
  unsigned m = 0;
  for (int K : {1, 2, 4}) {
    m = std::max(m, traversal(c, N, false, K).nblk);
    if (c->lat.LZ % 2 == 0) m = std::max(m, traversal(c, N, true, K).nblk);
  }
  return m;
}

struct EOArgs {
  const Field *own;
  int par, dmode;
  double hopc;
};

// IOR differs from IOI only for the <r2, out> epilogue (DOT_R2) with an FP16 r2
template <typename IOI, typename IOO, int RECON, bool EO, typename IOR = IOI>
static int launch(cvchip_ctx *c, Field &out, const Field &in, double mutm, int g5, DotMode dm,
                  const Field *r2, int *nblk_out, const EOArgs *eo = nullptr) {
  constexpr bool kSameR = std::is_same<IOR, IOI>::value;
  if (!kSameR && (dm != DOT_R2 || !r2)) {
    set_error("operator: a shadow residual of another precision serves only the <r2, out> epilogue");
    return CVCHIP_ERR_ARG;
  }
  using R = typename IOI::real;
  const int N = in.nrhs;
  DParams<R> P = make_params<R>(c, N, mutm);
  P.partner = (dm == DOT_TS && r2 != nullptr) ? 1 : 0;
  P.LZt = EO ? (uint32_t)c->lat.LZ / 2 : (uint32_t)c->lat.LZ;
  P.shLZt = log2_exact(P.LZt);
  P.par = eo ? eo->par : 0;
  P.dmode = eo ? eo->dmode : 0;
  P.hopc = (R)(eo ? eo->hopc : -0.5);
  P.dinv = (R)(1.0 / (0.25 / (c->kappa * c->kappa) + mutm * mutm));
  const int B = block_for(N);
  {
    const int S = B / N, KLs = RECON == 18 ? 9 : 6;   // sites per block, link entries per hop
    P.invS = 1.0f / (float)S;
    P.invKS = 1.0f / (float)(KLs * S);
  }
  const Traversal tr = traversal(c, N, EO, IOI::K);
  P.csz = tr.csz;
  P.bpc = tr.bpc;
  P.mode = tr.mode;
  P.yb = tr.yb;
  P.runlen = tr.runlen;
  P.bpr = tr.bpr;
  P.nrun = tr.nrun;
  P.bt = tr.bt;
  P.bx = tr.bx;
  P.ncx = tr.ncx;
  P.bpb = tr.bpb;
  P.bpcol = tr.bpcol;
  P.zb = tr.zb;
  P.ncy = tr.ncy;
  P.ncz = tr.ncz;
  P.lzb = tr.lzb;
  P.lyz = tr.lyz;
  P.lA = tr.lA;
  P.colsz = tr.colsz;
  P.sXt = (uint32_t)c->lat.LY * P.LZt;
  P.sTt = (uint32_t)c->lat.LX * P.sXt;
  P.mS = tr.mS;
  P.mM = tr.mM;
  P.mXP = tr.mXP;
  P.mYP = tr.mYP;
  P.mCP = tr.mCP;
  P.mNR = tr.mNR;
  P.mNPX = tr.mNPX;
  P.mXcd = tr.mXcd;
  P.done = c->op_mask;
  P.dstride = c->op_mask_stride;
  // streaming stores only for outputs that cannot stay in the 256 MB MALL
  // (bit 3: for every output)
  P.nt = (((c->dslash_policy >> 2) & 1) && out.bytes >= ((size_t)256 << 20)) || ((c->dslash_policy >> 3) & 1);
  const unsigned nblk = tr.nblk;
  double *part = nullptr;
  if (dm != DOT_NONE) {
    CVC_TRY(ensure_red(c, sizeof(double) * ((size_t)nblk * 3 * N + 64 * 256)));
    part = (double *)c->red.ptr;
  }
  const cplx<R> *U;
  if constexpr (sizeof(R) == 8) U = (RECON == 18) ? c->gauge64 : c->gauge64r12;
  else U = c->gauge32r12;
  IOO o = make_io<IOO>(out);
  IOI ii = make_io<IOI>(in);
  IOR rr;
  if constexpr (kSameR) rr = r2 ? make_io<IOI>(*r2) : ii;
  else rr = make_io<IOR>(*r2);
  IOI ow = (eo && eo->own) ? make_io<IOI>(*eo->own) : ii;
  dim3 g(nblk), b(B);
  if (!EO && tr.mode == 6) {
    using CR = cplx<R>;
    constexpr int KSM = (RECON == 12 && sizeof(R) == 4) ? 9 : (RECON == 18 ? 9 : 6);
    // bits 24-27: extra LDS per block in 8 KiB units (caps the blocks per CU)
    const size_t lds = (size_t)kMarchSlots * KSM * tr.mS * sizeof(CR) + (size_t)((c->dslash_policy >> 24) & 15) * 8192;
    dim3 bm(tr.threads);
#define CVC_MLAUNCH(G5, DMV) \
    hipLaunchKernelGGL((k_dslash_march<IOI, IOO, RECON, G5, DMV, IOR>), g, bm, lds, c->stream, o, ii, U, P, rr, part)
    if constexpr (!kSameR) {
      if (g5) CVC_MLAUNCH(true, DOT_R2);
      else CVC_MLAUNCH(false, DOT_R2);
    } else if (g5) {
      if (dm == DOT_NONE) CVC_MLAUNCH(true, DOT_NONE);
      else if (dm == DOT_TS) CVC_MLAUNCH(true, DOT_TS);
      else if (dm == DOT_NN) CVC_MLAUNCH(true, DOT_NN);
      else CVC_MLAUNCH(true, DOT_R2);
    } else {
      if (dm == DOT_NONE) CVC_MLAUNCH(false, DOT_NONE);
      else if (dm == DOT_TS) CVC_MLAUNCH(false, DOT_TS);
      else if (dm == DOT_NN) CVC_MLAUNCH(false, DOT_NN);
      else CVC_MLAUNCH(false, DOT_R2);
    }
#undef CVC_MLAUNCH
    CVC_HIP(hipGetLastError());
    if (nblk_out) *nblk_out = (int)nblk;
    return 0;
  }
  const bool multi = B / N <= LDS_SITES;
  const bool ldsl = multi && (c->dslash_policy & 1);
  const bool ldss = !EO && multi && (c->dslash_policy & 2) && tr.mode != 5;
  // bits 24-27: extra LDS per block in 8 KiB units (caps the blocks per CU)
  const size_t pad = (size_t)((c->dslash_policy >> 24) & 15) * 8192;
  const size_t dyn = (ldss ? (size_t)(B / N + 2) * N * 12 * sizeof(cplx<R>) : 0) + pad;
#define CVC_LAUNCH(G5, DMV)                                                                                     \
  do {                                                                                                          \
    if constexpr (EO) {                                                                                         \
      if (ldsl) hipLaunchKernelGGL((k_dslash<IOI, IOO, RECON, G5, DMV, true, false, true, IOR>), g, b, dyn, c->stream, o, ii, ow, U, P, rr, part); \
      else hipLaunchKernelGGL((k_dslash<IOI, IOO, RECON, G5, DMV, false, false, true, IOR>), g, b, dyn, c->stream, o, ii, ow, U, P, rr, part); \
    } else {                                                                                                    \
      if (ldsl && ldss) hipLaunchKernelGGL((k_dslash<IOI, IOO, RECON, G5, DMV, true, true, false, IOR>), g, b, dyn, c->stream, o, ii, ii, U, P, rr, part); \
      else if (ldsl) hipLaunchKernelGGL((k_dslash<IOI, IOO, RECON, G5, DMV, true, false, false, IOR>), g, b, dyn, c->stream, o, ii, ii, U, P, rr, part); \
      else if (ldss) hipLaunchKernelGGL((k_dslash<IOI, IOO, RECON, G5, DMV, false, true, false, IOR>), g, b, dyn, c->stream, o, ii, ii, U, P, rr, part); \
      else hipLaunchKernelGGL((k_dslash<IOI, IOO, RECON, G5, DMV, false, false, false, IOR>), g, b, dyn, c->stream, o, ii, ii, U, P, rr, part); \
    }                                                                                                           \
  } while (0)
  if constexpr (!kSameR) {
    if (g5) CVC_LAUNCH(true, DOT_R2);
    else CVC_LAUNCH(false, DOT_R2);
  } else if (g5) {
    if (dm == DOT_NONE) CVC_LAUNCH(true, DOT_NONE);
    else if (dm == DOT_TS) CVC_LAUNCH(true, DOT_TS);
    else if (dm == DOT_NN) CVC_LAUNCH(true, DOT_NN);
    else CVC_LAUNCH(true, DOT_R2);
  } else {
    if (dm == DOT_NONE) CVC_LAUNCH(false, DOT_NONE);
    else if (dm == DOT_TS) CVC_LAUNCH(false, DOT_TS);
    else if (dm == DOT_NN) CVC_LAUNCH(false, DOT_NN);
    else CVC_LAUNCH(false, DOT_R2);
  }
#undef CVC_LAUNCH
  CVC_HIP(hipGetLastError());
  if (nblk_out) *nblk_out = (int)nblk;
  return 0;
}

int dslash(cvchip_ctx *c, Field &out, const Field &in, double mutm, int gamma5_out, DotMode dm,
           const Field *r2, int *nblocks_out) {
  if (!c->gauge_set) {
    set_error("gauge field not set (call cvchip_set_gauge first)");
    return CVCHIP_ERR_STATE;
  }
  // r2 may be FP16 for FP32 fields in the <r2, out> epilogue (the FP32 mixed solver's shadow residual)
  const bool r2_half = r2 && dm == DOT_R2 && in.prec == CVCHIP_PREC_F32 && r2->prec == CVCHIP_PREC_F16;
  if (out.nrhs != in.nrhs || out.V != in.V ||
      (r2 && (r2->nrhs != in.nrhs || r2->V != in.V || (r2->prec != in.prec && !r2_half)))) {
    set_error("operator fields do not match (nrhs %d vs %d)", out.nrhs, in.nrhs);
    return CVCHIP_ERR_ARG;
  }
  if (out.data == in.data) {
    set_error("operator output must not alias its input");
    return CVCHIP_ERR_ARG;
  }
  if (in.prec == CVCHIP_PREC_F64 && out.prec == CVCHIP_PREC_F64) {
    if (c->link_recon == 12)
      return launch<IOPlain<double>, IOPlain<double>, 12, false>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out);
    return launch<IOPlain<double>, IOPlain<double>, 18, false>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out);
  }
  if (in.prec == CVCHIP_PREC_F32 && out.prec == CVCHIP_PREC_F32) {
    if (r2_half)
      return launch<IOPlain<float>, IOPlain<float>, 12, false, IOHalf>(c, out, in, mutm, gamma5_out, dm, r2,
                                                                       nblocks_out);
    return launch<IOPlain<float>, IOPlain<float>, 12, false>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out);
  }
  if (in.prec == CVCHIP_PREC_F16 && out.prec == CVCHIP_PREC_F16)
    return launch<IOHalf, IOHalf, 12, false>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out);
  set_error("unsupported operator precision pair %d -> %d", in.prec, out.prec);
  return CVCHIP_ERR_ARG;
}

// Even/odd operator pieces on half-lattice fields (V/2 sites, half-site order
// of DParams): out_par = dmode form of (own, hopc x hopping sum of in), in at
// the opposite parity.  With H the hopping sum and A = 1/(2 kappa) + i mutm g5,
// D = A - H/2 splits as D_ee = A, D_eo = -H_eo/2 (Q_phi.c:58-174 restricted to
// one parity of the output).
int dslash_eo(cvchip_ctx *c, Field &out, const Field &in, const Field *own, double mutm, int par, int dmode,
              double hopc, int gamma5_out, DotMode dm, const Field *r2, int *nblocks_out) {
  if (!c->gauge_set) {
    set_error("gauge field not set (call cvchip_set_gauge first)");
    return CVCHIP_ERR_STATE;
  }
  if (c->lat.LZ % 2) {
    set_error("even/odd operator needs an even LZ (got %d)", c->lat.LZ);
    return CVCHIP_ERR_ARG;
  }
  const int64_t Vh = c->lat.V / 2;
  const bool r2_half = r2 && dm == DOT_R2 && in.prec == CVCHIP_PREC_F32 && r2->prec == CVCHIP_PREC_F16;
  if (out.V != Vh || in.V != Vh || out.nrhs != in.nrhs || out.prec != in.prec ||
      (dmode != 1 && (!own || own->V != Vh || own->nrhs != in.nrhs || own->prec != in.prec)) ||
      (r2 && (r2->V != Vh || r2->nrhs != in.nrhs || (r2->prec != in.prec && !r2_half))) || dmode < 0 || dmode > 3 ||
      (par != 0 && par != 1)) {
    set_error("even/odd operator: fields must be half-lattice sets of one size and precision");
    return CVCHIP_ERR_ARG;
  }
  if (out.data == in.data) {
    set_error("operator output must not alias its input");
    return CVCHIP_ERR_ARG;
  }
  const EOArgs eo{own, par, dmode, hopc};
  if (in.prec == CVCHIP_PREC_F64) {
    if (c->link_recon == 12)
      return launch<IOPlain<double>, IOPlain<double>, 12, true>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out, &eo);
    return launch<IOPlain<double>, IOPlain<double>, 18, true>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out, &eo);
  }
  if (in.prec == CVCHIP_PREC_F32) {
    if (r2_half)
      return launch<IOPlain<float>, IOPlain<float>, 12, true, IOHalf>(c, out, in, mutm, gamma5_out, dm, r2,
                                                                      nblocks_out, &eo);
    return launch<IOPlain<float>, IOPlain<float>, 12, true>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out, &eo);
  }
  return launch<IOHalf, IOHalf, 12, true>(c, out, in, mutm, gamma5_out, dm, r2, nblocks_out, &eo);
}

// ------------------------------------------------------- reductions

// Two fixed-order stages: G blocks per (q, rhs) row each sum a contiguous
// segment of the row (four independent accumulators per thread, so several
// loads are in flight), then one block per row sums the G segment totals.
__global__ void __launch_bounds__(256) k_reduce_seg(const double *__restrict__ part, int nblk, int G,
                                                    double *__restrict__ seg) {
  __shared__ double sh[256];
  const int row = blockIdx.x / G, g = blockIdx.x - row * G;
  const int len = (nblk + G - 1) / G;
  const int b0 = g * len, b1 = min(nblk, b0 + len);
  const double *p = part + (size_t)row * nblk;
  double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
  int b = b0 + (int)threadIdx.x;
  for (; b + 768 < b1; b += 1024) {
    a0 += p[b];
    a1 += p[b + 256];
    a2 += p[b + 512];
    a3 += p[b + 768];
  }
  for (; b < b1; b += 256) a0 += p[b];
  sh[threadIdx.x] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) seg[blockIdx.x] = sh[0];
}

__global__ void k_reduce_partials(const double *__restrict__ part, int nblk, int Q, int N,
                                  double *__restrict__ res) {
  __shared__ double sh[256];
  (void)Q;
  (void)N;
  const double *row = part + (size_t)blockIdx.x * nblk;   // (q, rhs) = (blockIdx / N, blockIdx % N)
  double a = 0.;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += row[b];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) res[blockIdx.x] = sh[0];
}

int reduce_partials(cvchip_ctx *c, const double *part, int nblk, int Q, int N, double *res) {
  const int rows = Q * N;
  // segments of at least 4096 partials, at most 64 per row
  int G = nblk / 4096;
  if (G > 64) G = 64;
  if (G >= 2) {
    CVC_TRY(ensure_red2(c, sizeof(double) * (size_t)rows * G));
    double *seg = (double *)c->red2.ptr;
    hipLaunchKernelGGL(k_reduce_seg, dim3(rows * G), dim3(256), 0, c->stream, part, nblk, G, seg);
    hipLaunchKernelGGL(k_reduce_partials, dim3(rows), dim3(256), 0, c->stream, seg, G, Q, N, res);
  } else {
    hipLaunchKernelGGL(k_reduce_partials, dim3(rows), dim3(256), 0, c->stream, part, nblk, Q, N, res);
  }
  CVC_HIP(hipGetLastError());
  return 0;
}

}  // namespace cvchip

using namespace cvchip;

extern "C" int cvchip_field_dslash(cvchip_ctx *c, cvchip_field *out, const cvchip_field *in,
                                   double mutm, int gamma5_out) {
  return cvchip::guarded("cvchip_field_dslash", [&]() -> int {
    if (!c || !out || !in) return CVCHIP_ERR_NULL;
    CVC_TRY(dslash(c, out->f, in->f, mutm, gamma5_out));
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_dslash_timed(cvchip_ctx *c, cvchip_field *out, const cvchip_field *in,
                                         double mutm, int reps, double *ms) {
  return cvchip::guarded("cvchip_field_dslash_timed", [&]() -> int {
    if (!c || !out || !in || !ms) return CVCHIP_ERR_NULL;
    if (reps < 1) {
      set_error("reps must be positive");
      return CVCHIP_ERR_ARG;
    }
    CVC_HIP(hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < reps; k++) CVC_TRY(dslash(c, out->f, in->f, mutm, 0));
    CVC_HIP(hipEventRecord(c->ev1, c->stream));
    CVC_HIP(hipEventSynchronize(c->ev1));
    float t = 0.f;
    CVC_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = (double)t / reps;
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_Q_phi(cvchip_ctx *c, double *xi, const double *phi, double mutm, int gamma5_out) {
  return cvchip::guarded("cvchip_Q_phi", [&]() -> int {
    if (!c || !xi || !phi) {
      set_error("NULL argument to cvchip_Q_phi");
      return CVCHIP_ERR_NULL;
    }
    Field a, b;
    int rc = field_alloc(c, a, 1, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, b, 1, CVCHIP_PREC_F64);
    if (!rc) rc = upload_aos(c, a, phi, 0, 1);
    if (!rc) rc = dslash(c, b, a, mutm, gamma5_out);
    if (!rc) rc = download_aos(c, b, xi, 0, 1);
    field_free(a);
    field_free(b);
    return rc;
  });
}
