// Gradient of L = sum_mu sum_{q in Q} Re Pi_{mu mu}(q), Q = {q : |q| <= qmax}, with
// respect to the twisted mass mu and to every gauge link, by adjoint solves.
// No reference counterpart (the reference stops at Pi(x)).
//
// With w(x) = sum_{q in Q} cos(q.(x - y)) and c(x) = -w(x)/4,
//   L = Re sum_x sum_mu c(x) sum_k [T1 + T2 + T3 + T4]_{mu}(x)  -  |Q| Re sum_nu contact_nu,
// where T1..T4 are the four sink-current terms of the contraction (see
// contract.hip) with nu = mu.  L depends on
//   (i)  the 120 propagators S = D_f^{-1} B (f = up: D(mu), down: D(-mu)),
//   (ii) explicitly on the sink links U_mu(x) and the source links U_nu(y).
// For (i), dL = sum_sets Re<Xi, dS> with the adjoint sources Xi = dL/dS*;
// dS = -D^{-1} dD S gives dL = -sum Re<Lambda, dD S>, Lambda = D_f^{-dagger} Xi
// = gamma5 D_{-f}^{-1} gamma5 Xi (gamma5-hermiticity: one solve per column
// with the other flavour).  Then
//   dL/dmu = -Re sum [<Lambda_up, i g5 S_up> - <Lambda_dn, i g5 S_dn>]
// and for a link variation U_rho(z) -> exp(i eps T) U_rho(z):
//   dL/deps = Re[i tr(T K)],  K = 1/2 sum_cols (M1 - M2) + explicit terms,
//   M1 = sum_spin ((1 - g_rho) U S(z+rho)) (x) Lambda(z)^dagger,
//   M2 = sum_spin ((1 + g_rho) S(z)) (x) (U Lambda(z+rho))^dagger   (colour outer products).
// The reported force is F = traceless part of i (K - K^dagger)/2, so that
// dL/deps = tr(T F) for every traceless Hermitian T.
#include <math.h>
#include <string.h>
#include <complex>
#include <vector>

#include "contract_common.h"

namespace cvchip {

int solve_source_props(cvchip_ctx *c, const int src[4], double sp, int niter_max, int prec, double delta,
                       Field &up, Field &dn, cvchip_stats *tot);
int source_links(cvchip_ctx *c, const int src[4], SrcLinks *out);
int launch_psi(cvchip_ctx *c, double2 *psi, const Field &dn, const SrcLinks &L, int nu);
int contract_fields(cvchip_ctx *c, double *conn, double contact[8], const Field &up, const Field &dn,
                    const int src[4]);
int invert_her_fields(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, int prec,
                      double delta, unsigned flags, cvchip_stats *stats);

struct AdjParams {
  ConnParams P;
  uint32_t y;
};

__device__ __forceinline__ void neighbours(const ConnParams &P, uint32_t z, int mu, uint32_t &zm, uint32_t &zp) {
  uint32_t rem = z;
  const uint32_t ct = rem / P.sT; rem -= ct * P.sT;
  const uint32_t cx = rem / P.sX; rem -= cx * P.sX;
  const uint32_t cy = rem / (uint32_t)P.LZ;
  const uint32_t cz = rem - cy * (uint32_t)P.LZ;
  uint32_t coord, L, stride;
  if (mu == 0) { coord = ct; L = P.T; stride = P.sT; }
  else if (mu == 1) { coord = cx; L = P.LX; stride = P.sX; }
  else if (mu == 2) { coord = cy; L = P.LY; stride = P.LZ; }
  else { coord = cz; L = P.LZ; stride = 1; }
  zp = (coord == L - 1) ? z - (L - 1) * stride : z + stride;
  zm = (coord == 0) ? z + (L - 1) * stride : z - stride;
}

__device__ __forceinline__ void link_at(double2 u[9], const double2 *__restrict__ U, const ConnParams &P, int mu,
                                        uint32_t z) {
  const double2 ph = make_double2(P.ph_re[mu], P.ph_im[mu]);
#pragma unroll
  for (int q = 0; q < 9; q++) u[q] = cmul(U[(size_t)(mu * 9 + q) * P.V + z], ph);
}

__device__ __forceinline__ void load_col(double2 v[12], const double2 *__restrict__ f, uint32_t V, int ncol,
                                         uint32_t z, int col) {
  const size_t VN = (size_t)V * ncol;
#pragma unroll
  for (int c = 0; c < 12; c++) v[c] = f[(size_t)c * VN + (size_t)z * ncol + col];
}

// (1 + sgn gamma_mu) w  for runtime mu
__device__ __forceinline__ void one_pm_gamma(double2 v[12], const double2 w[12], int mu, double sgn) {
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int b = GCOL[mu][a];
    const double2 e = unit_val(GVAL[mu][a]);
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      const double2 g = cmul(e, w[3 * b + cc]);
      v[3 * a + cc] = make_double2(w[3 * a + cc].x + sgn * g.x, w[3 * a + cc].y + sgn * g.y);
    }
  }
}

// K += s * sum_spin y_a (x) conj(x_a)   (3x3 colour outer product, K[3*i+j])
__device__ __forceinline__ void outer_acc(double2 K[9], const double2 y[12], const double2 x[12], double s) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const double2 p = cmulc(x[3 * a + j], y[3 * a + i]);   // y_i conj(x_j)
        K[3 * i + j].x += s * p.x;
        K[3 * i + j].y += s * p.y;
      }
}

// c(x) = -1/4 sum_{q in Q} cos(q.(x - y)); q given as double[4] radians
__global__ void k_weight(double *__restrict__ cw, const double *__restrict__ qs, int nq, AdjParams A, int y0,
                         int y1, int y2, int y3) {
  const uint32_t z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= A.P.V) return;
  uint32_t rem = z;
  const int ct = (int)(rem / A.P.sT); rem -= ct * A.P.sT;
  const int cx = (int)(rem / A.P.sX); rem -= cx * A.P.sX;
  const int cy = (int)(rem / (uint32_t)A.P.LZ);
  const int cz = (int)(rem - cy * (uint32_t)A.P.LZ);
  const double d0 = ct - y0, d1 = cx - y1, d2 = cy - y2, d3 = cz - y3;
  double w = 0.;
  for (int q = 0; q < nq; q++) w += cos(qs[4 * q] * d0 + qs[4 * q + 1] * d1 + qs[4 * q + 2] * d2 + qs[4 * q + 3] * d3);
  cw[z] = -0.25 * w;
}

// Adjoint sources of the up propagators for one nu (= mu of the diagonal):
//   Xi_u4,k(z) += c(z-nu) U(z-nu)^+ Gm^+ Psi1_k(z-nu) + c(z) U(z) Gp^+ Psi1_k(z+nu)
//   Xi_unu,k(z) = same with Psi2
__global__ void __launch_bounds__(256) k_xi_up(double2 *__restrict__ xiu, const double2 *__restrict__ psi,
                                               const double *__restrict__ cw, const double2 *__restrict__ U,
                                               AdjParams A, int nu) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.P.V * 12u) return;
  const uint32_t z = i / 12u;
  const int k = (int)(i % 12u);
  uint32_t zm, zp;
  neighbours(A.P, z, nu, zm, zp);
  double2 um[9], uz[9];
  link_at(um, U, A.P, nu, zm);
  link_at(uz, U, A.P, nu, z);
  const double cm = cw[zm], cz = cw[z];
  const size_t VN60 = (size_t)A.P.V * 60;
  for (int part = 0; part < 2; part++) {
    double2 p[12], v[12], w1[12], w2[12];
    load_col(p, psi, A.P.V, 24, zm, 12 * part + k);
    sink_vertex_dag(v, p, nu, -1.0);
    su3_mul(w1, um, v, true);
    load_col(p, psi, A.P.V, 24, zp, 12 * part + k);
    sink_vertex_dag(v, p, nu, 1.0);
    su3_mul(w2, uz, v, false);
    const int col = part == 0 ? 48 + k : 12 * nu + k;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const size_t o = (size_t)c * VN60 + (size_t)z * 60 + col;
      double2 a = make_double2(cm * w1[c].x + cz * w2[c].x, cm * w1[c].y + cz * w2[c].y);
      if (part == 0) { a.x += xiu[o].x; a.y += xiu[o].y; }
      xiu[o] = a;
    }
  }
}

// V_k(z) = c(z) Gm U(z) S_k(z+nu) + c(z-nu) Gp U(z-nu)^+ S_k(z-nu)
__device__ __forceinline__ void vsink(double2 out[12], const double2 *__restrict__ up, const AdjParams &A, int nu,
                                      int col, uint32_t z, uint32_t zm, uint32_t zp, const double2 um[9],
                                      const double2 uz[9], double cz, double cm) {
  double2 s[12], w[12], v1[12], v2[12];
  load_col(s, up, A.P.V, 60, zp, col);
  su3_mul(w, uz, s, false);
  sink_vertex(v1, w, nu, -1.0);
  load_col(s, up, A.P.V, 60, zm, col);
  su3_mul(w, um, s, true);
  sink_vertex(v2, w, nu, 1.0);
#pragma unroll
  for (int c = 0; c < 12; c++) out[c] = make_double2(cz * v1[c].x + cm * v2[c].x, cz * v1[c].y + cm * v2[c].y);
}

// Adjoint sources of the down propagators for one nu, and the source-vertex
// bilinears B_{l,k} = sum_z chi_l(z)^+ V_k(z) (block partials, 24 reals per l).
// Thread per (z, l), l = 3 j + ib; the two rows ir of the vertex with column j
// are ir_a (gamma_nu gamma5) and ir_b = j (gamma5).
__global__ void __launch_bounds__(240) k_xi_dn(double2 *__restrict__ xid, double *__restrict__ part,
                                               const double2 *__restrict__ up, const double2 *__restrict__ dn,
                                               const double *__restrict__ cw, const double2 *__restrict__ U,
                                               SrcLinks Ls, AdjParams A, int nu) {
  __shared__ double sh[240];
  const int t = threadIdx.x;
  const uint32_t z = blockIdx.x * 20u + (uint32_t)(t / 12);
  const int l = t % 12;
  const int j = l / 3, ib = l % 3;
  const int ira = GCOL[nu][j], irb = j;
  double2 ea = unit_val(GVAL[nu][ira]);
  if (g5sign(j) < 0) { ea.x = -ea.x; ea.y = -ea.y; }
  const double s5 = (double)g5sign(j);
  double bil[24];
#pragma unroll
  for (int q = 0; q < 24; q++) bil[q] = 0.;
  if (z < A.P.V) {
    uint32_t zm, zp;
    neighbours(A.P, z, nu, zm, zp);
    double2 um[9], uz[9];
    link_at(um, U, A.P, nu, zm);
    link_at(uz, U, A.P, nu, z);
    const double cz = cw[z], cm = cw[zm];
    const double2 *Us = Ls.u[nu];
    double2 x1[12], x2[12], chi1[12], chi2[12];
#pragma unroll
    for (int c = 0; c < 12; c++) x1[c] = x2[c] = make_double2(0, 0);
    load_col(chi1, dn, A.P.V, 60, z, 12 * nu + l);
    load_col(chi2, dn, A.P.V, 60, z, 48 + l);
    for (int w = 0; w < 2; w++) {
      const int ir = w == 0 ? ira : irb;
      const double2 e1 = w == 0 ? ea : make_double2(-s5, 0.);   // (gamma_nu g5 - g5)_{ir, j}
      const double2 e2 = w == 0 ? ea : make_double2(s5, 0.);    // (gamma_nu g5 + g5)_{ir, j}
      for (int ia = 0; ia < 3; ia++) {
        const int k = 3 * ir + ia;
        double2 v[12];
        vsink(v, up, A, nu, 48 + k, z, zm, zp, um, uz, cz, cm);      // part 1: up from y
        const double2 g1 = cmul(e1, Us[3 * ia + ib]);
        const double2 d1 = dot12(chi1, v);
#pragma unroll
        for (int c = 0; c < 12; c++) x1[c] = cfma(x1[c], g1, v[c]);
        bil[2 * (3 * w + ia)] = d1.x;
        bil[2 * (3 * w + ia) + 1] = d1.y;
        vsink(v, up, A, nu, 12 * nu + k, z, zm, zp, um, uz, cz, cm); // part 2: up from y + nu
        const double2 g2 = cmulc(Us[3 * ib + ia], e2);                // e2 conj(Us[ib][ia])
        const double2 d2 = dot12(chi2, v);
#pragma unroll
        for (int c = 0; c < 12; c++) x2[c] = cfma(x2[c], g2, v[c]);
        bil[12 + 2 * (3 * w + ia)] = d2.x;
        bil[12 + 2 * (3 * w + ia) + 1] = d2.y;
      }
    }
    const size_t VN60 = (size_t)A.P.V * 60;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const size_t o1 = (size_t)c * VN60 + (size_t)z * 60 + 12 * nu + l;
      const size_t o2 = (size_t)c * VN60 + (size_t)z * 60 + 48 + l;
      xid[o1] = x1[c];
      double2 a = xid[o2];
      a.x += x2[c].x;
      a.y += x2[c].y;
      xid[o2] = a;
    }
  }
  // fixed-order tree over the 20 sites of the block for each of the 24 reals of l
  const int srow = t / 12;
  for (int q = 0; q < 24; q++) {
    sh[t] = bil[q];
    __syncthreads();
    for (int active = 20; active > 1;) {
      const int half = (active + 1) >> 1;
      if (srow < active - half) sh[t] += sh[t + half * 12];
      __syncthreads();
      active = half;
    }
    if (srow == 0) part[((size_t)(l * 24 + q)) * gridDim.x + blockIdx.x] = sh[t];
    __syncthreads();
  }
}

// Explicit sink-current links for rho = nu: K(z, nu) += c(z) (M3 - M4) for both parts,
// M3 = sum_k sum_spin (Gm U S_k(z+nu)) (x) Psi_k(z)^+, M4 = sum_k sum_spin (Gp S_k(z)) (x) (U Psi_k(z+nu))^+
__global__ void __launch_bounds__(256) k_force_sink(double2 *__restrict__ K, const double2 *__restrict__ up,
                                                    const double2 *__restrict__ psi, const double *__restrict__ cw,
                                                    const double2 *__restrict__ U, AdjParams A, int nu) {
  const uint32_t z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= A.P.V) return;
  uint32_t zm, zp;
  neighbours(A.P, z, nu, zm, zp);
  double2 uz[9];
  link_at(uz, U, A.P, nu, z);
  double2 M[9];
#pragma unroll
  for (int q = 0; q < 9; q++) M[q] = make_double2(0, 0);
  for (int part = 0; part < 2; part++) {
    const int s = part == 0 ? 4 : nu;
    for (int k = 0; k < 12; k++) {
      double2 a[12], b[12], y[12];
      load_col(a, up, A.P.V, 60, zp, 12 * s + k);
      su3_mul(b, uz, a, false);
      sink_vertex(y, b, nu, -1.0);
      load_col(a, psi, A.P.V, 24, z, 12 * part + k);
      outer_acc(M, y, a, 1.0);
      load_col(a, up, A.P.V, 60, z, 12 * s + k);
      sink_vertex(y, a, nu, 1.0);
      load_col(a, psi, A.P.V, 24, zp, 12 * part + k);
      su3_mul(b, uz, a, false);
      outer_acc(M, y, b, -1.0);
    }
  }
  const double cz = cw[z];
  double2 *Kz = K + ((size_t)z * 4 + nu) * 9;
#pragma unroll
  for (int q = 0; q < 9; q++) Kz[q] = make_double2(Kz[q].x + cz * M[q].x, Kz[q].y + cz * M[q].y);
}

// Implicit link force: K(z, rho) += 1/2 sum over the 120 columns of (M1 - M2).
__global__ void __launch_bounds__(256) k_force_impl(double2 *__restrict__ K, const double2 *__restrict__ up,
                                                    const double2 *__restrict__ dn, const double2 *__restrict__ lu,
                                                    const double2 *__restrict__ ld, const double2 *__restrict__ U,
                                                    AdjParams A) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.P.V * 4u) return;
  const uint32_t z = i / 4u;
  const int rho = (int)(i % 4u);
  uint32_t zm, zp;
  neighbours(A.P, z, rho, zm, zp);
  double2 uz[9];
  link_at(uz, U, A.P, rho, z);
  double2 M[9];
#pragma unroll
  for (int q = 0; q < 9; q++) M[q] = make_double2(0, 0);
  for (int f = 0; f < 2; f++) {
    const double2 *S = f == 0 ? up : dn;
    const double2 *Lm = f == 0 ? lu : ld;
    for (int col = 0; col < 60; col++) {
      double2 a[12], b[12], y[12];
      load_col(a, S, A.P.V, 60, zp, col);
      su3_mul(b, uz, a, false);
      one_pm_gamma(y, b, rho, -1.0);                 // (1 - gamma) U S(z+rho)
      load_col(a, Lm, A.P.V, 60, z, col);
      outer_acc(M, y, a, 0.5);
      load_col(a, S, A.P.V, 60, z, col);
      one_pm_gamma(y, a, rho, 1.0);                  // (1 + gamma) S(z)
      load_col(a, Lm, A.P.V, 60, zp, col);
      su3_mul(b, uz, a, false);                      // U Lambda(z+rho)
      outer_acc(M, y, b, -0.5);
    }
  }
  double2 *Kz = K + ((size_t)z * 4 + rho) * 9;
#pragma unroll
  for (int q = 0; q < 9; q++) Kz[q] = make_double2(Kz[q].x + M[q].x, Kz[q].y + M[q].y);
}

// sum over sites and 60 columns of <Lambda, gamma5 S> (complex), block partials (N = 1)
__global__ void __launch_bounds__(256) k_dot_g5(double *__restrict__ part, const double2 *__restrict__ lam,
                                                const double2 *__restrict__ S, uint32_t VN) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double dv[2] = {0., 0.};
  if (i < VN) {
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const double2 a = lam[(size_t)c * VN + i], b = S[(size_t)c * VN + i];
      const double s = c < 6 ? 1.0 : -1.0;
      dv[0] += s * (a.x * b.x + a.y * b.y);
      dv[1] += s * (a.x * b.y - a.y * b.x);
    }
  }
  block_reduce_store<2>(dv, part, 1);
}

// 12-column set <-> columns col0.. of a 60-column set, optionally times gamma5
__global__ void k_cols(double2 *__restrict__ dst, const double2 *__restrict__ src, uint32_t V, int to60, int col0,
                       int g5) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;   // over 12 comps * V * 12 cols
  if (i >= 144u * V) return;
  const uint32_t r = i % 12u;
  const uint32_t cs = i / 12u;                                 // comp * V + site
  const uint32_t comp = cs / V;
  const size_t o60 = (size_t)cs * 60 + col0 + r;
  double2 v = to60 ? src[i] : src[o60];
  if (g5 && comp >= 6) { v.x = -v.x; v.y = -v.y; }
  if (to60) dst[o60] = v;
  else dst[i] = v;
}

// one site's 12 x 60 block of a 60-column set: out[c][col]
__global__ void k_gather_site(double2 *__restrict__ out, const double2 *__restrict__ f, uint32_t V, uint32_t z) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 720) return;
  const int c = t / 60, col = t % 60;
  out[t] = f[(size_t)c * V * 60 + (size_t)z * 60 + col];
}

// add a 12 (comp) x 12 (col) block at one site into columns col0.. of a 60-column set
__global__ void k_add_site_block(double2 *__restrict__ f, const double2 *__restrict__ blk, uint32_t V, uint32_t z,
                                 int col0) {
  const int t = threadIdx.x;
  if (t >= 144) return;
  const int c = t / 12, col = t % 12;
  double2 &d = f[(size_t)c * V * 60 + (size_t)z * 60 + col0 + col];
  d.x += blk[t].x;
  d.y += blk[t].y;
}

}  // namespace cvchip

using namespace cvchip;
typedef std::complex<double> cd;

namespace {

cd gmat(int mu, int a, int b) {   // (gamma_mu)_{ab}
  if (GCOL[mu][a] != b) return 0.;
  const int u = GVAL[mu][a];
  return u == 0 ? cd(1, 0) : u == 1 ? cd(-1, 0) : u == 2 ? cd(0, 1) : cd(0, -1);
}
double g5d(int a) { return a < 2 ? 1. : -1.; }

}  // namespace

extern "C" int cvchip_cvc_gradient(cvchip_ctx *c, const int src[4], double qmax, double sp, int niter_max, int prec,
                                   double delta, double *L_out, double *dmu_out, double *force,
                                   cvchip_stats *stats) {
  return cvchip::guarded("cvchip_cvc_gradient", [&]() -> int {
    if (!c || !src || !L_out || !dmu_out) {
      set_error("NULL argument to cvchip_cvc_gradient");
      return CVCHIP_ERR_NULL;
    }
    if (!c->gauge_set) {
      set_error("gauge field not set (call cvchip_set_gauge first)");
      return CVCHIP_ERR_STATE;
    }
    const Lattice &Lt = c->lat;
    const uint32_t V = (uint32_t)Lt.V;
    const int Ls[4] = {Lt.T, Lt.LX, Lt.LY, Lt.LZ};
    // momentum set Q: n_mu in (-L/2, L/2], |q| <= qmax
    std::vector<double> qs;
    for (int n0 = -(Ls[0] - 1) / 2; n0 <= Ls[0] / 2; n0++)
      for (int n1 = -(Ls[1] - 1) / 2; n1 <= Ls[1] / 2; n1++)
        for (int n2 = -(Ls[2] - 1) / 2; n2 <= Ls[2] / 2; n2++)
          for (int n3 = -(Ls[3] - 1) / 2; n3 <= Ls[3] / 2; n3++) {
            const double q[4] = {2 * M_PI * n0 / Ls[0], 2 * M_PI * n1 / Ls[1], 2 * M_PI * n2 / Ls[2],
                                 2 * M_PI * n3 / Ls[3]};
            if (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] <= qmax * qmax + 1e-12)
              for (int m = 0; m < 4; m++) qs.push_back(q[m]);
          }
    const int nq = (int)(qs.size() / 4);
    AdjParams A;
    A.P = conn_params(c);
    A.y = site_of(c, src[0], src[1], src[2], src[3]);
    uint32_t ysp[4];
    for (int nu = 0; nu < 4; nu++) {
      int d[4] = {0, 0, 0, 0};
      d[nu] = 1;
      ysp[nu] = site_of(c, src[0] + d[0], src[1] + d[1], src[2] + d[2], src[3] + d[3]);
    }

    Field up, dn, xu, xd, b12, x12;
    double *cw = nullptr, *dqs = nullptr;
    double2 *psi = nullptr, *K = nullptr, *blk = nullptr;
    std::vector<double> conn((size_t)32 * V);
    double contact[8];
    cvchip_stats tot;
    memset(&tot, 0, sizeof(tot));
    int rc = field_alloc(c, up, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, dn, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, xu, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, xd, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, b12, 12, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, x12, 12, CVCHIP_PREC_F64);
    auto hipok = [&](hipError_t e, const char *w) {
      if (e != hipSuccess && !rc) rc = hip_fail(e, w, __FILE__, __LINE__);
    };
    if (!rc) {
      hipok(hipMalloc(&cw, sizeof(double) * V), "weights");
      hipok(hipMalloc(&dqs, sizeof(double) * (qs.size() + 4)), "momenta");
      hipok(hipMalloc(&psi, sizeof(double2) * 288 * (size_t)V), "psi");
      hipok(hipMalloc(&K, sizeof(double2) * 36 * (size_t)V), "force");
      hipok(hipMalloc(&blk, sizeof(double2) * 720), "site block");
    }
    const int nblk_dn = (int)((V + 19) / 20);
    if (!rc) rc = ensure_red(c, sizeof(double) * ((size_t)288 * nblk_dn + 64 * 256));
    // ---- forward: propagators, conn, L
    if (!rc) rc = solve_source_props(c, src, sp, niter_max, prec, delta, up, dn, &tot);
    if (!rc) rc = contract_fields(c, conn.data(), contact, up, dn, src);
    SrcLinks hL;
    if (!rc) rc = source_links(c, src, &hL);
    if (!rc) {
      hipok(hipMemcpyAsync(dqs, qs.data(), sizeof(double) * qs.size(), hipMemcpyHostToDevice, c->stream), "momenta");
      hipLaunchKernelGGL(k_weight, dim3((V + 255) / 256), dim3(256), 0, c->stream, cw, (const double *)dqs, nq, A,
                         src[0], src[1], src[2], src[3]);
      hipok(hipMemsetAsync(xu.data, 0, xu.bytes, c->stream), "xi");
      hipok(hipMemsetAsync(xd.data, 0, xd.bytes, c->stream), "xi");
      hipok(hipMemsetAsync(K, 0, sizeof(double2) * 36 * (size_t)V, c->stream), "force");
    }
    std::vector<double> hcw(V);
    if (!rc) {
      hipok(hipMemcpyAsync(hcw.data(), cw, sizeof(double) * V, hipMemcpyDeviceToHost, c->stream), "weights");
      hipok(hipStreamSynchronize(c->stream), "weights");
    }
    double Lval = 0.;
    if (!rc) {
      for (uint32_t x = 0; x < V; x++)
        for (int mu = 0; mu < 4; mu++) Lval += -4.0 * hcw[x] * conn[2 * ((size_t)16 * x + 5 * mu)];
      for (int mu = 0; mu < 4; mu++) Lval -= nq * contact[2 * mu];
    }
    // ---- adjoint sources, explicit sink force, source bilinears (per nu = mu)
    cd Ksrc[4][9];
    for (int nu = 0; nu < 4; nu++)
      for (int q = 0; q < 9; q++) Ksrc[nu][q] = 0.;
    std::vector<double> bil(288);
    double *dres = nullptr;
    if (!rc) hipok(hipMalloc(&dres, sizeof(double) * 288), "bilinears");
    for (int nu = 0; nu < 4 && !rc; nu++) {
      rc = launch_psi(c, psi, dn, hL, nu);
      if (rc) break;
      hipLaunchKernelGGL(k_xi_up, dim3((V * 12u + 255) / 256), dim3(256), 0, c->stream, (double2 *)xu.data,
                         (const double2 *)psi, (const double *)cw, (const double2 *)c->gauge64, A, nu);
      hipLaunchKernelGGL(k_xi_dn, dim3(nblk_dn), dim3(240), 0, c->stream, (double2 *)xd.data, (double *)c->red.ptr,
                         (const double2 *)up.data, (const double2 *)dn.data, (const double *)cw,
                         (const double2 *)c->gauge64, hL, A, nu);
      hipok(hipGetLastError(), "adjoint sources");
      if (!rc) rc = reduce_partials(c, (const double *)c->red.ptr, nblk_dn, 288, 1, dres);
      hipLaunchKernelGGL(k_force_sink, dim3((V + 255) / 256), dim3(256), 0, c->stream, K, (const double2 *)up.data,
                         (const double2 *)psi, (const double *)cw, (const double2 *)c->gauge64, A, nu);
      hipok(hipGetLastError(), "sink force");
      hipok(hipMemcpyAsync(bil.data(), dres, sizeof(double) * 288, hipMemcpyDeviceToHost, c->stream), "bilinears");
      hipok(hipStreamSynchronize(c->stream), "bilinears");
      if (rc) break;
      // source-link derivative through the vertices: K += Us N1 - N2 Us^+
      const double2 *Us = hL.u[nu];
      cd U3[9];
      for (int q = 0; q < 9; q++) U3[q] = cd(Us[q].x, Us[q].y);
      cd N1[9], N2[9];
      for (int q = 0; q < 9; q++) N1[q] = N2[q] = 0.;
      for (int l = 0; l < 12; l++) {
        const int j = l / 3, ib = l % 3;
        const int ira = GCOL[nu][j];
        const cd ea = gmat(nu, ira, j) * g5d(j);
        const double s5 = g5d(j);
        for (int w = 0; w < 2; w++) {
          const cd e1 = w == 0 ? ea : cd(-s5, 0), e2 = w == 0 ? ea : cd(s5, 0);
          for (int ia = 0; ia < 3; ia++) {
            const cd b1(bil[l * 24 + 2 * (3 * w + ia)], bil[l * 24 + 2 * (3 * w + ia) + 1]);
            const cd b2(bil[l * 24 + 12 + 2 * (3 * w + ia)], bil[l * 24 + 12 + 2 * (3 * w + ia) + 1]);
            N1[3 * ib + ia] += e1 * b1;
            N2[3 * ib + ia] += e2 * b2;
          }
        }
      }
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          cd s1 = 0., s2 = 0.;
          for (int m = 0; m < 3; m++) {
            s1 += U3[3 * a + m] * N1[3 * m + b];
            s2 += N2[3 * a + m] * std::conj(U3[3 * b + m]);
          }
          Ksrc[nu][3 * a + b] += s1 - s2;
        }
    }
    // ---- contact term: adjoint sources at y+nu (up from y) and y (up from y+nu), explicit source-link part
    if (!rc) {
      std::vector<double2> su4(720), suy(720);
      for (int nu = 0; nu < 4 && !rc; nu++) {
        hipLaunchKernelGGL(k_gather_site, dim3(3), dim3(256), 0, c->stream, blk, (const double2 *)up.data, V, ysp[nu]);
        hipok(hipMemcpyAsync(su4.data(), blk, sizeof(double2) * 720, hipMemcpyDeviceToHost, c->stream), "site");
        hipLaunchKernelGGL(k_gather_site, dim3(3), dim3(256), 0, c->stream, blk, (const double2 *)up.data, V, A.y);
        hipok(hipMemcpyAsync(suy.data(), blk, sizeof(double2) * 720, hipMemcpyDeviceToHost, c->stream), "site");
        hipok(hipStreamSynchronize(c->stream), "site");
        if (rc) break;
        const double2 *Us = hL.u[nu];
        cd U3[9];
        for (int q = 0; q < 9; q++) U3[q] = cd(Us[q].x, Us[q].y);
        // Xi_u4,k(y+nu) += nQ/2 Us^+ (gamma_nu - 1) e_k ; Xi_unu,k(y) += -nQ/2 Us (gamma_nu + 1) e_k
        std::vector<double2> add1(144), add2(144);
        for (int k = 0; k < 12; k++) {
          const int ir = k / 3, ia = k % 3;
          for (int a = 0; a < 4; a++)
            for (int col = 0; col < 3; col++) {
              const cd gm = gmat(nu, a, ir) - (a == ir ? 1. : 0.);   // ((gamma_nu - 1) e_k)_a = (g - 1)_{a, ir}
              const cd gp = gmat(nu, a, ir) + (a == ir ? 1. : 0.);
              const cd v1 = 0.5 * nq * gm * std::conj(U3[3 * ia + col]);   // (Us^+)_{col, ia}
              const cd v2 = -0.5 * nq * gp * U3[3 * col + ia];             // (Us)_{col, ia}
              add1[(3 * a + col) * 12 + k] = make_double2(v1.real(), v1.imag());
              add2[(3 * a + col) * 12 + k] = make_double2(v2.real(), v2.imag());
            }
        }
        double2 *dadd = blk;   // reuse the 720-entry buffer (144 needed)
        hipok(hipMemcpyAsync(dadd, add1.data(), sizeof(double2) * 144, hipMemcpyHostToDevice, c->stream), "contact");
        hipLaunchKernelGGL(k_add_site_block, dim3(1), dim3(256), 0, c->stream, (double2 *)xu.data,
                           (const double2 *)dadd, V, ysp[nu], 48);
        hipok(hipStreamSynchronize(c->stream), "contact");
        hipok(hipMemcpyAsync(dadd, add2.data(), sizeof(double2) * 144, hipMemcpyHostToDevice, c->stream), "contact");
        hipLaunchKernelGGL(k_add_site_block, dim3(1), dim3(256), 0, c->stream, (double2 *)xu.data,
                           (const double2 *)dadd, V, A.y, 12 * nu);
        hipok(hipStreamSynchronize(c->stream), "contact");
        // explicit: K_src += nQ/2 (C1 + C2)
        //   C1_{j,ia} = sum_{ir,a} (g-1)_{ir,a} X_{(ir,ia)}[a][j],  X_k = Us Su4_k(y+nu)
        //   C2_{d,c}  = sum_{k,a} (g+1)_{ir,a} Y_k[a][d] (Us^+)_{ia,c},  Y_k = Sunu_k(y)
        cd C1[9], C2[9];
        for (int q = 0; q < 9; q++) C1[q] = C2[q] = 0.;
        for (int k = 0; k < 12; k++) {
          const int ir = k / 3, ia = k % 3;
          for (int a = 0; a < 4; a++) {
            const cd gm = gmat(nu, ir, a) - (a == ir ? 1. : 0.);
            const cd gp = gmat(nu, ir, a) + (a == ir ? 1. : 0.);
            for (int jj = 0; jj < 3; jj++) {
              cd x = 0.;
              for (int m = 0; m < 3; m++) {
                const double2 s4 = su4[(3 * a + m) * 60 + 48 + k];
                x += U3[3 * jj + m] * cd(s4.x, s4.y);
              }
              C1[3 * jj + ia] += gm * x;
              const double2 sy = suy[(3 * a + jj) * 60 + 12 * nu + k];
              for (int cc = 0; cc < 3; cc++) C2[3 * jj + cc] += gp * cd(sy.x, sy.y) * std::conj(U3[3 * cc + ia]);
            }
          }
        }
        for (int q = 0; q < 9; q++) Ksrc[nu][q] += 0.5 * nq * (C1[q] + C2[q]);
      }
    }
    // ---- adjoint solves: Lambda_up = g5 D(-mu)^-1 g5 Xi_up, Lambda_dn = g5 D(mu)^-1 g5 Xi_dn (in place)
    const double mu0 = c->mu;
    for (int f = 0; f < 2 && !rc; f++) {
      Field &xi = f == 0 ? xu : xd;
      for (int s = 0; s < 5 && !rc; s++) {
        hipLaunchKernelGGL(k_cols, dim3((144u * V + 255) / 256), dim3(256), 0, c->stream, (double2 *)b12.data,
                           (const double2 *)xi.data, V, 0, 12 * s, 1);
        hipok(hipMemsetAsync(x12.data, 0, x12.bytes, c->stream), "adjoint");
        c->mu = f == 0 ? -mu0 : mu0;
        cvchip_stats st;
        memset(&st, 0, sizeof(st));
        if (!rc) rc = invert_her_fields(c, x12, b12, sp, niter_max, prec, delta, 0, &st);
        c->mu = mu0;
        if (rc > 0) rc = 0;
        if (rc) break;
        tot.iterations += st.iterations;
        tot.converged += st.converged;
        tot.reliable_updates += st.reliable_updates;
        tot.dslash_calls += st.dslash_calls;
        tot.seconds += st.seconds;
        hipLaunchKernelGGL(k_cols, dim3((144u * V + 255) / 256), dim3(256), 0, c->stream, (double2 *)xi.data,
                           (const double2 *)x12.data, V, 1, 12 * s, 1);
        hipok(hipGetLastError(), "adjoint columns");
      }
    }
    // ---- dL/dmu = Im<Lambda_up, g5 S_up> - Im<Lambda_dn, g5 S_dn>
    double dmu = 0.;
    if (!rc) {
      const uint32_t VN = V * 60u;
      const int nb = (int)((VN + 255) / 256);
      rc = ensure_red(c, sizeof(double) * ((size_t)2 * nb + 64 * 256));
      double r2[2][2];
      for (int f = 0; f < 2 && !rc; f++) {
        hipLaunchKernelGGL(k_dot_g5, dim3(nb), dim3(256), 0, c->stream, (double *)c->red.ptr,
                           (const double2 *)(f == 0 ? xu.data : xd.data), (const double2 *)(f == 0 ? up.data : dn.data),
                           VN);
        if (!rc) rc = reduce_partials(c, (const double *)c->red.ptr, nb, 2, 1, dres);
        hipok(hipMemcpyAsync(r2[f], dres, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream), "dmu");
        hipok(hipStreamSynchronize(c->stream), "dmu");
      }
      dmu = r2[0][1] - r2[1][1];
    }
    // ---- implicit link force and projection
    if (!rc && force) {
      hipLaunchKernelGGL(k_force_impl, dim3((V * 4u + 255) / 256), dim3(256), 0, c->stream, K, (const double2 *)up.data,
                         (const double2 *)dn.data, (const double2 *)xu.data, (const double2 *)xd.data,
                         (const double2 *)c->gauge64, A);
      hipok(hipGetLastError(), "force");
      std::vector<double2> hK((size_t)36 * V);
      if (!rc) rc = copy_to_host(c, hK.data(), K, sizeof(double2) * 36 * (size_t)V);
      if (!rc) {
        for (int nu = 0; nu < 4; nu++)
          for (int q = 0; q < 9; q++) {
            double2 &e = hK[((size_t)A.y * 4 + nu) * 9 + q];
            e.x += Ksrc[nu][q].real();
            e.y += Ksrc[nu][q].imag();
          }
        for (size_t zl = 0; zl < (size_t)4 * V; zl++) {
          cd Km[9], F[9];
          for (int q = 0; q < 9; q++) Km[q] = cd(hK[zl * 9 + q].x, hK[zl * 9 + q].y);
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) F[3 * a + b] = cd(0, 0.5) * (Km[3 * a + b] - std::conj(Km[3 * b + a]));
          const cd tr = (F[0] + F[4] + F[8]) / 3.0;
          F[0] -= tr;
          F[4] -= tr;
          F[8] -= tr;
          for (int q = 0; q < 9; q++) {
            force[zl * 18 + 2 * q] = F[q].real();
            force[zl * 18 + 2 * q + 1] = F[q].imag();
          }
        }
      }
    }
    *L_out = Lval;
    *dmu_out = dmu;
    if (stats) *stats = tot;
    field_free(up);
    field_free(dn);
    field_free(xu);
    field_free(xd);
    field_free(b12);
    field_free(x12);
    if (cw) (void)hipFree(cw);
    if (dqs) (void)hipFree(dqs);
    if (psi) (void)hipFree(psi);
    if (K) (void)hipFree(K);
    if (blk) (void)hipFree(blk);
    if (dres) (void)hipFree(dres);
    return rc;
  });
}
