// Conserved-vector-current contraction Pi_{mu nu}(x), contact term, Ward
// identity, Pi(q) and the whole-source pipeline.
//
// The reference (cvc_exact2_xspace.c:577-861) sweeps the volume 4608 times
// per source.  Written with 12x12 propagator matrices S(x) (rows: sink
// spin-colour, columns: source spin-colour k = 3*spin + colour) the same sum is
//
//   conn_{mu nu}(x) = -1/4 sum_{k,l} { G1_{kl} [ Sd_nu(x)_l^+  Gm_mu U~_mu(x)   Su_4(x+mu)_k
//                                              + Sd_nu(x+mu)_l^+ Gp_mu U~_mu(x)^+ Su_4(x)_k ]
//                                    + G2_{kl} [ Sd_4(x)_l^+   Gm_mu U~_mu(x)   Su_nu(x+mu)_k
//                                              + Sd_4(x+mu)_l^+ Gp_mu U~_mu(x)^+ Su_nu(x)_k ] }
//
// with sink vertices Gm_mu = gamma5 (gamma_mu - 1), Gp_mu = gamma5 (gamma_mu + 1),
// source vertices G1 = (gamma_nu gamma5 - gamma5) (x) Us_nu and
// G2 = (gamma_nu gamma5 + gamma5) (x) Us_nu^+ (spin (x) colour, Us_nu = U~_nu(y)),
// Su_s / Sd_s the up / down propagators from y + s^ (s = 4: y).  Here the
// reference's gperm / gperm_sign / isimag tables (:336-383) are the non-zero
// entries of gamma_nu gamma5 and gamma5, and the conj(Us[ib][ia]) of part 2 is
// (Us^+)[ia][ib].
//
// GPU schedule, per nu:
//   k_psi:   Psi1_k(x) = sum_l conj(G1_kl) Sd_nu_l(x),  Psi2_k(x) = sum_l conj(G2_kl) Sd_4_l(x)
//            (one thread per (x, part, k); stored as a 24-column field)
//   k_conn:  one thread per (x, part, k) with its own Psi and up column in
//            registers for the four mu: the terms as <(gamma_mu -+ 1) gamma5 Psi,
//            U phi> (the sink vertex applied to Psi, the link to phi two spin
//            rows at a time), summed over (part, k) in LDS in a fixed order,
//            written as conn_{mu nu}(x) * (-1/4).
// FP64 throughout.
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <vector>

#include "contract_common.h"

namespace cvchip {

int invert_her_fields(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, int prec,
                      double delta, unsigned flags, cvchip_stats *stats);

// Psi fields for one nu.  up/dn: 60-column sets (12*s + k); psi: 24-column set.
// A block of 240 threads handles 10 sites: the 24 down-propagator columns the
// block needs (12 from y + nu, 12 from y, all 12 components) are staged in LDS
// with coalesced loads, each read once; then one thread per (x, part, k)
// (k fastest): part 0 is Psi1 from the columns of y + nu, part 1 Psi2 from y.
constexpr int kPsiSites = 10;
__global__ void __launch_bounds__(240) k_psi(double2 *__restrict__ psi, const double2 *__restrict__ dn,
                                             SrcLinks L, int nu, uint32_t V) {
  __shared__ double2 sd[kPsiSites][12][24];
  const uint32_t x0 = blockIdx.x * (uint32_t)kPsiSites;
  const uint32_t VN = V * 60u, VP = V * 24u;
  for (uint32_t e = threadIdx.x; e < (uint32_t)kPsiSites * 12u * 24u; e += blockDim.x) {
    const uint32_t col = e % 24u, rem = e / 24u, c = rem % 12u, xs = rem / 12u;
    const uint32_t x = x0 + xs;
    if (x < V) sd[xs][c][col] = dn[(size_t)c * VN + (size_t)x * 60u + (col < 12u ? 12u * (uint32_t)nu + col : 36u + col)];
  }
  __syncthreads();
  const uint32_t xs = threadIdx.x / 24u, x = x0 + xs;
  if (x >= V) return;
  const int j24 = (int)(threadIdx.x % 24u), part = j24 / 12, k = j24 % 12;
  const int ir = k / 3, ia = k % 3;
  // (gamma_nu gamma5)_{ir, j}: j = GCOL[nu][ir], value = GVAL[nu][ir] * g5sign(j)
  const int j = GCOL[nu][ir];
  double2 e = unit_val(GVAL[nu][ir]);
  if (g5sign(j) < 0) { e.x = -e.x; e.y = -e.y; }
  const double s5 = (double)g5sign(ir);
  const double2 *Us = L.u[nu];
  double2 p[12];
#pragma unroll
  for (int c = 0; c < 12; c++) p[c] = make_double2(0, 0);
  for (int ib = 0; ib < 3; ib++) {
    double2 wa, wb;
    if (part == 0) {
      // conj(G1_{k,l}) with G1 = (e d_{l,j} - s5 d_{l,ir}) Us[ia][ib]
      const double2 u1 = Us[3 * ia + ib];
      wa = cmulc(u1, cmulc(e, make_double2(1, 0)));      // conj(e * u1)
      wb = make_double2(-s5 * u1.x, s5 * u1.y);          // conj(-s5 * u1)
    } else {
      // G2 = (e d_{l,j} + s5 d_{l,ir}) conj(Us[ib][ia]); conj(G2) = conj(e) Us[ib][ia], s5 Us[ib][ia]
      const double2 u2 = Us[3 * ib + ia];
      wa = cmulc(e, u2);
      wb = make_double2(s5 * u2.x, s5 * u2.y);
    }
    const int la = 12 * part + 3 * j + ib, lb = 12 * part + 3 * ir + ib;
#pragma unroll
    for (int c = 0; c < 12; c++) p[c] = cfma(cfma(p[c], wa, sd[xs][c][la]), wb, sd[xs][c][lb]);
#pragma unroll
    for (int c = 0; c < 12; c++) pin(p[c]);
  }
#pragma unroll
  for (int c = 0; c < 12; c++) psi[(size_t)c * VP + (size_t)x * 24u + j24] = p[c];
}

// acc += <Psi, G U phi> for one term, G = gamma5 (gamma_MU + SGN) the sink
// vertex, computed as <(gamma_MU + SGN) gamma5 Psi, U phi> (the vertex moved
// onto Psi; DAG: U^dagger).  gamma_MU couples the spin rows in two pairs
// (a, GCOL[MU][a]), done one pair at a time.  OWN_PSI: Psi is the thread's own
// (registers) and phi is read at nb (plane stride VS), else phi is the own and
// Psi is read at nb.
template <int MU, int SGN, bool DAG, bool OWN_PSI>
__device__ __forceinline__ void conn_term(double2 &acc, const double2 own[12], const double2 *__restrict__ nb,
                                          uint32_t VS, const double2 u[9]) {
#pragma unroll
  for (int pr = 0; pr < 2; pr++) {
    const int a0 = pr, a1 = GCOL[MU][pr];
    double2 p0[3], p1[3], f0[3], f1[3];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      if (OWN_PSI) {
        p0[cc] = own[3 * a0 + cc];
        p1[cc] = own[3 * a1 + cc];
        f0[cc] = nb[(size_t)(3 * a0 + cc) * VS];
        f1[cc] = nb[(size_t)(3 * a1 + cc) * VS];
      } else {
        p0[cc] = nb[(size_t)(3 * a0 + cc) * VS];
        p1[cc] = nb[(size_t)(3 * a1 + cc) * VS];
        f0[cc] = own[3 * a0 + cc];
        f1[cc] = own[3 * a1 + cc];
      }
    }
    // v_a = e_a (gamma5 Psi)_{GCOL a} + SGN (gamma5 Psi)_a  for a = a0, a1
    const double2 e0 = unit_val(GVAL[MU][a0]), e1 = unit_val(GVAL[MU][a1]);
    const double s0 = (double)g5sign(a0), s1 = (double)g5sign(a1);
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
      const double2 g0 = make_double2(s0 * p0[ii].x, s0 * p0[ii].y), g1 = make_double2(s1 * p1[ii].x, s1 * p1[ii].y);
      const double2 q0 = cmul(e0, g1), q1 = cmul(e1, g0);
      const double2 v0 = make_double2(q0.x + SGN * g0.x, q0.y + SGN * g0.y);
      const double2 v1 = make_double2(q1.x + SGN * g1.x, q1.y + SGN * g1.y);
      double2 w0 = make_double2(0, 0), w1 = make_double2(0, 0);
#pragma unroll
      for (int jj = 0; jj < 3; jj++) {
        if (DAG) {
          w0 = cfmac(w0, u[3 * jj + ii], f0[jj]);
          w1 = cfmac(w1, u[3 * jj + ii], f1[jj]);
        } else {
          w0 = cfma(w0, u[3 * ii + jj], f0[jj]);
          w1 = cfma(w1, u[3 * ii + jj], f1[jj]);
        }
      }
      acc = cfmac(acc, v0, w0);
      acc = cfmac(acc, v1, w1);
    }
    pin(acc);
  }
}

// The two terms of one mu for the thread's own Psi (ops) and up column (oph) at
// site x: <Psi(x), Gm U phi(x+mu)> + <Psi(x+mu), Gp U^dagger phi(x)>.
template <int MU>
__device__ __forceinline__ void conn_mu(double2 &acc, const double2 ops[12], const double2 oph[12],
                                        const double2 *__restrict__ up, const double2 *__restrict__ psi,
                                        const double2 *__restrict__ U, const ConnParams &P, uint32_t x,
                                        uint32_t coord, uint32_t col_up, uint32_t col_psi) {
  const uint32_t L = (uint32_t)(MU == 0 ? P.T : MU == 1 ? P.LX : MU == 2 ? P.LY : P.LZ);
  const uint32_t stride = MU == 0 ? P.sT : MU == 1 ? P.sX : MU == 2 ? (uint32_t)P.LZ : 1u;
  const uint32_t xp = (coord == L - 1) ? x - (L - 1) * stride : x + stride;
  double2 u[9];
  const double2 ph = make_double2(P.ph_re[MU], P.ph_im[MU]);
#pragma unroll
  for (int q = 0; q < 9; q++) u[q] = cmul(U[(size_t)(MU * 9 + q) * P.V + x], ph);
  conn_term<MU, -1, false, true>(acc, ops, up + (size_t)xp * 60u + col_up, P.V * 60u, u);
  conn_term<MU, +1, true, false>(acc, oph, psi + (size_t)xp * 24u + col_psi, P.V * 24u, u);
}

// One thread per (x, part, k) (k fastest), blocks of 240 = 10 sites x 24: the
// thread keeps its own Psi_part(x)_k and up column phi_s(x)_k (s = y for part 1,
// y + nu for part 2) in registers for the four mu and reads only the neighbour
// columns at x + mu.  The 24 values of each (x, mu) are summed in LDS in a
// fixed order and written as conn_{mu nu}(x) * (-1/4).
__global__ void __launch_bounds__(240) k_conn(double2 *__restrict__ conn, const double2 *__restrict__ up,
                                              const double2 *__restrict__ psi, const double2 *__restrict__ U,
                                              ConnParams P, int nu) {
  __shared__ double2 red[4][240];
  const int t = threadIdx.x;
  const int j24 = t % 24, part = j24 / 12, k = j24 % 12;
  const uint32_t x = blockIdx.x * 10u + (uint32_t)(t / 24);
  double2 acc[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  if (x < P.V) {
    uint32_t rem = x;
    const uint32_t ct = rem / P.sT; rem -= ct * P.sT;
    const uint32_t cx = rem / P.sX; rem -= cx * P.sX;
    const uint32_t cy = rem / (uint32_t)P.LZ;
    const uint32_t cz = rem - cy * (uint32_t)P.LZ;
    const uint32_t VN = P.V * 60u, VP = P.V * 24u;
    // part 1: up from y (s = 4), Psi1; part 2: up from y + nu, Psi2
    const uint32_t col_up = (part == 0 ? 48u : 12u * (uint32_t)nu) + (uint32_t)k;
    const uint32_t col_psi = (uint32_t)j24;
    double2 ops[12], oph[12];
#pragma unroll
    for (int c = 0; c < 12; c++) ops[c] = psi[(size_t)c * VP + (size_t)x * 24u + col_psi];
#pragma unroll
    for (int c = 0; c < 12; c++) oph[c] = up[(size_t)c * VN + (size_t)x * 60u + col_up];
    // z first: the z neighbour's columns were read moments ago by the next site's threads
    conn_mu<3>(acc[3], ops, oph, up, psi, U, P, x, cz, col_up, col_psi);
    conn_mu<2>(acc[2], ops, oph, up, psi, U, P, x, cy, col_up, col_psi);
    conn_mu<1>(acc[1], ops, oph, up, psi, U, P, x, cx, col_up, col_psi);
    conn_mu<0>(acc[0], ops, oph, up, psi, U, P, x, ct, col_up, col_psi);
  }
#pragma unroll
  for (int mu = 0; mu < 4; mu++) red[mu][t] = acc[mu];
  __syncthreads();
  if (j24 < 4 && x < P.V) {
    const int mu = j24, t0 = t - j24;
    double2 s = make_double2(0, 0);
    for (int q = 0; q < 24; q++) { s.x += red[mu][t0 + q].x; s.y += red[mu][t0 + q].y; }
    conn[(size_t)x * 16 + 4 * mu + nu] = make_double2(-0.25 * s.x, -0.25 * s.y);
  }
}

// contact term pieces: thread (nu, ir, ia) -> part-1 and part-2 values
__global__ void k_contact(double2 *__restrict__ out, const double2 *__restrict__ up, SrcLinks L,
                          uint32_t ys, const uint32_t ysp[4], uint32_t V) {
  const int t = threadIdx.x;
  if (t >= 48) return;
  const int nu = t / 12, k = t % 12;
  const uint32_t VN = V * 60u;
  double2 phi[12], w[12];
  const double2 *Us = L.u[nu];
  // part 1: (gamma_nu - 1) Us phi(y + nu), phi = up from y (s = 4), column k, row k
  for (int c = 0; c < 12; c++) phi[c] = up[(size_t)c * VN + (size_t)ysp[nu] * 60u + 48 + k];
  su3_mul(w, Us, phi, false);
  const int a = k / 3, cc = k % 3;
  {
    const int b = GCOL[nu][a];
    double2 g = cmul(unit_val(GVAL[nu][a]), w[3 * b + cc]);
    out[2 * t] = make_double2(-0.5 * (g.x - w[k].x), -0.5 * (g.y - w[k].y));
  }
  // part 2: (gamma_nu + 1) Us^+ phi(y), phi = up from y + nu (s = nu)
  for (int c = 0; c < 12; c++) phi[c] = up[(size_t)c * VN + (size_t)ys * 60u + 12 * nu + k];
  su3_mul(w, Us, phi, true);
  {
    const int b = GCOL[nu][a];
    double2 g = cmul(unit_val(GVAL[nu][a]), w[3 * b + cc]);
    out[2 * t + 1] = make_double2(0.5 * (g.x + w[k].x), 0.5 * (g.y + w[k].y));
  }
}

__global__ void k_src_links(SrcLinks *out, const double2 *__restrict__ U, uint32_t V, uint32_t y,
                            ConnParams P) {
  const int t = threadIdx.x;
  if (t >= 36) return;
  const int nu = t / 9, q = t % 9;
  out->u[nu][q] = cmul(U[(size_t)(nu * 9 + q) * V + y], make_double2(P.ph_re[nu], P.ph_im[nu]));
}

// Source links Us_nu = U_nu(y) * phase_nu on the host (for kernels taking them by value).
int source_links(cvchip_ctx *c, const int src[4], SrcLinks *out) {
  SrcLinks *dL = nullptr;
  CVC_HIP(hipMalloc(&dL, sizeof(SrcLinks)));
  const uint32_t y = site_of(c, src[0], src[1], src[2], src[3]);
  hipLaunchKernelGGL(k_src_links, dim3(1), dim3(64), 0, c->stream, dL, (const double2 *)c->gauge64,
                     (uint32_t)c->lat.V, y, conn_params(c));
  hipError_t e = hipMemcpyAsync(out, dL, sizeof(SrcLinks), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dL);
  if (e != hipSuccess) return hip_fail(e, "source links", __FILE__, __LINE__);
  return 0;
}

// Psi1 / Psi2 of one nu (24-column set psi) from the down propagators
int launch_psi(cvchip_ctx *c, double2 *psi, const Field &dn, const SrcLinks &L, int nu) {
  const uint32_t V = (uint32_t)c->lat.V;
  hipLaunchKernelGGL(k_psi, dim3((V + kPsiSites - 1) / kPsiSites), dim3(240), 0, c->stream, psi,
                     (const double2 *)dn.data, L, nu, V);
  CVC_HIP(hipGetLastError());
  return 0;
}

// conn (host [V][16][2]) and contact from device up / dn sets of 60 columns
int contract_fields(cvchip_ctx *c, double *conn, double contact[8], const Field &up, const Field &dn,
                    const int src[4]) {
  if (!c->gauge_set) {
    set_error("gauge field not set (call cvchip_set_gauge first)");
    return CVCHIP_ERR_STATE;
  }
  if (up.nrhs != 60 || dn.nrhs != 60 || up.prec != CVCHIP_PREC_F64 || dn.prec != CVCHIP_PREC_F64) {
    set_error("contraction needs FP64 up and down sets of 60 propagator columns");
    return CVCHIP_ERR_ARG;
  }
  const uint32_t V = (uint32_t)c->lat.V;
  const ConnParams P = conn_params(c);
  const uint32_t y = site_of(c, src[0], src[1], src[2], src[3]);
  uint32_t ysp[4] = {site_of(c, src[0] + 1, src[1], src[2], src[3]), site_of(c, src[0], src[1] + 1, src[2], src[3]),
                     site_of(c, src[0], src[1], src[2] + 1, src[3]), site_of(c, src[0], src[1], src[2], src[3] + 1)};
  SrcLinks *dL = nullptr, hL;
  double2 *dconn = nullptr, *psi = nullptr, *dct = nullptr;
  uint32_t *dysp = nullptr;
  int rc = 0;
  auto fail = [&](hipError_t e, const char *w) { return hip_fail(e, w, __FILE__, __LINE__); };
  const size_t conn_bytes = sizeof(double2) * 16 * (size_t)V, psi_bytes = sizeof(double2) * 24 * 12 * (size_t)V;
  hipError_t e = hipMalloc(&dL, sizeof(SrcLinks));
  if (e == hipSuccess) e = hipMalloc(&dct, sizeof(double2) * 96);
  if (e == hipSuccess) e = hipMalloc(&dysp, sizeof(uint32_t) * 4);
  if (e != hipSuccess) rc = fail(e, "contraction buffers");
  if (!rc) rc = pool_get(c, (void **)&dconn, conn_bytes);   // the large buffers come from the context's pool
  if (!rc) rc = pool_get(c, (void **)&psi, psi_bytes);
  if (!rc) {
    hipLaunchKernelGGL(k_src_links, dim3(1), dim3(64), 0, c->stream, dL, (const double2 *)c->gauge64, V, y, P);
    e = hipMemcpyAsync(&hL, dL, sizeof(SrcLinks), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dysp, ysp, sizeof(ysp), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(e, "source links");
  }
  for (int nu = 0; nu < 4 && !rc; nu++) {
    hipLaunchKernelGGL(k_psi, dim3((V + kPsiSites - 1) / kPsiSites), dim3(240), 0, c->stream, psi,
                       (const double2 *)dn.data, hL, nu, V);
    hipLaunchKernelGGL(k_conn, dim3((V + 9u) / 10u), dim3(240), 0, c->stream, dconn, (const double2 *)up.data,
                       (const double2 *)psi, (const double2 *)c->gauge64, P, nu);
    e = hipGetLastError();
    if (e != hipSuccess) rc = fail(e, "contraction kernels");
  }
  if (!rc) {
    hipLaunchKernelGGL(k_contact, dim3(1), dim3(64), 0, c->stream, dct, (const double2 *)up.data, hL, y,
                       (const uint32_t *)dysp, V);
    double2 hct[96];
    rc = copy_to_host(c, conn, dconn, sizeof(double2) * 16 * (size_t)V);
    if (!rc) {
      e = hipMemcpyAsync(hct, dct, sizeof(hct), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) rc = fail(e, "contraction results");
    }
    if (!rc) {
      // accumulate in the reference's order: part 1 over (ir, ia), then part 2
      for (int nu = 0; nu < 4; nu++) {
        double re = 0., im = 0.;
        for (int k = 0; k < 12; k++) { re += hct[2 * (12 * nu + k)].x; im += hct[2 * (12 * nu + k)].y; }
        for (int k = 0; k < 12; k++) { re += hct[2 * (12 * nu + k) + 1].x; im += hct[2 * (12 * nu + k) + 1].y; }
        contact[2 * nu] = re;
        contact[2 * nu + 1] = im;
      }
    }
  }
  if (dL) (void)hipFree(dL);
  pool_put(c, dconn, conn_bytes);
  pool_put(c, psi, psi_bytes);
  if (dct) (void)hipFree(dct);
  if (dysp) (void)hipFree(dysp);
  return rc;
}

// ------------------------------------------------------------ Ward identity
__global__ void k_wi(double2 *__restrict__ wi, const double2 *__restrict__ conn, ConnParams P) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= P.V) return;
  uint32_t rem = x;
  const uint32_t ct = rem / P.sT; rem -= ct * P.sT;
  const uint32_t cx = rem / P.sX; rem -= cx * P.sX;
  const uint32_t cy = rem / (uint32_t)P.LZ;
  const uint32_t cz = rem - cy * (uint32_t)P.LZ;
  const uint32_t xm[4] = {ct == 0 ? x + (P.T - 1) * P.sT : x - P.sT, cx == 0 ? x + (P.LX - 1) * P.sX : x - P.sX,
                          cy == 0 ? x + (P.LY - 1) * P.LZ : x - P.LZ, cz == 0 ? x + (P.LZ - 1) : x - 1};
  for (int nu = 0; nu < 4; nu++) {
    double re = 0., im = 0.;
    for (int mu = 0; mu < 4; mu++) {
      re += conn[(size_t)x * 16 + 4 * mu + nu].x;
      im += conn[(size_t)x * 16 + 4 * mu + nu].y;
    }
    for (int mu = 0; mu < 4; mu++) {
      re -= conn[(size_t)xm[mu] * 16 + 4 * mu + nu].x;
      im -= conn[(size_t)xm[mu] * 16 + 4 * mu + nu].y;
    }
    wi[(size_t)x * 4 + nu] = make_double2(re, im);
  }
}

// -------------------------------------------------- point sources, copies
// b[c][site][r] = delta(site, ys) delta(c, r) for r = 0..11 (12-column set)
__global__ void k_point_sources(double2 *__restrict__ b, uint32_t V, uint32_t ys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;   // over 12 * V * 12
  if (i >= 144u * V) return;
  const uint32_t r = i % 12u;
  const uint32_t c = i / (12u * V);
  const uint32_t site = (i / 12u) % V;
  b[i] = make_double2((site == ys && c == r) ? 1.0 : 0.0, 0.0);
}

// dst (ND columns) [.., col0 + r] = src (12 columns) [.., r]
__global__ void k_put_columns(double2 *__restrict__ dst, const double2 *__restrict__ src, uint32_t V, int ND,
                              int col0) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;   // over 12 * V * 12
  if (i >= 144u * V) return;
  const uint32_t r = i % 12u;
  const uint32_t cs = i / 12u;                                 // c * V + site
  dst[(size_t)cs * ND + col0 + r] = src[i];
}

// dst columns col0 .. col0 + 11 of an nd-column FP64 set = the 12 columns of src
int put_columns(cvchip_ctx *c, Field &dst, const Field &src, int col0) {
  const uint32_t V = (uint32_t)c->lat.V;
  hipLaunchKernelGGL(k_put_columns, dim3((144u * V + 255) / 256), dim3(256), 0, c->stream, (double2 *)dst.data,
                     (const double2 *)src.data, V, dst.nrhs, col0);
  CVC_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------------ Pi(q)
// One pass of a length-L DFT along a direction of stride `stride` (in sites):
// out[.. n ..] = sum_x w[(n x) mod L] in[.. x ..],  w[k] = exp(-2 pi i k / L).
// 16 components per site, one thread per (site, component).
__global__ void k_dft_dir(double2 *__restrict__ out, const double2 *__restrict__ in,
                          const double2 *__restrict__ w, uint32_t V, int L, uint32_t stride) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * 16u) return;
  const uint32_t site = i / 16u, comp = i % 16u;
  const uint32_t n = (site / stride) % (uint32_t)L;
  const uint32_t base = site - n * stride;
  double2 acc = make_double2(0, 0);
  uint32_t k = 0;
  for (int x = 0; x < L; x++) {
    acc = cfma(acc, w[k], in[(size_t)(base + x * stride) * 16u + comp]);
    k += n;
    if (k >= (uint32_t)L) k -= L;
  }
  out[i] = acc;
}

// Pi_{mu nu}(q) = exp(i q.y) exp(-i (q_mu - q_nu)/2) F_{mu nu}(q) - delta_{mu nu} contact_nu,
// F = DFT of conn; tables per direction d: ey[d][n] = exp(2 pi i n y_d / L_d),
// eh[d][n] = exp(-i pi n / L_d).
struct PiqParams {
  uint32_t V, sX, sT;
  int LZ, L[4];
  double2 contact[4];
};
__global__ void k_piq_phase(double2 *__restrict__ piq, const double2 *__restrict__ ey,
                            const double2 *__restrict__ eh, PiqParams P) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.V * 16u) return;
  const uint32_t site = i / 16u, comp = i % 16u;
  const int mu = comp / 4, nu = comp % 4;
  uint32_t rem = site;
  uint32_t n[4];
  n[0] = rem / P.sT; rem -= n[0] * P.sT;
  n[1] = rem / P.sX; rem -= n[1] * P.sX;
  n[2] = rem / (uint32_t)P.LZ;
  n[3] = rem - n[2] * (uint32_t)P.LZ;
  double2 f = make_double2(1, 0);
  for (int d = 0; d < 4; d++) f = cmul(f, ey[64 * d + n[d]]);
  f = cmul(f, eh[64 * mu + n[mu]]);
  f = cmulc(eh[64 * nu + n[nu]], f);
  double2 v = cmul(f, piq[i]);
  if (mu == nu) { v.x -= P.contact[nu].x; v.y -= P.contact[nu].y; }
  piq[i] = v;
}

int pi_q_device(cvchip_ctx *c, double2 *dpiq, double2 *tmp, const double contact[8], const int src[4]) {
  const Lattice &Lt = c->lat;
  const uint32_t V = (uint32_t)Lt.V;
  const int Ls[4] = {Lt.T, Lt.LX, Lt.LY, Lt.LZ};
  for (int d = 0; d < 4; d++)
    if (Ls[d] > 64) {
      set_error("Pi(q) supports extents up to 64 (got %d)", Ls[d]);
      return CVCHIP_ERR_ARG;
    }
  const uint32_t strides[4] = {(uint32_t)(Lt.LX * Lt.LY * Lt.LZ), (uint32_t)(Lt.LY * Lt.LZ), (uint32_t)Lt.LZ, 1u};
  double2 hw[4][64], hy[4][64], hh[4][64];
  for (int d = 0; d < 4; d++)
    for (int k = 0; k < 64; k++) {
      const double a = 2.0 * M_PI * (double)(k % Ls[d]) / (double)Ls[d];
      hw[d][k] = make_double2(cos(a), -sin(a));
      const double ay = 2.0 * M_PI * (double)(((int64_t)k * src[d]) % Ls[d]) / (double)Ls[d];
      hy[d][k] = make_double2(cos(ay), sin(ay));
      const double ah = M_PI * (double)k / (double)Ls[d];
      hh[d][k] = make_double2(cos(ah), -sin(ah));
    }
  double2 *tabs = nullptr;
  CVC_HIP(hipMalloc(&tabs, sizeof(double2) * 3 * 256));
  hipError_t e = hipMemcpyAsync(tabs, hw, sizeof(hw), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(tabs + 256, hy, sizeof(hy), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(tabs + 512, hh, sizeof(hh), hipMemcpyHostToDevice, c->stream);
  const dim3 g((V * 16u + 255) / 256), b(256);
  double2 *a = dpiq, *o = tmp;
  for (int d = 0; d < 4 && e == hipSuccess; d++) {
    hipLaunchKernelGGL(k_dft_dir, g, b, 0, c->stream, o, (const double2 *)a, (const double2 *)(tabs + 64 * d), V,
                       Ls[d], strides[d]);
    double2 *t = a; a = o; o = t;
  }
  // after 4 passes the result is back in dpiq
  PiqParams P;
  P.V = V;
  P.sX = strides[1];
  P.sT = strides[0];
  P.LZ = Lt.LZ;
  for (int d = 0; d < 4; d++) P.L[d] = Ls[d];
  for (int nu = 0; nu < 4; nu++) P.contact[nu] = make_double2(contact[2 * nu], contact[2 * nu + 1]);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_piq_phase, g, b, 0, c->stream, dpiq, (const double2 *)(tabs + 256),
                       (const double2 *)(tabs + 512), P);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(tabs);
  if (e != hipSuccess) return hip_fail(e, "Pi(q)", __FILE__, __LINE__);
  return 0;
}

}  // namespace cvchip

using namespace cvchip;

extern "C" int cvchip_field_cvc_contract(cvchip_ctx *c, double *conn, double contact[8],
                                         const cvchip_field *up, const cvchip_field *dn, const int src[4]) {
  return cvchip::guarded("cvchip_field_cvc_contract", [&]() -> int {
    if (!c || !conn || !contact || !up || !dn || !src) {
      set_error("NULL argument to cvchip_field_cvc_contract");
      return CVCHIP_ERR_NULL;
    }
    return contract_fields(c, conn, contact, up->f, dn->f, src);
  });
}

extern "C" int cvchip_cvc_contract(cvchip_ctx *c, double *conn, double contact[8], const double *const *props,
                                   const int src[4]) {
  return cvchip::guarded("cvchip_cvc_contract", [&]() -> int {
    if (!c || !conn || !contact || !props || !src) {
      set_error("NULL argument to cvchip_cvc_contract");
      return CVCHIP_ERR_NULL;
    }
    for (int k = 0; k < 120; k++)
      if (!props[k]) {
        set_error("propagator %d is NULL", k);
        return CVCHIP_ERR_NULL;
      }
    Field up, dn;
    int rc = field_alloc(c, up, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, dn, 60, CVCHIP_PREC_F64);
    for (int k = 0; k < 120 && !rc; k++) rc = upload_aos(c, k < 60 ? up : dn, props[k], k % 60, 1);
    if (!rc) rc = contract_fields(c, conn, contact, up, dn, src);
    field_free(up);
    field_free(dn);
    return rc;
  });
}

extern "C" int cvchip_wi(cvchip_ctx *c, double *wi, const double *conn) {
  return cvchip::guarded("cvchip_wi", [&]() -> int {
    if (!c || !wi || !conn) {
      set_error("NULL argument to cvchip_wi");
      return CVCHIP_ERR_NULL;
    }
    const size_t V = (size_t)c->lat.V;
    double2 *dc = nullptr, *dw = nullptr;
    CVC_HIP(hipMalloc(&dc, sizeof(double2) * 16 * V));
    hipError_t e = hipMalloc(&dw, sizeof(double2) * 4 * V);
    int rc = e == hipSuccess ? 0 : hip_fail(e, "Ward identity buffers", __FILE__, __LINE__);
    if (!rc) rc = copy_to_device(c, dc, conn, sizeof(double2) * 16 * V);
    if (!rc) {
      hipLaunchKernelGGL(k_wi, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, c->stream, dw, (const double2 *)dc,
                         conn_params(c));
      e = hipGetLastError();
      if (e != hipSuccess) rc = hip_fail(e, "Ward identity", __FILE__, __LINE__);
    }
    if (!rc) rc = copy_to_host(c, wi, dw, sizeof(double2) * 4 * V);
    (void)hipStreamSynchronize(c->stream);   // nothing in flight before the buffers go
    (void)hipFree(dc);
    if (dw) (void)hipFree(dw);
    return rc;
  });
}

extern "C" int cvchip_pi_q(cvchip_ctx *c, double *piq, const double *conn, const double contact[8],
                           const int src[4]) {
  return cvchip::guarded("cvchip_pi_q", [&]() -> int {
    if (!c || !piq || !conn || !contact || !src) {
      set_error("NULL argument to cvchip_pi_q");
      return CVCHIP_ERR_NULL;
    }
    const size_t V = (size_t)c->lat.V;
    double2 *a = nullptr, *t = nullptr;
    CVC_HIP(hipMalloc(&a, sizeof(double2) * 16 * V));
    hipError_t e = hipMalloc(&t, sizeof(double2) * 16 * V);
    int rc = e == hipSuccess ? 0 : hip_fail(e, "Pi(q) buffers", __FILE__, __LINE__);
    if (!rc) rc = copy_to_device(c, a, conn, sizeof(double2) * 16 * V);
    if (!rc) rc = pi_q_device(c, a, t, contact, src);
    if (!rc) rc = copy_to_host(c, piq, a, sizeof(double2) * 16 * V);
    (void)hipStreamSynchronize(c->stream);   // nothing in flight before the buffers go
    (void)hipFree(a);
    if (t) (void)hipFree(t);
    return rc;
  });
}

namespace cvchip {

// 5 x 12 up (mu) and 5 x 12 down (-mu) propagators from point sources at y + s^
// (s = 0..3) and y (s = 4) into the 60-column sets up / dn (column 12 s + k),
// with the gamma5-hermitian BiCGStab; statistics summed into tot.
int solve_source_props(cvchip_ctx *c, const int src[4], double sp, int niter_max, int prec, double delta,
                       Field &up, Field &dn, cvchip_stats *tot) {
  const uint32_t V = (uint32_t)c->lat.V;
  Field b, x;
  int rc = field_alloc(c, b, 12, CVCHIP_PREC_F64);
  if (!rc) rc = field_alloc(c, x, 12, CVCHIP_PREC_F64);
  const double mu0 = c->mu;
  for (int s = 0; s < 5 && !rc; s++) {
    const int d[4] = {s == 0, s == 1, s == 2, s == 3};
    const uint32_t ys = site_of(c, src[0] + d[0], src[1] + d[1], src[2] + d[2], src[3] + d[3]);
    for (int fl = 0; fl < 2 && !rc; fl++) {
      hipLaunchKernelGGL(k_point_sources, dim3((144u * V + 255) / 256), dim3(256), 0, c->stream,
                         (double2 *)b.data, V, ys);
      if (hipMemsetAsync(x.data, 0, x.bytes, c->stream) != hipSuccess) rc = CVCHIP_ERR_HIP;
      c->mu = fl == 0 ? mu0 : -mu0;              // down flavour: mu -> -mu (record 1)
      cvchip_stats st;
      memset(&st, 0, sizeof(st));
      if (!rc) rc = invert_her_fields(c, x, b, sp, niter_max, prec, delta, 0, &st);
      c->mu = mu0;
      if (rc > 0) rc = 0;
      if (rc) break;
      if (tot) {
        tot->iterations += st.iterations;
        tot->converged += st.converged;
        tot->reliable_updates += st.reliable_updates;
        tot->dslash_calls += st.dslash_calls;
        tot->seconds += st.seconds;
        if (st.true_res2_max > tot->true_res2_max) tot->true_res2_max = st.true_res2_max;
      }
      Field &dst = fl == 0 ? up : dn;
      hipLaunchKernelGGL(k_put_columns, dim3((144u * V + 255) / 256), dim3(256), 0, c->stream, (double2 *)dst.data,
                         (const double2 *)x.data, V, 60, 12 * s);
      if (hipGetLastError() != hipSuccess) rc = CVCHIP_ERR_HIP;
    }
  }
  field_free(b);
  field_free(x);
  return rc;
}

}  // namespace cvchip

// The whole path for one source point: 60 up and 60 down solves, then the
// contraction.
extern "C" int cvchip_cvc_source(cvchip_ctx *c, double *conn, double contact[8], const int src[4], double sp,
                                 int niter_max, int prec, double delta, cvchip_stats *stats) {
  return cvchip::guarded("cvchip_cvc_source", [&]() -> int {
    if (!c || !conn || !contact || !src) {
      set_error("NULL argument to cvchip_cvc_source");
      return CVCHIP_ERR_NULL;
    }
    Field up, dn;
    int rc = field_alloc(c, up, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, dn, 60, CVCHIP_PREC_F64);
    cvchip_stats tot;
    memset(&tot, 0, sizeof(tot));
    if (!rc) rc = solve_source_props(c, src, sp, niter_max, prec, delta, up, dn, &tot);
    if (!rc) rc = contract_fields(c, conn, contact, up, dn, src);
    if (stats) *stats = tot;
    field_free(up);
    field_free(dn);
    return rc;
  });
}

// The propagators of cvchip_cvc_source, downloaded into the reference layout.
extern "C" int cvchip_cvc_propagators(cvchip_ctx *c, double *const *props, const int src[4], double sp, int niter_max,
                                      int prec, double delta, cvchip_stats *stats) {
  return cvchip::guarded("cvchip_cvc_propagators", [&]() -> int {
    if (!c || !props || !src) {
      set_error("NULL argument to cvchip_cvc_propagators");
      return CVCHIP_ERR_NULL;
    }
    for (int k = 0; k < 120; k++)
      if (!props[k]) {
        set_error("propagator %d is NULL", k);
        return CVCHIP_ERR_NULL;
      }
    Field up, dn;
    int rc = field_alloc(c, up, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, dn, 60, CVCHIP_PREC_F64);
    cvchip_stats tot;
    memset(&tot, 0, sizeof(tot));
    if (!rc) rc = solve_source_props(c, src, sp, niter_max, prec, delta, up, dn, &tot);
    for (int k = 0; k < 120 && !rc; k++) rc = download_aos(c, k < 60 ? up : dn, props[k], k % 60, 1);
    if (stats) *stats = tot;
    field_free(up);
    field_free(dn);
    return rc;
  });
}
