// P -> gamma gamma in position space (reference p2gg_xspace.c): the
// conserved-vector-current contraction of cvc_exact2_xspace.c with the 60
// propagators of one flavour replaced by sequential propagators.
//
// For each of the 60 point-source columns S_s,k of the flavour f =
// propagator_position (0: up, D(mu); 1: down, D(-mu)) the sequential source
//   b(x) = delta_{x_0, t_seq} e^{i sgn 2 pi (x p_x / LX + y p_y / LY + z p_z / LZ)} Gamma S_s,k(x),
// t_seq = (y_0 + seq_timeslice) mod T, sgn = 1 - 2 f (p2gg_xspace.c:517-548),
// is inverted with D_f (the reference's tmLQCD_invert with op_id = f,
// :572-581), and the contraction (:615-899, the same loops as
// cvc_exact2_xspace.c:577-861) takes the sequential propagators in place of the
// records of flavour f and the point propagators of the other flavour in place
// of its second records.
#include <math.h>
#include <string.h>

#include "contract_common.h"

namespace cvchip {

int solve_source_props(cvchip_ctx *c, const int src[4], double sp, int niter_max, int prec, double delta,
                       Field &up, Field &dn, cvchip_stats *tot);
int contract_fields(cvchip_ctx *c, double *conn, double contact[8], const Field &up, const Field &dn,
                    const int src[4]);
int invert_her_fields(cvchip_ctx *c, Field &x, const Field &b, double sp, int niter_max, int prec,
                      double delta, unsigned flags, cvchip_stats *stats);
int put_columns(cvchip_ctx *c, Field &dst, const Field &src, int col0);

struct SeqParams {
  int col[4];
  double re[4], im[4];
  uint32_t V, VNi, VNo;
  int Ni, No, col0;
  uint32_t tseq, sT, sX;
  int LZ;
  double kx, ky, kz;   // sgn 2 pi p_i / L_i
};

// out(c, site, r) for r < No from in(c, site, col0 + r): one thread per (site, r)
__global__ void __launch_bounds__(256) k_seq_source(double2 *__restrict__ out, const double2 *__restrict__ in,
                                                    SeqParams P) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.V * (uint32_t)P.No) return;
  const uint32_t site = i / (uint32_t)P.No, r = i - site * (uint32_t)P.No;
  const uint32_t t = site / P.sT;
  if (t != P.tseq) {
#pragma unroll
    for (int c = 0; c < 12; c++) out[(size_t)c * P.VNo + i] = make_double2(0., 0.);
    return;
  }
  const uint32_t rem = site - t * P.sT;
  const uint32_t x = rem / P.sX, y = (rem % P.sX) / (uint32_t)P.LZ, z = rem % (uint32_t)P.LZ;
  const double ph = x * P.kx + y * P.ky + z * P.kz;
  double sn, cs;
  sincos(ph, &sn, &cs);
  const double2 w = make_double2(cs, sn);
  const size_t ii = (size_t)site * (uint32_t)P.Ni + (uint32_t)P.col0 + r;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const double2 g = make_double2((double)P.re[a], (double)P.im[a]);
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      const double2 v = cmul(g, in[(size_t)(3 * P.col[a] + cc) * P.VNi + ii]);
      out[(size_t)(3 * a + cc) * P.VNo + i] = cmul(v, w);
    }
  }
}

// out (No columns) = sequential sources of in's columns col0 .. col0 + No - 1
static int seq_sources(cvchip_ctx *c, Field &out, const Field &in, int col0, int tseq, int gid, const int mom[3],
                       double sgn) {
  const Lattice &L = c->lat;
  if (in.prec != CVCHIP_PREC_F64 || out.prec != CVCHIP_PREC_F64 || in.V != L.V || out.V != L.V || col0 < 0 ||
      col0 + out.nrhs > in.nrhs) {
    set_error("sequential sources: FP64 sets on the context's lattice, columns inside the input set");
    return CVCHIP_ERR_ARG;
  }
  if (gid < 0 || gid > 15 || tseq < 0 || tseq >= L.T) {
    set_error("sequential sources: gamma 0..15 and time slice 0..T-1 (got %d, %d)", gid, tseq);
    return CVCHIP_ERR_ARG;
  }
  SeqParams P;
  int re[4], im[4];
  gamma_matrix_rows(gid, P.col, re, im);
  for (int a = 0; a < 4; a++) {
    P.re[a] = re[a];
    P.im[a] = im[a];
  }
  P.V = (uint32_t)L.V;
  P.Ni = in.nrhs;
  P.No = out.nrhs;
  P.col0 = col0;
  P.VNi = (uint32_t)(L.V * in.nrhs);
  P.VNo = (uint32_t)(L.V * out.nrhs);
  P.tseq = (uint32_t)tseq;
  P.sX = (uint32_t)(L.LY * L.LZ);
  P.sT = (uint32_t)L.LX * P.sX;
  P.LZ = L.LZ;
  const int m0 = mom ? mom[0] : 0, m1 = mom ? mom[1] : 0, m2 = mom ? mom[2] : 0;
  P.kx = sgn * 2. * M_PI * m0 / (double)L.LX;
  P.ky = sgn * 2. * M_PI * m1 / (double)L.LY;
  P.kz = sgn * 2. * M_PI * m2 / (double)L.LZ;
  const uint32_t n = P.V * (uint32_t)P.No;
  hipLaunchKernelGGL(k_seq_source, dim3((n + 255) / 256), dim3(256), 0, c->stream, (double2 *)out.data,
                     (const double2 *)in.data, P);
  CVC_HIP(hipGetLastError());
  return 0;
}

}  // namespace cvchip

using namespace cvchip;

extern "C" int cvchip_field_sequential_sources(cvchip_ctx *c, cvchip_field *out, const cvchip_field *in, int col0,
                                               int tseq, int gamma_id, const int mom[3], double sgn) {
  return cvchip::guarded("cvchip_field_sequential_sources", [&]() -> int {
    if (!c || !out || !in) {
      set_error("NULL argument to cvchip_field_sequential_sources");
      return CVCHIP_ERR_NULL;
    }
    CVC_TRY(seq_sources(c, out->f, in->f, col0, tseq, gamma_id, mom, sgn));
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}

// The whole P -> gamma gamma source point: 120 point propagators, 60
// sequential sources and solves (12 at a time), the contraction.
extern "C" int cvchip_p2gg_source(cvchip_ctx *c, double *conn, double contact[8], const int src[4], int seq_timeslice,
                                  int seq_gamma_id, const int seq_mom[3], int propagator_position, double sp,
                                  int niter_max, int prec, double delta, cvchip_stats *stats) {
  return cvchip::guarded("cvchip_p2gg_source", [&]() -> int {
    if (!c || !conn || !contact || !src) {
      set_error("NULL argument to cvchip_p2gg_source");
      return CVCHIP_ERR_NULL;
    }
    if (propagator_position != 0 && propagator_position != 1) {
      set_error("propagator_position must be 0 (up) or 1 (down), got %d", propagator_position);
      return CVCHIP_ERR_ARG;
    }
    if (seq_gamma_id < 0 || seq_gamma_id > 15) {
      set_error("sequential source gamma must be 0..15, got %d", seq_gamma_id);
      return CVCHIP_ERR_ARG;
    }
    const int T = c->lat.T;
    const int tseq = (((src[0] + seq_timeslice) % T) + T) % T;   // p2gg_xspace.c:446
    const double sgn = 1. - 2. * propagator_position;
    Field up, dn, seq, b, x;
    int rc = field_alloc(c, up, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, dn, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, seq, 60, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, b, 12, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, x, 12, CVCHIP_PREC_F64);
    cvchip_stats tot;
    memset(&tot, 0, sizeof(tot));
    if (!rc) rc = solve_source_props(c, src, sp, niter_max, prec, delta, up, dn, &tot);
    Field &P = propagator_position == 0 ? up : dn;
    Field &Q = propagator_position == 0 ? dn : up;
    const double mu0 = c->mu;
    for (int s = 0; s < 5 && !rc; s++) {
      rc = seq_sources(c, b, P, 12 * s, tseq, seq_gamma_id, seq_mom, sgn);
      if (!rc && hipMemsetAsync(x.data, 0, x.bytes, c->stream) != hipSuccess) rc = CVCHIP_ERR_HIP;
      if (rc) break;
      c->mu = propagator_position == 0 ? mu0 : -mu0;
      cvchip_stats st;
      memset(&st, 0, sizeof(st));
      rc = invert_her_fields(c, x, b, sp, niter_max, prec, delta, 0, &st);
      c->mu = mu0;
      if (rc > 0) rc = 0;
      if (rc) break;
      tot.iterations += st.iterations;
      tot.converged += st.converged;
      tot.reliable_updates += st.reliable_updates;
      tot.dslash_calls += st.dslash_calls;
      tot.seconds += st.seconds;
      if (st.true_res2_max > tot.true_res2_max) tot.true_res2_max = st.true_res2_max;
      rc = put_columns(c, seq, x, 12 * s);
    }
    if (!rc) rc = contract_fields(c, conn, contact, seq, Q, src);
    if (stats) *stats = tot;
    field_free(up);
    field_free(dn);
    field_free(seq);
    field_free(b);
    field_free(x);
    return rc;
  });
}
