// Drop-in replacement of the reference's operator and solver translation
// units (Q_phi.c / invert_Qtm.c) for the twisted-mass path.  A reference
// program links this file and libcvchip.so instead of Q_phi.o and
// invert_Qtm.o; the functions keep the reference's names, C++ linkage in
// namespace cvc, argument meaning, return codes and stdout lines, and read the
// same globals (global.h:91-154), which the program defines as before.
//
// Replaced entry points (reference file:line):
//   void Q_phi_tbc(double *xi, double *phi)                  Q_phi.h:10, Q_phi.c:38-188
//   void Q_phi(double *xi, double *phi, const double mutm)   Q_phi.h:28, Q_phi.c:1092-1242
//   void Qf5(double *xi, double *phi, double mutm)           Q_phi.h:26, Q_phi.c:789-934
//   void g5_phi(double *phi)                                 Q_phi.h:37, Q_phi.c:350-385
//   void spinor_scalar_product_co(complex*, double*, double*, int)   invert_Qtm.h:8, invert_Qtm.c:28-58
//   void spinor_scalar_product_re(double*, double*, double*, int)    invert_Qtm.h:9, invert_Qtm.c:60-84
//   int  invert_Qtm(double *xi, double *phi, int kwork)              invert_Qtm.h:10, invert_Qtm.c:86-270
//   int  invert_Qtm_her(double *xi, double *phi, int kwork)          invert_Qtm.h:11, invert_Qtm.c:276-478
//
// Differences a caller can observe: the operator needs no halo exchange (one
// GPU holds the whole lattice, so xchange_field() is not needed around it);
// reductions are tree-ordered, so the last digits of dot products and the
// Krylov trajectory differ from the sequential reference sums.  The solvers
// keep the reference's lagged return iterate (invert_Qtm.c:199-211).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <thread>
#include <vector>

#include "cvc_dropin.h"
#include "cvchip.h"

namespace cvc {

// the reference's complex type (cvc_complex.h:16-18): same layout and name
typedef struct {
  double re, im;
} complex;

// globals defined by the reference program's MAIN_PROGRAM unit (global.h:91-154)
extern int T, LX, LY, LZ, VOLUME, T_global;
extern double **g_spinor_field;
extern double *g_gauge_field;
extern double g_kappa, g_mu;
extern int g_cart_id;
extern double BCangle[4];
extern double solver_precision;
extern int niter_max;

}  // namespace cvc

namespace {

cvchip_ctx *g_ctx = nullptr;
int g_dims[4] = {0, 0, 0, 0};
double g_bc[4] = {0, 0, 0, 0};
const double *g_gauge_seen = nullptr;
uint64_t g_gauge_hash = 0;
bool g_gauge_stale = false;   // set by cvchip_gauge_changed()
double g_kappa_seen = 0., g_mu_seen = 0.;

void die(int rc, const char *where) {
  fprintf(stderr, "[cvc_dropin] %s failed (%d): %s\n", where, rc, cvchip_last_error());
  exit(1);
}

const uint64_t kHashMul = 0x9e3779b97f4a7c15ULL;

// one block of words: four independent multiply-xorshift lanes, folded
uint64_t block_hash(const double *g, long n, uint64_t seed) {
  uint64_t h[4] = {seed + 1, seed + 2, seed + 3, seed + 4};
  long i = 0;
  for (; i + 4 <= n; i += 4) {
    for (int l = 0; l < 4; l++) {
      uint64_t w;
      memcpy(&w, g + i + l, sizeof(w));
      h[l] = (h[l] ^ w) * kHashMul;
      h[l] ^= h[l] >> 29;
    }
  }
  for (; i < n; i++) {
    uint64_t w;
    memcpy(&w, g + i, sizeof(w));
    h[0] = (h[0] ^ w) * kHashMul;
    h[0] ^= h[0] >> 29;
  }
  uint64_t r = (uint64_t)n;
  for (int l = 0; l < 4; l++) r = (r ^ h[l]) * kHashMul;
  return r ^ (r >> 31);
}

// Hash of every 64-bit word of the gauge field: 256 fixed blocks hashed by up
// to 16 host threads (the blocks, and so the value, do not depend on the
// number of threads), the block values folded in block order.  Every step is a
// bijection of the state in the word or block value it takes in, so a change
// of one word (one link entry, also a single bit) always changes the hash;
// changes of several words are missed only by a chance collision of the
// 64-bit value.  In place edits of g_gauge_field (a local gauge
// transformation as in apply_gt_gauge, cvc_utils.c:2044-2068, or a one-link
// finite difference) are seen by the next operator or solver call.
uint64_t gauge_hash(const double *g, long n) {
  constexpr int kBlocks = 256;
  uint64_t hb[kBlocks];
  const long per = (n + kBlocks - 1) / kBlocks;
  auto blocks = [&](unsigned first, unsigned step) {
    for (unsigned b = first; b < (unsigned)kBlocks; b += step) {
      const long lo = (long)b * per < n ? (long)b * per : n;
      const long hi = lo + per < n ? lo + per : n;
      hb[b] = block_hash(g + lo, hi - lo, (uint64_t)b << 8);
    }
  };
  unsigned nt = std::thread::hardware_concurrency();
  if (nt < 1 || n < (1L << 20)) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nt; t++) {
    try {
      pool.emplace_back(blocks, t, nt);
    } catch (...) {   // no thread to be had: its blocks on the calling thread
      blocks(t, nt);
    }
  }
  blocks(0, nt);
  for (auto &th : pool) th.join();
  uint64_t r = (uint64_t)n;
  for (int b = 0; b < kBlocks; b++) {
    r = (r ^ hb[b]) * kHashMul;
    r ^= r >> 29;
  }
  return r ^ (r >> 31);
}

// Context for the current globals; (re)uploads the gauge field when the
// pointer or the hash of its contents changed since the last call, or when the
// program announced a change with cvchip_gauge_changed().
cvchip_ctx *ctx() {
  using namespace cvc;
  const int dims[4] = {T, LX, LY, LZ};
  bool fresh = !g_ctx || memcmp(dims, g_dims, sizeof(dims)) != 0 || memcmp(BCangle, g_bc, sizeof(g_bc)) != 0;
  if (fresh) {
    if (g_ctx) cvchip_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    int rc = cvchip_ctx_create(&g_ctx, dims, BCangle, g_kappa, g_mu);
    if (rc) die(rc, "cvchip_ctx_create");
    memcpy(g_dims, dims, sizeof(dims));
    memcpy(g_bc, BCangle, sizeof(g_bc));
    g_gauge_seen = nullptr;
    g_kappa_seen = g_kappa;
    g_mu_seen = g_mu;
  }
  if (g_kappa != g_kappa_seen || g_mu != g_mu_seen) {
    int rc = cvchip_ctx_set_params(g_ctx, g_kappa, g_mu);
    if (rc) die(rc, "cvchip_ctx_set_params");
    g_kappa_seen = g_kappa;
    g_mu_seen = g_mu;
  }
  const long n = 72L * VOLUME;
  const uint64_t hash = g_gauge_field ? gauge_hash(g_gauge_field, n) : 0;
  if (g_gauge_field && (g_gauge_stale || g_gauge_field != g_gauge_seen || hash != g_gauge_hash)) {
    int rc = cvchip_set_gauge(g_ctx, g_gauge_field);
    if (rc) die(rc, "cvchip_set_gauge");
    g_gauge_seen = g_gauge_field;
    g_gauge_hash = hash;
    g_gauge_stale = false;
  }
  return g_ctx;
}

}  // namespace

// No counterpart in the reference (its operator reads g_gauge_field on every
// call): a program that changed g_gauge_field in place calls this to have the
// next operator or solver call upload it again, whatever the hash says.
extern "C" void cvchip_gauge_changed(void) { g_gauge_stale = true; }

namespace cvc {

void Q_phi_tbc(double *xi, double *phi) {
  int rc = cvchip_Q_phi(ctx(), xi, phi, g_mu, 0);
  if (rc) die(rc, "Q_phi_tbc");
}

void Q_phi(double *xi, double *phi, const double mutm) {
  int rc = cvchip_Q_phi(ctx(), xi, phi, mutm, 0);
  if (rc) die(rc, "Q_phi");
}

void Qf5(double *xi, double *phi, double mutm) {
  int rc = cvchip_Q_phi(ctx(), xi, phi, mutm, 1);
  if (rc) die(rc, "Qf5");
}

// gamma5 = diag(1,1,-1,-1) in the reference basis: negate spins 2 and 3
void g5_phi(double *phi) {
  for (long ix = 0; ix < (long)VOLUME; ix++)
    for (int k = 12; k < 24; k++) phi[24 * ix + k] = -phi[24 * ix + k];
}

void spinor_scalar_product_co(complex *w, double *xi, double *phi, int V) {
  if (V != VOLUME) {
    fprintf(stderr, "[cvc_dropin] scalar product over %d sites, lattice has %d\n", V, VOLUME);
    exit(1);
  }
  double r[2];
  int rc = cvchip_scalar_product_co(ctx(), r, xi, phi);
  if (rc) die(rc, "spinor_scalar_product_co");
  w->re = r[0];
  w->im = r[1];
}

void spinor_scalar_product_re(double *r, double *xi, double *phi, int V) {
  if (V != VOLUME) {
    fprintf(stderr, "[cvc_dropin] scalar product over %d sites, lattice has %d\n", V, VOLUME);
    exit(1);
  }
  int rc = cvchip_scalar_product_re(ctx(), r, xi, phi);
  if (rc) die(rc, "spinor_scalar_product_re");
}

static int solve(double *xi, double *phi, int kwork, bool her) {
  // the reference tests r1, r2, s, t, p (g_spinor_field[kwork..kwork+4]), x and
  // phi for NULL and returns -2 (invert_Qtm.c:112-113, 304-305); the device
  // solver keeps its own workspace, the check is kept for the return code
  if (!g_spinor_field || !xi || !phi) return -2;
  for (int k = 0; k < 5; k++)
    if (!g_spinor_field[kwork + k]) return -2;
  double normb = 0.;
  spinor_scalar_product_re(&normb, phi, phi, VOLUME);
  if (g_cart_id == 0) fprintf(stdout, "# norm of r.-h. side: %e\n", normb);
  cvchip_stats st;
  memset(&st, 0, sizeof(st));
  int rc = her ? cvchip_invert_Qtm_her(ctx(), xi, phi, 1, solver_precision, niter_max, CVCHIP_PREC_F64, 0.1,
                                       CVCHIP_SOLVE_LAGGED, &st)
               : cvchip_invert_Qtm(ctx(), xi, phi, 1, solver_precision, niter_max, CVCHIP_PREC_F64, 0.1,
                                   CVCHIP_SOLVE_LAGGED, &st);
  if (rc == CVCHIP_ERR_NOCONV) {
    if (g_cart_id == 0)
      fprintf(stdout, "# No convergence in BiCGStab; after %d steps relative residuum is %e\n", st.iterations,
              st.true_res2_max);
    return -3;
  }
  if (rc < 0) die(rc, her ? "invert_Qtm_her" : "invert_Qtm");
  if (g_cart_id == 0 && rc > 0) {
    fprintf(stdout, "# BiCGStab converged after %d steps\n", rc);
    fprintf(stdout, "# true relative squared residuum is %e\n", st.true_res2_max);
  }
  if (g_cart_id == 0 && rc == 0) fprintf(stdout, "start spinor solves to requested precision\n");
  return rc;
}

int invert_Qtm(double *xi, double *phi, int kwork) { return solve(xi, phi, kwork, false); }

int invert_Qtm_her(double *xi, double *phi, int kwork) { return solve(xi, phi, kwork, true); }

}  // namespace cvc
