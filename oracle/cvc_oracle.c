/*
 * cvc_oracle.c -- CPU restatement of CVC's propagator-and-contraction path.
 *
 * TEST INFRASTRUCTURE ONLY (see cvc_oracle.h).  Built with -ffp-contract=off.
 * The small helpers below restate the reference's algebra macros with the
 * same operand order; the file:line of each macro is given beside it.
 */
#define _GNU_SOURCE
#include "cvc_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* ---------------------------------------------------------------- gammas */

/* Dirac matrices of cvc_utils.c:709-745 (chiral basis, gamma5 = diag(1,1,-1,-1)).
 * Entry encoding: re + i*im with re, im in {-1, 0, 1}. */
typedef struct { int re, im; } ci;

static const ci GAM[4][4][4] = {
  /* gamma_0 */
  {{{0,0},{0,0},{-1,0},{0,0}}, {{0,0},{0,0},{0,0},{-1,0}},
   {{-1,0},{0,0},{0,0},{0,0}}, {{0,0},{-1,0},{0,0},{0,0}}},
  /* gamma_1 */
  {{{0,0},{0,0},{0,0},{0,-1}}, {{0,0},{0,0},{0,-1},{0,0}},
   {{0,0},{0,1},{0,0},{0,0}},  {{0,1},{0,0},{0,0},{0,0}}},
  /* gamma_2 */
  {{{0,0},{0,0},{0,0},{-1,0}}, {{0,0},{0,0},{1,0},{0,0}},
   {{0,0},{1,0},{0,0},{0,0}},  {{-1,0},{0,0},{0,0},{0,0}}},
  /* gamma_3 */
  {{{0,0},{0,0},{0,-1},{0,0}}, {{0,0},{0,0},{0,0},{0,1}},
   {{0,1},{0,0},{0,0},{0,0}},  {{0,0},{0,-1},{0,0},{0,0}}},
};

static void mat_mul4(ci c[4][4], const ci a[4][4], const ci b[4][4]) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      int re = 0, im = 0;
      for (int k = 0; k < 4; k++) {
        re += a[i][k].re * b[k][j].re - a[i][k].im * b[k][j].im;
        im += a[i][k].re * b[k][j].im + a[i][k].im * b[k][j].re;
      }
      c[i][j].re = re; c[i][j].im = im;
    }
}

/* Index order of cvc_utils.c:743-744:
 * 0,1,2,3 = gamma_mu; 4 = id; 5 = gamma5; 6..9 = gamma_mu gamma5;
 * 10..15 = gamma_0 gamma_1, 0 2, 0 3, 1 2, 1 3, 2 3.
 * Each matrix has one non-zero (a unit in {+-1, +-i}) per row; it becomes a
 * signed permutation of the 24 reals (s[k] = t[perm[k]] * sign[k]). */
void orc_gamma_tables(int perm[16][24], int sign[16][24]) {
  ci g[16][4][4];
  memset(g, 0, sizeof(g));
  for (int m = 0; m < 4; m++) memcpy(g[m], GAM[m], sizeof(GAM[m]));
  for (int a = 0; a < 4; a++) g[4][a][a].re = 1;
  for (int a = 0; a < 4; a++) g[5][a][a].re = (a < 2) ? 1 : -1;
  for (int m = 0; m < 4; m++) mat_mul4(g[6 + m], g[m], g[5]);
  static const int pairs[6][2] = {{0,1},{0,2},{0,3},{1,2},{1,3},{2,3}};
  for (int p = 0; p < 6; p++) mat_mul4(g[10 + p], g[pairs[p][0]], g[pairs[p][1]]);

  for (int n = 0; n < 16; n++)
    for (int a = 0; a < 4; a++) {
      int b = -1;
      for (int k = 0; k < 4; k++) if (g[n][a][k].re || g[n][a][k].im) b = k;
      const ci u = g[n][a][b];
      for (int c = 0; c < 3; c++) {
        int kr = 6 * a + 2 * c, kb = 6 * b + 2 * c;
        if (u.im == 0) {            /* +-1: re <- +-re, im <- +-im */
          perm[n][kr] = kb;     sign[n][kr] = u.re;
          perm[n][kr + 1] = kb + 1; sign[n][kr + 1] = u.re;
        } else {                    /* +-i: re <- -+im, im <- +-re */
          perm[n][kr] = kb + 1; sign[n][kr] = -u.im;
          perm[n][kr + 1] = kb;   sign[n][kr + 1] = u.im;
        }
      }
    }
}

/* -------------------------------------------------------------- geometry */

static long lex(const orc_lat *L, int t, int x, int y, int z) {
  /* get_index, cvc_geometry.c:17-26 (serial wrap) */
  t = (t + L->T) % L->T;  x = (x + L->LX) % L->LX;
  y = (y + L->LY) % L->LY; z = (z + L->LZ) % L->LZ;
  return ((long)(t * L->LX + x) * L->LY + y) * L->LZ + z;
}

int orc_lat_init(orc_lat *L, int T, int LX, int LY, int LZ,
                 double kappa, double mu, const double bc[4]) {
  memset(L, 0, sizeof(*L));
  L->nthreads = 1;
  L->T = T; L->LX = LX; L->LY = LY; L->LZ = LZ;
  L->V = (long)T * LX * LY * LZ;
  L->kappa = kappa; L->mu = mu;
  const int Ls[4] = {T, LX, LY, LZ};
  for (int m = 0; m < 4; m++) {
    L->bc[m] = bc[m];
    /* cvc_geometry.c:452-459 */
    L->phase[m][0] = cos(bc[m] * ORC_PI / (double)Ls[m]);
    L->phase[m][1] = sin(bc[m] * ORC_PI / (double)Ls[m]);
  }
  orc_gamma_tables(L->perm, L->sign);
  L->iup = (int *)malloc(sizeof(int) * 4 * L->V);
  L->idn = (int *)malloc(sizeof(int) * 4 * L->V);
  if (!L->iup || !L->idn) return -1;
  /* geometry, cvc_geometry.c:243-251 */
  for (int t = 0; t < T; t++) for (int x = 0; x < LX; x++)
  for (int y = 0; y < LY; y++) for (int z = 0; z < LZ; z++) {
    long ix = lex(L, t, x, y, z);
    L->iup[4*ix+0] = (int)lex(L, t+1, x, y, z);
    L->iup[4*ix+1] = (int)lex(L, t, x+1, y, z);
    L->iup[4*ix+2] = (int)lex(L, t, x, y+1, z);
    L->iup[4*ix+3] = (int)lex(L, t, x, y, z+1);
    L->idn[4*ix+0] = (int)lex(L, t-1, x, y, z);
    L->idn[4*ix+1] = (int)lex(L, t, x-1, y, z);
    L->idn[4*ix+2] = (int)lex(L, t, x, y-1, z);
    L->idn[4*ix+3] = (int)lex(L, t, x, y, z-1);
  }
  return 0;
}

void orc_lat_free(orc_lat *L) {
  free(L->iup); free(L->idn);
  L->iup = L->idn = NULL;
}

/* ------------------------------------------------------- algebra helpers */

/* _cm_eq_cm_ti_co, cvc_linalg.h:271-289: A = B * c */
static void cm_ti_co(double *A, const double *B, const double *c) {
  for (int k = 0; k < 9; k++) {
    A[2*k]   = B[2*k]   * c[0] - B[2*k+1] * c[1];
    A[2*k+1] = B[2*k+1] * c[0] + B[2*k]   * c[1];
  }
}

/* _fv_eq_cm_ti_fv, cvc_linalg.h:1111-1219: s = A t per spin */
static void cm_ti_fv(double *s, const double *A, const double *t) {
  for (int sp = 0; sp < 4; sp++) {
    const double *u = t + 6 * sp;
    double *o = s + 6 * sp;
    for (int i = 0; i < 3; i++) {
      const double *a = A + 6 * i;
      o[2*i] = a[0]*u[0] - a[1]*u[1] + a[2]*u[2] - a[3]*u[3] + a[4]*u[4] - a[5]*u[5];
      o[2*i+1] = a[0]*u[1] + a[1]*u[0] + a[2]*u[3] + a[3]*u[2] + a[4]*u[5] + a[5]*u[4];
    }
  }
}

/* _fv_eq_cm_dag_ti_fv, cvc_linalg.h:1227-1335: s = A^dagger t per spin */
static void cm_dag_ti_fv(double *s, const double *A, const double *t) {
  for (int sp = 0; sp < 4; sp++) {
    const double *u = t + 6 * sp;
    double *o = s + 6 * sp;
    for (int i = 0; i < 3; i++) {
      const double *a0 = A + 2 * i, *a1 = A + 6 + 2 * i, *a2 = A + 12 + 2 * i;
      o[2*i] = a0[0]*u[0] + a0[1]*u[1] + a1[0]*u[2] + a1[1]*u[3] + a2[0]*u[4] + a2[1]*u[5];
      o[2*i+1] = a0[0]*u[1] - a0[1]*u[0] + a1[0]*u[3] - a1[1]*u[2] + a2[0]*u[5] - a2[1]*u[4];
    }
  }
}

/* _fv_eq_gamma_ti_fv, cvc_linalg.h:1344-1368 */
static void gamma_ti_fv(const orc_lat *L, double *s, int g, const double *t) {
  for (int k = 0; k < 24; k++) s[k] = t[L->perm[g][k]] * L->sign[g][k];
}

/* _co_eq_fv_dag_ti_fv, cvc_linalg.h:669-695: c = s^dagger t */
static void fv_dag_ti_fv(double c[2], const double *s, const double *t) {
  double re = 0., im = 0.;
  re = s[0]*t[0] + s[1]*t[1];
  im = s[0]*t[1] - s[1]*t[0];
  for (int k = 2; k < 24; k += 2) {
    re = re + s[k]*t[k] + s[k+1]*t[k+1];
    im = im + s[k]*t[k+1] - s[k+1]*t[k];
  }
  c[0] = re; c[1] = im;
}

/* ------------------------------------------------------------ operators */

/* Body of Q_phi_tbc / Q_phi / Qf5, Q_phi.c:47-175: directions in the order
 * -t,+t,-x,+x,-y,+y,-z,+z, then *(-1/2), then +phi/(2 kappa), then
 * +i mutm gamma5 phi.  g5out applies the trailing gamma5 of Qf5 (Q_phi.c:927). */
static void hop_site(const orc_lat *L, const double *gauge, double *out,
                     const double *phi, long ix, double mutm, int g5out) {
  double xi[24], s1[24], s2[24], U[18];
  const double k2 = 0.5 / L->kappa;
  memset(xi, 0, sizeof(xi));
  for (int mu = 0; mu < 4; mu++) {
    /* negative direction: (1 + gamma_mu) U_mu(x-mu)^dagger phi(x-mu) */
    const double *pn = phi + 24L * L->idn[4*ix+mu];
    gamma_ti_fv(L, s1, mu, pn);
    for (int k = 0; k < 24; k++) s1[k] += pn[k];
    cm_ti_co(U, gauge + 18L * (4L * L->idn[4*ix+mu] + mu), L->phase[mu]);
    cm_dag_ti_fv(s2, U, s1);
    for (int k = 0; k < 24; k++) xi[k] += s2[k];
    /* positive direction: (1 - gamma_mu) U_mu(x) phi(x+mu) */
    const double *pp = phi + 24L * L->iup[4*ix+mu];
    gamma_ti_fv(L, s1, mu, pp);
    for (int k = 0; k < 24; k++) s1[k] = -s1[k];
    for (int k = 0; k < 24; k++) s1[k] += pp[k];
    cm_ti_co(U, gauge + 18L * (4L * ix + mu), L->phase[mu]);
    cm_ti_fv(s2, U, s1);
    for (int k = 0; k < 24; k++) xi[k] += s2[k];
  }
  for (int k = 0; k < 24; k++) xi[k] *= -0.5;
  const double *p0 = phi + 24L * ix;
  for (int k = 0; k < 24; k++) s1[k] = p0[k] * k2;
  for (int k = 0; k < 24; k++) xi[k] += s1[k];
  gamma_ti_fv(L, s1, 5, p0);
  for (int k = 0; k < 24; k += 2) {      /* _fv_eq_fv_ti_im, cvc_linalg.h:1023 */
    s2[k] = -s1[k+1] * mutm;
    s2[k+1] = s1[k] * mutm;
  }
  for (int k = 0; k < 24; k++) xi[k] += s2[k];
  if (g5out) gamma_ti_fv(L, out, 5, xi);
  else memcpy(out, xi, sizeof(xi));
}

void orc_set_threads(orc_lat *L, int n) { L->nthreads = n > 0 ? n : 1; }

/* the site loop of Q_phi_tbc; with threads as the reference's OpenMP
 * version (Q_phi.c:197-206), cyclic over sites */
static void hop_all(const orc_lat *L, const double *gauge, double *xi, const double *phi, double mutm, int g5) {
  const int nt = L->nthreads > 1 ? L->nthreads : 1;
  if (nt == 1) {
    for (long ix = 0; ix < L->V; ix++) hop_site(L, gauge, xi + 24*ix, phi, ix, mutm, g5);
    return;
  }
#pragma omp parallel num_threads(nt)
  {
    const long k = omp_get_thread_num(), n = omp_get_num_threads();
    for (long ix = k; ix < L->V; ix += n) hop_site(L, gauge, xi + 24*ix, phi, ix, mutm, g5);
  }
}

void orc_Q_phi(const orc_lat *L, const double *gauge, double *xi,
               const double *phi, double mutm) {
  hop_all(L, gauge, xi, phi, mutm, 0);
}

void orc_Qf5(const orc_lat *L, const double *gauge, double *xi,
             const double *phi, double mutm) {
  hop_all(L, gauge, xi, phi, mutm, 1);
}

void orc_g5_phi(const orc_lat *L, double *phi) {
  double s[24];
  for (long ix = 0; ix < L->V; ix++) {
    gamma_ti_fv(L, s, 5, phi + 24*ix);
    memcpy(phi + 24*ix, s, sizeof(s));
  }
}

/* ---------------------------------------------------------- reductions */

void orc_dot_co(const orc_lat *L, double w[2], const double *xi, const double *phi) {
  double re = 0., im = 0., p[2];
  for (long ix = 0; ix < L->V; ix++) {
    fv_dag_ti_fv(p, xi + 24*ix, phi + 24*ix);
    re += p[0]; im += p[1];
  }
  w[0] = re; w[1] = im;
}

double orc_dot_re(const orc_lat *L, const double *xi, const double *phi) {
  double w = 0., p[2];
  for (long ix = 0; ix < L->V; ix++) {
    fv_dag_ti_fv(p, xi + 24*ix, phi + 24*ix);
    w += p[0];
  }
  return w;
}

/* ------------------------------------------------------------ BiCGStab */

/* _co_eq_co_ti_co_inv, cvc_complex.h:25-28: c = a / b */
static void co_div(double c[2], const double a[2], const double b[2]) {
  double d = b[0]*b[0] + b[1]*b[1];
  double re = (a[0]*b[0] + a[1]*b[1]) / d;
  double im = (a[1]*b[0] - a[0]*b[1]) / d;
  c[0] = re; c[1] = im;
}

/* _fv_eq_fv_ti_co, cvc_linalg.h:1053: s = t * c */
static void fv_ti_co(double *s, const double *t, const double c[2]) {
  for (int k = 0; k < 24; k += 2) {
    double re = t[k] * c[0] - t[k+1] * c[1];
    double im = t[k+1] * c[0] + t[k] * c[1];
    s[k] = re; s[k+1] = im;
  }
}

typedef void (*orc_op)(const orc_lat *, const double *, double *, const double *, void *);

/* The BiCGStab loop shared by invert_Qtm (op = Q_phi_tbc) and
 * invert_Qtm_her (op = Qf5(mu) Qf5(-mu)), invert_Qtm.c:126-237, 317-431. */
static int bicgstab_core(const orc_lat *L, const double *gauge, orc_op op, void *opctx,
                         double *x, const double *rhs, double normb, double sp,
                         int niter_max, int lagged, double *work, double *norm_out) {
  const long n = 24 * L->V;
  double *r1 = work, *r2 = work + n, *s = work + 2*n, *t = work + 3*n;
  double *p = work + 4*n, *p2 = work + 5*n;
  double s1[24], s2[24];
  double alpha[2] = {0, 0}, beta[2] = {0, 0}, omega[2] = {0, 0};
  double w[2], w2[2], w3[2], r0rn[2], u, norm;
  int niter;

  /* p = rhs - op x */
  op(L, gauge, p, x, opctx);
  for (long i = 0; i < n; i++) p[i] = rhs[i] - p[i];
  norm = orc_dot_re(L, p, p);
  *norm_out = norm;
  if (norm <= sp * normb) return 0;
  memcpy(r1, p, sizeof(double) * n);
  memcpy(r2, p, sizeof(double) * n);
  op(L, gauge, p2, p, opctx);

  for (niter = 0; niter <= niter_max; niter++) {
    orc_dot_co(L, r0rn, r2, r1);
    orc_dot_co(L, w, r2, p2);
    co_div(alpha, r0rn, w);
    for (long ix = 0; ix < L->V; ix++) {
      fv_ti_co(s1, p2 + 24*ix, alpha);
      for (int k = 0; k < 24; k++) s[24*ix+k] = r1[24*ix+k] - s1[k];
    }
    op(L, gauge, t, s, opctx);
    orc_dot_co(L, w, t, s);
    u = orc_dot_re(L, t, t);
    omega[0] = w[0] * (1. / u); omega[1] = w[1] * (1. / u);
    for (long ix = 0; ix < L->V; ix++) {
      fv_ti_co(s1, t + 24*ix, omega);
      for (int k = 0; k < 24; k++) r1[24*ix+k] = s[24*ix+k] - s1[k];
    }
    norm = orc_dot_re(L, r1, r1);
    if (norm <= sp * normb && lagged) break;           /* invert_Qtm.c:201 */
    for (long ix = 0; ix < L->V; ix++) {
      fv_ti_co(s1, s + 24*ix, omega);
      for (int k = 0; k < 24; k++) x[24*ix+k] += s1[k];
      fv_ti_co(s1, p + 24*ix, alpha);
      for (int k = 0; k < 24; k++) x[24*ix+k] += s1[k];
    }
    if (norm <= sp * normb) break;                     /* non-lagged variant */
    orc_dot_co(L, w, r2, r1);
    co_div(w2, w, r0rn);
    co_div(w3, alpha, omega);
    beta[0] = w2[0]*w3[0] - w2[1]*w3[1];
    beta[1] = w2[0]*w3[1] + w2[1]*w3[0];
    r0rn[0] = w[0]; r0rn[1] = w[1];
    for (long ix = 0; ix < L->V; ix++) {
      fv_ti_co(s1, p2 + 24*ix, omega);
      for (int k = 0; k < 24; k++) s1[k] = p[24*ix+k] - s1[k];
      fv_ti_co(s2, s1, beta);
      for (int k = 0; k < 24; k++) p[24*ix+k] = r1[24*ix+k] + s2[k];
    }
    op(L, gauge, p2, p, opctx);
  }
  *norm_out = norm;
  if (norm <= sp * normb && niter <= niter_max) return niter;
  return -3;
}

static void op_Q(const orc_lat *L, const double *g, double *o, const double *i, void *c) {
  (void)c;
  orc_Q_phi(L, g, o, i, L->mu);
}

static void op_her(const orc_lat *L, const double *g, double *o, const double *i, void *c) {
  double *aux = (double *)c;                 /* invert_Qtm.c:325-327 */
  orc_Qf5(L, g, aux, i, -L->mu);
  orc_Qf5(L, g, o, aux, L->mu);
}

int orc_invert_Qtm(const orc_lat *L, const double *gauge, double *x,
                   const double *b, double sp, int niter_max, int lagged,
                   double *true_res2) {
  if (!x || !b || !gauge) return -2;
  const long n = 24 * L->V;
  double *work = (double *)malloc(sizeof(double) * 6 * n);
  if (!work) return -2;
  double normb = orc_dot_re(L, b, b), norm;
  int it = bicgstab_core(L, gauge, op_Q, NULL, x, b, normb, sp, niter_max, lagged, work, &norm);
  if (true_res2) {                                   /* invert_Qtm.c:257-267 */
    double *p = work;
    orc_Q_phi(L, gauge, p, x, L->mu);
    for (long i = 0; i < n; i++) p[i] -= b[i];
    *true_res2 = orc_dot_re(L, p, p) / normb;
  }
  free(work);
  return it;
}

int orc_invert_Qtm_her(const orc_lat *L, const double *gauge, double *x,
                       const double *b, double sp, int niter_max,
                       double *true_res2) {
  if (!x || !b || !gauge) return -2;
  const long n = 24 * L->V;
  double *work = (double *)malloc(sizeof(double) * 8 * n);
  if (!work) return -2;
  double *aux = work + 6*n, *g5b = work + 7*n;
  double normb = orc_dot_re(L, b, b), norm;
  memcpy(g5b, b, sizeof(double) * n);              /* invert_Qtm.c:328-332 */
  orc_g5_phi(L, g5b);
  int it = bicgstab_core(L, gauge, op_her, aux, x, g5b, normb, sp, niter_max, 1, work, &norm);
  /* x <- gamma5 D(-mu) y, invert_Qtm.c:441-446 (done whether or not converged) */
  orc_Qf5(L, gauge, aux, x, -L->mu);
  memcpy(x, aux, sizeof(double) * n);
  if (true_res2) {
    double *p = work;
    orc_Q_phi(L, gauge, p, x, L->mu);
    for (long i = 0; i < n; i++) p[i] -= b[i];
    *true_res2 = orc_dot_re(L, p, p) / normb;
  }
  free(work);
  return it;
}

/* ---------------------------------------------------------- contraction */

void orc_cvc_contract(const orc_lat *L, const double *gauge,
                      const double *const *props, const int src[4],
                      double *conn, double contact[8]) {
  const long V = L->V;
  double Usrc[4][18], U[18], s1[24], s2[24], w[2], w1[2];
  int gperm[5][4], isimag[4];
  double gsign[5][4];
  const long sloc = lex(L, src[0], src[1], src[2], src[3]);

  memset(conn, 0, sizeof(double) * 32 * V);
  memset(contact, 0, sizeof(double) * 8);
  /* Usource[nu] = U_nu(y) * phase, cvc_exact2_xspace.c:300-303 */
  for (int nu = 0; nu < 4; nu++) cm_ti_co(Usrc[nu], gauge + 18L * (4L * sloc + nu), L->phase[nu]);
  /* gperm tables, cvc_exact2_xspace.c:336-367 (index 4 = gamma5) */
  for (int r = 0; r < 4; r++) {
    gperm[4][r] = L->perm[5][6*r] / 6;
    gsign[4][r] = L->sign[5][6*r];
  }
  for (int nu = 0; nu < 4; nu++) {
    for (int r = 0; r < 4; r++) {
      gperm[nu][r] = L->perm[6+nu][6*r] / 6;
      gsign[nu][r] = L->sign[6+nu][6*r];
    }
    isimag[nu] = L->perm[6+nu][0] % 2;
  }

  /* part 1, cvc_exact2_xspace.c:577-716: up from y, down from y+nu, Usource */
  /* part 2, cvc_exact2_xspace.c:723-847: up from y+nu, down from y, conj */
  for (int part = 0; part < 2; part++) {
    for (int nu = 0; nu < 4; nu++) {
      for (int ir = 0; ir < 4; ir++) {
        for (int sub = 0; sub < 2; sub++) {  /* sub 0: gamma_nu gamma5, sub 1: gamma5 */
          const int gi = sub ? 4 : nu;
          for (int ia = 0; ia < 3; ia++) {
            const double *phi = props[(part ? nu : 4) * 12 + 3*ir + ia];
            for (int ib = 0; ib < 3; ib++) {
              const double *chi = props[60 + (part ? 4 : nu) * 12 + 3*gperm[gi][ir] + ib];
              /* source link element: Usource[nu][ia][ib] (part 1), conj Usource[nu][ib][ia] (part 2) */
              const double *uc = part ? Usrc[nu] + 2*(3*ib+ia) : Usrc[nu] + 2*(3*ia+ib);
              for (int mu = 0; mu < 4; mu++) {
                const int imunu = 4*mu + nu;
                for (int dir = 0; dir < 2; dir++) {
                  for (long ix = 0; ix < V; ix++) {
                    cm_ti_co(U, gauge + 18L * (4L * ix + mu), L->phase[mu]);
                    if (dir == 0) {
                      /* gamma5 (gamma_mu - 1) U phi(x+mu), dotted with chi(x) */
                      cm_ti_fv(s1, U, phi + 24L * L->iup[4*ix+mu]);
                      gamma_ti_fv(L, s2, mu, s1);
                      for (int k = 0; k < 24; k++) s2[k] -= s1[k];
                      gamma_ti_fv(L, s1, 5, s2);
                      fv_dag_ti_fv(w, chi + 24*ix, s1);
                    } else {
                      /* gamma5 (gamma_mu + 1) U^dagger phi(x), dotted with chi(x+mu) */
                      cm_dag_ti_fv(s1, U, phi + 24*ix);
                      gamma_ti_fv(L, s2, mu, s1);
                      for (int k = 0; k < 24; k++) s2[k] += s1[k];
                      gamma_ti_fv(L, s1, 5, s2);
                      fv_dag_ti_fv(w, chi + 24L * L->iup[4*ix+mu], s1);
                    }
                    if (part == 0) {               /* _co_eq_co_ti_co */
                      w1[0] = w[0]*uc[0] - w[1]*uc[1];
                      w1[1] = w[0]*uc[1] + w[1]*uc[0];
                    } else {                       /* _co_eq_co_ti_co_conj */
                      w1[0] = w[0]*uc[0] + w[1]*uc[1];
                      w1[1] = -w[0]*uc[1] + w[1]*uc[0];
                    }
                    double *c = conn + 2*(16*ix + imunu);
                    if (sub == 0) {
                      const double sg = gsign[nu][ir];
                      if (!isimag[nu]) { c[0] += sg * w1[0]; c[1] += sg * w1[1]; }
                      else             { c[0] += sg * w1[1]; c[1] -= sg * w1[0]; }
                    } else {
                      /* -gamma5 term: subtracted in part 1, added in part 2 */
                      const double sg = gsign[4][ir];
                      if (part == 0) { c[0] -= sg * w1[0]; c[1] -= sg * w1[1]; }
                      else           { c[0] += sg * w1[0]; c[1] += sg * w1[1]; }
                    }
                  }
                }
              }
            }
            if (sub == 1) {
              /* contact term, cvc_exact2_xspace.c:697-704 and 837-844 */
              if (part == 0) {
                cm_ti_fv(s1, Usrc[nu], phi + 24L * L->iup[4*sloc+nu]);
                gamma_ti_fv(L, s2, nu, s1);
                for (int k = 0; k < 24; k++) s2[k] -= s1[k];
                contact[2*nu]   += -0.5 * s2[2*(3*ir+ia)];
                contact[2*nu+1] += -0.5 * s2[2*(3*ir+ia)+1];
              } else {
                cm_dag_ti_fv(s1, Usrc[nu], phi + 24*sloc);
                gamma_ti_fv(L, s2, nu, s1);
                for (int k = 0; k < 24; k++) s2[k] += s1[k];
                contact[2*nu]   += 0.5 * s2[2*(3*ir+ia)];
                contact[2*nu+1] += 0.5 * s2[2*(3*ir+ia)+1];
              }
            }
          }
        }
      }
    }
  }
  /* cvc_exact2_xspace.c:858-861 */
  for (long i = 0; i < 32 * V; i++) conn[i] *= -0.25;
}

void orc_wi(const orc_lat *L, const double *conn, double *wi) {
  for (long ix = 0; ix < L->V; ix++)
    for (int nu = 0; nu < 4; nu++)
      for (int c = 0; c < 2; c++) {
        double v = conn[2*(16*ix + 0 + nu) + c] + conn[2*(16*ix + 4 + nu) + c]
                 + conn[2*(16*ix + 8 + nu) + c] + conn[2*(16*ix + 12 + nu) + c]
                 - conn[2*(16L*L->idn[4*ix+0] + 0 + nu) + c]
                 - conn[2*(16L*L->idn[4*ix+1] + 4 + nu) + c]
                 - conn[2*(16L*L->idn[4*ix+2] + 8 + nu) + c]
                 - conn[2*(16L*L->idn[4*ix+3] + 12 + nu) + c];
        wi[2*(4*ix+nu)+c] = v;
      }
}

/* _cm_eq_cm_ti_cm / _cm_eq_cm_ti_cm_dag / trace; cvc_utils.c:393-405 */
double orc_plaquette(const orc_lat *L, const double *gauge) {
  double pl = 0.;
  for (long ix = 0; ix < L->V; ix++)
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        const double *a = gauge + 72*ix + 18*mu, *b = gauge + 72L*L->iup[4*ix+mu] + 18*nu;
        const double *c = gauge + 72*ix + 18*nu, *d = gauge + 72L*L->iup[4*ix+nu] + 18*mu;
        double s[18], t[18], tr = 0.;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
          double sr = 0, si = 0, tr_ = 0, ti = 0;
          for (int k = 0; k < 3; k++) {
            sr += a[2*(3*i+k)]*b[2*(3*k+j)] - a[2*(3*i+k)+1]*b[2*(3*k+j)+1];
            si += a[2*(3*i+k)]*b[2*(3*k+j)+1] + a[2*(3*i+k)+1]*b[2*(3*k+j)];
            tr_ += c[2*(3*i+k)]*d[2*(3*k+j)] - c[2*(3*i+k)+1]*d[2*(3*k+j)+1];
            ti += c[2*(3*i+k)]*d[2*(3*k+j)+1] + c[2*(3*i+k)+1]*d[2*(3*k+j)];
          }
          s[2*(3*i+j)] = sr; s[2*(3*i+j)+1] = si; t[2*(3*i+j)] = tr_; t[2*(3*i+j)+1] = ti;
        }
        /* Re tr (s t^dagger) = sum_ij Re s_ij conj(t_ij) */
        for (int k = 0; k < 9; k++) tr += s[2*k]*t[2*k] + s[2*k+1]*t[2*k+1];
        pl += tr;
      }
  return pl / ((double)L->V * 18.);
}

/* ------------------------------------------------- meson two-point function */

void orc_contract_twopoint(const orc_lat *L, double *contr, int idsource, int idsink,
                           const double *const *chi, const double *const *phi, int n_c,
                           const int *snk_mom, const int src[4]) {
  /* cvc_utils.c:3084-3096: source column map, imaginary flag and signs */
  int psource[4];
  double ssource[4];
  for (int m = 0; m < 4; m++) {
    psource[m] = L->perm[idsource][6 * m] / 6;
    ssource[m] = L->sign[idsource][6 * m] * L->sign[5][L->perm[idsource][6 * m]];
  }
  const int isimag = L->perm[idsource][0] % 2;
  double px = 0., py = 0., pz = 0.;
  if (snk_mom) {
    px = 2. * ORC_PI * snk_mom[0] / L->LX;
    py = 2. * ORC_PI * snk_mom[1] / L->LY;
    pz = 2. * ORC_PI * snk_mom[2] / L->LZ;
  }
  double spinor1[24], spinor2[24];
  for (int x0 = 0; x0 < L->T; x0++) {
    for (int x1 = 0; x1 < L->LX; x1++)
    for (int x2 = 0; x2 < L->LY; x2++)
    for (int x3 = 0; x3 < L->LZ; x3++) {
      const long iix = lex(L, x0, x1, x2, x3);
      const double phase = (x1 - src[1]) * px + (x2 - src[2]) * py + (x3 - src[3]) * pz;
      const double cphase = cos(phase), sphase = sin(phase);
      double tmp[2] = {0., 0.};
      for (int mu = 0; mu < 4; mu++)
        for (int c = 0; c < n_c; c++) {
          const double *ph = phi[mu * n_c + c] + 24 * iix;
          const double *ch = chi[psource[mu] * n_c + c] + 24 * iix;
          /* spinor1 = Gamma_sink phi, spinor2 = gamma5 spinor1 (cvc_linalg.h:1344-1368) */
          for (int k = 0; k < 24; k++) spinor1[k] = ph[L->perm[idsink][k]] * L->sign[idsink][k];
          for (int k = 0; k < 24; k++) spinor2[k] = spinor1[L->perm[5][k]] * L->sign[5][k];
          /* w = chi^dagger spinor2 (cvc_linalg.h:669) */
          double wr = 0., wi = 0.;
          for (int k = 0; k < 12; k++) {
            wr += ch[2 * k] * spinor2[2 * k] + ch[2 * k + 1] * spinor2[2 * k + 1];
            wi += ch[2 * k] * spinor2[2 * k + 1] - ch[2 * k + 1] * spinor2[2 * k];
          }
          tmp[0] += ssource[mu] * wr;
          tmp[1] += ssource[mu] * wi;
        }
      const double re = tmp[0] * cphase - tmp[1] * sphase;
      const double im = tmp[0] * sphase + tmp[1] * cphase;
      if (!isimag) {
        contr[2 * x0] += re;
        contr[2 * x0 + 1] += im;
      } else {
        contr[2 * x0] += im;
        contr[2 * x0 + 1] += -re;
      }
    }
  }
}

/* Sequential source of p2gg_xspace.c:517-548 (serial build): out = 0 except on
 * the time slice tseq, where out(x) = e^{i phase(x)} Gamma_gid prop(x) with
 * phase = sgn 2 pi (x p_x / LX + y p_y / LY + z p_z / LZ); the reference's sgn
 * is 1 - 2 g_propagator_position.  Loop order and operations as there:
 * _fv_eq_gamma_ti_fv, then _fv_eq_fv_ti_co. */
void orc_seq_source(const orc_lat *L, const double *prop, double *out, int tseq, int gid,
                    const int mom[3], double sgn) {
  double spinor1[24], w[2];
  memset(out, 0, sizeof(double) * 24 * L->V);
  const int x0 = tseq;
  for (int x1 = 0; x1 < L->LX; x1++)
    for (int x2 = 0; x2 < L->LY; x2++)
      for (int x3 = 0; x3 < L->LZ; x3++) {
        const long ix = lex(L, x0, x1, x2, x3);
        const double phase = sgn * 2. * ORC_PI * (x1 * mom[0] / (double)L->LX + x2 * mom[1] / (double)L->LY +
                                                 x3 * mom[2] / (double)L->LZ);
        w[0] = cos(phase);
        w[1] = sin(phase);
        gamma_ti_fv(L, spinor1, gid, prop + 24 * ix);
        fv_ti_co(out + 24 * ix, spinor1, w);
      }
}

