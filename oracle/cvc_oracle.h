/*
 * cvc_oracle.h -- CPU restatement of CVC's propagator-and-contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is the checker for the HIP path: the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark are its only users.
 *
 * Every function follows a reference function loop for loop and in the same
 * floating-point accumulation order (built with -ffp-contract=off so that no
 * fused multiply-add changes the rounding).  The reference file:line each one
 * restates is given beside it.
 *
 * Conventions (reference global.h:55-61, cvc_geometry.c:17-26):
 *   site index   ix = ((t*LX + x)*LY + y)*LZ + z   (z fastest), periodic wrap
 *   spinor       24 doubles per site, 2*(3*spin + colour) + {0 re, 1 im}
 *   gauge        72 doubles per site, 18*mu + 2*(3*i + j) + {re, im}, mu = t,x,y,z
 *   contraction  conn[2*(16*ix + 4*mu + nu) + {re, im}]
 *
 * Parity pins (which test fixes which function) are listed in the header of
 * tests/test_oracle.py; the reference sources could not be compiled here (its
 * global.h unconditionally includes FFTW2's <fftw.h>, which the image lacks),
 * so each function is pinned by closed forms, invariants and brute force.
 */
#ifndef CVC_ORACLE_H
#define CVC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int T, LX, LY, LZ;
  long V;
  double kappa, mu;
  double bc[4];          /* BCangle[mu]                                    */
  double phase[4][2];    /* co_phase_up[mu] = exp(i pi BCangle[mu] / L_mu) */
  int perm[16][24];      /* gamma_permutation                              */
  int sign[16][24];      /* gamma_sign                                     */
  int *iup;              /* [V][4] forward neighbours                      */
  int *idn;              /* [V][4] backward neighbours                     */
  int nthreads;          /* OpenMP threads of the operator loop (default 1) */
} orc_lat;

int  orc_lat_init(orc_lat *L, int T, int LX, int LY, int LZ,
                  double kappa, double mu, const double bc[4]);
void orc_lat_free(orc_lat *L);
/* threads of orc_Q_phi / orc_Qf5 (the reference's OpenMP Q_phi_tbc,
 * Q_phi.c:190-206: thread k takes the sites k, k + n, k + 2n, ...); the
 * per-site arithmetic, and so every result, does not depend on it. */
void orc_set_threads(orc_lat *L, int n);
void orc_gamma_tables(int perm[16][24], int sign[16][24]);

/* xi = D(mutm) phi; Q_phi.c:1092-1242 (Q_phi_tbc = this with mutm = g_mu). */
void orc_Q_phi(const orc_lat *L, const double *gauge, double *xi,
               const double *phi, double mutm);
/* xi = gamma5 D(mutm) phi; Q_phi.c:789-934. */
void orc_Qf5(const orc_lat *L, const double *gauge, double *xi,
             const double *phi, double mutm);
/* phi <- gamma5 phi; Q_phi.c:350-361. */
void orc_g5_phi(const orc_lat *L, double *phi);

/* invert_Qtm.c:28-58 and 60-84 (sequential site order). */
void   orc_dot_co(const orc_lat *L, double w[2], const double *xi, const double *phi);
double orc_dot_re(const orc_lat *L, const double *xi, const double *phi);

/* BiCGStab, invert_Qtm.c:86-270.  x is in/out (initial guess in).
 * sp is the squared relative tolerance (solver_precision).  lagged=1 keeps the
 * reference's return of the iterate from before the last update; lagged=0
 * applies that update.  Returns the iteration count, -2 for NULL buffers or
 * -3 without convergence.  *true_res2 receives |D x - b|^2/|b|^2. */
int orc_invert_Qtm(const orc_lat *L, const double *gauge, double *x,
                   const double *b, double sp, int niter_max, int lagged,
                   double *true_res2);
/* BiCGStab on gamma5 D(mu) gamma5 D(-mu), invert_Qtm.c:276-478. */
int orc_invert_Qtm_her(const orc_lat *L, const double *gauge, double *x,
                       const double *b, double sp, int niter_max,
                       double *true_res2);

/* Conserved-vector-current contraction, cvc_exact2_xspace.c:293-304,336-383,
 * 577-861.  props[12*s + 3*spin + colour] (s = 0..3: source y+s^, s = 4: y)
 * are the up propagators, props[60 + ...] the down ones.  src = (t,x,y,z) of
 * y.  conn gets 32*V doubles (already multiplied by -1/4), contact 8 doubles. */
void orc_cvc_contract(const orc_lat *L, const double *gauge,
                      const double *const *props, const int src[4],
                      double *conn, double contact[8]);
/* Ward identity in position space, cvc_exact2_xspace.c:987-1006: wi gets
 * 8*V doubles, wi[2*(4*ix+nu)+c] = sum_mu conn_{mu nu}(x) - conn_{mu nu}(x-mu). */
void orc_wi(const orc_lat *L, const double *conn, double *wi);

/* Average plaquette, cvc_utils.c:384-405. */
double orc_plaquette(const orc_lat *L, const double *gauge);

/* Meson two-point function (cvc_utils.c:3075-3140 contract_twopoint; with a
 * sink momentum cvc_utils.c:3143-3235 contract_twopoint_snk_momentum, serial
 * variant): contr[T][2] += sum_x e^{i p.(x - y)} sum_{mu, c} ssource[mu]
 * chi[psource[mu] n_c + c](x)^dagger gamma5 Gamma_sink phi[mu n_c + c](x),
 * times -i for an imaginary source gamma; snk_mom NULL for p = 0, src the
 * source point (its spatial part enters the phase). */
void orc_contract_twopoint(const orc_lat *L, double *contr, int idsource, int idsink,
                           const double *const *chi, const double *const *phi, int n_c,
                           const int *snk_mom, const int src[4]);

/* Sequential source of p2gg_xspace.c:517-548: on time slice tseq,
 * out(x) = e^{i sgn 2 pi (x p_x / LX + y p_y / LY + z p_z / LZ)} Gamma_gid prop(x),
 * zero elsewhere (sgn = 1 - 2 g_propagator_position in the reference). */
void orc_seq_source(const orc_lat *L, const double *prop, double *out, int tseq, int gid,
                    const int mom[3], double sgn);

#ifdef __cplusplus
}
#endif
#endif
