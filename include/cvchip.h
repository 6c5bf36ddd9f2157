/*
 * cvchip.h -- C ABI of the MI355X (gfx950) implementation of CVC's
 * propagator-and-contraction path.
 *
 * Each entry point replaces a reference function (file:line given beside it).
 * Conventions shared by all entry points:
 *   - host arrays use the reference's layouts (global.h:55-61):
 *       spinor  double[V][24]      2*(3*spin+colour)+{re,im}, site ((t*LX+x)*LY+y)*LZ+z
 *       gauge   double[V][4][18]   18*mu + 2*(3*i+j)+{re,im}, mu = t,x,y,z
 *       conn    double[V][16][2]   2*(16*ix + 4*mu + nu)+{re,im}
 *     nrhs fields are passed as nrhs consecutive V*24 arrays;
 *   - the boundary phase exp(i*pi*bc_angle[mu]/L_mu) multiplies every link
 *     (cvc_geometry.c:452-459), resolved inside the kernels; no halo exchange;
 *   - return 0 on success; -2 for NULL buffers and -3 for a solve that did not
 *     converge, as invert_Qtm (invert_Qtm.c:112-113, 250); values <= -100
 *     are argument / HIP errors, described by cvchip_last_error();
 *   - calls are synchronous at the boundary and not thread-safe per context.
 */
#ifndef CVCHIP_H
#define CVCHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVCHIP_OK            0
#define CVCHIP_ERR_NULL     (-2)
#define CVCHIP_ERR_NOCONV   (-3)
#define CVCHIP_ERR_ARG      (-100)
#define CVCHIP_ERR_HIP      (-101)
#define CVCHIP_ERR_OOM      (-102)
#define CVCHIP_ERR_STATE    (-103)

/* arithmetic / storage precision of a field or of a solve */
#define CVCHIP_PREC_F64   0
#define CVCHIP_PREC_F32   1
#define CVCHIP_PREC_F16   2   /* half storage + per-site float norm, FP32 arithmetic */

/* solver flags */
#define CVCHIP_SOLVE_LAGGED   1u  /* return the iterate from before the last update,
                                     as invert_Qtm.c:199-211 does */
#define CVCHIP_SOLVE_VERBOSE  2u  /* print the reference's per-iteration residual lines */
#define CVCHIP_SOLVE_EO       4u  /* even/odd preconditioning: the solver runs on the Schur
                                     complement of the odd sites (half-lattice vectors) and
                                     rebuilds the even sites; same solution and true-residual
                                     check, fewer and cheaper iterations than the reference's
                                     unpreconditioned BiCGStab (needs an even LZ) */

typedef struct cvchip_ctx cvchip_ctx;
typedef struct cvchip_field cvchip_field;

typedef struct {
  int    iterations;        /* BiCGStab iterations (largest over the right-hand sides)     */
  int    converged;         /* number of right-hand sides that met the tolerance           */
  int    reliable_updates;  /* FP64 residual recomputations (mixed precision)              */
  int    dslash_calls;      /* operator applications in all precisions                     */
  double true_res2_max;     /* max over rhs of |D x - b|^2 / |b|^2, in FP64                 */
  double seconds;           /* wall time inside the call                                   */
} cvchip_stats;

/* ------------------------------------------------------------ context */

/* Lattice geometry and parameters: init_geometry (cvc_geometry.c:343-465) plus
 * the g_kappa / g_mu / BCangle globals (global.h:91-206).  dims = T,LX,LY,LZ. */
int cvchip_ctx_create(cvchip_ctx **ctx, const int dims[4], const double bc_angle[4],
                      double kappa, double mu);
int cvchip_ctx_destroy(cvchip_ctx *ctx);
int cvchip_ctx_set_params(cvchip_ctx *ctx, double kappa, double mu);
/* link storage of the FP64 operator: 18 (full) or 12 (two rows, third rebuilt) */
int cvchip_ctx_set_link_recon(cvchip_ctx *ctx, int recon);
/* solver flags OR-ed into every solve of the context, also those inside
 * cvchip_cvc_source and cvchip_cvc_gradient (for instance CVCHIP_SOLVE_EO).
 * No reference counterpart; default 0. */
int cvchip_ctx_set_solver_flags(cvchip_ctx *ctx, unsigned flags);
/* operator tuning.  Bits 0-1, staging of the multi-rhs operator (N >= 8): bit 0
 * keeps a block's links in LDS, bit 1 its own and z-neighbour spinors; bit 2
 * writes outputs of 256 MB and more with streaming (non-temporal) stores, bit 3
 * every output.  Bits 4-7,
 * block-to-site traversal: 0 automatic, 1 time slices cut into 8 XCD chunks,
 * 2 one time slice per XCD walked in y blocks, 3 linear, 4 bricks of bt x bx x
 * yb x LZ sites marched along t per XCD, 5 columns of bx x yb x zb sites
 * (powers of two) swept along t per XCD, 6 t march (multi-rhs operator on the
 * whole lattice: each workgroup walks its columns through consecutive time
 * slices and carries the backward t hop in registers; bits 8-15 the march
 * length (0: 16), bits 16-19 and 20-23 the x and y extents of the patches of
 * columns that one XCD works on together (0: 8)).  Bits 8-15: the y-block height of
 * traversals 2, 4 and 5 (0: the largest of 16, 8, 4, 2 dividing LY; for the
 * automatic bricks the largest of 8, 4, 2 with yb * N <= 96; 4 for traversal
 * 5); bits 16-19 bt of traversal 4 (0: 4) or log2 zb of traversal 5 (0: zb 8)
 * and 20-23 bx of traversals 4 and 5 (0: 4); bits 24-27 extra LDS per block in
 * 8 KiB units (caps the resident blocks).  Default 5: links in LDS,
 * streaming stores, automatic traversal (for N >= 8 bricks of 4 slices x 4
 * x-planes x yb y-rows, yb the largest of 16, 8, 4, 2 dividing LY with
 * yb * N <= 96 K, K = 1, 2, 4 for FP64, FP32, FP16+norm spinors; for the FP32
 * operator with 8 <= N <= 20 the t march over 4 slices in patches of 8x8
 * columns).  The operator output
 * does not depend on the traversal up to rounding (the t march sums the
 * diagonal term first); the fused dot products of the solvers differ at
 * rounding level. */
int cvchip_ctx_set_dslash_policy(cvchip_ctx *ctx, int policy);
const char *cvchip_last_error(void);

/* Upload g_gauge_field (read_lime_gauge_field_doubleprec, io.c:202-347, fills it
 * on the reference side).  Must be called before any operator. */
int cvchip_set_gauge(cvchip_ctx *ctx, const double *gauge);

/* ------------------------------------------------- host-array operators */

/* xi = D(mutm) phi with D = 1/(2 kappa) + i mutm gamma5 - 1/2 sum_mu [...]:
 * Q_phi_tbc (Q_phi.c:38-188) with mutm = mu, Q_phi (Q_phi.c:1092-1242) for any
 * mutm; gamma5_out = 1 gives Qf5 (Q_phi.c:789-934): xi = gamma5 D(mutm) phi. */
int cvchip_Q_phi(cvchip_ctx *ctx, double *xi, const double *phi, double mutm, int gamma5_out);

/* spinor_scalar_product_co / _re (invert_Qtm.c:28-84): sum_x xi(x)^dagger phi(x). */
int cvchip_scalar_product_co(cvchip_ctx *ctx, double w[2], const double *xi, const double *phi);
int cvchip_scalar_product_re(cvchip_ctx *ctx, double *r, const double *xi, const double *phi);

/* invert_Qtm (invert_Qtm.c:86-270): BiCGStab for D(mu) x = b on nrhs right-hand
 * sides solved together.  x is in/out (initial guess in).  sp is the squared
 * relative tolerance solver_precision (|r|^2 <= sp |b|^2).  prec selects the
 * FP64 solver or the mixed solver (FP32 / FP16 inner iteration with FP64
 * reliable updates whenever |r| falls below delta times its value at the last
 * update).  Returns the iteration count (largest over rhs), -2 or -3. */
int cvchip_invert_Qtm(cvchip_ctx *ctx, double *x, const double *b, int nrhs, double sp,
                      int niter_max, int prec, double delta, unsigned flags,
                      cvchip_stats *stats);

/* invert_Qtm_her (invert_Qtm.c:276-478): BiCGStab on gamma5 D(mu) gamma5 D(-mu)
 * y = gamma5 b (= D(-mu)^dagger D(-mu) y, Hermitian positive definite),
 * returning x = gamma5 D(-mu) y, which solves D(mu) x = b.  x is the initial
 * guess for y on entry, as in the reference.  prec / delta / flags as for
 * cvchip_invert_Qtm (the mixed solver runs on the same Hermitian system; the
 * reference returns the lagged iterate here too: CVCHIP_SOLVE_LAGGED). */
int cvchip_invert_Qtm_her(cvchip_ctx *ctx, double *x, const double *b, int nrhs, double sp,
                          int niter_max, int prec, double delta, unsigned flags,
                          cvchip_stats *stats);

/* ---------------------------------------------- device-resident fields */

/* A set of nrhs spinor fields kept in HBM in the library's layout. */
int cvchip_field_create(cvchip_ctx *ctx, int nrhs, int prec, cvchip_field **f);
int cvchip_field_destroy(cvchip_field *f);
int cvchip_field_upload(cvchip_ctx *ctx, cvchip_field *f, const double *host, int rhs0, int n);
int cvchip_field_download(cvchip_ctx *ctx, const cvchip_field *f, double *host, int rhs0, int n);
/* Fill an FP64 set with deterministic uniform(-1,1) values from (seed, index);
 * synthetic operator inputs that never leave HBM.  No reference counterpart. */
int cvchip_field_fill_random(cvchip_ctx *ctx, cvchip_field *f, uint64_t seed);
/* dst = src, converting precision (FP64 / FP32 / FP16+norm); equal nrhs. */
int cvchip_field_copy(cvchip_ctx *ctx, cvchip_field *dst, const cvchip_field *src);
/* f = 0 (every component, and the FP16 scales). */
int cvchip_field_zero(cvchip_ctx *ctx, cvchip_field *f);
/* Point sources of nsrc source points coords[4 k .. 4 k + 3] = (t, x, y, z) into an
 * FP64 set of 12 nsrc columns (nsrc <= 21): column 12 k + 3 spin + colour is the
 * unit vector source[_GSI(site_k) + 2 (3 spin + colour)] = 1 of
 * cvc_exact2_xspace.c:521-524. */
int cvchip_field_point_sources(cvchip_ctx *ctx, cvchip_field *f, int nsrc, const int *coords);
/* dst columns dst_col0 .. dst_col0 + n - 1 = src columns src_col0 .. (FP64 sets of
 * one volume; ranges inside the sets and not overlapping).  No reference
 * counterpart: it selects one source's 12 columns of a batched set. */
int cvchip_field_copy_columns(cvchip_ctx *ctx, cvchip_field *dst, int dst_col0, const cvchip_field *src,
                              int src_col0, int n);
/* out = D(mutm) in (gamma5_out: gamma5 D in) on all rhs of the set; precision
 * of the sets selects the kernel. */
int cvchip_field_dslash(cvchip_ctx *ctx, cvchip_field *out, const cvchip_field *in,
                        double mutm, int gamma5_out);
/* Launches the same operator reps times back to back on the library's stream
 * and returns the average duration per launch from HIP events on that stream. */
int cvchip_field_dslash_timed(cvchip_ctx *ctx, cvchip_field *out, const cvchip_field *in,
                              double mutm, int reps, double *ms_per_launch);
/* Device-resident invert_Qtm: solves D(mu) x = b for all rhs of b into x. */
int cvchip_field_invert(cvchip_ctx *ctx, cvchip_field *x, const cvchip_field *b, double sp,
                        int niter_max, int prec, double delta, unsigned flags,
                        cvchip_stats *stats);

/* Device-resident invert_Qtm_her. */
int cvchip_field_invert_her(cvchip_ctx *ctx, cvchip_field *x, const cvchip_field *b, double sp,
                            int niter_max, int prec, double delta, unsigned flags,
                            cvchip_stats *stats);

/* ---------------------------------------------------------- observables */

/* Conserved-vector-current contraction of cvc_exact2_xspace.c:577-861 for the
 * source point src = (t,x,y,z).  props: 120 host spinor fields ordered as the
 * reference's g_spinor_field (cvc_exact2_xspace.c:401-430): index
 * 12*s + 3*spin + colour for the up flavour, 60 + ... for the down flavour,
 * with s = 0..3 the source y + s^ and s = 4 the source y.  conn receives the
 * normalised (-1/4, :861) field, contact the 4 complex contact terms
 * (contact_term, :697-704, 837-844). */
int cvchip_cvc_contract(cvchip_ctx *ctx, double *conn, double contact[8],
                        const double *const *props, const int src[4]);
/* The same from device-resident propagators: up and dn each hold 60 rhs in the
 * reference's order (12*s + 3*spin + colour). */
int cvchip_field_cvc_contract(cvchip_ctx *ctx, double *conn, double contact[8],
                              const cvchip_field *up, const cvchip_field *dn,
                              const int src[4]);
/* Ward identity of cvc_exact2_xspace.c:987-1006: wi[V][4][2],
 * wi_nu(x) = sum_mu conn_{mu nu}(x) - conn_{mu nu}(x - mu). */
int cvchip_wi(cvchip_ctx *ctx, double *wi, const double *conn);

/* Momentum-space polarisation tensor (no reference counterpart; the reference
 * writes Pi(x) only, cvc_exact2_xspace.c:876-882):
 *   Pi_{mu nu}(q) = sum_x exp(-i q.(x - y)) exp(-i (q_mu - q_nu)/2) conn_{mu nu}(x)
 *                   - delta_{mu nu} contact_nu,
 * q_mu = 2 pi n_mu / L_mu, n_mu in [0, L_mu), extents up to 64.  The half-link
 * phases and the contact term make it transverse exactly,
 * sum_mu 2 sin(q_mu/2) Pi_{mu nu}(q) = 0, which is the momentum-space form of
 * the position-space identity
 * sum_mu conn_{mu nu}(x) - conn_{mu nu}(x-mu) = contact_nu (delta_{x,y} - delta_{x,y+nu}).
 * piq: double[V][16][2] with n in the lexicographic site order. */
int cvchip_pi_q(cvchip_ctx *ctx, double *piq, const double *conn, const double contact[8],
                const int src[4]);

/* The scalar vacuum polarisation on qhat^2 orbits (no reference counterpart:
 * the reference leaves this to analyse_piq.c / make_q2orbits.c, which it does
 * not ship).  From piq of cvchip_pi_q: Pi(qhat^2) = -sum_mu Re Pi_mu_mu(q) /
 * (3 qhat^2), qhat_mu = 2 sin(pi n_mu / L_mu), averaged over the momenta of
 * equal qhat^2 (relative 1e-12); q = 0 is left out.  Orbits in increasing
 * qhat^2: qhat2[k], pi[k], the standard deviation pi_sd[k] within the orbit and
 * count[k]; *n_orbits of them (error if more than max_orbits).  Host only. */
int cvchip_piq_orbits(const int dims[4], const double *piq, int max_orbits, double *qhat2, double *pi,
                      double *pi_sd, int *count, int *n_orbits);

/* The whole path for one source point: 60 up + 60 down solves (mu and -mu)
 * from point sources at y and y + nu^, then the contraction.  Replaces the
 * propagator reads and contraction of cvc_exact2_xspace.c:401-861 with
 * propagators computed on the GPU. */
int cvchip_cvc_source(cvchip_ctx *ctx, double *conn, double contact[8], const int src[4],
                      double sp, int niter_max, int prec, double delta,
                      cvchip_stats *stats);
/* The 120 propagators of one source point into host fields props (the
 * g_spinor_field order of cvchip_cvc_contract): the inversions of
 * cvc_exact2_xspace.c:437-541 (tmLQCD_invert of a point source at y + s^ with
 * the up operator, op_id 0, and the down operator, op_id 1), done on the GPU
 * with the same solves as cvchip_cvc_source. */
int cvchip_cvc_propagators(cvchip_ctx *ctx, double *const *props, const int src[4], double sp, int niter_max,
                           int prec, double delta, cvchip_stats *stats);

/* gamma_permutation / gamma_sign (global.h:143, filled by init_gamma,
 * cvc_utils.c:707-1520): the 16 gamma matrices as signed permutations of the
 * 24 reals of a spinor, index 0-3 gamma_mu, 4 identity, 5 gamma5, 6-9
 * gamma_mu gamma5, 10-15 gamma_0gamma_1, _0_2, _0_3, _1_2, _1_3, _2_3. */
int cvchip_gamma_tables(int perm[16][24], int sign[16][24]);

/* contract_twopoint / contract_twopoint_snk_momentum (cvc_utils.c:3075-3235):
 * contr[T][2] += sum_x e^{i p.(x - y)} sum_{mu, c} s_mu chi[p(mu) n_c + c]^dagger
 * gamma5 Gamma_idsink phi[mu n_c + c] per time slice (times -i for an
 * imaginary source gamma); chi, phi: 4 n_c spinor fields double[V][24];
 * snk_mom NULL for zero momentum, else 2 pi snk_mom / L; src: the source
 * point, whose spatial part enters the phase (g_source_location). */
int cvchip_contract_twopoint(cvchip_ctx *ctx, double *contr, int idsource, int idsink,
                             const double *const *chi, const double *const *phi, int n_c,
                             const int snk_mom[3], const int src[4]);
/* The same on device-resident FP64 sets of 4 n_c columns (column mu n_c + c). */
int cvchip_field_contract_twopoint(cvchip_ctx *ctx, double *contr, int idsource, int idsink,
                                   const cvchip_field *chi, const cvchip_field *phi, int n_c,
                                   const int snk_mom[3], const int src[4]);
/* No counterpart in the reference: cvchip_field_contract_twopoint for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).
 * Adds C(t) to contr_dev, device memory of 2 T doubles owned by the caller,
 * and allocates nothing: the per-block partial sums live in a file-scope
 * device array of 262144 doubles (2 MiB) per GPU, so the lattice must satisfy
 * T * ceil(LX LY LZ / 256) <= 131072 (else CVCHIP_ERR_ARG).  Calls that overlap
 * on one GPU share that array and race. */
int cvchip_field_contract_twopoint_noalloc(cvchip_ctx *ctx, double *contr_dev, int idsource, int idsink,
                                           const cvchip_field *chi, const cvchip_field *phi, int n_c,
                                           const int snk_mom[3], const int src[4]);

/* P -> gamma gamma (p2gg_xspace.c).  Sequential sources of p2gg_xspace.c:517-548
 * on device sets: out column r = the source of in column col0 + r,
 * out(x) = delta_{x_0, tseq} e^{i sgn 2 pi (x p_x / LX + y p_y / LY + z p_z / LZ)}
 * Gamma_gamma_id in(x) (mom NULL: p = 0; the reference's sgn is
 * 1 - 2 g_propagator_position). */
int cvchip_field_sequential_sources(cvchip_ctx *ctx, cvchip_field *out, const cvchip_field *in, int col0,
                                    int tseq, int gamma_id, const int mom[3], double sgn);
/* One P -> gamma gamma source point (p2gg_xspace.c:393-899): the 120 point
 * propagators from y and y + nu (as cvchip_cvc_source), the 60 sequential
 * sources of the flavour propagator_position (0 up, 1 down) on the time slice
 * (y_0 + seq_timeslice) mod T with gamma seq_gamma_id and momentum seq_mom
 * (p2gg_xspace.c:517-548), their solves with that flavour's operator
 * (:572-581), and the conserved-vector-current contraction of the sequential
 * propagators with the point propagators of the other flavour (:615-899):
 * conn[V][16][2] (-1/4 normalised, 4 mu + nu) and the contact term. */
int cvchip_p2gg_source(cvchip_ctx *ctx, double *conn, double contact[8], const int src[4], int seq_timeslice,
                       int seq_gamma_id, const int seq_mom[3], int propagator_position, double sp, int niter_max,
                       int prec, double delta, cvchip_stats *stats);

/* ------------------------------------------------ LIME / SciDAC / ILDG files
 * Host-only readers and writers of the reference's file formats (no GPU).
 * checksum, when not NULL, receives the SciDAC checksum (suma, sumb) of the
 * binary data (reference dml.c:41-53). */

/* write_lime_contraction (contractions_io.c:355-420): records
 * "cvc_contraction-format" (XML: type, precision, components, lx, ly, lz, lt,
 * nconf, source, date), "scidac-binary-data" (per site, t x y z order with z
 * fastest, N complex numbers, big-endian, prec 64 or 32) and
 * "scidac-checksum".  s: double[V][N][2] (conn layout for N = 16). */
int cvchip_lime_write_contraction(const char *filename, const double *s, int N, int prec, const int dims[4],
                                  const char *type, int nconf, int source, uint32_t checksum[2]);
/* read_lime_contraction (contractions_io.c:430-505): the position-th
 * scidac-binary-data record (0-based), precision from its length; an error if
 * a following scidac-checksum record disagrees with the data. */
int cvchip_lime_read_contraction(const char *filename, double *s, int N, int position, const int dims[4],
                                 uint32_t checksum[2]);
/* read_lime_gauge_field_doubleprec / _singleprec (io.c:202-347, 350-481): the
 * first ildg-binary-data record (sites t z y x with x fastest, links x y z t,
 * big-endian, 64 or 32 bit) into gauge[V][4][18] (the reference layout);
 * an error if a scidac-checksum record disagrees. */
int cvchip_lime_read_gauge(const char *filename, const int dims[4], double *gauge, uint32_t checksum[2]);
/* An ILDG file of gauge[V][4][18] that the reader above and the reference's
 * read_lime_gauge_field_doubleprec accept ("ildg-format", "ildg-binary-data",
 * "scidac-checksum").  No reference counterpart (the path reads gauges only). */
int cvchip_lime_write_gauge(const char *filename, const int dims[4], const double *gauge, int prec,
                            uint32_t checksum[2]);
/* read_lime_spinor (propagator_io.c:771-840, with read_binary_spinor_data
 * 344-405): the position-th scidac-binary-data record (0: up, 1: down in a
 * two-flavour file; sites t z y x with x fastest, 24 big-endian reals, 64 or 32
 * bit by the record length) into s[V][24] (the reference layout); an error if a
 * following scidac-checksum record disagrees.  A file without such a record is
 * read as CMI (read_cmi, propagator_io.c:842-892: host-order floats, sites
 * x y z t with t fastest, times 2 kappa; kappa is used only there). */
int cvchip_lime_read_spinor(const char *filename, const int dims[4], int position, double kappa, double *s,
                            uint32_t checksum[2]);
/* write_lime_spinor (propagator_io.c:582-647): "scidac-binary-data" (prec 64
 * or 32) and "scidac-checksum" (write_checksum, 410-466), appended (append = 1)
 * or in a new file. */
int cvchip_lime_write_spinor(const char *filename, const int dims[4], const double *s, int append, int prec,
                             uint32_t checksum[2]);
/* write_propagator (propagator_io.c:125-133): a new file with the
 * "etmc-propagator-format" record (write_propagator_format, 522-576, one
 * flavour) followed by write_lime_spinor. */
int cvchip_lime_write_propagator(const char *filename, const int dims[4], const double *s, int prec,
                                 uint32_t checksum[2]);

/* Gradient of L = sum_mu sum_{q in Q} Re Pi_{mu mu}(q) (Pi(q) as in cvchip_pi_q,
 * Q = {q : |q| <= qmax}, q_mu = 2 pi n_mu / L_mu with n_mu in (-L_mu/2, L_mu/2])
 * with respect to the twisted mass mu and every gauge link, by adjoint solves:
 * 120 forward and 120 adjoint gamma5-hermitian BiCGStab solves (prec / delta
 * as for cvchip_invert_Qtm), adjoint sources from the contraction, and
 * link-derivative outer products.  No reference counterpart.
 * Outputs: *L, *dL_dmu, and (if force != NULL) force[V][4][18]: per link the
 * traceless Hermitian 3x3 matrix F (row-major complex) with
 * dL/d eps = tr(T F) for U_mu(x) -> exp(i eps T) U_mu(x), T traceless Hermitian. */
int cvchip_cvc_gradient(cvchip_ctx *ctx, const int src[4], double qmax, double sp, int niter_max,
                        int prec, double delta, double *L, double *dL_dmu, double *force,
                        cvchip_stats *stats);

/* -------------------------------------------------------------- inputs */

/* Seeded gauge field U_mu(x) = exp(i eps H) in SU(3), H traceless Hermitian
 * from complex Gaussian entries drawn with Philox4x32-10 keyed by (seed, site,
 * mu).  Host-only; deterministic across runs and hosts of this image.  No
 * reference counterpart (the reference's generators use rand(),
 * cvc_utils.c:2275-2680). */
int cvchip_gauge_random(const int dims[4], uint64_t seed, double eps, double *gauge);

#ifdef __cplusplus
}
#endif
#endif
