// Seeded SU(3) gauge field U_mu(x) = exp(i eps H_mu(x)).
//
// H is the traceless Hermitian part of a 3x3 matrix A whose entries have
// independent N(0,1) real and imaginary parts: H = (A + A^dagger)/2 - tr/3.
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11) with key (seed lo, seed hi) and counter
// (site lo, site hi, 16*mu + block, 0x61756765); Box-Muller turns uniform pairs into normals.
// exp is evaluated by scaling and squaring of a degree-18 Taylor polynomial,
// then the rows are re-orthonormalised and the third row is set to
// conj(row0 x row1), so every link is in SU(3) to rounding.
//
// Compiled for the host with -ffp-contract=off: the output is identical on
// every host of this image.  OpenMP splits sites; results do not depend on
// the thread count.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/cvchip.h"
#include "entry_guard.h"

namespace {

struct C { double re, im; };

inline C cm(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C ca(C a, C b) { return {a.re + b.re, a.im + b.im}; }
inline C cconj(C a) { return {a.re, -a.im}; }

inline void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * ctr[0];
    uint64_t p1 = (uint64_t)M1 * ctr[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ ctr[1] ^ k0;
    uint32_t n2 = hi0 ^ ctr[3] ^ k1;
    ctr[0] = n0; ctr[1] = lo1; ctr[2] = n2; ctr[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

inline double u01(uint32_t v) { return ((double)v + 0.5) * (1.0 / 4294967296.0); }

void mat_mul(C *o, const C *a, const C *b) {
  C t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C s = {0., 0.};
      for (int k = 0; k < 3; k++) s = ca(s, cm(a[3 * i + k], b[3 * k + j]));
      t[3 * i + j] = s;
    }
  memcpy(o, t, sizeof(t));
}

void one_link(uint64_t seed, int64_t site, int mu, double eps, double *out18) {
  // 18 normals from 20 uniforms (5 Philox blocks)
  uint32_t u[20];
  for (int b = 0; b < 5; b++) {
    uint32_t ctr[4] = {(uint32_t)site, (uint32_t)((uint64_t)site >> 32),
                       (uint32_t)(16 * mu + b), 0x61756765u};
    philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
    for (int k = 0; k < 4; k++) u[4 * b + k] = ctr[k];
  }
  double g[18];
  const double twopi = 6.283185307179586476925286766559;
  for (int k = 0; k < 9; k++) {
    double r = sqrt(-2.0 * log(u01(u[2 * k])));
    double th = twopi * u01(u[2 * k + 1]);
    g[2 * k] = r * cos(th);
    g[2 * k + 1] = r * sin(th);
  }
  C A[9], H[9];
  for (int k = 0; k < 9; k++) A[k] = {g[2 * k], g[2 * k + 1]};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C a = A[3 * i + j], b = cconj(A[3 * j + i]);
      H[3 * i + j] = {0.5 * (a.re + b.re), 0.5 * (a.im + b.im)};
    }
  double tr = (H[0].re + H[4].re + H[8].re) / 3.0;
  H[0].re -= tr; H[4].re -= tr; H[8].re -= tr;
  // X = i eps H
  C X[9];
  double nrm = 0.;
  for (int k = 0; k < 9; k++) {
    X[k] = {-eps * H[k].im, eps * H[k].re};
    nrm += X[k].re * X[k].re + X[k].im * X[k].im;
  }
  nrm = sqrt(nrm);
  int sq = 0;
  double scale = 1.0;
  while (nrm * scale > 0.25) { scale *= 0.5; sq++; }
  for (int k = 0; k < 9; k++) { X[k].re *= scale; X[k].im *= scale; }
  // Taylor: E = sum_{n=0}^{18} X^n / n!  (Horner form)
  C E[9];
  for (int k = 0; k < 9; k++) E[k] = {0., 0.};
  E[0].re = E[4].re = E[8].re = 1.0;
  for (int n = 18; n >= 1; n--) {
    C T[9];
    mat_mul(T, X, E);
    for (int k = 0; k < 9; k++) { T[k].re /= (double)n; T[k].im /= (double)n; }
    for (int k = 0; k < 9; k++) E[k] = T[k];
    E[0].re += 1.0; E[4].re += 1.0; E[8].re += 1.0;
  }
  for (int s = 0; s < sq; s++) mat_mul(E, E, E);
  // re-orthonormalise rows 0 and 1, row 2 = conj(row0 x row1)
  C *r0 = E, *r1 = E + 3, *r2 = E + 6;
  double n0 = 0.;
  for (int j = 0; j < 3; j++) n0 += r0[j].re * r0[j].re + r0[j].im * r0[j].im;
  n0 = 1.0 / sqrt(n0);
  for (int j = 0; j < 3; j++) { r0[j].re *= n0; r0[j].im *= n0; }
  C d = {0., 0.};
  for (int j = 0; j < 3; j++) d = ca(d, cm(cconj(r0[j]), r1[j]));
  for (int j = 0; j < 3; j++) {
    C p = cm(d, r0[j]);
    r1[j].re -= p.re; r1[j].im -= p.im;
  }
  double n1 = 0.;
  for (int j = 0; j < 3; j++) n1 += r1[j].re * r1[j].re + r1[j].im * r1[j].im;
  n1 = 1.0 / sqrt(n1);
  for (int j = 0; j < 3; j++) { r1[j].re *= n1; r1[j].im *= n1; }
  r2[0] = cconj({cm(r0[1], r1[2]).re - cm(r0[2], r1[1]).re, cm(r0[1], r1[2]).im - cm(r0[2], r1[1]).im});
  r2[1] = cconj({cm(r0[2], r1[0]).re - cm(r0[0], r1[2]).re, cm(r0[2], r1[0]).im - cm(r0[0], r1[2]).im});
  r2[2] = cconj({cm(r0[0], r1[1]).re - cm(r0[1], r1[0]).re, cm(r0[0], r1[1]).im - cm(r0[1], r1[0]).im});
  for (int k = 0; k < 9; k++) { out18[2 * k] = E[k].re; out18[2 * k + 1] = E[k].im; }
}

}  // namespace

extern "C" int cvchip_gauge_random(const int dims[4], uint64_t seed, double eps, double *gauge) {
  return cvchip::guarded("cvchip_gauge_random", [&]() -> int {
    if (!dims || !gauge) return CVCHIP_ERR_NULL;
    for (int m = 0; m < 4; m++)
      if (dims[m] <= 0) return CVCHIP_ERR_ARG;
    const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  #pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < V; s++)
      for (int mu = 0; mu < 4; mu++) one_link(seed, s, mu, eps, gauge + 72 * s + 18 * mu);
    return CVCHIP_OK;
  });
}
