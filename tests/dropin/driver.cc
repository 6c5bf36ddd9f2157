// Test driver for the drop-in shim: plays the part of a reference program
// (defines the globals of the reference's MAIN_PROGRAM unit and calls the
// reference entry points by name).  Inputs and outputs are raw FP64 files.
//
//   driver <dir> T LX LY LZ kappa mu
//     reads  <dir>/gauge.bin (V*72), <dir>/phi.bin (V*24)
//     writes <dir>/qphi.bin (Q_phi_tbc phi), <dir>/qf5.bin (Qf5(phi, -mu)),
//            <dir>/x.bin (invert_Qtm_her solution of D x = phi), <dir>/ret.txt,
//            <dir>/qphi_edit.bin (Q_phi_tbc phi after one link entry was changed
//            in place, no notice), <dir>/qphi_edit2.bin (after a second in-place
//            change announced by cvchip_gauge_changed())
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

namespace cvc {
typedef struct {
  double re, im;
} complex;
int T, LX, LY, LZ, VOLUME, T_global;
double **g_spinor_field;
double *g_gauge_field;
double g_kappa, g_mu;
int g_cart_id = 0;
double BCangle[4] = {1., 0., 0., 0.};
double solver_precision = 1.e-20;
int niter_max = 5000;
void Q_phi_tbc(double *xi, double *phi);
void Qf5(double *xi, double *phi, double mutm);
void spinor_scalar_product_co(complex *w, double *xi, double *phi, int V);
int invert_Qtm_her(double *xi, double *phi, int kwork);
}  // namespace cvc
#include "cvc_dropin.h"

static std::vector<double> readf(const char *dir, const char *name, size_t n) {
  char path[1024];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  std::vector<double> v(n);
  FILE *f = fopen(path, "rb");
  if (!f || fread(v.data(), sizeof(double), n, f) != n) { fprintf(stderr, "cannot read %s\n", path); exit(2); }
  fclose(f);
  return v;
}

static void writef(const char *dir, const char *name, const double *p, size_t n) {
  char path[1024];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, sizeof(double), n, f) != n) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
  fclose(f);
}

int main(int argc, char **argv) {
  using namespace cvc;
  if (argc != 8) { fprintf(stderr, "usage: driver dir T LX LY LZ kappa mu\n"); return 2; }
  const char *dir = argv[1];
  T = atoi(argv[2]); LX = atoi(argv[3]); LY = atoi(argv[4]); LZ = atoi(argv[5]);
  T_global = T;
  VOLUME = T * LX * LY * LZ;
  g_kappa = atof(argv[6]);
  g_mu = atof(argv[7]);
  std::vector<double> gauge = readf(dir, "gauge.bin", 72UL * VOLUME);
  std::vector<double> phi = readf(dir, "phi.bin", 24UL * VOLUME);
  g_gauge_field = gauge.data();
  std::vector<double> out(24UL * VOLUME), x(24UL * VOLUME, 0.);
  std::vector<std::vector<double>> work(8, std::vector<double>(24UL * VOLUME));
  std::vector<double *> ptrs(8);
  for (int k = 0; k < 8; k++) ptrs[k] = work[k].data();
  g_spinor_field = ptrs.data();
  Q_phi_tbc(out.data(), phi.data());
  writef(dir, "qphi.bin", out.data(), out.size());
  Qf5(out.data(), phi.data(), -g_mu);
  writef(dir, "qf5.bin", out.data(), out.size());
  complex w;
  spinor_scalar_product_co(&w, phi.data(), out.data(), VOLUME);
  int ret = invert_Qtm_her(x.data(), phi.data(), 0);
  writef(dir, "x.bin", x.data(), x.size());
  // in-place edits of one link entry, U_1(site 5)[0][0] (real part), as a
  // one-link finite difference would make; the second one is announced
  g_gauge_field[72 * 5 + 18 * 1 + 0] += 1e-3;
  Q_phi_tbc(out.data(), phi.data());
  writef(dir, "qphi_edit.bin", out.data(), out.size());
  g_gauge_field[72 * 5 + 18 * 1 + 1] -= 1e-3;
  cvchip_gauge_changed();
  Q_phi_tbc(out.data(), phi.data());
  writef(dir, "qphi_edit2.bin", out.data(), out.size());
  // a missing workspace field gives -2 as in the reference
  g_spinor_field[3] = NULL;
  int ret2 = invert_Qtm_her(x.data(), phi.data(), 0);
  char path[1024];
  snprintf(path, sizeof(path), "%s/ret.txt", dir);
  FILE *f = fopen(path, "w");
  fprintf(f, "%d %d %.17g %.17g\n", ret, ret2, w.re, w.im);
  fclose(f);
  return 0;
}
