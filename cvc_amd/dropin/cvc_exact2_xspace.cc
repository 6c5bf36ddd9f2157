// Drop-in for the reference program cvc_exact2_xspace (cvc_exact2_xspace.c):
// the conserved-vector-current Pi_mu_nu(x) of one point source from 120
// propagators, read from files or computed on the GPU.
//
//   cvc_exact2_xspace [-f input] [-w] [-h]
//
// -f  input file (default cvc.input), the reference's "key = value" lines
//     (read_input_parser_cvc.l); used keys: T, L, LX, LY, LZ, kappa, mu,
//     BCangleT/X/Y/Z, Nconf, gaugefilename_prefix, outfile_prefix,
//     source_location, solver_precision, niter_max, read_propagator,
//     write_propagator, format, filename_prefix, filename_prefix2; and two keys
//     of this program: gpu_precision (64, 32 or 16: the inner solver; default
//     64) and gpu_even_odd (0/1: even/odd preconditioned solves; default 0).
//     Other keys are accepted and ignored.
//
// read_propagator = yes (the reference's default) reads the up propagator of
// shift s and spin-colour a from record 0 and the down propagator from record
// 1 of the file get_filename(s, a, +1) names (cvc_exact2_xspace.c:395-431,
// cvc_utils.c:2143-2197, formats 0, 2, 3 and 4).  read_propagator = no
// computes them on the GPU (the reference's tmLQCD_invert branch,
// cvc_exact2_xspace.c:437-541); with write_propagator = yes they are also
// written to those files (up: write_propagator, down appended), so that a
// later run with read_propagator = yes reads them back.
// -w  write the position-space Ward identity to WI_X.<Nconf>.0000
//     (cvc_exact2_xspace.c:928-1008).
//
// The gauge field is read from <gaugefilename_prefix>.<Nconf> (ILDG, as
// read_lime_gauge_field_doubleprec), or set to unit links for "identity".  The
// result is written as cvc2_v_x.<Nconf> (or <outfile_prefix>/cvc2_v_x.<Nconf>)
// in the format of write_lime_contraction, 64 bit, 16 components.  Standard
// output carries the reference's lines (plaquette, source coordinates,
// timings, the contact term).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <complex>
#include <map>
#include <string>
#include <vector>

#include "../../include/cvchip.h"

namespace {

using cd = std::complex<double>;

struct Input {
  int T = 0, LX = 0, LY = 0, LZ = 0, Nconf = 0, source_location = 0, niter_max = 1000;
  double kappa = 0., mu = 0., bc[4] = {1., 0., 0., 0.}, solver_precision = 1e-12;
  std::string gauge_prefix = "identity", out_prefix = "NA";
  std::string prop_prefix = "prop.", prop_prefix2 = "prop.";   // default_input_values.h:28-29
  int format = 0, read_propagator = 1, write_propagator = 0;   // default_input_values.h:23, 135-136
  int gpu_precision = 64, gpu_even_odd = 0;
};

bool read_input(const char *fn, Input &in) {
  FILE *f = fopen(fn, "r");
  if (!f) return false;
  char line[4096];
  while (fgets(line, sizeof(line), f)) {
    char *h = strchr(line, '#');
    if (h) *h = 0;
    char *eq = strchr(line, '=');
    if (!eq) continue;
    *eq = 0;
    char key[256], val[4096];
    if (sscanf(line, " %255s", key) != 1 || sscanf(eq + 1, " %4095s", val) != 1) continue;
    const std::string k = key;
    if (k == "T") in.T = atoi(val);
    else if (k == "L") in.LX = in.LY = in.LZ = atoi(val);
    else if (k == "LX") in.LX = atoi(val);
    else if (k == "LY") in.LY = atoi(val);
    else if (k == "LZ") in.LZ = atoi(val);
    else if (k == "kappa") in.kappa = atof(val);
    else if (k == "mu") in.mu = atof(val);
    else if (k == "BCangleT") in.bc[0] = atof(val);
    else if (k == "BCangleX") in.bc[1] = atof(val);
    else if (k == "BCangleY") in.bc[2] = atof(val);
    else if (k == "BCangleZ") in.bc[3] = atof(val);
    else if (k == "Nconf") in.Nconf = atoi(val);
    else if (k == "gaugefilename_prefix") in.gauge_prefix = val;
    else if (k == "outfile_prefix") in.out_prefix = val;
    else if (k == "source_location") in.source_location = atoi(val);
    else if (k == "solver_precision") in.solver_precision = atof(val);
    else if (k == "niter_max") in.niter_max = atoi(val);
    else if (k == "format") in.format = atoi(val);
    else if (k == "filename_prefix") in.prop_prefix = val;
    else if (k == "filename_prefix2") in.prop_prefix2 = val;
    else if (k == "read_propagator" && (!strcmp(val, "yes") || !strcmp(val, "no"))) in.read_propagator = !strcmp(val, "yes");
    else if (k == "write_propagator" && (!strcmp(val, "yes") || !strcmp(val, "no"))) in.write_propagator = !strcmp(val, "yes");
    else if (k == "gpu_precision") in.gpu_precision = atoi(val);
    else if (k == "gpu_even_odd") in.gpu_even_odd = atoi(val);
  }
  fclose(f);
  return true;
}

// plaquette (reference cvc_utils.c:384-411): sum over x and mu < nu of
// Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dagger U_nu(x)^dagger, over 18 V
double plaquette(const std::vector<double> &g, const int d[4]) {
  const long V = (long)d[0] * d[1] * d[2] * d[3];
  auto link = [&](long ix, int mu, int i, int j) {
    const double *p = g.data() + (ix * 4 + mu) * 18 + 2 * (3 * i + j);
    return cd(p[0], p[1]);
  };
  auto up = [&](long ix, int mu) {
    long c[4], r = ix;
    for (int m = 3; m >= 0; m--) {
      c[m] = r % d[m];
      r /= d[m];
    }
    c[mu] = (c[mu] + 1) % d[mu];
    return ((c[0] * d[1] + c[1]) * d[2] + c[2]) * d[3] + c[3];
  };
  double sum = 0.;
  for (long ix = 0; ix < V; ix++)
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        const long xm = up(ix, mu), xn = up(ix, nu);
        cd s[3][3], t[3][3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            s[i][j] = t[i][j] = 0.;
            for (int k = 0; k < 3; k++) {
              s[i][j] += link(ix, mu, i, k) * link(xm, nu, k, j);
              t[i][j] += link(ix, nu, i, k) * link(xn, mu, k, j);
            }
          }
        for (int i = 0; i < 3; i++)
          for (int k = 0; k < 3; k++) sum += (s[i][k] * std::conj(t[i][k])).real();
      }
  return sum / ((double)V * 18.);
}

// get_filename(filename, nu, sc, +1) (reference cvc_utils.c:2143-2197): the
// propagator file of source shift nu (0-3: y + nu^, 4: y) and spin-colour sc;
// empty for a format the reference does not name a file for
std::string prop_filename(const Input &in, int nu, int sc) {
  int isx[4] = {in.source_location / (in.LX * in.LY * in.LZ), (in.source_location % (in.LX * in.LY * in.LZ)) / (in.LY * in.LZ),
                (in.source_location % (in.LY * in.LZ)) / in.LZ, in.source_location % in.LZ};
  const int L[4] = {in.T, in.LX, in.LY, in.LZ};
  char b[8400];
  const char *p = in.prop_prefix.c_str(), *p2 = in.prop_prefix2.c_str();
  switch (in.format) {
    case 0: snprintf(b, sizeof(b), "%s.%.4d.%.2d.%.2d.inverted", p, in.Nconf, nu, sc); break;
    case 2:
    case 3:
      if (nu != 4) isx[nu] = (isx[nu] + 1) % L[nu];
      if (in.format == 2)
        snprintf(b, sizeof(b), "%s.%.4d.t%.2dx%.2dy%.2dz%.2ds%.2dc%.2d.pmass.%s.inverted", p, in.Nconf, isx[0], isx[1],
                 isx[2], isx[3], sc / 3, sc % 3, p2);
      else
        snprintf(b, sizeof(b), "%s.%.4d.t%.2dx%.2dy%.2dz%.2d.pmass.%s.%.2d.inverted", p, in.Nconf, isx[0], isx[1], isx[2],
                 isx[3], p2, sc);
      break;
    case 4: snprintf(b, sizeof(b), "%s.%.4d.%.2d.inverted", p, in.Nconf, sc); break;
    default: return std::string();
  }
  return b;
}

double now() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void usage() {
  fprintf(stdout, "Code to perform cvc correlator conn. contractions\n");
  fprintf(stdout, "Usage:    [options]\n");
  fprintf(stdout, "         -f input filename [default cvc.input]\n");
  fprintf(stdout, "         -w check the Ward identity in position space\n");
}

}  // namespace

int main(int argc, char **argv) {
  const char *fn = "cvc.input";
  int check_wi = 0, c;
  while ((c = getopt(argc, argv, "wh?f:")) != -1) {
    switch (c) {
      case 'f': fn = optarg; break;
      case 'w': check_wi = 1; fprintf(stdout, "\n# [cvc_exact2_xspace] will check Ward identity in position space\n"); break;
      default: usage(); return 0;
    }
  }
  fprintf(stdout, "# [cvc_exact2_xspace] Reading input from file %s\n", fn);
  Input in;
  if (!read_input(fn, in)) {
    fprintf(stderr, "[cvc_exact2_xspace] could not read input file %s\n", fn);
    return 1;
  }
  if (in.T == 0 || in.LX == 0 || in.LY == 0 || in.LZ == 0) {
    fprintf(stderr, "[cvc_exact2_xspace] T and L's must be set\n");
    return 2;
  }
  const int dims[4] = {in.T, in.LX, in.LY, in.LZ};
  const long V = (long)in.T * in.LX * in.LY * in.LZ;
  std::vector<double> g((size_t)V * 72, 0.);
  if (in.gauge_prefix == "identity") {
    fprintf(stdout, "\n# [cvc_exact] initializing unit matrices\n");
    for (long ix = 0; ix < V * 4; ix++)
      for (int i = 0; i < 3; i++) g[(size_t)ix * 18 + 2 * (3 * i + i)] = 1.;
  } else {
    char name[4200];
    snprintf(name, sizeof(name), "%s.%.4d", in.gauge_prefix.c_str(), in.Nconf);
    fprintf(stdout, "# [cvc_exact2_xspace] reading gauge field from file %s\n", name);
    uint32_t ck[2];
    if (cvchip_lime_read_gauge(name, dims, g.data(), ck) != CVCHIP_OK) {
      fprintf(stderr, "[cvc_exact2_xspace] %s\n", cvchip_last_error());
      return 3;
    }
    fprintf(stdout, "# [cvc_exact2_xspace] gauge field checksum %#x %#x\n", ck[0], ck[1]);
  }
  fprintf(stdout, "# [cvc_exact2_xspace] measured plaquette value: %25.16e\n", plaquette(g, dims));

  const int sl = in.source_location;
  const int src[4] = {sl / (in.LX * in.LY * in.LZ), (sl % (in.LX * in.LY * in.LZ)) / (in.LY * in.LZ),
                      (sl % (in.LY * in.LZ)) / in.LZ, sl % in.LZ};
  fprintf(stdout, "# [cvc_exact2_xspace] global source coordinates: (%3d,%3d,%3d,%3d)\n", src[0], src[1], src[2],
          src[3]);

  cvchip_ctx *ctx = nullptr;
  int rc = cvchip_ctx_create(&ctx, dims, in.bc, in.kappa, in.mu);
  if (!rc) rc = cvchip_set_gauge(ctx, g.data());
  if (!rc && in.gpu_even_odd) rc = cvchip_ctx_set_solver_flags(ctx, CVCHIP_SOLVE_EO);
  const int prec = in.gpu_precision == 16 ? CVCHIP_PREC_F16 : in.gpu_precision == 32 ? CVCHIP_PREC_F32 : CVCHIP_PREC_F64;
  // reliable-update factors: FP16 inner 0.1 (0.02 on the better conditioned even/odd system), FP32 inner 0.01
  const double delta = prec == CVCHIP_PREC_F16 ? (in.gpu_even_odd ? 0.02 : 0.1) : 0.01;
  std::vector<double> conn((size_t)V * 32);
  double contact[8];
  cvchip_stats st;
  memset(&st, 0, sizeof(st));
  const double t0 = now();
  if (!rc && !in.read_propagator && !in.write_propagator) {
    // propagators stay in HBM from the solves to the contraction
    rc = cvchip_cvc_source(ctx, conn.data(), contact, src, in.solver_precision, in.niter_max, prec, delta, &st);
  } else if (!rc) {
    std::vector<std::vector<double>> store(120, std::vector<double>((size_t)V * 24));
    std::vector<double *> props(120);
    for (int k = 0; k < 120; k++) props[k] = store[k].data();
    if (in.read_propagator) {
      for (int k = 0; k < 120 && !rc; k++) {
        const std::string name = prop_filename(in, (k % 60) / 12, k % 12);
        if (name.empty()) {
          fprintf(stderr, "[cvc_exact2_xspace] no propagator file name for format %d\n", in.format);
          rc = CVCHIP_ERR_ARG;
          break;
        }
        uint32_t ck[2];
        fprintf(stdout, "# [read_lime_spinor] Reading Dirac-fermion field in LIME format from %s\n", name.c_str());
        rc = cvchip_lime_read_spinor(name.c_str(), dims, k < 60 ? 0 : 1, in.kappa, props[k], ck);
        if (!rc) fprintf(stdout, "[read_binary_spinor_data] The final checksum is %#x %#x\n", ck[0], ck[1]);
      }
    } else {
      rc = cvchip_cvc_propagators(ctx, props.data(), src, in.solver_precision, in.niter_max, prec, delta, &st);
      for (int k = 0; k < 60 && !rc; k++) {
        const std::string name = prop_filename(in, k / 12, k % 12);
        if (name.empty()) {
          fprintf(stderr, "[cvc_exact2_xspace] no propagator file name for format %d\n", in.format);
          rc = CVCHIP_ERR_ARG;
          break;
        }
        uint32_t ck[2];
        rc = cvchip_lime_write_propagator(name.c_str(), dims, props[k], 64, ck);
        if (!rc) rc = cvchip_lime_write_spinor(name.c_str(), dims, props[60 + k], 1, 64, ck);
      }
    }
    if (!rc) rc = cvchip_cvc_contract(ctx, conn.data(), contact, props.data(), src);
  }
  if (rc) {
    fprintf(stderr, "[cvc_exact2_xspace] error %d: %s\n", rc, cvchip_last_error());
    if (ctx) cvchip_ctx_destroy(ctx);
    return 4;
  }
  const double t1 = now();
  fprintf(stdout, "# [cvc_exact2_xspace] reading / invert and contractions in %e seconds (%d iterations summed, %d rhs converged)\n",
          t1 - t0, st.iterations, st.converged);
  fprintf(stdout, "# [cvc_exact2_xspace] contact term:\n");
  for (int nu = 0; nu < 4; nu++)
    fprintf(stdout, "\t%d\t%25.16e%25.16e\n", nu, contact[2 * nu], contact[2 * nu + 1]);

  char name[4200];
  if (in.out_prefix == "NA") snprintf(name, sizeof(name), "cvc2_v_x.%.4d", in.Nconf);
  else snprintf(name, sizeof(name), "%s/cvc2_v_x.%.4d", in.out_prefix.c_str(), in.Nconf);
  uint32_t ck[2];
  rc = cvchip_lime_write_contraction(name, conn.data(), 16, 64, dims, "cvc - cvc in position space, all 16 components",
                                     in.Nconf, 0, ck);
  if (rc) {
    fprintf(stderr, "[cvc_exact2_xspace] %s\n", cvchip_last_error());
    cvchip_ctx_destroy(ctx);
    return 5;
  }
  fprintf(stdout, "# Final check sum is (%#x  %#x)\n", ck[0], ck[1]);
  fprintf(stdout, "# [cvc_exact2_xspace] saved position space results in %e seconds\n", now() - t1);

  if (check_wi) {
    std::vector<double> wi((size_t)V * 8);
    rc = cvchip_wi(ctx, wi.data(), conn.data());
    if (!rc) {
      snprintf(name, sizeof(name), "WI_X.%.4d.%.4d", in.Nconf, 0);
      FILE *ofs = fopen(name, "w");
      fprintf(stdout, "\n# [cvc_exact2_xspace] checking Ward identity in position space ...\n");
      for (int x0 = 0; x0 < in.T; x0++)
        for (int x1 = 0; x1 < in.LX; x1++)
          for (int x2 = 0; x2 < in.LY; x2++)
            for (int x3 = 0; x3 < in.LZ; x3++) {
              const long ix = ((long)(x0 * in.LX + x1) * in.LY + x2) * in.LZ + x3;
              if (ofs) fprintf(ofs, "# t=%2d x=%2d y=%2d z=%2d\n", x0, x1, x2, x3);
              for (int nu = 0; nu < 4; nu++)
                if (ofs) fprintf(ofs, "\t%3d%25.16e%25.16e\n", nu, wi[(size_t)ix * 8 + 2 * nu], wi[(size_t)ix * 8 + 2 * nu + 1]);
            }
      if (ofs) fclose(ofs);
    }
  }
  cvchip_ctx_destroy(ctx);
  return rc ? 6 : 0;
}
