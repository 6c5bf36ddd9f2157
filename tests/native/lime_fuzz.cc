// Host-only parts of libcvchip under AddressSanitizer and UBSan: LIME
// write/read round trips in both precisions, then every truncation, single-byte
// corruptions and random files fed to every reader (each must return, with an
// error or not, without a sanitizer report), the q-hat^2 orbits and the gauge
// generator.  Built and run by tests/test_native_sanitizers.py.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <vector>
#include <random>
#include "cvchip.h"
namespace cvchip {
// the library's error channel lives in ctx.hip; the readers only need a sink
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  char b[512];
  vsnprintf(b, sizeof b, fmt, ap);
  va_end(ap);
}
}
static std::vector<unsigned char> slurp(const char *p) {
  FILE *f = fopen(p, "rb"); std::vector<unsigned char> v; if (!f) return v;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  v.resize(n > 0 ? (size_t)n : 0);
  if (n > 0 && fread(v.data(), 1, (size_t)n, f) != (size_t)n) v.clear();
  fclose(f);
  return v;
}
static void dump(const char *p, const std::vector<unsigned char> &v) {
  FILE *f = fopen(p, "wb");
  if (!v.empty()) fwrite(v.data(), 1, v.size(), f);
  fclose(f);
}
static char P_C[4096], P_G[4096], P_S[4096], P_P[4096], P_X[4096];
int main(int argc, char **argv) {
  const char *dir = argc > 1 ? argv[1] : ".";
  snprintf(P_C, sizeof P_C, "%s/c", dir);
  snprintf(P_G, sizeof P_G, "%s/g", dir);
  snprintf(P_S, sizeof P_S, "%s/s", dir);
  snprintf(P_P, sizeof P_P, "%s/p", dir);
  snprintf(P_X, sizeof P_X, "%s/x", dir);
  const int dims[4] = {4, 2, 2, 2};
  const int V = 32;
  std::mt19937_64 rng(5);
  std::uniform_real_distribution<double> u(-1, 1);
  std::vector<double> conn(32 * V), back(32 * V), gauge(72 * V), gb(72 * V), sp(24 * V), sb(24 * V);
  for (auto &x : conn) x = u(rng);
  for (auto &x : sp) x = u(rng);
  if (cvchip_gauge_random(dims, 3, 0.3, gauge.data())) return 1;
  uint32_t ck[2];
  for (int prec : {64, 32}) {
    if (cvchip_lime_write_contraction(P_C, conn.data(), 16, prec, dims, "cvc", 1, 0, ck)) return 2;
    if (cvchip_lime_read_contraction(P_C, back.data(), 16, 0, dims, ck)) return 3;
    if (cvchip_lime_write_gauge(P_G, dims, gauge.data(), prec, ck)) return 4;
    if (cvchip_lime_read_gauge(P_G, dims, gb.data(), ck)) return 5;
    if (cvchip_lime_write_spinor(P_S, dims, sp.data(), 0, prec, ck)) return 6;
    if (cvchip_lime_write_spinor(P_S, dims, sp.data(), 1, prec, ck)) return 7;
    if (cvchip_lime_read_spinor(P_S, dims, 1, 0.0, sb.data(), ck)) return 8;
    if (cvchip_lime_write_propagator(P_P, dims, sp.data(), prec, ck)) return 9;
  }
  // damaged files: every truncation length and many single-byte flips must not crash
  for (const char *name : {P_C, P_G, P_S, P_P}) {
    std::vector<unsigned char> orig = slurp(name);
    for (size_t cut = 0; cut < orig.size(); cut += (cut < 400 ? 1 : 97)) {
      std::vector<unsigned char> v(orig.begin(), orig.begin() + cut);
      dump(P_X, v);
      cvchip_lime_read_contraction(P_X, back.data(), 16, 0, dims, ck);
      cvchip_lime_read_gauge(P_X, dims, gb.data(), ck);
      cvchip_lime_read_spinor(P_X, dims, 0, 0.0, sb.data(), ck);
      cvchip_lime_read_spinor(P_X, dims, 1, 0.125, sb.data(), ck);
    }
    std::uniform_int_distribution<size_t> pos(0, orig.size() - 1);
    for (int k = 0; k < 3000; k++) {
      std::vector<unsigned char> v = orig;
      v[pos(rng)] ^= (unsigned char)(1u << (k % 8));
      if (k % 3 == 0) v[pos(rng)] = (unsigned char)rng();
      dump(P_X, v);
      cvchip_lime_read_contraction(P_X, back.data(), 16, 0, dims, ck);
      cvchip_lime_read_gauge(P_X, dims, gb.data(), ck);
      cvchip_lime_read_spinor(P_X, dims, 0, 0.0, sb.data(), ck);
      cvchip_lime_read_spinor(P_X, dims, 1, 0.125, sb.data(), ck);
    }
  }
  // random garbage files
  for (int k = 0; k < 500; k++) {
    std::vector<unsigned char> v(rng() % 4000);
    for (auto &b : v) b = (unsigned char)rng();
    if (v.size() >= 4) { v[0] = 0x45; v[1] = 0x67; v[2] = 0x89; v[3] = 0xab; }
    dump(P_X, v);
    cvchip_lime_read_contraction(P_X, back.data(), 16, 0, dims, ck);
    cvchip_lime_read_gauge(P_X, dims, gb.data(), ck);
    cvchip_lime_read_spinor(P_X, dims, 0, 0.0, sb.data(), ck);
  }
  // orbits on a random Pi(q)
  std::vector<double> piq(32 * V);
  for (auto &x : piq) x = u(rng);
  std::vector<double> q2(64), pi(64), sd(64); std::vector<int> cnt(64); int no = 0;
  int rc = cvchip_piq_orbits(dims, piq.data(), 64, q2.data(), pi.data(), sd.data(), cnt.data(), &no);
  if (rc != 0 || no < 1) return 10;
  printf("host readers under sanitizers: done\n");
  return 0;
}
