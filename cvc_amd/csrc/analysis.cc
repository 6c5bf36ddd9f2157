// Pi(q) analysis: the scalar vacuum polarisation on orbits of equal qhat^2.
//
// The reference writes Pi_mu_nu(x) and leaves the momentum-space analysis to
// programs it does not ship (its README names analyse_piq.c and
// make_q2orbits.c, "which mainly computes the lattice momenta qhat^2").  This
// is that step, defined here: with the transverse form
//   Pi_mu_nu(q) = (qhat_mu qhat_nu - delta_mu_nu qhat^2) Pi(qhat^2),
//   qhat_mu = 2 sin(q_mu / 2),  q_mu = 2 pi n_mu / L_mu,
// the trace gives Pi(qhat^2) = - sum_mu Re Pi_mu_mu(q) / (3 qhat^2) for every
// q != 0; momenta with equal qhat^2 (relative 1e-12) form one orbit, averaged.
// Host code: the input is the Pi(q) array of cvchip_pi_q.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "../../include/cvchip.h"
#include "entry_guard.h"

namespace cvchip {
void set_error(const char *fmt, ...);
}

extern "C" int cvchip_piq_orbits(const int dims[4], const double *piq, int max_orbits, double *qhat2,
                                 double *pi, double *pi_sd, int *count, int *n_orbits) {
  return cvchip::guarded("cvchip_piq_orbits", [&]() -> int {
    if (!dims || !piq || !qhat2 || !pi || !pi_sd || !count || !n_orbits) {
      cvchip::set_error("NULL argument to cvchip_piq_orbits");
      return CVCHIP_ERR_NULL;
    }
    for (int m = 0; m < 4; m++)
      if (dims[m] < 1) {
        cvchip::set_error("invalid lattice extent %d", dims[m]);
        return CVCHIP_ERR_ARG;
      }
    const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
    std::vector<double> s[4];
    for (int m = 0; m < 4; m++) {
      s[m].resize(dims[m]);
      for (int n = 0; n < dims[m]; n++) {
        const double h = 2.0 * sin(M_PI * n / dims[m]);
        s[m][n] = h * h;
      }
    }
    std::vector<std::pair<double, int64_t>> q(V);
    int64_t k = 0;
    for (int t = 0; t < dims[0]; t++)
      for (int x = 0; x < dims[1]; x++)
        for (int y = 0; y < dims[2]; y++)
          for (int z = 0; z < dims[3]; z++, k++) q[k] = {s[0][t] + s[1][x] + s[2][y] + s[3][z], k};
    std::sort(q.begin(), q.end());
    int no = 0;
    int64_t i = 0;
    while (i < V && q[i].first < 1e-14) i++;   // q = 0: Pi(qhat^2) is not defined there
    while (i < V) {
      int64_t j = i;
      const double q0 = q[i].first;
      while (j < V && q[j].first - q0 <= 1e-12 * q0) j++;
      if (no >= max_orbits) {
        cvchip::set_error("more than %d qhat^2 orbits", max_orbits);
        return CVCHIP_ERR_ARG;
      }
      // Re Pi_00, Pi_11, Pi_22, Pi_33 sit at offsets 0, 10, 20, 30 of a momentum's 32 doubles
      auto value = [&](int64_t l) {
        const double *p = piq + q[l].second * 32;
        return -(p[0] + p[10] + p[20] + p[30]) / (3.0 * q[l].first);
      };
      double sum = 0., qs = 0.;
      for (int64_t l = i; l < j; l++) {
        sum += value(l);
        qs += q[l].first;
      }
      const double c = (double)(j - i), mean = sum / c;
      double dev = 0.;
      for (int64_t l = i; l < j; l++) dev += (value(l) - mean) * (value(l) - mean);
      qhat2[no] = qs / c;
      pi[no] = mean;
      pi_sd[no] = c > 1 ? sqrt(dev / (c - 1)) : 0.;
      count[no] = (int)(j - i);
      no++;
      i = j;
    }
    *n_orbits = no;
    return CVCHIP_OK;
  });
}
