// Device helpers shared by the contraction and the adjoint (gradient)
// kernels: source links, Dirac-vertex application, SU(3) x spinor products.
#pragma once

#include "spinor_io.h"

namespace cvchip {

struct SrcLinks {
  double2 u[4][9];   // Us_nu = U_nu(y) * phase_nu
};

// gamma_nu gamma5 (nu = 0..3) and gamma5 as (column, unit value) per row.
// gamma_nu gamma5 row a: gamma_nu[a][GCOL] * gamma5[GCOL][GCOL].
__device__ __forceinline__ double2 unit_val(int code) {
  return code == 0 ? make_double2(1, 0) : code == 1 ? make_double2(-1, 0)
         : code == 2 ? make_double2(0, 1) : make_double2(0, -1);
}
__device__ __forceinline__ int g5sign(int spin) { return spin < 2 ? 1 : -1; }

// v = gamma5 (gamma_mu -+ 1) w  (sgn = -1: Gm, +1: Gp), for runtime mu
__device__ __forceinline__ void sink_vertex(double2 v[12], const double2 w[12], int mu, double sgn) {
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int b = GCOL[mu][a];
    const double2 e = unit_val(GVAL[mu][a]);
    const double s5 = (double)g5sign(a);
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      double2 g = cmul(e, w[3 * b + cc]);
      g.x += sgn * w[3 * a + cc].x;
      g.y += sgn * w[3 * a + cc].y;
      v[3 * a + cc] = make_double2(s5 * g.x, s5 * g.y);
    }
  }
}

__device__ __forceinline__ void su3_mul(double2 o[12], const double2 u[9], const double2 p[12], bool dag) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double2 s = make_double2(0, 0);
#pragma unroll
      for (int j = 0; j < 3; j++) s = dag ? cfmac(s, u[3 * j + i], p[3 * a + j]) : cfma(s, u[3 * i + j], p[3 * a + j]);
      o[3 * a + i] = s;
    }
}

__device__ __forceinline__ double2 dot12(const double2 a[12], const double2 b[12]) {
  double2 s = make_double2(0, 0);
#pragma unroll
  for (int c = 0; c < 12; c++) s = cfmac(s, a[c], b[c]);
  return s;
}

// Bounds the compiler's load hoisting: the value is taken as read and written
// here and memory as clobbered, so loads that follow cannot move above it (the
// contraction kernels otherwise issue every load of a thread first and run at
// one wave per SIMD).
__device__ __forceinline__ void pin(double2 &v) { asm volatile("" : "+v"(v.x), "+v"(v.y) : : "memory"); }

struct ConnParams {
  uint32_t V, sX, sT;
  int T, LX, LY, LZ;
  double ph_re[4], ph_im[4];
};

inline ConnParams conn_params(const cvchip_ctx *c) {
  ConnParams P;
  P.V = (uint32_t)c->lat.V;
  P.T = c->lat.T;
  P.LX = c->lat.LX;
  P.LY = c->lat.LY;
  P.LZ = c->lat.LZ;
  P.sX = (uint32_t)(c->lat.LY * c->lat.LZ);
  P.sT = (uint32_t)(c->lat.LX * c->lat.LY * c->lat.LZ);
  for (int m = 0; m < 4; m++) {
    P.ph_re[m] = c->ph.re[m];
    P.ph_im[m] = c->ph.im[m];
  }
  return P;
}

inline uint32_t site_of(const cvchip_ctx *c, int t, int x, int y, int z) {
  const Lattice &L = c->lat;
  t = ((t % L.T) + L.T) % L.T;
  x = ((x % L.LX) + L.LX) % L.LX;
  y = ((y % L.LY) + L.LY) % L.LY;
  z = ((z % L.LZ) + L.LZ) % L.LZ;
  return (uint32_t)(((int64_t)(t * L.LX + x) * L.LY + y) * L.LZ + z);
}


// v = (gamma_mu + sgn) gamma5 w  (the hermitian conjugate of gamma5 (gamma_mu + sgn))
__device__ __forceinline__ void sink_vertex_dag(double2 v[12], const double2 w[12], int mu, double sgn) {
  double2 g5w[12];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      const double s5 = (double)g5sign(a);
      g5w[3 * a + cc] = make_double2(s5 * w[3 * a + cc].x, s5 * w[3 * a + cc].y);
    }
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int b = GCOL[mu][a];
    const double2 e = unit_val(GVAL[mu][a]);
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      double2 g = cmul(e, g5w[3 * b + cc]);
      g.x += sgn * g5w[3 * a + cc].x;
      g.y += sgn * g5w[3 * a + cc].y;
      v[3 * a + cc] = g;
    }
  }
}

}  // namespace cvchip
