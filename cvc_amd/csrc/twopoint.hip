// Meson two-point functions on propagators resident in HBM (reference
// cvc_utils.c:3075-3235, contract_twopoint and contract_twopoint_snk_momentum):
//   C(t) += sum_x e^{i p.(x - y)} sum_{mu, c} s_mu chi_{p(mu) n_c + c}(x)^dagger gamma5 Gamma_snk phi_{mu n_c + c}(x)
// (times -i for an imaginary source gamma), with the source gamma acting
// through its column map p(mu) and signs s_mu, as the reference's permutation
// tables do.  The 16 gamma matrices are built from the chiral-basis gamma_mu
// (GCOL/GVAL), each a unit-per-row matrix, in the reference's index order:
// 0-3 gamma_mu, 4 identity, 5 gamma5, 6-9 gamma_mu gamma5, 10-15 gamma_0 gamma_1,
// gamma_0 gamma_2, gamma_0 gamma_3, gamma_1 gamma_2, gamma_1 gamma_3, gamma_2 gamma_3.
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "spinor_io.h"

namespace cvchip {

namespace {

// a unit-per-row 4x4 matrix: row a has g[a] = (re, im) at column col[a]
struct UMat {
  int col[4];
  int re[4], im[4];
};

UMat umul(const UMat &A, const UMat &B) {
  UMat C;
  for (int a = 0; a < 4; a++) {
    const int k = A.col[a];
    C.col[a] = B.col[k];
    C.re[a] = A.re[a] * B.re[k] - A.im[a] * B.im[k];
    C.im[a] = A.re[a] * B.im[k] + A.im[a] * B.re[k];
  }
  return C;
}

void gamma_mats(UMat g[16]) {
  static const int ure[4] = {1, -1, 0, 0}, uim[4] = {0, 0, 1, -1};   // unit codes +1 -1 +i -i
  for (int m = 0; m < 4; m++)
    for (int a = 0; a < 4; a++) {
      g[m].col[a] = GCOL[m][a];
      g[m].re[a] = ure[GVAL[m][a]];
      g[m].im[a] = uim[GVAL[m][a]];
    }
  for (int a = 0; a < 4; a++) {
    g[4].col[a] = a, g[4].re[a] = 1, g[4].im[a] = 0;
    g[5].col[a] = a, g[5].re[a] = a < 2 ? 1 : -1, g[5].im[a] = 0;
  }
  for (int m = 0; m < 4; m++) g[6 + m] = umul(g[m], g[5]);
  static const int pr[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int p = 0; p < 6; p++) g[10 + p] = umul(g[pr[p][0]], g[pr[p][1]]);
}

struct TwoPt {
  int q[12];            // gamma5 Gamma_snk: output component k takes input component q[k] ...
  double ua[12], ub[12];   // ... times (ua + i ub)
  int psource[4];
  double ssource[4];
  int isimag, n_c, N;
  double px, py, pz;
  int src[4];
  int T, LX, LY, LZ;
  uint32_t VN, vol3;
  int nchunk;
};

}  // namespace

// row a of gamma matrix id (reference index order): the entry re[a] + i im[a] at column col[a]
void gamma_matrix_rows(int id, int col[4], int re[4], int im[4]) {
  UMat g[16];
  gamma_mats(g);
  for (int a = 0; a < 4; a++) {
    col[a] = g[id].col[a];
    re[a] = g[id].re[a];
    im[a] = g[id].im[a];
  }
}

__global__ void __launch_bounds__(256) k_twopoint(const double2 *__restrict__ chi, const double2 *__restrict__ phi,
                                                  TwoPt P, double *__restrict__ part) {
  const int t = blockIdx.x / P.nchunk, chunk = blockIdx.x - t * P.nchunk;
  const uint32_t s3 = (uint32_t)chunk * 256u + threadIdx.x;
  double vr = 0., vi = 0.;
  if (s3 < P.vol3) {
    const uint32_t z = s3 % (uint32_t)P.LZ, y = (s3 / (uint32_t)P.LZ) % (uint32_t)P.LY,
                   x = s3 / ((uint32_t)P.LZ * (uint32_t)P.LY);
    const uint32_t site = (uint32_t)t * P.vol3 + s3;
    double tr = 0., ti = 0.;
    for (int mu = 0; mu < 4; mu++)
      for (int c = 0; c < P.n_c; c++) {
        const uint32_t ip = site * (uint32_t)P.N + (uint32_t)(mu * P.n_c + c);
        const uint32_t ic = site * (uint32_t)P.N + (uint32_t)(P.psource[mu] * P.n_c + c);
        double wr = 0., wi = 0.;
#pragma unroll
        for (int k = 0; k < 12; k++) {
          const double2 f = phi[(size_t)P.q[k] * P.VN + ip];
          const double sr = P.ua[k] * f.x - P.ub[k] * f.y, si = P.ua[k] * f.y + P.ub[k] * f.x;
          const double2 h = chi[(size_t)k * P.VN + ic];
          wr += h.x * sr + h.y * si;
          wi += h.x * si - h.y * sr;
        }
        tr += P.ssource[mu] * wr;
        ti += P.ssource[mu] * wi;
      }
    const double ph = ((double)x - P.src[1]) * P.px + ((double)y - P.src[2]) * P.py + ((double)z - P.src[3]) * P.pz;
    double sp, cp;
    sincos(ph, &sp, &cp);
    const double re = tr * cp - ti * sp, im = tr * sp + ti * cp;
    if (!P.isimag) {
      vr = re;
      vi = im;
    } else {
      vr = im;
      vi = -re;
    }
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = vr;
  sh[1][threadIdx.x] = vi;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + h];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// Per time slice t: contr[2 t], contr[2 t + 1] += the sums of the block partials
// of slice t, in chunk order (the host loop of twopoint_fields, on the device).
__global__ void k_twopoint_final(const double *__restrict__ part, double *__restrict__ contr, int T, int nchunk) {
  for (int t = threadIdx.x; t < T; t += blockDim.x) {
    double re = 0., im = 0.;
    for (int k = 0; k < nchunk; k++) {
      re += part[2 * ((size_t)t * nchunk + k)];
      im += part[2 * ((size_t)t * nchunk + k) + 1];
    }
    contr[2 * t] += re;
    contr[2 * t + 1] += im;
  }
}

// Scratch of the allocation-free entry point cvchip_field_contract_twopoint_noalloc:
// the per-block partial sums of k_twopoint (two doubles per block), at most
// kTwoPtScratch doubles (2 MiB, 131072 blocks: T * ceil(LX LY LZ / 256) of them).
// Each GPU holds its own copy until the process ends; its device address is
// taken once per device when a context is created there.
constexpr size_t kTwoPtScratch = 262144;
static __device__ __align__(256) double scratch[kTwoPtScratch];
constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mutex;

int twopoint_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d beyond the %d devices the library supports", device, kMaxDevices);
    return CVCHIP_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return 0;
  double *ptr = nullptr;
  CVC_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = ptr;
  return 0;
}

namespace {

int twopoint_params(cvchip_ctx *c, int idsource, int idsink, const Field &chi, const Field &phi, int n_c,
                    const int *snk_mom, const int src[4], TwoPt &P, size_t &nblk) {
  if (idsource < 0 || idsource > 15 || idsink < 0 || idsink > 15 || n_c < 1 || n_c > 3 ||
      chi.nrhs != 4 * n_c || phi.nrhs != 4 * n_c || chi.prec != CVCHIP_PREC_F64 || phi.prec != CVCHIP_PREC_F64 ||
      chi.V != c->lat.V || phi.V != c->lat.V) {
    set_error("two-point function: gammas 0..15, n_c 1..3, FP64 sets of 4 n_c propagator columns on the "
              "context's lattice");
    return CVCHIP_ERR_ARG;
  }
  const Lattice &L = c->lat;
  UMat g[16];
  gamma_mats(g);
  // gamma5 Gamma_snk
  const UMat M = umul(g[5], g[idsink]);
  for (int a = 0; a < 4; a++)
    for (int col = 0; col < 3; col++) {
      const int k = 3 * a + col;
      P.q[k] = 3 * M.col[a] + col;
      P.ua[k] = M.re[a];
      P.ub[k] = M.im[a];
    }
  // source: column map, sign (the gamma5 sign of the source spin included), imaginary flag
  const UMat &S = g[idsource];
  P.isimag = S.im[0] != 0;
  for (int m = 0; m < 4; m++) {
    P.psource[m] = S.col[m];
    const double s = P.isimag ? -S.im[m] : S.re[m];
    P.ssource[m] = s * (S.col[m] < 2 ? 1. : -1.);
  }
  P.n_c = n_c;
  P.N = 4 * n_c;
  P.px = snk_mom ? 2. * M_PI * snk_mom[0] / L.LX : 0.;
  P.py = snk_mom ? 2. * M_PI * snk_mom[1] / L.LY : 0.;
  P.pz = snk_mom ? 2. * M_PI * snk_mom[2] / L.LZ : 0.;
  for (int m = 0; m < 4; m++) P.src[m] = src[m];
  P.T = L.T;
  P.LX = L.LX;
  P.LY = L.LY;
  P.LZ = L.LZ;
  P.VN = (uint32_t)(L.V * P.N);
  P.vol3 = (uint32_t)(L.LX * L.LY * L.LZ);
  P.nchunk = (int)((P.vol3 + 255) / 256);
  nblk = (size_t)L.T * P.nchunk;
  return 0;
}

int twopoint_fields(cvchip_ctx *c, double *contr, int idsource, int idsink, const Field &chi, const Field &phi,
                    int n_c, const int *snk_mom, const int src[4]) {
  TwoPt P;
  size_t nblk = 0;
  CVC_TRY(twopoint_params(c, idsource, idsink, chi, phi, n_c, snk_mom, src, P, nblk));
  const Lattice &L = c->lat;
  CVC_TRY(ensure_red2(c, sizeof(double) * 2 * nblk));
  double *part = (double *)c->red2.ptr;
  hipLaunchKernelGGL(k_twopoint, dim3((unsigned)nblk), dim3(256), 0, c->stream, (const double2 *)chi.data,
                     (const double2 *)phi.data, P, part);
  CVC_HIP(hipGetLastError());
  std::vector<double> h(2 * nblk);
  CVC_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, c->stream));
  CVC_HIP(hipStreamSynchronize(c->stream));
  for (int t = 0; t < L.T; t++) {
    double re = 0., im = 0.;
    for (int k = 0; k < P.nchunk; k++) {
      re += h[2 * ((size_t)t * P.nchunk + k)];
      im += h[2 * ((size_t)t * P.nchunk + k) + 1];
    }
    contr[2 * t] += re;
    contr[2 * t + 1] += im;
  }
  return 0;
}

}  // namespace
}  // namespace cvchip

using namespace cvchip;

extern "C" int cvchip_gamma_tables(int perm[16][24], int sign[16][24]) {
  return cvchip::guarded("cvchip_gamma_tables", [&]() -> int {
    if (!perm || !sign) return CVCHIP_ERR_NULL;
    UMat g[16];
    gamma_mats(g);
    for (int n = 0; n < 16; n++)
      for (int a = 0; a < 4; a++)
        for (int col = 0; col < 3; col++) {
          const int kr = 6 * a + 2 * col, kb = 6 * g[n].col[a] + 2 * col;
          if (g[n].im[a] == 0) {   // +-1: re <- +-re, im <- +-im
            perm[n][kr] = kb, sign[n][kr] = g[n].re[a];
            perm[n][kr + 1] = kb + 1, sign[n][kr + 1] = g[n].re[a];
          } else {                 // +-i: re <- -+im, im <- +-re
            perm[n][kr] = kb + 1, sign[n][kr] = -g[n].im[a];
            perm[n][kr + 1] = kb, sign[n][kr + 1] = g[n].im[a];
          }
        }
    return CVCHIP_OK;
  });
}

extern "C" int cvchip_field_contract_twopoint(cvchip_ctx *c, double *contr, int idsource, int idsink,
                                              const cvchip_field *chi, const cvchip_field *phi, int n_c,
                                              const int snk_mom[3], const int src[4]) {
  return cvchip::guarded("cvchip_field_contract_twopoint", [&]() -> int {
    if (!c || !contr || !chi || !phi || !src) {
      set_error("NULL argument to cvchip_field_contract_twopoint");
      return CVCHIP_ERR_NULL;
    }
    return twopoint_fields(c, contr, idsource, idsink, chi->f, phi->f, n_c, snk_mom, src);
  });
}

extern "C" int cvchip_contract_twopoint(cvchip_ctx *c, double *contr, int idsource, int idsink,
                                        const double *const *chi, const double *const *phi, int n_c,
                                        const int snk_mom[3], const int src[4]) {
  return cvchip::guarded("cvchip_contract_twopoint", [&]() -> int {
    if (!c || !contr || !chi || !phi || !src) {
      set_error("NULL argument to cvchip_contract_twopoint");
      return CVCHIP_ERR_NULL;
    }
    if (n_c < 1 || n_c > 3) {
      set_error("n_c must be 1..3, got %d", n_c);
      return CVCHIP_ERR_ARG;
    }
    Field fc, fp;
    const int N = 4 * n_c;
    int rc = field_alloc(c, fc, N, CVCHIP_PREC_F64);
    if (!rc) rc = field_alloc(c, fp, N, CVCHIP_PREC_F64);
    for (int k = 0; k < N && !rc; k++) {
      if (!chi[k] || !phi[k]) {
        set_error("NULL propagator column %d", k);
        rc = CVCHIP_ERR_NULL;
        break;
      }
      rc = upload_aos(c, fc, chi[k], k, 1);
      if (!rc) rc = upload_aos(c, fp, phi[k], k, 1);
    }
    if (!rc) rc = twopoint_fields(c, contr, idsource, idsink, fc, fp, n_c, snk_mom, src);
    field_free(fc);
    field_free(fp);
    return rc;
  });
}

// No counterpart in the reference.  cvchip_field_contract_twopoint for callers
// that cannot allocate: the block partials go to the file-scope scratch array
// (through its per-device address), the per-slice sums are formed on the device
// and added to contr_dev, device memory of 2 T doubles that the caller owns.
// Calls no hipMalloc, hipFree or allocating routine.  Calls that overlap on one
// GPU share its scratch and race.
extern "C" int cvchip_field_contract_twopoint_noalloc(cvchip_ctx *c, double *contr_dev, int idsource, int idsink,
                                                      const cvchip_field *chi, const cvchip_field *phi, int n_c,
                                                      const int snk_mom[3], const int src[4]) {
  return cvchip::guarded("cvchip_field_contract_twopoint_noalloc", [&]() -> int {
    if (!c || !contr_dev || !chi || !phi || !src) {
      set_error("NULL argument to cvchip_field_contract_twopoint_noalloc");
      return CVCHIP_ERR_NULL;
    }
    const Lattice &L = c->lat;
    const size_t need = 2 * (size_t)L.T * (size_t)((L.LX * L.LY * L.LZ + 255) / 256);
    if (need > kTwoPtScratch) {
      set_error("two-point function without allocation: %zu partial sums exceed the scratch of %zu "
                "(T * ceil(LX LY LZ / 256) must stay at or below %zu)", need, kTwoPtScratch, kTwoPtScratch / 2);
      return CVCHIP_ERR_ARG;
    }
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device]) {
      set_error("two-point scratch of device %d not set up (create the context on that device)", c->device);
      return CVCHIP_ERR_STATE;
    }
    double *part = g_scratch_ptr[c->device];
    TwoPt P;
    size_t nblk = 0;
    CVC_TRY(twopoint_params(c, idsource, idsink, chi->f, phi->f, n_c, snk_mom, src, P, nblk));
    hipLaunchKernelGGL(k_twopoint, dim3((unsigned)nblk), dim3(256), 0, c->stream, (const double2 *)chi->f.data,
                       (const double2 *)phi->f.data, P, part);
    hipLaunchKernelGGL(k_twopoint_final, dim3(1), dim3(256), 0, c->stream, (const double *)part, contr_dev, L.T,
                       P.nchunk);
    CVC_HIP(hipGetLastError());
    CVC_HIP(hipStreamSynchronize(c->stream));
    return CVCHIP_OK;
  });
}
