// Conserved-vector-current contraction, Ward identity, Pi(q) and the
// whole-source pipeline.
#include "spinor_io.h"

using namespace cvchip;

extern "C" int cvchip_cvc_contract(cvchip_ctx *c, double *conn, double contact[8], const double *const *props,
                                   const int src[4]) {
  (void)c; (void)conn; (void)contact; (void)props; (void)src;
  set_error("cvchip_cvc_contract: not built yet");
  return CVCHIP_ERR_STATE;
}

extern "C" int cvchip_field_cvc_contract(cvchip_ctx *c, double *conn, double contact[8],
                                         const cvchip_field *up, const cvchip_field *dn, const int src[4]) {
  (void)c; (void)conn; (void)contact; (void)up; (void)dn; (void)src;
  set_error("cvchip_field_cvc_contract: not built yet");
  return CVCHIP_ERR_STATE;
}

extern "C" int cvchip_wi(cvchip_ctx *c, double *wi, const double *conn) {
  (void)c; (void)wi; (void)conn;
  set_error("cvchip_wi: not built yet");
  return CVCHIP_ERR_STATE;
}

extern "C" int cvchip_pi_q(cvchip_ctx *c, double *piq, const double *conn, const double contact[8],
                           const int src[4]) {
  (void)c; (void)piq; (void)conn; (void)contact; (void)src;
  set_error("cvchip_pi_q: not built yet");
  return CVCHIP_ERR_STATE;
}

extern "C" int cvchip_cvc_source(cvchip_ctx *c, double *conn, double contact[8], const int src[4], double sp,
                                 int niter_max, int prec, double delta, cvchip_stats *stats) {
  (void)c; (void)conn; (void)contact; (void)src; (void)sp; (void)niter_max; (void)prec; (void)delta;
  (void)stats;
  set_error("cvchip_cvc_source: not built yet");
  return CVCHIP_ERR_STATE;
}
