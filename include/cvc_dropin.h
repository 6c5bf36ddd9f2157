/* Entry points of the drop-in shim (cvc_amd/dropin/cvc_dropin.o) beyond the
 * reference's own symbols.  A reference program links the shim and
 * libcvchip.so in place of Q_phi.o / invert_Qtm.o; the shim defines the
 * reference's functions with their C++ names in namespace cvc (Q_phi_tbc,
 * Q_phi, Qf5, g5_phi, spinor_scalar_product_co/_re, invert_Qtm,
 * invert_Qtm_her) and, with C linkage, the functions declared here. */
#ifndef CVC_DROPIN_H
#define CVC_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

/* No counterpart in the reference: its Q_phi_tbc (Q_phi.c:38-188) reads
 * g_gauge_field on every call.  The shim keeps the gauge field on the GPU and
 * uploads it again when the pointer g_gauge_field or a hash of all its words
 * changed since the last call; a program that edits g_gauge_field in place
 * (e.g. apply_gt_gauge, cvc_utils.c:2044-2068) may call this to force the
 * upload on the next operator or solver call. */
void cvchip_gauge_changed(void);

#ifdef __cplusplus
}
#endif
#endif
