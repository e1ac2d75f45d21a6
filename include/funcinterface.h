/* funcinterface.h -- C++ drop-in for the reference's cudasymmfft/funcinterface.h.
 *
 * Same names, argument order and meaning (C++ linkage, so the mangled names
 * match code built against the reference header).  Arrays are device
 * pointers to padded (NX+2)*(NY+2)*(NZ+2) FP64 volumes, x fastest.
 *
 * Behaviour vs the reference (cudasymmfft/funcinterface.cu):
 *  - do_transform: in place when d_out == NULL or d_out == d_data, which
 *    requires NX == NY == NZ (:26-33); otherwise out of place with the result
 *    in d_out.  d_data is left unchanged (the reference clobbers it).
 *  - every call synchronises the device before returning (the reference
 *    syncs after each axis, :111-116).
 *  - errors are printed like the reference's messages but never exit();
 *    the call returns and chz_last_error() (include/cuhelmholtz.h) has the
 *    reason.
 *  - freeMemory_cubic: plans and twiddle tables are cached per device, so
 *    there is nothing to destroy per call; chz_release() frees the caches.
 */
#ifndef __FUNC_INTERFACE__
#define __FUNC_INTERFACE__
#include "global.h"

/* replaces cudasymmfft/funcinterface.h:5-12 */
void do_transform(double *d_data,
                  double *d_out,
                  int NX,
                  int NY,
                  int NZ,
                  transform_kind X_TRANS,
                  transform_kind Y_TRANS,
                  transform_kind Z_TRANS);

/* replaces cudasymmfft/funcinterface.h:14-21 (out of place, any even sizes) */
void do_transform_nocubic(double *d_data,
                          double *d_out,
                          int NX,
                          int NY,
                          int NZ,
                          transform_kind X_TRANS,
                          transform_kind Y_TRANS,
                          transform_kind Z_TRANS);

/* replaces cudasymmfft/funcinterface.h:23-29 (in place) */
void do_transform_cubic(double *d_data,
                        int NX,
                        int NY,
                        int NZ,
                        transform_kind X_TRANS,
                        transform_kind Y_TRANS,
                        transform_kind Z_TRANS);

/* replaces cudasymmfft/funcinterface.h:32 */
void freeMemory_cubic();

#endif /* __FUNC_INTERFACE__ */
