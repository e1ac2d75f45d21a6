/* cuhelmholtz.h -- C ABI of the MI355X (gfx950) transform and Helmholtz library.
 *
 * Plain C: pointers, sizes and status codes only.  Every array argument is a
 * device pointer owned by the caller.  Calls are stream-ordered on `stream`
 * (a hipStream_t passed as void*, NULL = the default stream) and return
 * before the device work completes; synchronise the stream to read results.
 *
 * Status: 0 = OK, -1 = invalid argument, -2 = unsupported size,
 * -3 = device / launch error, -4 = device allocation failure.
 * chz_last_error() returns the message of the last failure on this thread.
 *
 * Supported sizes: even N with N/2 = 2^a 3^b, N <= 4096 per axis.
 */
#ifndef CUHELMHOLTZ_H
#define CUHELMHOLTZ_H

#ifdef __cplusplus
extern "C" {
#endif

/* Transform kind ordinals, identical to enum transform_kind
 * (reference cudasymmfft/global.h:6-15). */
#define CHZ_DCT_1 0
#define CHZ_DCT_2 1
#define CHZ_DCT_3 2
#define CHZ_DST_1 3
#define CHZ_DST_2 4
#define CHZ_DST_3 5
#define CHZ_DFT_R2C 6
#define CHZ_DFT_C2R 7

/* 3D separable transform of a padded (NX+2)(NY+2)(NZ+2) FP64 volume (x
 * fastest), X then Y then Z.  d_out == NULL or d_out == d_data: in place (any
 * shape); otherwise the result is written to d_out and d_data is unchanged.
 * Replaces do_transform / do_transform_cubic / do_transform_nocubic
 * (reference cudasymmfft/funcinterface.cu:25-40, :42-198, :200-334). */
int chz_do_transform(double *d_data, double *d_out, int NX, int NY, int NZ,
                     int x_kind, int y_kind, int z_kind, void *stream);

/* Same in FP32 (twiddles generated in FP64, rounded).  No reference
 * counterpart: the reference is FP64 only; used by the mixed-precision solve. */
int chz_do_transform_f32(float *d_data, float *d_out, int NX, int NY, int NZ,
                         int x_kind, int y_kind, int z_kind, void *stream);

/* Message of the last failed call on this thread ("" if none). */
const char *chz_last_error(void);

/* Free cached twiddle tables and workspaces of the current device.
 * Replaces freeMemory_cubic (reference cudasymmfft/funcinterface.cu:337-342). */
void chz_release(void);

/* Launch configuration chosen for an axis of N cells: elements per thread,
 * threads per line, lines per workgroup and the radix plan (one byte per
 * stage).  prec = 64 or 32; rows = 1 for the contiguous (x) axis. */
int chz_axis_config(int N, int prec, int rows, int *ept, int *tpl, int *lines,
                    unsigned long long *plan);

int chz_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CUHELMHOLTZ_H */
