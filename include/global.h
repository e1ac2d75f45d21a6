/* global.h -- drop-in for the reference's cudasymmfft/global.h.
 *
 * Replaces: cudasymmfft/global.h:6-15 (enum transform_kind).  The ordinals
 * are part of the ABI and are kept: DCT_1=0, DCT_2=1, DCT_3=2, DST_1=3,
 * DST_2=4, DST_3=5, DFT_R2C=6, DFT_C2R=7.  The reference's cuFFT result-code
 * string table (:17-33) has no counterpart: no vendor FFT is used.
 */
#ifndef __GLOBAL__
#define __GLOBAL__

enum transform_kind {
    DCT_1,
    DCT_2,
    DCT_3,
    DST_1,
    DST_2,
    DST_3,
    DFT_R2C,
    DFT_C2R
};

#endif /* __GLOBAL__ */
