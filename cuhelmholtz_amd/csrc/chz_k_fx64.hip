// FP64 fixed-shape pass kernels for N = 512 (elements/thread 8, 32 threads per
// line, 8 or 16 lines per tile): every role of the solve and the transform.
// (16 points per thread with radix 16 x 16 measured slower in every role.)
#include <cstdlib>

#include "chz_axis_kernels.h"
namespace chz {

// radix plan 8 x 8 x 4 of the 256-point FFT (one byte per stage, first stage low)
constexpr unsigned long long P256_884 = 0x040808ull;

#define FX(T, PREC, OP, N, EPT, TPL, L, PLAN) {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_fixed<T, EPT, OP, N, TPL, L, PLAN>}

static const FixedEntry kF64[] = {
    FX(double, 64, B_ROWS, 512, 8, 32, 8, P256_884),
    FX(double, 64, B_ROWS, 512, 8, 32, 16, P256_884),
    FX(double, 64, B_COLS, 512, 8, 32, 8, P256_884),
    FX(double, 64, B_COLS, 512, 8, 32, 16, P256_884),
    FX(double, 64, B_XFWD, 512, 8, 32, 8, P256_884),
    FX(double, 64, B_XFWD, 512, 8, 32, 16, P256_884),
    FX(double, 64, B_ZMID, 512, 8, 32, 8, P256_884),
    FX(double, 64, B_ZMID, 512, 8, 32, 16, P256_884),
};

const FixedEntry *fixed_table_f64(int *n)
{
    *n = (int)(sizeof(kF64) / sizeof(kF64[0]));
    return kF64;
}

}  // namespace chz
