// Fixed-shape batched pass kernels (FP64, part 1 of 2): the launch configurations
// of the batched size classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt.
#include "chz_axis_kernels.h"

namespace chz {

#define BX(T, PREC, OP, N, EPT, TPL, L, PLAN) \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_batch_fixed<T, EPT, OP, N, TPL, L, PLAN>}

static const FixedEntry kTab[] = {
    BX(double, 64, B_ROWS, 32, 4, 8, 16, 0x404ull),
    BX(double, 64, B_ROWS, 64, 8, 4, 32, 0x408ull),
    BX(double, 64, B_ROWS, 128, 8, 8, 16, 0x808ull),
    BX(double, 64, B_COLS, 48, 8, 4, 16, 0x308ull),
    BX(double, 64, B_COLS, 96, 6, 8, 16, 0x2020206ull),
    BX(double, 64, B_XFWD, 32, 4, 8, 12, 0x404ull),
    BX(double, 64, B_XFWD, 64, 8, 4, 32, 0x408ull),
    BX(double, 64, B_XFWD, 128, 8, 9, 14, 0x808ull),
    BX(double, 64, B_ZMID, 48, 8, 4, 16, 0x308ull),
    BX(double, 64, B_ZMID, 96, 6, 8, 8, 0x2020206ull),
};

const FixedEntry *batch_fixed_table_64_0(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
