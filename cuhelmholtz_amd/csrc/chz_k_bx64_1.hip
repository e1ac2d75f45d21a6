// Fixed-shape batched pass kernels (FP64, part 2 of 2): the launch configurations
// of the batched size classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt.
#include "chz_axis_kernels.h"

namespace chz {

#define BX(T, PREC, OP, N, EPT, TPL, L, PLAN) \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_batch_fixed<T, EPT, OP, N, TPL, L, PLAN>}

static const FixedEntry kTab[] = {
    BX(double, 64, B_ROWS, 48, 8, 4, 16, 0x308ull),
    BX(double, 64, B_ROWS, 96, 6, 8, 8, 0x2020206ull),
    BX(double, 64, B_COLS, 32, 4, 8, 8, 0x404ull),
    BX(double, 64, B_COLS, 64, 8, 4, 16, 0x408ull),
    BX(double, 64, B_COLS, 128, 8, 8, 16, 0x808ull),
    BX(double, 64, B_XFWD, 48, 6, 4, 16, 0x20206ull),
    BX(double, 64, B_XFWD, 96, 6, 8, 16, 0x2020206ull),
    BX(double, 64, B_ZMID, 32, 4, 4, 32, 0x404ull),
    BX(double, 64, B_ZMID, 64, 8, 4, 16, 0x408ull),
    BX(double, 64, B_ZMID, 128, 8, 8, 16, 0x808ull),
};

const FixedEntry *batch_fixed_table_64_1(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
