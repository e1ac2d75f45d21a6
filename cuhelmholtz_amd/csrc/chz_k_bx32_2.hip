// Fixed-shape batched pass kernels (FP32, part 3 of 3): the launch configurations
// of the batched size classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt.
#include "chz_axis_kernels.h"

namespace chz {

#define BX(T, PREC, OP, N, EPT, TPL, L, PLAN) \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_batch_fixed<T, EPT, OP, N, TPL, L, PLAN>}

static const FixedEntry kTab[] = {
    BX(float, 32, B_XUPD, 64, 8, 4, 64, 0x408ull),
    BX(float, 32, B_COLS, 32, 4, 5, 24, 0x404ull),
    BX(float, 32, B_COLS, 96, 12, 4, 16, 0x40cull),
    BX(float, 32, B_XRES, 32, 4, 8, 12, 0x404ull),
    BX(float, 32, B_XFWD, 64, 8, 4, 16, 0x408ull),
    BX(float, 32, B_XRES, 96, 6, 8, 8, 0x2020206ull),
    BX(float, 32, B_ZMID, 32, 4, 4, 32, 0x404ull),
    BX(float, 32, B_ZMID, 96, 12, 4, 16, 0x40cull),
};

const FixedEntry *batch_fixed_table_32_2(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
