// Fixed-shape batched pass kernels (FP32, part 1 of 3): the launch configurations
// of the batched size classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt.
#include "chz_axis_kernels.h"

namespace chz {

#define BX(T, PREC, OP, N, EPT, TPL, L, PLAN) \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_batch_fixed<T, EPT, OP, N, TPL, L, PLAN>}

static const FixedEntry kTab[] = {
    BX(float, 32, B_XUPD, 32, 5, 8, 8, 0x404ull),
    BX(float, 32, B_XUPD, 96, 12, 4, 32, 0x40cull),
    BX(float, 32, B_COLS, 48, 8, 4, 64, 0x308ull),
    BX(float, 32, B_COLS, 128, 8, 8, 16, 0x808ull),
    BX(float, 32, B_XFWD, 48, 6, 8, 8, 0x406ull),
    BX(float, 32, B_XRES, 64, 8, 4, 16, 0x408ull),
    BX(float, 32, B_XFWD, 128, 8, 8, 8, 0x808ull),
    BX(float, 32, B_ZMID, 48, 8, 4, 16, 0x308ull),
    BX(float, 32, B_ZMID, 128, 8, 8, 16, 0x808ull),
};

const FixedEntry *batch_fixed_table_32_0(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
