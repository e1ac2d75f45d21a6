// FP32 batched pass kernels with lean tile I/O (part 1 of 3)
// (chz_line.h, k_line_batch): the launch configurations of the batched size
// classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt with an even number of
// threads per line.
#include "chz_line.h"
namespace chz {

#define LB(T, PREC, OP, N, EPT, TPL, L, PLAN)                                                        \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_line_batch<T, OP, N, TPL, L, PLAN, EPT>,        \
     line_lds_extra<T, N, OP>()}

static const FixedEntry kTab[] = {
    LB(float, 32, B_XUPD, 32, 5, 8, 8, 0x404ull),
    LB(float, 32, B_XUPD, 48, 8, 4, 16, 0x308ull),
    LB(float, 32, B_XUPD, 64, 8, 4, 64, 0x408ull),
    LB(float, 32, B_XUPD, 96, 12, 4, 32, 0x40cull),
    LB(float, 32, B_XUPD, 128, 8, 8, 32, 0x808ull),
    LB(float, 32, B_COLS, 32, 4, 4, 32, 0x404ull),
    LB(float, 32, B_COLS, 48, 8, 4, 64, 0x308ull),
    LB(float, 32, B_COLS, 64, 8, 4, 16, 0x408ull),
    LB(float, 32, B_COLS, 96, 12, 4, 16, 0x40cull)
};

const FixedEntry *batch_line_table_32_0(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
