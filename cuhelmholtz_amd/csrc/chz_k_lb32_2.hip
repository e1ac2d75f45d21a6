// FP32 batched pass kernels with lean tile I/O (part 3 of 3)
// (chz_line.h, k_line_batch): the launch configurations of the batched size
// classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt with an even number of
// threads per line.
#include "chz_line.h"
namespace chz {

#define LB(T, PREC, OP, N, EPT, TPL, L, PLAN)                                                        \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_line_batch<T, OP, N, TPL, L, PLAN, EPT>,        \
     line_lds_extra<T, N, OP>()}

static const FixedEntry kTab[] = {
    LB(float, 32, B_XRES, 64, 8, 4, 16, 0x408ull),
    LB(float, 32, B_XRES, 96, 6, 8, 8, 0x2020206ull),
    LB(float, 32, B_XRES, 128, 8, 8, 8, 0x808ull),
    LB(float, 32, B_ZMID, 32, 4, 4, 32, 0x404ull),
    LB(float, 32, B_ZMID, 48, 8, 4, 16, 0x308ull),
    LB(float, 32, B_ZMID, 64, 8, 4, 16, 0x408ull),
    LB(float, 32, B_ZMID, 96, 12, 4, 16, 0x40cull),
    LB(float, 32, B_ZMID, 128, 8, 8, 16, 0x808ull)
};

const FixedEntry *batch_line_table_32_2(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
