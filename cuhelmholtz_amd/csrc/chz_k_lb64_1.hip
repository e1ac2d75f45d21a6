// FP64 batched pass kernels with lean tile I/O (part 2 of 2)
// (chz_line.h, k_line_batch): the launch configurations of the batched size
// classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt with an even number of
// threads per line.
#include "chz_line.h"
namespace chz {

#define LB(T, PREC, OP, N, EPT, TPL, L, PLAN)                                                        \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_line_batch<T, OP, N, TPL, L, PLAN, EPT>,        \
     line_lds_extra<T, N, OP>()}

static const FixedEntry kTab[] = {
    LB(double, 64, B_XFWD, 32, 4, 8, 12, 0x404ull),
    LB(double, 64, B_XFWD, 48, 6, 4, 16, 0x20206ull),
    LB(double, 64, B_XFWD, 64, 8, 4, 32, 0x408ull),
    LB(double, 64, B_XFWD, 96, 6, 8, 16, 0x2020206ull),
    LB(double, 64, B_XFWD, 128, 8, 8, 16, 0x808ull),
    LB(double, 64, B_ZMID, 32, 4, 4, 32, 0x404ull),
    LB(double, 64, B_ZMID, 48, 8, 4, 32, 0x308ull),
    LB(double, 64, B_ZMID, 64, 8, 4, 32, 0x408ull),
    LB(double, 64, B_ZMID, 96, 6, 8, 8, 0x2020206ull),
    LB(double, 64, B_ZMID, 128, 8, 8, 16, 0x808ull)
};

const FixedEntry *batch_line_table_64_1(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
