// FP64 batched pass kernels with lean tile I/O (part 1 of 2)
// (chz_line.h, k_line_batch): the launch configurations of the batched size
// classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt with an even number of
// threads per line.
#include "chz_line.h"
namespace chz {

#define LB(T, PREC, OP, N, EPT, TPL, L, PLAN)                                                        \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_line_batch<T, OP, N, TPL, L, PLAN, EPT>,        \
     line_lds_extra<T, N, OP>()}

static const FixedEntry kTab[] = {
    LB(double, 64, B_ROWS, 32, 4, 8, 16, 0x404ull),
    LB(double, 64, B_ROWS, 48, 8, 4, 16, 0x308ull),
    LB(double, 64, B_ROWS, 64, 8, 4, 32, 0x408ull),
    LB(double, 64, B_ROWS, 96, 6, 8, 8, 0x2020206ull),
    LB(double, 64, B_ROWS, 128, 8, 8, 16, 0x808ull),
    LB(double, 64, B_COLS, 32, 4, 8, 8, 0x404ull),
    LB(double, 64, B_COLS, 48, 8, 4, 16, 0x308ull),
    LB(double, 64, B_COLS, 64, 8, 4, 16, 0x408ull),
    LB(double, 64, B_COLS, 96, 6, 8, 16, 0x2020206ull),
    LB(double, 64, B_COLS, 128, 8, 8, 16, 0x808ull)
};

const FixedEntry *batch_line_table_64_0(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
