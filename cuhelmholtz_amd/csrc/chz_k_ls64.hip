// FP64 fixed-shape pass kernels with lean tile I/O for N = 32..128
// (chz_line.h, k_line): single-problem passes at the launch configurations
// of the size classes N = 32, 48, 64, 96, 128 in tuning/gfx950.txt (the
// shapes of the batched kernels, so a batched problem and its single solve
// run the same tile code).
#include "chz_line.h"
namespace chz {

#define LS(T, PREC, OP, N, EPT, TPL, L, PLAN)                                                        \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_line<T, OP, N, TPL, L, PLAN, EPT>,              \
     line_lds_extra<T, N, OP>()}

static const FixedEntry kTab[] = {
    LS(double, 64, B_ROWS, 32, 4, 8, 16, 0x404ull),
    LS(double, 64, B_ROWS, 48, 8, 4, 16, 0x308ull),
    LS(double, 64, B_ROWS, 64, 8, 4, 32, 0x408ull),
    LS(double, 64, B_ROWS, 96, 6, 8, 8, 0x2020206ull),
    LS(double, 64, B_ROWS, 128, 8, 8, 16, 0x808ull),
    LS(double, 64, B_COLS, 32, 4, 8, 8, 0x404ull),
    LS(double, 64, B_COLS, 48, 8, 4, 16, 0x308ull),
    LS(double, 64, B_COLS, 64, 8, 4, 16, 0x408ull),
    LS(double, 64, B_COLS, 96, 6, 8, 16, 0x2020206ull),
    LS(double, 64, B_COLS, 128, 8, 8, 16, 0x808ull),
    LS(double, 64, B_XFWD, 32, 4, 8, 12, 0x404ull),
    LS(double, 64, B_XFWD, 48, 6, 4, 16, 0x20206ull),
    LS(double, 64, B_XFWD, 64, 8, 4, 32, 0x408ull),
    LS(double, 64, B_XFWD, 96, 6, 8, 16, 0x2020206ull),
    LS(double, 64, B_XFWD, 128, 8, 8, 16, 0x808ull),
    LS(double, 64, B_ZMID, 32, 4, 4, 32, 0x404ull),
    LS(double, 64, B_ZMID, 48, 8, 4, 32, 0x308ull),
    LS(double, 64, B_ZMID, 64, 8, 4, 32, 0x408ull),
    LS(double, 64, B_ZMID, 96, 6, 8, 8, 0x2020206ull),
    LS(double, 64, B_ZMID, 128, 8, 8, 16, 0x808ull)
};

const FixedEntry *line_small_table_f64(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
