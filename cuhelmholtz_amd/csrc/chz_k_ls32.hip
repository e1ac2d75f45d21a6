// FP32 fixed-shape pass kernels with lean tile I/O for N = 32..128
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
    LS(float, 32, B_XUPD, 32, 5, 8, 8, 0x404ull),
    LS(float, 32, B_XUPD, 48, 8, 4, 16, 0x308ull),
    LS(float, 32, B_XUPD, 64, 8, 4, 64, 0x408ull),
    LS(float, 32, B_XUPD, 96, 12, 4, 32, 0x40cull),
    LS(float, 32, B_XUPD, 128, 8, 8, 32, 0x808ull),
    LS(float, 32, B_COLS, 32, 4, 4, 32, 0x404ull),
    LS(float, 32, B_COLS, 48, 8, 4, 64, 0x308ull),
    LS(float, 32, B_COLS, 64, 8, 4, 16, 0x408ull),
    LS(float, 32, B_COLS, 96, 12, 4, 16, 0x40cull),
    LS(float, 32, B_COLS, 128, 8, 8, 16, 0x808ull),
    LS(float, 32, B_XFWD, 32, 4, 8, 12, 0x404ull),
    LS(float, 32, B_XFWD, 48, 6, 8, 8, 0x406ull),
    LS(float, 32, B_XFWD, 64, 8, 4, 16, 0x408ull),
    LS(float, 32, B_XFWD, 96, 6, 8, 8, 0x2020206ull),
    LS(float, 32, B_XFWD, 128, 8, 8, 8, 0x808ull),
    LS(float, 32, B_XRES, 32, 4, 8, 12, 0x404ull),
    LS(float, 32, B_XRES, 48, 6, 8, 8, 0x406ull),
    LS(float, 32, B_XRES, 64, 8, 4, 16, 0x408ull),
    LS(float, 32, B_XRES, 96, 6, 8, 8, 0x2020206ull),
    LS(float, 32, B_XRES, 128, 8, 8, 8, 0x808ull),
    LS(float, 32, B_ZMID, 32, 4, 4, 32, 0x404ull),
    LS(float, 32, B_ZMID, 48, 8, 4, 16, 0x308ull),
    LS(float, 32, B_ZMID, 64, 8, 4, 16, 0x408ull),
    LS(float, 32, B_ZMID, 96, 12, 4, 16, 0x40cull),
    LS(float, 32, B_ZMID, 128, 8, 8, 16, 0x808ull)
};

const FixedEntry *line_small_table_f32(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
