// FP32 fixed-shape pass kernels with lean tile I/O for N = 256, 384
// (chz_line.h, k_line): N = 256 and 384 (the 256^3 and 512x384x256 workloads),
// the shapes tuning/gfx950.txt selects and the alternatives measured against them.
#include "chz_line.h"
namespace chz {

#define LM(T, PREC, OP, N, EPT, TPL, L, PLAN)                                                        \
    {PREC, OP, N, EPT, TPL, L, PLAN, (const void *)k_line<T, OP, N, TPL, L, PLAN, EPT>,              \
     line_lds_extra<T, N, OP>()}

static const FixedEntry kTab[] = {
    LM(float, 32, B_XFWD, 256, 8, 32, 8, 0x20808ull),
    LM(float, 32, B_XRES, 256, 8, 32, 8, 0x20808ull),
    LM(float, 32, B_XUPD, 256, 8, 32, 8, 0x20808ull),
    LM(float, 32, B_COLS, 256, 8, 16, 32, 0x20808ull),
    LM(float, 32, B_COLS, 256, 16, 8, 32, 0x810ull),
    LM(float, 32, B_ZMID, 256, 8, 16, 32, 0x20808ull),
    LM(float, 32, B_ZMID, 256, 16, 8, 32, 0x810ull),
    LM(float, 32, B_XFWD, 384, 8, 32, 8, 0x30808ull),
    LM(float, 32, B_XRES, 384, 8, 32, 8, 0x30808ull),
    LM(float, 32, B_XUPD, 384, 8, 32, 8, 0x30808ull),
    LM(float, 32, B_COLS, 384, 8, 32, 16, 0x30808ull),
    LM(float, 32, B_COLS, 384, 16, 16, 32, 0xc10ull),
    LM(float, 32, B_ZMID, 384, 8, 32, 16, 0x30808ull),
    LM(float, 32, B_ZMID, 384, 16, 16, 32, 0xc10ull)
};

const FixedEntry *line_mid_table_f32(int *n)
{
    *n = (int)(sizeof(kTab) / sizeof(kTab[0]));
    return kTab;
}

}  // namespace chz
