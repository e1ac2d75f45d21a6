// FP32 fixed-shape pass kernels for N = 512 with lean tile I/O (chz_line.h):
// the passes of the mixed-precision solve, including the FP64 residual and
// update passes (8 points per thread, 32 threads per line, 8 or 16 lines per
// tile, radix plan 8 x 8 x 4), and fused z passes with 16 points per thread.
#include "chz_line.h"
namespace chz {

constexpr unsigned long long L256_884_32 = 0x040808ull;

#define LN32(OP, L) {32, OP, 512, 8, 32, L, L256_884_32, (const void *)k_line<float, OP, 512, 32, L, L256_884_32, 8>, line_lds_extra<float, 512, OP>()}

static const FixedEntry kLine32[] = {
    LN32(B_ROWS, 8), LN32(B_ROWS, 16), LN32(B_COLS, 8), LN32(B_COLS, 16),
    LN32(B_XFWD, 8), LN32(B_XFWD, 16), LN32(B_ZMID, 8), LN32(B_ZMID, 16),
    LN32(B_XRES, 8), LN32(B_XRES, 16), LN32(B_XUPD, 8), LN32(B_XUPD, 16),
    // fused z pass: 16 points per thread, 16 threads per line, radix plan 16 x 16,
    // 32 lines (whole 128-byte lines per z slot): 0.39 -> 0.30 ms at 512^3
    {32, B_COLS, 512, 16, 16, 32, 0x1010ull, (const void *)k_line<float, B_COLS, 512, 16, 32, 0x1010ull, 16>, 0u},
    {32, B_ZMID, 512, 16, 16, 32, 0x1010ull, (const void *)k_line<float, B_ZMID, 512, 16, 32, 0x1010ull, 16>,
     line_lds_extra<float, 512, B_ZMID>()},
};

const FixedEntry *line_table_f32(int *n)
{
    *n = (int)(sizeof(kLine32) / sizeof(kLine32[0]));
    return kLine32;
}

}  // namespace chz
