// FP32 fixed-shape pass kernels for N = 512 with lean tile I/O (chz_line.h):
// the passes of the mixed-precision solve, including the FP64 residual and
// update passes (8 points per thread, 32 threads per line, 8 or 16 lines per
// tile, radix plan 8 x 8 x 4).
#include "chz_line.h"
namespace chz {

constexpr unsigned long long L256_884_32 = 0x040808ull;

#define LN32(OP, L) {32, OP, 512, 8, 32, L, L256_884_32, (const void *)k_line<float, OP, 512, 32, L, L256_884_32, 8>, line_lds_extra<float, 512, OP>()}

static const FixedEntry kLine32[] = {
    LN32(B_ROWS, 8), LN32(B_ROWS, 16), LN32(B_COLS, 8), LN32(B_COLS, 16),
    LN32(B_XFWD, 8), LN32(B_XFWD, 16), LN32(B_ZMID, 8), LN32(B_ZMID, 16),
    LN32(B_XRES, 8), LN32(B_XRES, 16), LN32(B_XUPD, 8), LN32(B_XUPD, 16),
};

const FixedEntry *line_table_f32(int *n)
{
    *n = (int)(sizeof(kLine32) / sizeof(kLine32[0]));
    return kLine32;
}

}  // namespace chz
