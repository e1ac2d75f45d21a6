// FP64 fixed-shape pass kernels for N = 512 with lean tile I/O (chz_line.h):
// 8 points per thread, 32 threads per line, 8 or 16 lines per tile, radix
// plan 8 x 8 x 4 (the tuned shapes: x passes 8 lines, column and fused z
// passes 16 lines = whole 128-byte lines per slot on 512 threads); and, as
// the measured alternative for the column passes, 16 points per thread, 16
// threads per line, 16 lines, radix plan 16 x 16.
#include "chz_line.h"
namespace chz {

constexpr unsigned long long L256_884 = 0x040808ull;
constexpr unsigned long long L256_1616 = 0x1010ull;

#define LN(OP, L) {64, OP, 512, 8, 32, L, L256_884, (const void *)k_line<double, OP, 512, 32, L, L256_884, 8>, \
                   line_lds_extra<double, 512, OP>()}

static const FixedEntry kLine64[] = {
    LN(B_ROWS, 8), LN(B_ROWS, 16), LN(B_COLS, 8), LN(B_COLS, 16),
    LN(B_XFWD, 8), LN(B_XFWD, 16), LN(B_ZMID, 8), LN(B_ZMID, 16),
    {64, B_COLS, 512, 16, 16, 16, L256_1616, (const void *)k_line<double, B_COLS, 512, 16, 16, L256_1616, 16>, 0u},
    {64, B_ZMID, 512, 16, 16, 16, L256_1616, (const void *)k_line<double, B_ZMID, 512, 16, 16, L256_1616, 16>,
     line_lds_extra<double, 512, B_ZMID>()},
};

const FixedEntry *line_table_f64(int *n)
{
    *n = (int)(sizeof(kLine64) / sizeof(kLine64[0]));
    return kLine64;
}

}  // namespace chz
