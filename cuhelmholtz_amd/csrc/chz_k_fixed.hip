// Lookup of the fixed-shape pass kernels (tables in chz_k_fx64.hip, chz_k_fx32.hip).
#include <cstdlib>

#include "chz_kernels.h"

namespace chz {

const FixedEntry *fixed_table_f64(int *n);
const FixedEntry *line_table_f64(int *n);
const FixedEntry *line_table_f32(int *n);
const FixedEntry *line_small_table_f64(int *n);
const FixedEntry *line_small_table_f32(int *n);
const FixedEntry *line_mid_table_f64(int *n);
const FixedEntry *line_mid_table_f32(int *n);
const FixedEntry *fixed_table_f32(int *n);
const FixedEntry *batch_fixed_table_64_0(int *n);
const FixedEntry *batch_fixed_table_64_1(int *n);
const FixedEntry *batch_fixed_table_32_0(int *n);
const FixedEntry *batch_fixed_table_32_1(int *n);
const FixedEntry *batch_fixed_table_32_2(int *n);
const FixedEntry *batch_line_table_64_0(int *n);
const FixedEntry *batch_line_table_64_1(int *n);
const FixedEntry *batch_line_table_32_0(int *n);
const FixedEntry *batch_line_table_32_1(int *n);
const FixedEntry *batch_line_table_32_2(int *n);

// The lean-I/O kernels check the input, store and output windows only in the
// first and last slot chunks of a line: every window must contain 2..N-2.
bool line_windows_ok(const PassArgs &p, int N)
{
    auto in = [N](int lo, int hi) { return lo <= 2 && hi >= N - 2; };
    return in(p.in_lo, p.in_hi) && in(p.st_lo, p.st_hi) && in(p.out_lo, p.out_hi);
}

const FixedEntry *line_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan)
{
    // CHZ_NO_LINE / CHZ_NO_FIXED in the environment: none
    static const bool off = getenv("CHZ_NO_FIXED") != nullptr || getenv("CHZ_NO_LINE") != nullptr;
    if (off) return nullptr;
    typedef const FixedEntry *(*Tab)(int *);
    static const Tab t64[] = {line_table_f64, line_small_table_f64, line_mid_table_f64};
    static const Tab t32[] = {line_table_f32, line_small_table_f32, line_mid_table_f32};
    for (int q = 0; q < 3; q++) {
        int n = 0;
        const FixedEntry *t = (prec == 64 ? t64 : t32)[q](&n);
        for (int i = 0; i < n; i++) {
            const FixedEntry &e = t[i];
            if (e.op == op && e.N == N && e.ept == ept && e.tpl == tpl && e.lines == lines && e.plan == plan) return &e;
        }
    }
    return nullptr;
}

const FixedEntry *fixed_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan,
                               const PassArgs *p)
{
    static const bool off = getenv("CHZ_NO_FIXED") != nullptr;
    if (off) return nullptr;
    // the lean-I/O kernels first where the pass windows allow them
    if (p && line_windows_ok(*p, N))
        if (const FixedEntry *e = line_kernel(prec, op, N, ept, tpl, lines, plan)) return e;
    int n = 0;
    const FixedEntry *t = prec == 64 ? fixed_table_f64(&n) : fixed_table_f32(&n);
    for (int i = 0; i < n; i++) {
        const FixedEntry &e = t[i];
        if (e.op == op && e.N == N && e.ept == ept && e.tpl == tpl && e.lines == lines && e.plan == plan) return &e;
    }
    return nullptr;
}

const void *batch_fixed_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan)
{
    static const bool off = getenv("CHZ_NO_FIXED") != nullptr;
    if (off) return nullptr;
    typedef const FixedEntry *(*Tab)(int *);
    static const Tab t64[] = {batch_fixed_table_64_0, batch_fixed_table_64_1};
    static const Tab t32[] = {batch_fixed_table_32_0, batch_fixed_table_32_1, batch_fixed_table_32_2};
    const Tab *tabs = prec == 64 ? t64 : t32;
    const int ntab = prec == 64 ? 2 : 3;
    for (int q = 0; q < ntab; q++) {
        int n = 0;
        const FixedEntry *t = tabs[q](&n);
        for (int i = 0; i < n; i++) {
            const FixedEntry &e = t[i];
            if (e.op == op && e.N == N && e.ept == ept && e.tpl == tpl && e.lines == lines && e.plan == plan) return e.fn;
        }
    }
    return nullptr;
}

const FixedEntry *batch_line_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan)
{
    static const bool off = getenv("CHZ_NO_FIXED") != nullptr || getenv("CHZ_NO_LINE") != nullptr;
    if (off) return nullptr;
    typedef const FixedEntry *(*Tab)(int *);
    static const Tab t64[] = {batch_line_table_64_0, batch_line_table_64_1};
    static const Tab t32[] = {batch_line_table_32_0, batch_line_table_32_1, batch_line_table_32_2};
    const Tab *tabs = prec == 64 ? t64 : t32;
    const int ntab = prec == 64 ? 2 : 3;
    for (int q = 0; q < ntab; q++) {
        int n = 0;
        const FixedEntry *t = tabs[q](&n);
        for (int i = 0; i < n; i++) {
            const FixedEntry &e = t[i];
            if (e.op == op && e.N == N && e.ept == ept && e.tpl == tpl && e.lines == lines && e.plan == plan) return &e;
        }
    }
    return nullptr;
}

}  // namespace chz
