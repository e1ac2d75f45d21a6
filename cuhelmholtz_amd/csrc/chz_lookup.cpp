// Lookup of the compiled fixed-shape pass kernels.  The tables are generated
// at build time from the tuning tables (cuhelmholtz_amd/kernelgen.py:
// tuning/<arch>.txt, the chosen shapes, and tuning/<arch>.alt.txt, the
// alternatives the autotuner measures), reached through gen_table().
#include <cstdlib>

#include "chz_kernels.h"

namespace chz {

bool line_windows_ok(const PassArgs &p, int N)
{
    auto in = [N](int lo, int hi) { return lo <= 2 && hi >= N - 2; };
    return in(p.in_lo, p.in_hi) && in(p.st_lo, p.st_hi) && in(p.out_lo, p.out_hi);
}

int pass_kind_sel(int op, const PassArgs &p, const SolveArgs *sa)
{
    if (op != B_ZMID) return p.kind;
    if (!sa) return -1;
    static const int fwd[5] = {K_R2C, K_DST_1, K_DST_3, K_DCT_1, K_DCT_3};
    static const int inv[5] = {K_C2R, K_DST_1, K_DST_2, K_DCT_1, K_DCT_2};
    for (int c = 0; c < 5; c++)
        if (p.kind == fwd[c] && sa->kind2 == inv[c]) return c;
    return -1;
}

// the compiled kernel of one form, shape and kind selector, nullptr if none
static const FixedEntry *find_ks(int form, int prec, int op, int N, int ept, int tpl, int lines,
                                 unsigned long long plan, int ks)
{
    for (int q = 0; q < gen_table_count(); q++) {
        int n = 0;
        const FixedEntry *t = gen_table(q, &n);
        for (int i = 0; i < n; i++) {
            const FixedEntry &e = t[i];
            if (e.form == form && e.prec == prec && e.op == op && e.N == N && e.ept == ept && e.tpl == tpl &&
                e.lines == lines && e.plan == plan && e.ks == ks)
                return &e;
        }
    }
    return nullptr;
}

// the kind-specialised kernel of the shape if one is compiled, else the one for every kind
static const FixedEntry *find(int form, int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan,
                              int ks = -1)
{
    if (ks >= 0)
        if (const FixedEntry *e = find_ks(form, prec, op, N, ept, tpl, lines, plan, ks)) return e;
    return find_ks(form, prec, op, N, ept, tpl, lines, plan, -1);
}

// CHZ_NO_FIXED in the environment: no fixed-shape kernels at all; CHZ_NO_LINE:
// none with lean tile I/O (both for measuring the generic kernels)
static bool no_fixed()
{
    static const bool off = getenv("CHZ_NO_FIXED") != nullptr;
    return off;
}
static bool no_line()
{
    static const bool off = no_fixed() || getenv("CHZ_NO_LINE") != nullptr;
    return off;
}

const FixedEntry *line_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan, int ks)
{
    return no_line() ? nullptr : find(F_LINE, prec, op, N, ept, tpl, lines, plan, ks);
}

const FixedEntry *fixed_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan,
                               const PassArgs *p, int ks)
{
    if (no_fixed()) return nullptr;
    // the lean-I/O kernels first where the pass windows allow them
    if (p && line_windows_ok(*p, N))
        if (const FixedEntry *e = line_kernel(prec, op, N, ept, tpl, lines, plan, ks)) return e;
    return find(F_FIXED, prec, op, N, ept, tpl, lines, plan);
}

const void *batch_fixed_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan)
{
    if (no_fixed()) return nullptr;
    const FixedEntry *e = find(F_FIXED_BATCH, prec, op, N, ept, tpl, lines, plan);
    return e ? e->fn : nullptr;
}

const FixedEntry *batch_line_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan,
                                    int ks)
{
    return no_line() ? nullptr : find(F_LINE_BATCH, prec, op, N, ept, tpl, lines, plan, ks);
}

int compiled_kernels(int *rows, int max_rows)
{
    int k = 0;
    for (int q = 0; q < gen_table_count(); q++) {
        int n = 0;
        const FixedEntry *t = gen_table(q, &n);
        for (int i = 0; i < n; i++, k++) {
            if (!rows || k >= max_rows) continue;
            const FixedEntry &e = t[i];
            int *r = rows + 10 * k;
            r[0] = e.form; r[1] = e.prec; r[2] = e.op; r[3] = e.N; r[4] = e.ept; r[5] = e.tpl; r[6] = e.lines;
            r[7] = (int)(e.plan & 0xffffffffull);
            r[8] = (int)(e.plan >> 32);
            r[9] = e.ks;
        }
    }
    return k;
}

}  // namespace chz
