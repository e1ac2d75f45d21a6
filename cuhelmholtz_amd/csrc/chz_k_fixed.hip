// Lookup of the fixed-shape pass kernels (tables in chz_k_fx64.hip, chz_k_fx32.hip).
#include <cstdlib>

#include "chz_kernels.h"

namespace chz {

const FixedEntry *fixed_table_f64(int *n);
const FixedEntry *fixed_table_f32(int *n);

const FixedEntry *fixed_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan)
{
    static const bool off = getenv("CHZ_NO_FIXED") != nullptr;
    if (off) return nullptr;
    int n = 0;
    const FixedEntry *t = prec == 64 ? fixed_table_f64(&n) : fixed_table_f32(&n);
    for (int i = 0; i < n; i++) {
        const FixedEntry &e = t[i];
        if (e.op == op && e.N == N && e.ept == ept && e.tpl == tpl && e.lines == lines && e.plan == plan) return &e;
    }
    return nullptr;
}

}  // namespace chz
