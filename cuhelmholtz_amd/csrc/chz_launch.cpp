// chz_launch.cpp -- host-side launch of the axis and solver passes.
#include <cstdlib>

#include "chz_kernels.h"

namespace chz {

size_t axis_pass_lds(int prec, int M, int L, int TPL)
{
    const size_t esz = prec == 64 ? 8 : 4;
    // tile points + FFT twiddles (complex) + reduction scratch (real)
    return ((size_t)lds_points(M, L, (int)esz) + (size_t)M) * 2 * esz + (size_t)TPL * L * esz;
}

hipError_t launch_fixed(const FixedEntry *fx, const PassArgs &p, const SolveArgs *sa, const void *u_in, void *u_out,
                        int prec, int nb, hipStream_t s)
{
    const size_t lds = axis_pass_lds(prec, p.M, p.L, p.TPL) + fx->lds_extra;
    dim3 grid((unsigned)(((p.na + p.L - 1) / p.L) * (long long)nb), 1, 1);
    const void *fn = fx->fn;
    PassArgs a = p;
    SolveArgs b{};
    if (sa) b = *sa;
    const void *x = u_in;
    void *y = u_out;
    void *args[] = {&a, &b, &x, &y};
    return hipLaunchKernel(fn, grid, dim3((unsigned)(p.L * p.TPL)), args, lds, s);
}

hipError_t launch_axis_pass(const PassArgs &p, int prec, int ept, bool rows, int nb, hipStream_t s)
{
    if (const FixedEntry *fx = fixed_kernel(prec, rows ? B_ROWS : B_COLS, p.N, ept, p.TPL, p.L, p.plan, &p, p.kind))
        return launch_fixed(fx, p, nullptr, nullptr, nullptr, prec, nb, s);
    const void *fn = prec == 64 ? (rows ? axis_kernel_f64_rows(ept) : axis_kernel_f64_cols(ept))
                                : (rows ? axis_kernel_f32_rows(ept) : axis_kernel_f32_cols(ept));
    if (!fn) return hipErrorInvalidValue;
    dim3 grid((unsigned)(((p.na + p.L - 1) / p.L) * (long long)nb), 1, 1);
    PassArgs a = p;
    void *args[] = {&a};
    return hipLaunchKernel(fn, grid, dim3((unsigned)(p.L * p.TPL)), args, axis_pass_lds(prec, p.M, p.L, p.TPL), s);
}

hipError_t launch_solve_pass(const void *fn, const PassArgs &p, const SolveArgs &sa, int prec, int nb, hipStream_t s)
{
    if (!fn) return hipErrorInvalidValue;
    dim3 grid((unsigned)(((p.na + p.L - 1) / p.L) * (long long)nb), 1, 1);
    PassArgs a = p;
    SolveArgs b = sa;
    void *args[] = {&a, &b};
    return hipLaunchKernel(fn, grid, dim3((unsigned)(p.L * p.TPL)), args, axis_pass_lds(prec, p.M, p.L, p.TPL), s);
}

hipError_t launch_pass_ptrs(const void *fn, const PassArgs &p, const void *a0, const void *a1, int prec, int nb,
                            hipStream_t s)
{
    if (!fn) return hipErrorInvalidValue;
    dim3 grid((unsigned)(((p.na + p.L - 1) / p.L) * (long long)nb), 1, 1);
    PassArgs a = p;
    const void *x = a0, *y = a1;
    void *args[] = {&a, &x, &y};
    return hipLaunchKernel(fn, grid, dim3((unsigned)(p.L * p.TPL)), args, axis_pass_lds(prec, p.M, p.L, p.TPL), s);
}

hipError_t launch_pass_sa_ptr(const void *fn, const PassArgs &p, const SolveArgs &sa, const void *a0, int prec, int nb,
                              hipStream_t s)
{
    if (!fn) return hipErrorInvalidValue;
    dim3 grid((unsigned)(((p.na + p.L - 1) / p.L) * (long long)nb), 1, 1);
    PassArgs a = p;
    SolveArgs b = sa;
    const void *x = a0;
    void *args[] = {&a, &b, &x};
    return hipLaunchKernel(fn, grid, dim3((unsigned)(p.L * p.TPL)), args, axis_pass_lds(prec, p.M, p.L, p.TPL), s);
}

}  // namespace chz
