// chz_kernels.hip -- __global__ entry points of the axis passes and their
// launch table.  Compiled for gfx950 only.
#include "chz_kernels.h"
#include "chz_pass.h"

namespace chz {

// One fused axis pass over a tile of L lines: load -> fold -> FFT -> unfold -> store.
template <typename T, typename SRC, typename DST, int EPT, bool ROWS>
__global__ void __launch_bounds__(256) k_axis_pass(PassArgs p)
{
    typedef typename vec2_of<T>::type V;
    extern __shared__ __align__(16) unsigned char smem[];
    V *buf = reinterpret_cast<V *>(smem);
    const int L = p.L;
    const Tile<T> tl{buf, L + 1, (int)threadIdx.x % L};
    const int u = (int)threadIdx.x / L;
    T *red = reinterpret_cast<T *>(buf + lds_points(p.M, L));
    const int a_tile = p.a0 + (int)blockIdx.x * L;
    const int nlines = min(L, p.a0 + p.na - a_tile);
    const long long b = p.b0 + (long long)blockIdx.y;
    load_tile<T, ROWS>(p, tl, nlines, (long long)a_tile * p.s_sa + b * p.s_sb, reinterpret_cast<const SRC *>(p.src));
    __syncthreads();
    line_transform<EPT, T>(p.kind, tl, u, p.TPL, p.M, p.N, p.plan, reinterpret_cast<const V *>(p.tw), red);
    store_tile<T, ROWS>(p, tl, nlines, (long long)a_tile * p.d_sa + b * p.d_sb, reinterpret_cast<DST *>(p.dst));
}

template <typename T, typename SRC, typename DST, bool ROWS>
static const void *pick_pass(int ept)
{
    switch (ept) {
        case 1: return (const void *)k_axis_pass<T, SRC, DST, 1, ROWS>;
        case 2: return (const void *)k_axis_pass<T, SRC, DST, 2, ROWS>;
        case 3: return (const void *)k_axis_pass<T, SRC, DST, 3, ROWS>;
        case 4: return (const void *)k_axis_pass<T, SRC, DST, 4, ROWS>;
        case 6: return (const void *)k_axis_pass<T, SRC, DST, 6, ROWS>;
        case 8: return (const void *)k_axis_pass<T, SRC, DST, 8, ROWS>;
        case 12: return (const void *)k_axis_pass<T, SRC, DST, 12, ROWS>;
        case 16: return (const void *)k_axis_pass<T, SRC, DST, 16, ROWS>;
    }
    return nullptr;
}

const void *axis_pass_kernel(int prec, int ept, bool rows)
{
    if (prec == 64) return rows ? pick_pass<double, double, double, true>(ept) : pick_pass<double, double, double, false>(ept);
    return rows ? pick_pass<float, float, float, true>(ept) : pick_pass<float, float, float, false>(ept);
}

hipError_t launch_axis_pass(const PassArgs &p, int prec, int ept, bool rows, int nb, hipStream_t s)
{
    const void *fn = axis_pass_kernel(prec, ept, rows);
    if (!fn) return hipErrorInvalidValue;
    const int threads = p.L * p.TPL;
    const size_t esz = prec == 64 ? 8 : 4;
    const size_t lds = (size_t)lds_points(p.M, p.L) * 2 * esz + (size_t)p.TPL * p.L * esz;
    dim3 grid((unsigned)((p.na + p.L - 1) / p.L), (unsigned)nb, 1);
    PassArgs a = p;
    void *args[] = {&a};
    return hipLaunchKernel(fn, grid, dim3((unsigned)threads), args, lds, s);
}

size_t axis_pass_lds(int prec, int M, int L, int TPL)
{
    const size_t esz = prec == 64 ? 8 : 4;
    return (size_t)lds_points(M, L) * 2 * esz + (size_t)TPL * L * esz;
}

}  // namespace chz
