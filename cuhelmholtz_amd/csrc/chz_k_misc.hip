// Small helper kernels of the solves (periodic copies live with the FP64 axis pass).
#include "chz_axis_kernels.h"

namespace chz {

// weight of node index q on an axis with unknown window [lo, hi] and code bc:
// 1/2 at the Neumann-Neumann ends (the trapezoid weights of the DCT-I 0-mode)
__device__ __forceinline__ double mode0_weight(int bc, int q, int lo, int hi)
{
    return (bc == 3 && (q == lo || q == hi)) ? 0.5 : 1.0;
}

// acc[0] = 0 (a fill kernel rather than a memset so the sequence can be
// recorded into a graph)
__global__ void k_zero1(double *acc) { if (threadIdx.x == 0 && blockIdx.x == 0) acc[0] = 0.0; }

// acc += sum over the unknown box of w(i) w(j) w(k) v(i,j,k) with v = u, or
// (u == nullptr) the folded rhs of sa
__global__ void __launch_bounds__(256) k_weighted_sum(const double *u, SolveArgs sa, double *acc)
{
    __shared__ double part[256];
    const long long fx = sa.NX + 1, fy = sa.NY + 1;
    const long long nx = sa.e[0] - sa.s[0] + 1, ny = sa.e[1] - sa.s[1] + 1, nz = sa.e[2] - sa.s[2] + 1;
    const long long total = nx * ny * nz;
    double s = 0.0;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int i = sa.s[0] + (int)(idx % nx), j = sa.s[1] + (int)((idx / nx) % ny), k = sa.s[2] + (int)(idx / (nx * ny));
        const double w = mode0_weight(sa.bc[0], i, sa.s[0], sa.e[0]) * mode0_weight(sa.bc[1], j, sa.s[1], sa.e[1]) *
                         mode0_weight(sa.bc[2], k, sa.s[2], sa.e[2]);
        s += w * (u ? u[i + j * fx + k * fx * fy] : residual_at(sa, nullptr, i, j, k));
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc, part[0]);
}

// u -= acc / wsum over the unknown box
__global__ void __launch_bounds__(256) k_shift(double *u, SolveArgs sa, const double *acc, double inv_wsum)
{
    const long long fx = sa.NX + 1, fy = sa.NY + 1;
    const long long nx = sa.e[0] - sa.s[0] + 1, ny = sa.e[1] - sa.s[1] + 1, nz = sa.e[2] - sa.s[2] + 1;
    const long long total = nx * ny * nz;
    const double m = acc[0] * inv_wsum;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int i = sa.s[0] + (int)(idx % nx), j = sa.s[1] + (int)((idx / nx) % ny), k = sa.s[2] + (int)(idx / (nx * ny));
        u[i + j * fx + k * fx * fy] -= m;
    }
}

double mode0_inv_wsum(const SolveArgs &sa)
{
    double wsum = 1.0;
    for (int a = 0; a < 3; a++) {
        const int n = sa.e[a] - sa.s[a] + 1;
        wsum *= sa.bc[a] == 3 ? (double)(n - 1) : (double)n;
    }
    return 1.0 / wsum;
}

hipError_t launch_fold_sum(const SolveArgs &sa, double *acc, hipStream_t s)
{
    SolveArgs b = sa;
    b.pertrb = nullptr;
    double *aa = acc;
    void *z[] = {&aa};
    hipError_t e = hipLaunchKernel((const void *)k_zero1, dim3(1), dim3(64), z, 0, s);
    if (e != hipSuccess) return e;
    const double *nu = nullptr;
    void *a1[] = {&nu, &b, &aa};
    return hipLaunchKernel((const void *)k_weighted_sum, dim3(2048), dim3(256), a1, 0, s);
}

hipError_t launch_remove_mode0(double *u, const SolveArgs &sa, double *acc, hipStream_t s)
{
    const double invw = mode0_inv_wsum(sa);
    SolveArgs b = sa;
    double *uu = u, *aa = acc;
    void *z[] = {&aa};
    hipError_t e = hipLaunchKernel((const void *)k_zero1, dim3(1), dim3(64), z, 0, s);
    if (e != hipSuccess) return e;
    const double *cu = u;
    void *a1[] = {&cu, &b, &aa};
    if ((e = hipLaunchKernel((const void *)k_weighted_sum, dim3(2048), dim3(256), a1, 0, s)) != hipSuccess) return e;
    const double *ca = acc;
    double inv = invw;
    void *a2[] = {&uu, &b, &ca, &inv};
    return hipLaunchKernel((const void *)k_shift, dim3(2048), dim3(256), a2, 0, s);
}

}  // namespace chz
