// Kernels of the adjoint solve: the transposes of the periodic copies, of the
// Dirichlet pass-through and of the boundary folds around the solve, and the
// lambda-gradient reduction.
#include "chz_kernels.h"

namespace chz {

// weight of node q on an axis: 1/2 at the Neumann ends that are unknowns
// (code 3 both ends, code 2 the high end, code 4 the low end); the operator
// satisfies A^T = W A W^-1 with W the product of the axis weights.
__device__ __forceinline__ double axis_weight(int bc, int q, int N)
{
    if ((bc == 3 || bc == 4) && q == 0) return 0.5;
    if ((bc == 3 || bc == 2) && q == N) return 0.5;
    return 1.0;
}

// Transpose of the periodic copies and the pass-through: gf = accumulated
// output gradient on every node that is not a periodic duplicate (0 on the
// duplicates); g = gf / W on the unknown box, 0 elsewhere (the rhs of the
// adjoint solve, homogeneous boundary data).  Zeroes the face gradients.
// Problem blockIdx.y of a batch.
__global__ void __launch_bounds__(256) k_adj_prep(AdjArgs a, const double *__restrict__ gu, double *__restrict__ gf,
                                                  double *__restrict__ g)
{
    const long long y = blockIdx.y;
    gu += y * a.bstride;
    gf += y * a.bstride;
    g += y * a.bstride;
    for (int f = 0; f < 6; f++)
        if (a.face[f]) a.face[f] += y * a.fstride[f];
    const long long fx = a.N[0] + 1, fy = a.N[1] + 1, fz = a.N[2] + 1;
    const long long total = fx * fy * fz;
    const long long fsz[3] = {fy * fz, fx * fz, fx * fy};
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        for (int f = 0; f < 6; f++)
            if (a.face[f] && idx < fsz[f / 2]) a.face[f][idx] = 0.0;
        const int q[3] = {(int)(idx % fx), (int)((idx / fx) % fy), (int)(idx / (fx * fy))};
        const long long st[3] = {1, fx, fx * fy};
        bool dup = false, unk = true;
        int zero_mask = 0;
        for (int d = 0; d < 3; d++) {
            if (a.bc[d] == 0 && q[d] == a.N[d]) dup = true;
            if (a.bc[d] == 0 && q[d] == 0) zero_mask |= 1 << d;
            if (q[d] < a.s[d] || q[d] > a.e[d]) unk = false;
        }
        if (dup) {
            gf[idx] = 0.0;
            g[idx] = 0.0;
            continue;
        }
        // sum of the gradient over the node and its periodic images
        double acc = 0.0;
        for (int m = 0; m < 8; m++) {
            if (m & ~zero_mask) continue;
            long long off = idx;
            for (int d = 0; d < 3; d++)
                if (m & (1 << d)) off += (long long)a.N[d] * st[d];
            acc += gu[off];
        }
        gf[idx] = acc;
        if (unk) {
            const double w = axis_weight(a.bc[0], q[0], a.N[0]) * axis_weight(a.bc[1], q[1], a.N[1]) *
                             axis_weight(a.bc[2], q[2], a.N[2]);
            g[idx] = acc / w;
        } else {
            g[idx] = 0.0;
        }
    }
}

// r = W g on the unknown box (the adjoint solution): gf = r there; the fold
// transposes (Dirichlet boundary node -= r / h^2, Neumann face gradient
// +-2 r / h); part[block] = sum of r u (u = forward solution, may be null).
// Problem blockIdx.y of a batch (its partial sums at part + blockIdx.y * gridDim.x).
__global__ void __launch_bounds__(256) k_adj_finish(AdjArgs a, const double *__restrict__ g, const double *__restrict__ u,
                                                    double *__restrict__ gf, double *__restrict__ part)
{
    __shared__ double red[256];
    const long long y = blockIdx.y;
    g += y * a.bstride;
    if (u) u += y * a.bstride;
    gf += y * a.bstride;
    part += y * gridDim.x;
    for (int f = 0; f < 6; f++)
        if (a.face[f]) a.face[f] += y * a.fstride[f];
    const long long fx = a.N[0] + 1, fy = a.N[1] + 1;
    const long long nx = a.e[0] - a.s[0] + 1, ny = a.e[1] - a.s[1] + 1, nz = a.e[2] - a.s[2] + 1;
    const long long total = nx * ny * nz;
    double sum = 0.0;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int q[3] = {a.s[0] + (int)(idx % nx), a.s[1] + (int)((idx / nx) % ny), a.s[2] + (int)(idx / (nx * ny))};
        const long long st[3] = {1, fx, fx * fy};
        const long long c = q[0] + q[1] * fx + q[2] * fx * fy;
        const double w = axis_weight(a.bc[0], q[0], a.N[0]) * axis_weight(a.bc[1], q[1], a.N[1]) *
                         axis_weight(a.bc[2], q[2], a.N[2]);
        const double r = w * g[c];
        gf[c] = r;
        if (u) sum += r * u[c];
        for (int d = 0; d < 3; d++) {
            const int bc = a.bc[d];
            if (bc == 0) continue;
            // face index of the node on the faces normal to axis d
            const int p1 = d == 0 ? q[1] : q[0];
            const int p2 = d == 2 ? q[1] : q[2];
            const long long fi = (long long)p2 * (d == 0 ? fy : fx) + p1;
            if (q[d] == a.s[d]) {   // low end: Dirichlet node 0 (codes 1, 2) or Neumann face (3, 4)
                if (bc == 1 || bc == 2) gf[c - (long long)q[d] * st[d]] -= r * a.ih2[d];
                else if (a.face[2 * d]) a.face[2 * d][fi] = 2.0 * r * a.ih[d];
            }
            if (q[d] == a.e[d]) {   // high end: Dirichlet node N (codes 1, 4) or Neumann face (2, 3)
                if (bc == 1 || bc == 4) gf[c + (long long)(a.N[d] - q[d]) * st[d]] -= r * a.ih2[d];
                else if (a.face[2 * d + 1]) a.face[2 * d + 1][fi] = -2.0 * r * a.ih[d];
            }
        }
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[b] = -sum of part[b*n .. b*n + n) in a fixed order (block b: problem b of a batch)
__global__ void __launch_bounds__(256) k_adj_lambda(const double *__restrict__ part, int n, double *__restrict__ out)
{
    __shared__ double red[256];
    part += (long long)blockIdx.x * n;
    out += blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = -red[0];
}

// zero fill of n doubles
__global__ void __launch_bounds__(256) k_adj_zero(double *p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

const void *adj_kernel(int which)
{
    switch (which) {
        case 0: return (const void *)k_adj_prep;
        case 1: return (const void *)k_adj_finish;
        case 2: return (const void *)k_adj_lambda;
        case 3: return (const void *)k_adj_zero;
    }
    return nullptr;
}

}  // namespace chz
