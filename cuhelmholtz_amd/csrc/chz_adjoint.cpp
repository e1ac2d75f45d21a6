// chz_adjoint.cpp -- adjoint of the FFT-diagonalised Helmholtz solve.
//
// The forward map (cudahelmholtz/cuda_helmholtz_solver.cu:33-380) is linear
// in (f, Neumann face data): fold the boundary data into the rhs of the
// unknown box, solve A u = r there, keep the Dirichlet boundary values of f
// (pass-through), copy node 0 to node N along periodic axes.  Its transpose:
//   1  periodic copies and pass-through transposed (k_adj_prep): the output
//      gradient summed over each node's periodic images;
//   2  r_bar = A^-T g.  The 7-point operator with the codes 0-4 is
//      W-symmetric: A = W^-1 S with S symmetric and W = 1/2 at the Neumann
//      end nodes of each axis, so A^-T = W A^-1 W^-1 -- the same spectral
//      solve (transforms, divide, singular 0-mode) between two diagonal
//      scalings, run with homogeneous boundary data;
//   3  fold transposes (k_adj_finish): the gradient of each Dirichlet
//      boundary node gets -r_bar/h^2 of its unknown neighbour, the Neumann
//      face gradients +-2 r_bar/h;
//   4  d lambda = -<r_bar, u> over the unknown box (A u = r, dA/dlambda = I).
#include <algorithm>
#include <string>
#include <vector>

#include "chz_context.h"
#include "chz_solver.h"
#include "../../include/cuhelmholtz.h"

namespace chz {

int launch_spec(Context &ctx, const PassSpec &q, hipStream_t st);

static int adjoint(const double dom[6], const int N[3], const int bc[3], double lambda, const double *u,
                   const double *gu, double *gf, double *const gfaces[6], double *glambda, hipStream_t st)
{
    if (!gu || !gf) return fail(ERR_INVALID, "null gradient array");
    if (gu == gf) return fail(ERR_INVALID, "the output gradient must not alias the input gradient");
    if (glambda && !u) return fail(ERR_INVALID, "the lambda gradient needs the forward solution u");
    Context &ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex());
    WsLease ws(ctx, st);
    const long long fx = N[0] + 1, fy = N[1] + 1, fz = N[2] + 1;
    const long long nodes = fx * fy * fz;
    const long long maxface = std::max(fy * fz, std::max(fx * fz, fx * fy));
    // homogeneous boundary data for the adjoint solve: a zeroed face buffer
    double *zero = nodes > 0 ? reinterpret_cast<double *>(ws.get(sizeof(double) * (size_t)maxface, 5)) : nullptr;
    const double *faces[6] = {zero, zero, zero, zero, zero, zero};
    SolveGeom g;
    int rc = make_geom(g, dom, N, bc, lambda, faces);
    if (rc) return rc;
    if (!zero) return fail(ERR_ALLOC, "workspace allocation failed");
    const size_t exbytes = sizeof(double) * (size_t)ws_row(N[0], sizeof(double)) * (N[1] + 2) * (N[2] + 2);
    double *ex = reinterpret_cast<double *>(ws.get(exbytes, 0));
    double *gw = reinterpret_cast<double *>(ws.get(sizeof(double) * (size_t)nodes, 4));
    const int nblk = 1024;
    double *part = reinterpret_cast<double *>(ws.get(sizeof(double) * nblk, 6));
    if (!ex || !gw || !part) return fail(ERR_ALLOC, "workspace allocation failed");
    std::vector<PassSpec> v;
    if ((rc = solve_passes<double>(ctx, g, gw, ex, nullptr, nullptr, v)) || (rc = final_pass<double>(ctx, g, ex, nullptr, gw, v)))
        return rc;

    AdjArgs a{};
    for (int d = 0; d < 3; d++) {
        a.N[d] = g.N[d];
        a.bc[d] = g.bc[d];
        a.s[d] = g.s[d];
        a.e[d] = g.e[d];
        a.ih[d] = g.ih[d];
        a.ih2[d] = g.ih2[d];
    }
    for (int f = 0; f < 6; f++) {
        // face gradients exist only where the code reads face data
        const int d = f / 2;
        const bool used = f % 2 == 0 ? (bc[d] == 3 || bc[d] == 4) : (bc[d] == 2 || bc[d] == 3);
        a.face[f] = used && gfaces ? gfaces[f] : nullptr;
    }
    PassTimer timer(ctx, st);
    hipError_t e;
    {
        timer.mark("azero");
        double *z = zero;
        long long n = maxface;
        void *args[] = {&z, &n};
        if ((e = hipLaunchKernel(adj_kernel(3), dim3((unsigned)std::min<long long>((n + 255) / 256, 1024)), dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint zero fill: ") + hipGetErrorString(e));
    }
    {
        timer.mark("aprep");
        const double *cgu = gu;
        double *cgf = gf, *cgw = gw;
        void *args[] = {&a, &cgu, &cgf, &cgw};
        if ((e = hipLaunchKernel(adj_kernel(0), dim3((unsigned)std::min<long long>((nodes + 255) / 256, 4096)), dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint prep launch: ") + hipGetErrorString(e));
    }
    for (const PassSpec &q : v)
        if ((rc = launch_spec(ctx, q, st))) return rc;
    {
        timer.mark("afinish");
        const double *cgw = gw, *cu = u;
        double *cgf = gf, *cp = part;
        void *args[] = {&a, &cgw, &cu, &cgf, &cp};
        if ((e = hipLaunchKernel(adj_kernel(1), dim3(nblk), dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint finish launch: ") + hipGetErrorString(e));
    }
    if (glambda) {
        const double *cp = part;
        int n = nblk;
        double *out = glambda;
        void *args[] = {&cp, &n, &out};
        if ((e = hipLaunchKernel(adj_kernel(2), dim3(1), dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint lambda launch: ") + hipGetErrorString(e));
    }
    return OK;
}

// Adjoint of `count` problems of one geometry with per-problem lambdas:
// consecutive node arrays u, gu, gf, per-problem face gradient arrays and
// lambda gradients.  The node-array steps run once over the whole batch
// (blockIdx.y = problem); the solve is the batched solve of the library's
// plan for this geometry (one graph replay, one launch per pass and size class).
static int adjoint_batch(const double dom[6], const int N[3], const int bc[3], int count, const double *lambdas,
                         const double *u, const double *gu, double *gf, double *const gfaces[6], double *glambda,
                         hipStream_t st)
{
    if (count <= 0 || !lambdas) return fail(ERR_INVALID, "adjoint batch: empty batch or null lambdas");
    if (!gu || !gf) return fail(ERR_INVALID, "null gradient array");
    if (gu == gf) return fail(ERR_INVALID, "the output gradient must not alias the input gradient");
    if (glambda && !u) return fail(ERR_INVALID, "the lambda gradient needs the forward solution u");
    const long long fx = N[0] + 1, fy = N[1] + 1, fz = N[2] + 1;
    const long long nodes = fx * fy * fz;
    const long long fsz[3] = {fy * fz, fx * fz, fx * fy};
    const long long maxface = std::max(fsz[0], std::max(fsz[1], fsz[2]));
    const int nblk = 1024;
    Context &ctx = Context::get();
    // held through the launches (the batched solve below re-enters it), so the
    // workspace slots are ordered after their previous users
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex());
    WsLease ws(ctx, st);
    double *zero = reinterpret_cast<double *>(ws.get(sizeof(double) * (size_t)maxface, 5));
    double *gw = reinterpret_cast<double *>(ws.get(sizeof(double) * (size_t)(nodes * count), 4));
    double *part = reinterpret_cast<double *>(ws.get(sizeof(double) * (size_t)nblk * count, 6));
    if (!zero || !gw || !part) return fail(ERR_ALLOC, "workspace allocation failed");
    SolveGeom g;
    {
        const double *faces[6] = {zero, zero, zero, zero, zero, zero};
        int rc = make_geom(g, dom, N, bc, lambdas[0], faces);
        if (rc) return rc;
    }
    AdjArgs a{};
    for (int d = 0; d < 3; d++) {
        a.N[d] = g.N[d];
        a.bc[d] = g.bc[d];
        a.s[d] = g.s[d];
        a.e[d] = g.e[d];
        a.ih[d] = g.ih[d];
        a.ih2[d] = g.ih2[d];
    }
    a.bstride = nodes;
    for (int f = 0; f < 6; f++) {
        const int d = f / 2;
        const bool used = f % 2 == 0 ? (bc[d] == 3 || bc[d] == 4) : (bc[d] == 2 || bc[d] == 3);
        a.face[f] = used && gfaces ? gfaces[f] : nullptr;
        a.fstride[f] = fsz[d];
    }
    hipError_t e;
    {
        double *z = zero;
        long long n = maxface;
        void *args[] = {&z, &n};
        if ((e = hipLaunchKernel(adj_kernel(3), dim3((unsigned)std::min<long long>((n + 255) / 256, 1024)), dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint zero fill: ") + hipGetErrorString(e));
    }
    {
        const double *cgu = gu;
        double *cgf = gf, *cgw = gw;
        void *args[] = {&a, &cgu, &cgf, &cgw};
        if ((e = hipLaunchKernel(adj_kernel(0), dim3((unsigned)std::min<long long>((nodes + 255) / 256, 4096), (unsigned)count),
                                 dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint prep launch: ") + hipGetErrorString(e));
    }
    // the homogeneous solves of the batch on g (in gw)
    std::vector<chz_problem> probs((size_t)count);
    for (int k = 0; k < count; k++) {
        chz_problem &q = probs[(size_t)k];
        for (int i = 0; i < 6; i++) q.domain[i] = dom[i];
        for (int d = 0; d < 3; d++) {
            q.N[d] = N[d];
            q.bc[d] = bc[d];
        }
        q.lambda = lambdas[k];
        q.f = gw + nodes * k;
        for (int i = 0; i < 6; i++) q.faces[i] = g.faces[i] ? zero : nullptr;
    }
    int rc = chz_batch_solve_cached(probs.data(), count, 64, 0, st);
    if (rc) return rc;
    {
        const double *cgw = gw, *cu = u;
        double *cgf = gf, *cp = part;
        void *args[] = {&a, &cgw, &cu, &cgf, &cp};
        if ((e = hipLaunchKernel(adj_kernel(1), dim3(nblk, (unsigned)count), dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint finish launch: ") + hipGetErrorString(e));
    }
    if (glambda) {
        const double *cp = part;
        int n = nblk;
        double *out = glambda;
        void *args[] = {&cp, &n, &out};
        if ((e = hipLaunchKernel(adj_kernel(2), dim3((unsigned)count), dim3(256), args, 0, st)) != hipSuccess)
            return fail(ERR_DEVICE, std::string("adjoint lambda launch: ") + hipGetErrorString(e));
    }
    return OK;
}

}  // namespace chz

extern "C" int chz_helmholtz_adjoint_batch(double left, double right, int NX, int xbc, double bottom, double top, int NY,
                                           int ybc, double front, double back, int NZ, int zbc, int count,
                                           const double *lambdas, const double *u, const double *gu, double *gf,
                                           double *gbcl, double *gbcr, double *gbcb, double *gbct, double *gbcf,
                                           double *gbce, double *glambda, void *stream)
{
    const double dom[6] = {left, right, bottom, top, front, back};
    const int N[3] = {NX, NY, NZ}, bc[3] = {xbc, ybc, zbc};
    double *const faces[6] = {gbcl, gbcr, gbcb, gbct, gbcf, gbce};
    return chz::adjoint_batch(dom, N, bc, count, lambdas, u, gu, gf, faces, glambda, (hipStream_t)stream);
}

extern "C" int chz_helmholtz_adjoint(double left, double right, int NX, int xbc, double bottom, double top, int NY,
                                     int ybc, double front, double back, int NZ, int zbc, double lambda,
                                     const double *u, const double *gu, double *gf, double *gbcl, double *gbcr,
                                     double *gbcb, double *gbct, double *gbcf, double *gbce, double *glambda,
                                     void *stream)
{
    const double dom[6] = {left, right, bottom, top, front, back};
    const int N[3] = {NX, NY, NZ}, bc[3] = {xbc, ybc, zbc};
    double *const faces[6] = {gbcl, gbcr, gbcb, gbct, gbcf, gbce};
    return chz::adjoint(dom, N, bc, lambda, u, gu, gf, faces, glambda, (hipStream_t)stream);
}
