// chz_hw3crt.cpp -- the reference's CPU entry points, hw3crt_ and solver_3d
// (cudahelmholtz/hw3crt_wrapper.h:1-19, hw3crt_wrapper.cpp:5-33), served by the
// GPU solve.  Arrays are host arrays in FISHPACK's layout:
//   F(LDIMF, MDIMF, N+1)          node (i, j, k) at i + LDIMF (j + MDIMF k)
//   BDXS, BDXF (MDIMF, N+1)       x faces, (j, k) at j + MDIMF k
//   BDYS, BDYF (LDIMF, N+1)       y faces, (i, k) at i + LDIMF k
//   BDZS, BDZF (LDIMF, M+1)       z faces, (i, j) at i + LDIMF j
// The node box and the faces the codes use are packed, copied to the device,
// solved by the FP64 path and copied back into F; the rest of F is untouched.
// IERROR as documented for FISHPACK's hw3crt: 1 xs >= xf, 2 L < 3, 3 LBDCND
// outside 0..4, 4 ys >= yf, 5 M < 3, 6 MBDCND, 7 zs >= zf, 8 N < 3, 9 NBDCND,
// 10 LDIMF < L+1, 11 MDIMF < M+1, 12 ELMBDA > 0 (a warning: the solution is
// still computed); plus 13 for an odd cell count (the GPU transforms take even ones)
// (odd, or N/2 not of the form 2^a 3^b 5^c) and 20 for a device failure.
// PERTRB: for ELMBDA = 0 with periodic / Neumann codes only the problem is
// singular; PERTRB is the weighted mean of the folded rhs (the constant that
// makes it consistent) and the solution returned has a zero (0,0,0) mode.
// W (FISHPACK's workspace) is not used.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "chz_context.h"
#include "chz_solver.h"
#include "hw3crt_wrapper.h"

namespace chz {

static bool supported_cells(int n) { return n >= 2 && !(n & 1); }

// device buffers of one call, released on every exit path
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p) (void)hipFree(q);
    }
    double *get(size_t n)
    {
        void *q = nullptr;
        if (hipMalloc(&q, n * sizeof(double)) != hipSuccess) return nullptr;
        p.push_back(q);
        return reinterpret_cast<double *>(q);
    }
};

static int hw3crt_gpu(double xs, double xf, int l, int lbdcnd, const double *bdxs, const double *bdxf, double ys, double yf,
               int m, int mbdcnd, const double *bdys, const double *bdyf, double zs, double zf, int n, int nbdcnd,
               const double *bdzs, const double *bdzf, double elmbda, int ldimf, int mdimf, double *f, double *pertrb)
{
    *pertrb = 0.0;
    if (xs >= xf) return 1;
    if (l < 3) return 2;
    if (lbdcnd < 0 || lbdcnd > 4) return 3;
    if (ys >= yf) return 4;
    if (m < 3) return 5;
    if (mbdcnd < 0 || mbdcnd > 4) return 6;
    if (zs >= zf) return 7;
    if (n < 3) return 8;
    if (nbdcnd < 0 || nbdcnd > 4) return 9;
    if (ldimf < l + 1) return 10;
    if (mdimf < m + 1) return 11;
    if (!supported_cells(l) || !supported_cells(m) || !supported_cells(n)) {
        set_error("hw3crt_: cell counts must be even (the real-FFT split of the spectral solve)");
        return 13;
    }
    const int NX = l, NY = m, NZ = n;
    const long long fx = NX + 1, fy = NY + 1, fz = NZ + 1;
    const int bc[3] = {lbdcnd, mbdcnd, nbdcnd};
    // pack the node box of F
    std::vector<double> hf((size_t)(fx * fy * fz));
    for (long long k = 0; k < fz; k++)
        for (long long j = 0; j < fy; j++)
            memcpy(&hf[(size_t)(fx * (j + fy * k))], f + (size_t)(ldimf * (j + (long long)mdimf * k)), sizeof(double) * fx);
    // faces the codes use: low face for codes 3, 4; high face for codes 2, 3
    const double *src[6] = {bdxs, bdxf, bdys, bdyf, bdzs, bdzf};
    const long long n1[3] = {fy, fx, fx}, n2[3] = {fz, fz, fy}, ld[3] = {mdimf, ldimf, ldimf};
    std::vector<double> hface[6];
    for (int a = 0; a < 3; a++)
        for (int side = 0; side < 2; side++) {
            const bool need = side == 0 ? (bc[a] == 3 || bc[a] == 4) : (bc[a] == 2 || bc[a] == 3);
            if (!need) continue;
            const double *s = src[2 * a + side];
            if (!s) {
                set_error("hw3crt_: missing Neumann boundary array");
                return 20;
            }
            std::vector<double> &h = hface[2 * a + side];
            h.resize((size_t)(n1[a] * n2[a]));
            for (long long q = 0; q < n2[a]; q++) memcpy(&h[(size_t)(n1[a] * q)], s + (size_t)(ld[a] * q), sizeof(double) * n1[a]);
        }

    DevBufs d;
    double *df = d.get(hf.size());
    double *dface[6] = {};
    for (int i = 0; i < 6; i++)
        if (!hface[i].empty()) dface[i] = d.get(hface[i].size());
    double *dacc = d.get(4);
    if (!df || !dacc) {
        fail(ERR_ALLOC, "hw3crt_: device allocation failed");
        return 20;
    }
    hipStream_t st = nullptr;
    bool ok = hipMemcpy(df, hf.data(), sizeof(double) * hf.size(), hipMemcpyHostToDevice) == hipSuccess;
    for (int i = 0; i < 6; i++)
        if (ok && dface[i])
            ok = hipMemcpy(dface[i], hface[i].data(), sizeof(double) * hface[i].size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        fail(ERR_DEVICE, "hw3crt_: host to device copy failed");
        return 20;
    }

    const double dom[6] = {xs, xf, ys, yf, zs, zf};
    const int N[3] = {NX, NY, NZ};
    const double *faces[6] = {dface[0], dface[1], dface[2], dface[3], dface[4], dface[5]};
    SolveGeom g;
    if (make_geom(g, dom, N, bc, elmbda, faces) != OK) return 20;
    if (g.singular) {
        // PERTRB from the folded rhs, before the solve overwrites f
        SolveArgs sa;
        fill_solve_args(sa, g, df, nullptr);
        double acc = 0.0;
        if (launch_fold_sum(sa, dacc, st) != hipSuccess || hipMemcpy(&acc, dacc, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
            fail(ERR_DEVICE, "hw3crt_: PERTRB reduction failed");
            return 20;
        }
        *pertrb = acc * mode0_inv_wsum(sa);
    }
    if (helmholtz_fp64(dom, N, bc, faces, elmbda, df, st) != OK) return 20;
    if (hipMemcpy(hf.data(), df, sizeof(double) * hf.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("hw3crt_: device error");
        return 20;
    }
    for (long long k = 0; k < fz; k++)
        for (long long j = 0; j < fy; j++)
            memcpy(f + (size_t)(ldimf * (j + (long long)mdimf * k)), &hf[(size_t)(fx * (j + fy * k))], sizeof(double) * fx);
    return elmbda > 0.0 ? 12 : 0;
}

}  // namespace chz

extern "C" {

// hw3crt_wrapper.h:1-9 (Fortran calling convention: every argument by pointer)
void hw3crt_(double *xs, double *xf, int *l, int *lbdcnd, double *bdxs, double *bdxf, double *ys, double *yf, int *m,
             int *mbdcnd, double *bdys, double *bdyf, double *zs, double *zf, int *n, int *nbdcnd, double *bdzs,
             double *bdzf, double *elmbda, int *ldimf, int *mdimf, double *f, double *pertrb, int *ierror, double *w)
{
    (void)w;
    *ierror = chz::hw3crt_gpu(*xs, *xf, *l, *lbdcnd, bdxs, bdxf, *ys, *yf, *m, *mbdcnd, bdys, bdyf, *zs, *zf, *n,
                              *nbdcnd, bdzs, bdzf, *elmbda, *ldimf, *mdimf, f, pertrb);
}

}  // extern "C"

// hw3crt_wrapper.cpp:5-33: packed node array (LDIMF = NX+1, MDIMF = NY+1), prints the state code
void solver_3d(double xs, double xf, int NX, int xbc, double *bcxs, double *bcxf, double ys, double yf, int NY,
               int ybc, double *bcys, double *bcyf, double zs, double zf, int NZ, int zbc, double *bczs, double *bczf,
               double lambda, double *f)
{
    double pertrb = 0.0;
    const int ierror = chz::hw3crt_gpu(xs, xf, NX, xbc, bcxs, bcxf, ys, yf, NY, ybc, bcys, bcyf, zs, zf, NZ, zbc, bczs,
                                       bczf, lambda, NX + 1, NY + 1, f, &pertrb);
    printf("State code : %d\n", ierror);
}
