// chz_solver.h -- host side of the Helmholtz solves.
#pragma once
#include <hip/hip_runtime.h>

#include "chz_context.h"

namespace chz {

// Validated geometry of one solve (reference argument list, cuda_helmholtz_solver.h:5-11).
struct SolveGeom {
    int N[3], bc[3], s[3], e[3], fwd[3], inv[3];
    double ih[3], ih2[3];
    double lambda;
    bool singular;
    double ratio;
    const double *faces[6];   // bcl, bcr, bcb, bct, bcf, bce (null where unused)
};

long long ws_row(int NX, size_t esz);
int make_geom(SolveGeom &g, const double dom[6], const int N[3], const int bc[3], double lambda,
              const double *const faces[6]);
const double *solver_eig(Context &ctx, const SolveGeom &g);
void fill_solve_args(SolveArgs &sa, const SolveGeom &g, const double *f, const double *eig);

// passes 1-4 of a solve into workspace ex (rhs read from f)
template <typename T>
int solve_core(Context &ctx, const SolveGeom &g, double *f, T *ex, hipStream_t st, const double *U, const double *pertrb);
int periodic_copies(const SolveGeom &g, double *f, hipStream_t st);

int helmholtz_mixed(const double dom[6], const int N[3], const int bc[3], const double *const faces[6], double lambda,
                    double *f, int refine, hipStream_t st);
int helmholtz_fp64(const double dom[6], const int N[3], const int bc[3], const double *const faces[6], double lambda,
                   double *f, hipStream_t st);

}  // namespace chz
