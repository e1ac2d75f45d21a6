// chz_solver.h -- host side of the Helmholtz solves.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

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
    int stencil;              // 7 (the reference's operator) or 19 (its unlaunched compact variant)
};

long long ws_row(int NX, size_t esz);
int make_geom(SolveGeom &g, const double dom[6], const int N[3], const int bc[3], double lambda,
              const double *const faces[6]);
// slot eigenvalue tables x | y | z of a geometry (host FP64) and their cached device copy
std::vector<double> eig_tables(const SolveGeom &g);
// max|e + lambda| / min|e + lambda| over the modes of the 7-point operator (inf if singular)
double spectral_condition(const SolveGeom &g);
const double *solver_eig(Context &ctx, const SolveGeom &g);
void fill_solve_args(SolveArgs &sa, const SolveGeom &g, const double *f, const double *eig);

// One pass of a solve: the kernel (BatchOp), its launch configuration and
// its arguments; launched alone (launch_spec) or grouped with the same pass
// of other problems of the same size class (batched solves).
struct PassSpec {
    int op, prec, ept, nb;   // nb: line groups (tile index b)
    int N, role;             // cell count along the pass axis, configuration role
    const char *name;
    BatchDesc d;
};

// Device tables a solve's passes read: the twiddles of the x, y, z cell
// counts (TwOffsets layout, precision of the passes) and the slot eigenvalues
// x | y | z.  The builders take the context's cached tables when none are given.
struct SolveTables {
    const void *tw[3];
    const double *eig;
};

// passes 1-4 of a solve into workspace ex (rhs read from f; U: residual of U instead)
// (u_f32: the iterate U is FP32, see PassArgs::u_f32)
template <typename T>
int solve_passes(Context &ctx, const SolveGeom &g, double *f, T *ex, const double *U, const double *pertrb,
                 std::vector<PassSpec> &out, const SolveTables *tabs = nullptr, int u_f32 = 0);
// inverse x pass into out (FP64: the solution; FP32: out = Uin + delta)
// (u_f32: PassArgs::u_f32 bits, Uin / out FP32 where set)
template <typename T>
int final_pass(Context &ctx, const SolveGeom &g, const T *ex, const double *Uin, double *out, std::vector<PassSpec> &v,
               const SolveTables *tabs = nullptr, int u_f32 = 0);
// the FP32 residual and update passes of a mixed solve of g run on the
// lean-I/O kernels (single-problem or batched), the only ones that read and
// write FP32 iterates
bool iterate_f32_ok(Context &ctx, const SolveGeom &g, bool batched);
int launch_spec(Context &ctx, const PassSpec &q, hipStream_t st);
int periodic_copies(const SolveGeom &g, double *f, hipStream_t st);

// free the batched plans kept by chz_batch_solve_cached (waits for the device)
void release_cached_plans();

int helmholtz_mixed(const double dom[6], const int N[3], const int bc[3], const double *const faces[6], double lambda,
                    double *f, int refine, hipStream_t st);
int helmholtz_fp64(const double dom[6], const int N[3], const int bc[3], const double *const faces[6], double lambda,
                   double *f, hipStream_t st, int stencil = 7);

}  // namespace chz
