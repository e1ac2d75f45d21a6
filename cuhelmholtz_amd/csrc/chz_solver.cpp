// chz_solver.cpp -- FFT-diagonalised 7-point Helmholtz solve (plain FP64).
//
// The reference (cudahelmholtz/cuda_helmholtz_solver.cu:33-380) folds the
// boundary data into f, gathers the unknown box into a padded buffer it
// allocates per call, runs a forward do_transform (12 sweeps with its
// transposes), an eigenvalue divide kernel, an inverse do_transform, a store
// kernel and the periodic copies, synchronising after most steps.  Here the
// same arithmetic runs as 5 stream-ordered passes over a persistent workspace:
//   1  rows of f:   folds + gather + forward x  -> ex
//   2  y columns:   forward y                    (in place)
//   3  z columns:   forward z, divide, inverse z (in place, one pass)
//   4  y columns:   inverse y                    (in place)
//   5  rows of ex:  inverse x, x ratio          -> f (unknown box)
// plus face kernels for the periodic duplicate nodes: 80 bytes per cell.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "chz_context.h"
#include "chz_solver.h"
#include "chz_transform.h"

namespace chz {

// Row length of the solve workspace: the N+2 slots of the padded layout rounded
// up to whole 128-byte lines, so every row and plane starts on a cache line
// and the column tiles of the y/z passes touch whole lines only.
long long ws_row(int NX, size_t esz)
{
    const long long per_line = 128 / (long long)esz;
    return ((NX + 2 + per_line - 1) / per_line) * per_line;
}

// bc code -> unknown window and transform kinds (cuda_helmholtz_solver.cu:78-172)
static bool bc_kinds(int bc, int N, int *s, int *e, int *fwd, int *inv)
{
    switch (bc) {
        case 0: *s = 0; *e = N - 1; *fwd = K_R2C; *inv = K_C2R; return true;
        case 1: *s = 1; *e = N - 1; *fwd = K_DST_1; *inv = K_DST_1; return true;
        case 2: *s = 1; *e = N; *fwd = K_DST_3; *inv = K_DST_2; return true;
        case 3: *s = 0; *e = N; *fwd = K_DCT_1; *inv = K_DCT_1; return true;
        case 4: *s = 0; *e = N - 1; *fwd = K_DCT_3; *inv = K_DCT_2; return true;
    }
    return false;
}

int make_geom(SolveGeom &g, const double dom[6], const int N[3], const int bc[3], double lambda,
              const double *const faces[6])
{
    static const char *ax = "xyz";
    for (int a = 0; a < 3; a++) {
        g.N[a] = N[a];
        g.bc[a] = bc[a];
        if (!bc_kinds(bc[a], N[a], &g.s[a], &g.e[a], &g.fwd[a], &g.inv[a]))
            return fail(ERR_INVALID, std::string("bad boundary condition along ") + ax[a]);
        if (N[a] < 2 || (N[a] & 1)) return fail(ERR_INVALID, std::string("cell count must be even and >= 2 along ") + ax[a]);
        const double len = dom[2 * a + 1] - dom[2 * a];
        if (!(len > 0)) return fail(ERR_INVALID, std::string("empty domain along ") + ax[a]);
        g.ih[a] = N[a] / len;                         // cuda_helmholtz_solver.cu:44-46
        g.ih2[a] = g.ih[a] * g.ih[a];
        const bool need_lo = bc[a] == 3 || bc[a] == 4, need_hi = bc[a] == 2 || bc[a] == 3;
        if ((need_lo && !faces[2 * a]) || (need_hi && !faces[2 * a + 1]))
            return fail(ERR_INVALID, std::string("missing Neumann boundary array along ") + ax[a]);
        g.faces[2 * a] = need_lo ? faces[2 * a] : nullptr;
        g.faces[2 * a + 1] = need_hi ? faces[2 * a + 1] : nullptr;
    }
    g.stencil = 7;
    g.lambda = lambda;
    g.singular = lambda == 0.0 && (bc[0] == 0 || bc[0] == 3) && (bc[1] == 0 || bc[1] == 3) && (bc[2] == 0 || bc[2] == 3);
    // ratio = 1 / prod(N_a/2 * (2 if periodic)), cuda_helmholtz_solver.cu:211-215
    double r = 1.0;
    for (int a = 0; a < 3; a++) r *= (double)(N[a] / 2) * (bc[a] == 0 ? 2.0 : 1.0);
    g.ratio = 1.0 / r;
    return OK;
}

// Slot eigenvalue tables (2 cos(m theta + shift) - 2)/h^2 of the three axes,
// host FP64 (cuda_helmholtz_solver.cu:58-73, _cuda_helmholtz_.inl:413-418).
std::vector<double> eig_tables(const SolveGeom &g)
{
    const double pi = 3.14159265358979323846;
    std::vector<double> t;
    for (int a = 0; a < 3; a++) {
        const int N = g.N[a];
        double theta = pi / N, shift = 0;
        if (g.bc[a] == 0) theta *= 2;
        if (g.bc[a] == 2) shift = -0.5 * pi / N;
        if (g.bc[a] == 4) shift = 0.5 * pi / N;
        for (int sl = 0; sl < N + 2; sl++) {
            const int m = g.bc[a] == 0 ? sl / 2 : sl;
            t.push_back((2 * std::cos(m * theta + shift) - 2) * g.ih2[a]);
        }
    }
    return t;
}

// Spectral condition number max|e + lambda| / min|e + lambda| of the 7-point
// operator, e running over the sums e_x + e_y + e_z of the axis eigenvalues
// of every mode (the modes of each code: the same angles as eig_tables).
// The minimum comes from a nearest-value search of the sorted z eigenvalues
// for every (x, y) pair; infinite for a singular operator.
double spectral_condition(const SolveGeom &g)
{
    const double pi = 3.14159265358979323846;
    std::vector<double> e[3];
    for (int a = 0; a < 3; a++) {
        const int N = g.N[a];
        int m0 = 0, m1 = N;
        double scale = pi / N, shift = 0;
        switch (g.bc[a]) {
            case 0: m1 = N / 2; scale = 2 * pi / N; break;
            case 1: m0 = 1; m1 = N - 1; break;
            case 2: m0 = 1; shift = -0.5 * pi / N; break;
            case 4: m1 = N - 1; shift = 0.5 * pi / N; break;
            default: break;
        }
        for (int m = m0; m <= m1; m++) e[a].push_back((2 * std::cos(m * scale + shift) - 2) * g.ih2[a]);
        std::sort(e[a].begin(), e[a].end());
    }
    double lo = INFINITY;
    for (double x : e[0])
        for (double y : e[1]) {
            const double want = -(x + y + g.lambda);
            auto it = std::lower_bound(e[2].begin(), e[2].end(), want);
            if (it != e[2].end()) lo = std::min(lo, std::fabs(x + y + *it + g.lambda));
            if (it != e[2].begin()) lo = std::min(lo, std::fabs(x + y + *(it - 1) + g.lambda));
        }
    const double s0 = e[0].front() + e[1].front() + e[2].front() + g.lambda;
    const double s1 = e[0].back() + e[1].back() + e[2].back() + g.lambda;
    const double hi = std::max(std::fabs(s0), std::fabs(s1));
    return lo > 0 ? hi / lo : INFINITY;
}

const double *solver_eig(Context &ctx, const SolveGeom &g)
{
    std::vector<double> t = eig_tables(g);
    std::string key(reinterpret_cast<const char *>(t.data()), t.size() * sizeof(double));
    void *d = ctx.cached_blob(key, t.data(), t.size() * sizeof(double));
    return reinterpret_cast<const double *>(d);
}

void fill_solve_args(SolveArgs &sa, const SolveGeom &g, const double *f, const double *eig)
{
    memset(&sa, 0, sizeof(sa));
    sa.f = f;
    sa.bcl = g.faces[0]; sa.bcr = g.faces[1];
    sa.bcb = g.faces[2]; sa.bct = g.faces[3];
    sa.bcf = g.faces[4]; sa.bce = g.faces[5];
    sa.NX = g.N[0]; sa.NY = g.N[1]; sa.NZ = g.N[2];
    for (int a = 0; a < 3; a++) {
        sa.bc[a] = g.bc[a];
        sa.s[a] = g.s[a];
        sa.e[a] = g.e[a];
        sa.ih[a] = g.ih[a];
        sa.ih2[a] = g.ih2[a];
    }
    sa.eig = eig;
    sa.lambda = g.lambda;
    sa.singular = g.singular ? 1 : 0;
    if (g.stencil == 19) {
        // L19 = sum e_a + sum_{a<b} (h_a^2 + h_b^2)/12 e_a e_b, B = 1 + sum h_a^2/12 e_a
        const double hx = 1.0 / g.ih2[0], hy = 1.0 / g.ih2[1], hz = 1.0 / g.ih2[2];
        sa.stencil19 = 1;
        sa.c19[0] = (hx + hy) / 12.0;
        sa.c19[1] = (hx + hz) / 12.0;
        sa.c19[2] = (hy + hz) / 12.0;
        sa.c19[3] = hx / 12.0;
        sa.c19[4] = hy / 12.0;
        sa.c19[5] = hz / 12.0;
    }
}

static int check_spec(const PassSpec &q, long long svol, long long dvol, const char *what)
{
    const bool rows = q.op == B_ROWS || q.op == B_XFWD || q.op == B_XRES || q.op == B_XUPD;
    if (!pass_in_bounds(q.d.p, rows, q.nb, svol, dvol)) return fail(ERR_INVALID, std::string(what) + ": pass geometry outside the buffers");
    return OK;
}

// Passes 1-4 of a solve on workspace ex (precision of T, rows padded to
// 128 bytes): pass 1 reads f (node layout, FP64) and applies the boundary
// folds, or -- with U set -- forms the FP64 residual f_fold - A U (a
// refinement step of the mixed-precision solve).
template <typename T>
int solve_passes(Context &ctx, const SolveGeom &g, double *f, T *ex, const double *U, const double *pertrb,
                 std::vector<PassSpec> &out, const SolveTables *tabs, int u_f32)
{
    const int prec = sizeof(T) == 8 ? 64 : 32;
    const int NX = g.N[0], NY = g.N[1], NZ = g.N[2];
    AxisCfg cx, cy, cz;
    const int zrole = g.bc[2] == 0 ? ROLE_ZMID : ROLE_ZMIDN;
    if (!ctx.config(NX, prec, ROLE_XFWD, &cx) || !ctx.config(NY, prec, ROLE_COLS, &cy) || !ctx.config(NZ, prec, zrole, &cz))
        return fail(ERR_UNSUPPORTED, "no launch configuration for this cell count (LDS or thread limits)");
    const void *twx = tabs ? tabs->tw[0] : ctx.twiddles(NX, prec), *twy = tabs ? tabs->tw[1] : ctx.twiddles(NY, prec),
               *twz = tabs ? tabs->tw[2] : ctx.twiddles(NZ, prec);
    const double *eig = tabs ? tabs->eig : solver_eig(ctx, g);
    if (!twx || !twy || !twz || !eig) return fail(ERR_ALLOC, "table allocation failed");
    if (U && prec != 32) return fail(ERR_INVALID, "residual pass is FP32 only");

    const long long fx = NX + 1, fy = NY + 1, P0 = ws_row(NX, sizeof(T)), P1 = NY + 2;
    const long long fvol = fx * fy * (NZ + 1), vol = P0 * P1 * (NZ + 2);
    int sxl, sxh, syl, syh, szl, szh;   // spectral windows (forward outputs)
    out_window(g.fwd[0], NX, &sxl, &sxh);
    out_window(g.fwd[1], NY, &syl, &syh);
    out_window(g.fwd[2], NZ, &szl, &szh);
    SolveArgs sa;
    fill_solve_args(sa, g, f, eig);
    sa.pertrb = pertrb;
    sa.pert_scale = pertrb ? mode0_inv_wsum(sa) : 0.0;
    int rc;

    // pass 1: folds + forward x, f rows (j,k) of the unknown box -> ex
    {
        PassSpec q{};
        q.op = U ? B_XRES : B_XFWD;
        q.prec = prec; q.ept = cx.EPT; q.N = NX; q.role = ROLE_XFWD;
        q.name = U ? "xres" : "xfwd";
        PassArgs &p = q.d.p;
        fill_pass(p, cx, prec, g.fwd[0], twx);
        p.src = f; p.dst = ex;
        p.s_sa = fx; p.s_sb = fx * fy; p.s_sp = 1; p.s_rowlen = (int)fx;
        p.d_sa = P0; p.d_sb = P0 * P1; p.d_sp = 1; p.d_rowlen = (int)P0;
        p.a0 = g.s[1]; p.na = g.e[1] - g.s[1] + 1; p.b0 = g.s[2];
        p.in_lo = g.s[0]; p.in_hi = g.e[0];
        p.st_lo = sxl; p.st_hi = sxh;
        p.out_lo = sxl; p.out_hi = sxh;
        q.nb = g.e[2] - g.s[2] + 1;
        q.d.sa = sa;
        q.d.u_in = U;
        p.u_f32 = U ? (u_f32 & 1) : 0;
        // the residual reads the z-neighbour rows of U: step through z first so
        // they are the rows of the tiles running at the same time (L2 hits)
        p.bfast = U ? 1 : 0;
        if ((rc = check_spec(q, fvol, vol, "pass 1"))) return rc;
        out.push_back(q);
    }
    // pass 2: forward y, columns x in the x spectral window, z in the unknown window
    {
        PassSpec q{};
        q.op = B_COLS; q.prec = prec; q.ept = cy.EPT; q.N = NY; q.role = ROLE_COLS; q.name = "yfwd";
        PassArgs &p = q.d.p;
        fill_pass(p, cy, prec, g.fwd[1], twy);
        p.src = ex; p.dst = ex;
        p.s_sa = 1; p.s_sb = P0 * P1; p.s_sp = P0;
        p.d_sa = 1; p.d_sb = P0 * P1; p.d_sp = P0;
        p.a0 = sxl; p.na = sxh - sxl + 1; p.b0 = g.s[2];
        p.in_lo = g.s[1]; p.in_hi = g.e[1];
        p.st_lo = syl; p.st_hi = syh;
        p.out_lo = syl; p.out_hi = syh;
        q.nb = g.e[2] - g.s[2] + 1;
        if ((rc = check_spec(q, vol, vol, "pass 2"))) return rc;
        out.push_back(q);
    }
    // pass 3: forward z, divide, inverse z
    {
        PassSpec q{};
        q.op = B_ZMID; q.prec = prec; q.ept = cz.EPT; q.N = NZ; q.role = zrole; q.name = "zmid";
        PassArgs &p = q.d.p;
        fill_pass(p, cz, prec, g.fwd[2], twz);
        p.src = ex; p.dst = ex;
        p.s_sa = 1; p.s_sb = P0; p.s_sp = P0 * P1;
        p.d_sa = 1; p.d_sb = P0; p.d_sp = P0 * P1;
        p.a0 = sxl; p.na = sxh - sxl + 1; p.b0 = syl;
        p.in_lo = g.s[2]; p.in_hi = g.e[2];
        p.st_lo = g.s[2]; p.st_hi = g.e[2];
        p.out_lo = szl; p.out_hi = szh;
        q.d.sa = sa;
        q.d.sa.kind2 = g.inv[2];
        q.d.sa.out2_lo = g.s[2];
        q.d.sa.out2_hi = g.e[2];
        q.nb = syh - syl + 1;
        if ((rc = check_spec(q, vol, vol, "pass 3"))) return rc;
        out.push_back(q);
    }
    // pass 4: inverse y
    {
        PassSpec q{};
        q.op = B_COLS; q.prec = prec; q.ept = cy.EPT; q.N = NY; q.role = ROLE_COLS; q.name = "yinv";
        PassArgs &p = q.d.p;
        fill_pass(p, cy, prec, g.inv[1], twy);
        p.src = ex; p.dst = ex;
        p.s_sa = 1; p.s_sb = P0 * P1; p.s_sp = P0;
        p.d_sa = 1; p.d_sb = P0 * P1; p.d_sp = P0;
        p.a0 = sxl; p.na = sxh - sxl + 1; p.b0 = g.s[2];
        p.in_lo = syl; p.in_hi = syh;
        p.st_lo = g.s[1]; p.st_hi = g.e[1];
        p.out_lo = g.s[1]; p.out_hi = g.e[1];
        q.nb = g.e[2] - g.s[2] + 1;
        if ((rc = check_spec(q, vol, vol, "pass 4"))) return rc;
        out.push_back(q);
    }
    return OK;
}

// Last pass: inverse x of the workspace rows and the x ratio.  FP64: stored
// into f over the unknown box (out == f, Uin == nullptr).  FP32: out = Uin +
// ratio * delta in FP64 (Uin == nullptr: out = ratio * delta).
template <typename T>
int final_pass(Context &ctx, const SolveGeom &g, const T *ex, const double *Uin, double *out, std::vector<PassSpec> &v,
               const SolveTables *tabs, int u_f32)
{
    const int prec = sizeof(T) == 8 ? 64 : 32;
    const int NX = g.N[0], NY = g.N[1], NZ = g.N[2];
    AxisCfg cx;
    if (!ctx.config(NX, prec, ROLE_ROWS, &cx)) return fail(ERR_UNSUPPORTED, "unsupported cell count");
    const void *twx = tabs ? tabs->tw[0] : ctx.twiddles(NX, prec);
    if (!twx) return fail(ERR_ALLOC, "table allocation failed");
    const long long fx = NX + 1, fy = NY + 1, P0 = ws_row(NX, sizeof(T)), P1 = NY + 2;
    const long long fvol = fx * fy * (NZ + 1), vol = P0 * P1 * (NZ + 2);
    int sxl, sxh;
    out_window(g.fwd[0], NX, &sxl, &sxh);
    PassSpec q{};
    q.op = prec == 64 ? B_ROWS : B_XUPD;
    q.prec = prec; q.ept = cx.EPT; q.N = NX; q.role = ROLE_ROWS;
    q.name = Uin ? "xupd" : "xinv";
    PassArgs &p = q.d.p;
    fill_pass(p, cx, prec, g.inv[0], twx);
    p.src = ex; p.dst = out;
    p.s_sa = P0; p.s_sb = P0 * P1; p.s_sp = 1; p.s_rowlen = (int)P0;
    p.d_sa = fx; p.d_sb = fx * fy; p.d_sp = 1; p.d_rowlen = (int)fx;
    p.a0 = g.s[1]; p.na = g.e[1] - g.s[1] + 1; p.b0 = g.s[2];
    p.in_lo = sxl; p.in_hi = sxh;
    p.st_lo = g.s[0]; p.st_hi = g.e[0];
    p.out_lo = g.s[0]; p.out_hi = g.e[0];
    p.scale = g.ratio;
    p.u_f32 = u_f32;
    q.nb = g.e[2] - g.s[2] + 1;
    q.d.u_in = Uin;
    q.d.u_out = out;
    if (prec == 64 && Uin) return fail(ERR_INVALID, "the FP64 store pass writes the solution directly");
    int rc;
    if ((rc = check_spec(q, vol, fvol, "last pass"))) return rc;
    v.push_back(q);
    return OK;
}

// Launch one pass of a single solve with the single-problem kernels.
int launch_spec(Context &ctx, const PassSpec &q, hipStream_t st)
{
    if (ctx.timing()) ctx.timing_mark(q.name, st);
    const PassArgs &p = q.d.p;
    const bool f64 = q.prec == 64;
    hipError_t e = hipErrorInvalidValue;
    const FixedEntry *fx = fixed_kernel(q.prec, q.op, p.N, q.ept, p.TPL, p.L, p.plan, &p, pass_kind_sel(q.op, p, &q.d.sa));
    // an FP32 iterate is read / written by the lean-I/O kernels only
    if (p.u_f32 && (!fx || fx->form != F_LINE))
        return fail(ERR_INVALID, std::string("internal: FP32 iterate on the ") + q.name + " pass without a lean-I/O kernel");
    if (fx) {
        e = launch_fixed(fx, p, &q.d.sa, q.d.u_in, q.d.u_out, q.prec, q.nb, st);
        if (e != hipSuccess) return fail(ERR_DEVICE, std::string(q.name) + " pass launch: " + hipGetErrorString(e));
        return OK;
    }
    switch (q.op) {
        case B_ROWS:
        case B_COLS: e = launch_axis_pass(p, q.prec, q.ept, q.op == B_ROWS, q.nb, st); break;
        case B_XFWD:
            e = launch_solve_pass(f64 ? solve_xfwd_kernel_f64(q.ept) : solve_xfwd_kernel_f32(q.ept), p, q.d.sa, q.prec, q.nb, st);
            break;
        case B_ZMID:
            e = launch_solve_pass(f64 ? solve_zmid_kernel_f64(q.ept) : solve_zmid_kernel_f32(q.ept), p, q.d.sa, q.prec, q.nb, st);
            break;
        case B_XRES:
            if (!f64) e = launch_pass_sa_ptr(solve_xres_kernel_f32(q.ept), p, q.d.sa, q.d.u_in, q.prec, q.nb, st);
            break;
        case B_XUPD:
            if (!f64) e = launch_pass_ptrs(solve_xinv_upd_kernel_f32(q.ept), p, q.d.u_in, q.d.u_out, q.prec, q.nb, st);
            break;
    }
    if (e != hipSuccess) return fail(ERR_DEVICE, std::string(q.name) + " pass launch: " + hipGetErrorString(e));
    return OK;
}

static int run_specs(Context &ctx, const std::vector<PassSpec> &v, hipStream_t st)
{
    int rc;
    for (const PassSpec &q : v)
        if ((rc = launch_spec(ctx, q, st))) return rc;
    return OK;
}

template int solve_passes<double>(Context &, const SolveGeom &, double *, double *, const double *, const double *,
                                  std::vector<PassSpec> &, const SolveTables *, int);
template int solve_passes<float>(Context &, const SolveGeom &, double *, float *, const double *, const double *,
                                 std::vector<PassSpec> &, const SolveTables *, int);
template int final_pass<double>(Context &, const SolveGeom &, const double *, const double *, double *, std::vector<PassSpec> &,
                                const SolveTables *, int);
template int final_pass<float>(Context &, const SolveGeom &, const float *, const double *, double *, std::vector<PassSpec> &,
                               const SolveTables *, int);

bool iterate_f32_ok(Context &ctx, const SolveGeom &g, bool batched)
{
    AxisCfg cr, cu;
    const int N = g.N[0];
    if (!ctx.config(N, 32, ROLE_XFWD, &cr) || !ctx.config(N, 32, ROLE_ROWS, &cu)) return false;
    if (batched)
        return batch_line_kernel(32, B_XRES, N, cr.EPT, cr.TPL, cr.L, cr.plan) &&
               batch_line_kernel(32, B_XUPD, N, cu.EPT, cu.TPL, cu.L, cu.plan);
    return line_kernel(32, B_XRES, N, cr.EPT, cr.TPL, cr.L, cr.plan) && line_kernel(32, B_XUPD, N, cu.EPT, cu.TPL, cu.L, cu.plan);
}

int periodic_copies(const SolveGeom &g, double *f, hipStream_t st)
{
    const int NX = g.N[0], NY = g.N[1], NZ = g.N[2];
    for (int a = 0; a < 3; a++) {
        if (g.bc[a] != 0) continue;
        const long long n = a == 0 ? (long long)(NY + 1) * (NZ + 1) : a == 1 ? (long long)(NX + 1) * (NZ + 1)
                                                                             : (long long)(NX + 1) * (NY + 1);
        int nx = NX, ny = NY, nz = NZ, axis = a;
        void *args[] = {&f, &nx, &ny, &nz, &axis};
        hipError_t e = hipLaunchKernel(periodic_copy_kernel(), dim3((unsigned)((n + 255) / 256)), dim3(256), args, 0, st);
        if (e != hipSuccess) return fail(ERR_DEVICE, std::string("periodic copy launch: ") + hipGetErrorString(e));
    }
    return OK;
}

int helmholtz_fp64(const double dom[6], const int N[3], const int bc[3], const double *const faces[6], double lambda,
                   double *f, hipStream_t st, int stencil)
{
    if (!f) return fail(ERR_INVALID, "null solution array");
    if (stencil != 7 && stencil != 19) return fail(ERR_INVALID, "stencil must be 7 or 19");
    SolveGeom g;
    int rc = make_geom(g, dom, N, bc, lambda, faces);
    if (rc) return rc;
    g.stencil = stencil;
    Context &ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex());
    WsLease ws(ctx, st);
    const size_t bytes = sizeof(double) * (size_t)ws_row(N[0], sizeof(double)) * (N[1] + 2) * (N[2] + 2);
    double *ex = reinterpret_cast<double *>(ws.get(bytes, 0));
    if (!ex) return fail(ERR_ALLOC, "workspace allocation failed");
    std::vector<PassSpec> v;
    if ((rc = solve_passes<double>(ctx, g, f, ex, nullptr, nullptr, v)) || (rc = final_pass<double>(ctx, g, ex, nullptr, f, v)))
        return rc;
    PassTimer timer(ctx, st);
    if ((rc = run_specs(ctx, v, st))) return rc;
    timer.mark("periodic");
    return periodic_copies(g, f, st);
}

// FP32 transforms and divide with FP64 iterative refinement: an FP32 solve
// of the folded rhs, then `refine` times: FP64 residual -> FP32 correction
// solve -> FP64 update.  Twiddles and eigenvalues are generated in FP64 and
// rounded to FP32; the residual and the update are the only FP64 arithmetic.
int helmholtz_mixed(const double dom[6], const int N[3], const int bc[3], const double *const faces[6], double lambda,
                    double *f, int refine, hipStream_t st)
{
    if (!f) return fail(ERR_INVALID, "null solution array");
    if (refine < 0 || refine > 16) return fail(ERR_INVALID, "refinement count must be in 0..16");
    SolveGeom g;
    int rc = make_geom(g, dom, N, bc, lambda, faces);
    if (rc) return rc;
    Context &ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex());
    WsLease ws(ctx, st);
    const size_t wbytes = sizeof(float) * (size_t)ws_row(N[0], sizeof(float)) * (N[1] + 2) * (N[2] + 2);
    const size_t nodes = (size_t)(N[0] + 1) * (N[1] + 1) * (N[2] + 1);
    // the first iterate (the FP32 solve, ratio x FP32 values) is kept in FP32
    // where the lean-I/O residual and update kernels run: 4 bytes per node
    // less to write and to read back; the later iterates are FP64
    const bool u32 = refine > 0 && iterate_f32_ok(ctx, g, false);
    float *ex = reinterpret_cast<float *>(ws.get(wbytes, 1));
    float *U0 = u32 ? reinterpret_cast<float *>(ws.get(sizeof(float) * nodes, 7)) : nullptr;
    double *U = refine > 0 && (!u32 || refine > 1) ? reinterpret_cast<double *>(ws.get(sizeof(double) * nodes, 2)) : nullptr;
    if (!ex || (u32 && !U0) || (refine > 0 && (!u32 || refine > 1) && !U)) return fail(ERR_ALLOC, "workspace allocation failed");
    PassTimer timer(ctx, st);
    // singular case: PERTRB (weighted sum of the folded rhs) in FP64, once
    double *acc = nullptr;
    SolveArgs sa0;
    fill_solve_args(sa0, g, f, nullptr);
    if (g.singular && refine > 0) {
        acc = reinterpret_cast<double *>(ws.get(256, 3));
        if (!acc) return fail(ERR_ALLOC, "workspace allocation failed");
        if (ctx.timing()) ctx.timing_mark("pertrb", st);
        hipError_t e = launch_fold_sum(sa0, acc + 1, st);
        if (e != hipSuccess) return fail(ERR_DEVICE, std::string("PERTRB launch: ") + hipGetErrorString(e));
    }
    std::vector<PassSpec> v;
    double *first = refine == 0 ? f : u32 ? reinterpret_cast<double *>(U0) : U;
    if ((rc = solve_passes<float>(ctx, g, f, ex, nullptr, nullptr, v)) ||
        (rc = final_pass<float>(ctx, g, ex, nullptr, first, v, nullptr, u32 ? 2 : 0)))
        return rc;
    for (int it = 1; it <= refine; it++) {
        const bool from32 = u32 && it == 1;
        const double *Uin = from32 ? reinterpret_cast<const double *>(U0) : U;
        if ((rc = solve_passes<float>(ctx, g, f, ex, Uin, acc ? acc + 1 : nullptr, v, nullptr, from32 ? 1 : 0)) ||
            (rc = final_pass<float>(ctx, g, ex, Uin, it == refine ? f : U, v, nullptr, from32 ? 1 : 0)))
            return rc;
    }
    if ((rc = run_specs(ctx, v, st))) return rc;
    if (acc) {
        // the corrections never change the 0-mode of the FP32 solve (rounding
        // size): remove it so the result has the zero mode of the FP64 path
        if (ctx.timing()) ctx.timing_mark("mode0", st);
        hipError_t e = launch_remove_mode0(f, sa0, acc, st);
        if (e != hipSuccess) return fail(ERR_DEVICE, std::string("mode-0 projection launch: ") + hipGetErrorString(e));
    }
    timer.mark("periodic");
    return periodic_copies(g, f, st);
}

}  // namespace chz

extern "C" {

int chz_helmholtz_solve_mixed(double left, double right, int NX, int xbc, const double *bcl, const double *bcr,
                              double bottom, double top, int NY, int ybc, const double *bcb, const double *bct,
                              double front, double back, int NZ, int zbc, const double *bcf, const double *bce,
                              double lambda, double *f, int refine, void *stream)
{
    const double dom[6] = {left, right, bottom, top, front, back};
    const int N[3] = {NX, NY, NZ}, bc[3] = {xbc, ybc, zbc};
    const double *faces[6] = {bcl, bcr, bcb, bct, bcf, bce};
    return chz::helmholtz_mixed(dom, N, bc, faces, lambda, f, refine, (hipStream_t)stream);
}

int chz_helmholtz_solve(double left, double right, int NX, int xbc, const double *bcl, const double *bcr,
                        double bottom, double top, int NY, int ybc, const double *bcb, const double *bct,
                        double front, double back, int NZ, int zbc, const double *bcf, const double *bce,
                        double lambda, double *f, void *stream)
{
    const double dom[6] = {left, right, bottom, top, front, back};
    const int N[3] = {NX, NY, NZ}, bc[3] = {xbc, ybc, zbc};
    const double *faces[6] = {bcl, bcr, bcb, bct, bcf, bce};
    return chz::helmholtz_fp64(dom, N, bc, faces, lambda, f, (hipStream_t)stream);
}

int chz_helmholtz_solve19(double left, double right, int NX, int xbc, const double *bcl, const double *bcr,
                          double bottom, double top, int NY, int ybc, const double *bcb, const double *bct,
                          double front, double back, int NZ, int zbc, const double *bcf, const double *bce,
                          double lambda, double *f, void *stream)
{
    const double dom[6] = {left, right, bottom, top, front, back};
    const int N[3] = {NX, NY, NZ}, bc[3] = {xbc, ybc, zbc};
    const double *faces[6] = {bcl, bcr, bcb, bct, bcf, bce};
    return chz::helmholtz_fp64(dom, N, bc, faces, lambda, f, (hipStream_t)stream, 19);
}

}  // extern "C"

// C++ drop-in: cudahelmholtz/cuda_helmholtz_solver.h:5-11 (synchronous like the reference)
int cuda_helmholtz_solver(double left, double right, int NX, int xbc, double *bcl, double *bcr, double bottom,
                          double top, int NY, int ybc, double *bcb, double *bct, double front, double back, int NZ,
                          int zbc, double *bcf, double *bce, double lambda, double *f)
{
    int rc = chz_helmholtz_solve(left, right, NX, xbc, bcl, bcr, bottom, top, NY, ybc, bcb, bct, front, back, NZ, zbc,
                                 bcf, bce, lambda, f, nullptr);
    if (rc == 0) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = chz::fail(chz::ERR_DEVICE, std::string("device error: ") + hipGetErrorString(e));
    }
    if (rc != 0) printf("cuhelmholtz_amd: %s\n", chz::last_error());
    return rc;
}
