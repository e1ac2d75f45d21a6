// chz_noalloc.hip -- the FP64 Helmholtz solve without device allocation.
//
// chz_helmholtz_solve allocates (once, then caches) its spectral workspace and
// its twiddle and eigenvalue tables.  Callers that cannot allocate -- code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc -- use chz_helmholtz_solve_noalloc instead: the same passes and
// kernels, with the workspace and the tables in `scratch`, a fixed device array
// of CHZ_NOALLOC_SCRATCH_BYTES that every device holds from module load until
// the process ends.  The tables are generated on the device by one small
// kernel per call (FP64 sincospi / sinpi; the eigenvalues as
// -4 sin^2(angle/2) / h^2, which equals (2 cos(angle) - 2) / h^2 without its
// cancellation), so the call issues kernel launches only: no allocation, no
// copy, no memset.  Every call on a device shares its scratch: calls that
// overlap on one device (two streams, two threads) race.
#include <mutex>
#include <string>
#include <vector>

#include "chz_context.h"
#include "chz_solver.h"
#include "chz_transform.h"
#include "../../include/cuhelmholtz.h"

namespace chz {

static __device__ __align__(256) unsigned char scratch[CHZ_NOALLOC_SCRATCH_BYTES];

// Generator arguments: the tables of the three axes.
struct NoallocTables {
    double2 *tw[3];   // TwOffsets layout per axis
    double *eig;      // (NX+2) | (NY+2) | (NZ+2) slot eigenvalues
    int N[3], bc[3];
    double ih2[3];
};

// blockIdx.y = axis: its twiddle table (TwOffsets: half-angle pairs, FFT
// roots, reciprocal sines) and its slot eigenvalues (the solver's
// cuda_helmholtz_solver.cu:58-73 angles and shifts per code).
__global__ void __launch_bounds__(256) k_noalloc_tables(NoallocTables a)
{
    const int ax = (int)blockIdx.y;
    const int N = a.N[ax], M = N / 2, bc = a.bc[ax];
    const int nhalf = 2 * N + 1, ntw = nhalf + M + M + 1;
    double2 *tw = a.tw[ax];
    double *eig = a.eig + (ax >= 1 ? a.N[0] + 2 : 0) + (ax >= 2 ? a.N[1] + 2 : 0);
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < ntw + N + 2; i += (int)(gridDim.x * blockDim.x)) {
        if (i < nhalf) {   // (cos, sin)(i pi / 2N)
            double s, c;
            sincospi((double)i / (2.0 * N), &s, &c);
            tw[i] = make_double2(c, s);
        } else if (i < nhalf + M) {   // (cos, -sin)(2 pi j / M)
            const int j = i - nhalf;
            double s, c;
            sincospi(2.0 * (double)j / (double)M, &s, &c);
            tw[i] = make_double2(c, -s);
        } else if (i < ntw) {   // (1 / sin(k pi / N), 0), 0 at k = 0
            const int k = i - nhalf - M;
            tw[i] = make_double2(k == 0 ? 0.0 : 1.0 / sinpi((double)k / (double)N), 0.0);
        } else {
            const int sl = i - ntw, m = bc == 0 ? sl / 2 : sl;
            // angle / pi: 2m/N periodic, m/N codes 1 and 3, (m -+ 1/2)/N codes 2 and 4
            const double x = bc == 0 ? 2.0 * m / N : bc == 2 ? (m - 0.5) / N : bc == 4 ? (m + 0.5) / N : (double)m / N;
            const double s = sinpi(0.5 * x);
            eig[sl] = -4.0 * s * s * a.ih2[ax];
        }
    }
}

// Device address of `scratch` per device, looked up once with the device current.
static void *scratch_of(int dev)
{
    static std::mutex mu;
    static void *ptr[64];
    static bool have[64];
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!have[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        ptr[dev] = p;
        have[dev] = true;
    }
    return ptr[dev];
}

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

static int solve_noalloc(const double dom[6], const int N[3], const int bc[3], const double *const faces[6],
                         double lambda, double *f, hipStream_t st)
{
    if (!f) return fail(ERR_INVALID, "null solution array");
    SolveGeom g;
    int rc = make_geom(g, dom, N, bc, lambda, faces);
    if (rc) return rc;
    // scratch layout: workspace | twiddles x, y, z | eigenvalues
    size_t off = al256(sizeof(double) * (size_t)ws_row(N[0], sizeof(double)) * (N[1] + 2) * (N[2] + 2));
    size_t tw_off[3];
    for (int a = 0; a < 3; a++) {
        tw_off[a] = off;
        off += al256(sizeof(double2) * (size_t)TwOffsets::total(N[a]));
    }
    const size_t eig_off = off;
    off += al256(sizeof(double) * (size_t)(N[0] + N[1] + N[2] + 6));
    if (off > (size_t)CHZ_NOALLOC_SCRATCH_BYTES)
        return fail(ERR_INVALID, "the allocation-free solve needs " + std::to_string(off) + " bytes of scratch for " +
                                     std::to_string(N[0]) + "x" + std::to_string(N[1]) + "x" + std::to_string(N[2]) +
                                     "; it holds " + std::to_string((size_t)CHZ_NOALLOC_SCRATCH_BYTES) +
                                     " (CHZ_NOALLOC_SCRATCH_BYTES): use chz_helmholtz_solve");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(ERR_DEVICE, "no current device");
    char *base = reinterpret_cast<char *>(scratch_of(dev));
    if (!base) return fail(ERR_DEVICE, "scratch symbol lookup failed");
    NoallocTables nt;
    SolveTables tabs;
    for (int a = 0; a < 3; a++) {
        nt.tw[a] = reinterpret_cast<double2 *>(base + tw_off[a]);
        nt.N[a] = N[a];
        nt.bc[a] = bc[a];
        nt.ih2[a] = g.ih2[a];
        tabs.tw[a] = nt.tw[a];
    }
    nt.eig = reinterpret_cast<double *>(base + eig_off);
    tabs.eig = nt.eig;
    double *ex = reinterpret_cast<double *>(base);

    Context &ctx = Context::get();
    std::lock_guard<std::recursive_mutex> lock(ctx.mutex());
    std::vector<PassSpec> v;
    if ((rc = solve_passes<double>(ctx, g, f, ex, nullptr, nullptr, v, &tabs)) ||
        (rc = final_pass<double>(ctx, g, ex, nullptr, f, v, &tabs)))
        return rc;
    {
        int longest = 0;
        for (int a = 0; a < 3; a++) longest = std::max(longest, TwOffsets::total(N[a]) + N[a] + 2);
        void *args[] = {&nt};
        hipError_t e = hipLaunchKernel(reinterpret_cast<const void *>(&k_noalloc_tables),
                                       dim3((unsigned)((longest + 255) / 256), 3), dim3(256), args, 0, st);
        if (e != hipSuccess) return fail(ERR_DEVICE, std::string("table generation launch: ") + hipGetErrorString(e));
    }
    for (const PassSpec &q : v) {
        // launch_spec adds timing events only when timing is on; they are
        // recorded into events the context created earlier or creates now
        // (event objects, no device memory)
        if ((rc = launch_spec(ctx, q, st))) return rc;
    }
    return periodic_copies(g, f, st);
}

}  // namespace chz

extern "C" int chz_helmholtz_solve_noalloc(double left, double right, int NX, int xbc, const double *bcl,
                                           const double *bcr, double bottom, double top, int NY, int ybc,
                                           const double *bcb, const double *bct, double front, double back, int NZ,
                                           int zbc, const double *bcf, const double *bce, double lambda, double *f,
                                           void *stream)
{
    const double dom[6] = {left, right, bottom, top, front, back};
    const int N[3] = {NX, NY, NZ}, bc[3] = {xbc, ybc, zbc};
    const double *faces[6] = {bcl, bcr, bcb, bct, bcf, bce};
    return chz::solve_noalloc(dom, N, bc, faces, lambda, f, (hipStream_t)stream);
}
