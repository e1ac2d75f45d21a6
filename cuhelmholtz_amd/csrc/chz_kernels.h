// chz_kernels.h -- host-visible declarations shared by the kernels and the
// host orchestration: kind ordinals, twiddle-table layout, pass arguments and
// the launch entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace chz {

// largest workgroup of a pass kernel (lines per tile x threads per line)
#define CHZ_MAX_THREADS 512

enum Kind { K_DCT_1 = 0, K_DCT_2, K_DCT_3, K_DST_1, K_DST_2, K_DST_3, K_R2C, K_C2R };

__host__ __device__ constexpr bool kind_is_z2d(int k) { return k == K_DST_1 || k == K_DCT_1 || k == K_DCT_2 || k == K_DST_2 || k == K_C2R; }

// Twiddle table of one N (host-generated in FP64, rounded to T), all as V:
//   half[n] = (cos(n pi/2N), sin(n pi/2N)),  n = 0..2N     (offset 0)
//   fftw[j] = (cos(2 pi j/M), -sin(2 pi j/M)), j = 0..M-1   (offset 2N+1)
//   rsin[n] = (1/sin(n pi/N), 0),            n = 0..M      (offset 2N+1+M; rsin[0] = 0)
struct TwOffsets {
    __host__ __device__ static int half(int) { return 0; }
    __host__ __device__ static int fftw(int N) { return 2 * N + 1; }
    __host__ __device__ static int rsin(int N) { return 2 * N + 1 + N / 2; }
    __host__ __device__ static int total(int N) { return 2 * N + 1 + N / 2 + N / 2 + 1; }
};

struct PassArgs {
    const void *src;
    void *dst;
    long long s_sa, s_sb, s_sp;   // source strides (elements) of line index a, b and slot
    long long d_sa, d_sb, d_sp;   // destination strides
    int s_rowlen, d_rowlen;       // row tiles: slots per source / destination row (= s_sa / d_sa)
    int a0, na, b0;               // lines a in [a0, a0+na) tiled by L; b = b0 + blockIdx.y
    int N, M, kind;
    int in_lo, in_hi;             // slots loaded (others 0)
    int st_lo, st_hi;             // slots stored
    int out_lo, out_hi;           // valid output slots (others stored as 0)
    int L, TPL;
    int xcds;                     // XCD count for the tile swizzle (1 = dispatch order)
    int bfast;                    // 1: consecutive tiles step through b first (neighbour planes in time)
    int u_f32;                    // residual / update passes of the mixed solve (lean-I/O kernels only):
                                  // bit 0 the iterate read is FP32, bit 1 the iterate written is FP32
    unsigned long long plan;
    const void *tw;
    double scale;
};

// Helmholtz-solve extras of a pass: boundary data for the folds of pass 1,
// eigenvalue tables and the inverse kind of the fused middle z pass.
struct SolveArgs {
    const double *f;                                   // node array (NX+1)(NY+1)(NZ+1)
    const double *bcl, *bcr, *bcb, *bct, *bcf, *bce;   // Neumann face data (may be null)
    int NX, NY, NZ;
    int bc[3], s[3], e[3];                             // codes and unknown windows per axis
    double ih[3], ih2[3];                              // 1/h, 1/h^2
    const double *eig;                                 // slot eigenvalue tables x | y | z
    double lambda;
    int singular;                                      // zero the (0,0,0) mode
    int kind2, out2_lo, out2_hi;                       // inverse kind of the middle pass and its window
    const double *pertrb;                              // singular case: weighted sum of the folded rhs (or null)
    double pert_scale;                                 // 1 / (sum of the 0-mode weights)
    // 19-point operator (0: the 7-point one): divide by (L19 + lambda) / B with
    // L19 = sum e_a + sum_{a<b} c_ab e_a e_b, B = 1 + sum_a b_a e_a; c19 = {c_xy, c_xz, c_yz, b_x, b_y, b_z}
    int stencil19;
    double c19[6];
};

// Point stride of a tile of L lines in LDS: FP64 tiles are unpadded (a
// smaller tile fits one more workgroup per CU, measured faster), FP32 tiles
// carry one padding point per point row (measured faster).
__host__ __device__ constexpr int tile_ls(int L, int esz) { return L + (esz == 8 ? 0 : 1); }
__host__ __device__ inline int lds_points(int M, int L, int esz) { return (M + 1) * tile_ls(L, esz); }

// Pass kernels of a solve (also the op of a batched pass).
enum BatchOp { B_ROWS = 0, B_COLS = 1, B_XFWD = 2, B_ZMID = 3, B_XRES = 4, B_XUPD = 5 };

// Descriptor of one pass of one problem of a batched solve.
struct BatchDesc {
    PassArgs p;
    SolveArgs sa;          // folds / eigenvalues (passes 1 and 3)
    const double *u_in;    // residual pass: current solution; update pass: solution added to (or null)
    double *u_out;         // update pass: destination
};
// one work item of a batched pass: a tile of one problem (index into the descriptors of the launch)
struct BatchItem {
    int prob, tile;
};
// one periodic face copy of a batched solve
struct BatchCopy {
    double *f;
    int NX, NY, NZ, axis;
};



// Geometry of an adjoint solve (node array (NX+1)(NY+1)(NZ+1), x fastest).
struct AdjArgs {
    int N[3], bc[3], s[3], e[3];
    double ih[3], ih2[3];
    double *face[6];   // gradient face arrays (null where not wanted)
    // batch of problems of this geometry (blockIdx.y = problem): node-array
    // and face-array strides between consecutive problems (0: one problem)
    long long bstride, fstride[6];
};
// adjoint kernels: 0 prep, 1 finish, 2 lambda reduction, 3 zero fill
const void *adj_kernel(int which);

// kernel getters (one translation unit per precision)
const void *axis_kernel_f64_rows(int ept);
const void *axis_kernel_f64_cols(int ept);
const void *axis_kernel_f32_rows(int ept);
const void *axis_kernel_f32_cols(int ept);
const void *solve_xfwd_kernel_f64(int ept);
const void *solve_xfwd_kernel_f32(int ept);
const void *solve_zmid_kernel_f64(int ept);
const void *solve_zmid_kernel_f32(int ept);
const void *solve_xres_kernel_f32(int ept);
const void *solve_xinv_upd_kernel_f32(int ept);
const void *periodic_copy_kernel();
// batched pass kernels: op 0 rows, 1 columns, 2 solve pass 1, 3 fused z pass,
// 4 residual pass (FP32), 5 update pass (FP32); nullptr where not built
const void *batch_kernel_f64(int op, int ept);
const void *batch_kernel_f32(int op, int ept);
const void *batch_periodic_kernel();
// forms of a compiled fixed-shape kernel: single-problem k_line / batched
// k_line_batch (lean tile I/O, even threads per line), single k_fixed /
// batched k_batch_fixed (odd threads per line)
enum FixedForm { F_LINE = 0, F_LINE_BATCH = 1, F_FIXED = 2, F_FIXED_BATCH = 3 };
// a compiled fixed-shape pass kernel: (precision, op, N, elements/thread, threads/line, lines, plan)
struct FixedEntry {
    int prec, op, N, ept, tpl, lines;
    unsigned long long plan;
    const void *fn;
    unsigned lds_extra;   // LDS bytes beyond axis_pass_lds (tables the kernel stages)
    int form;             // FixedForm
    int ks;               // -1: any kind (read from the pass); else the kind of the pass (the z code of a fused z pass)
};
// kind selector of a pass for the kind-specialised kernels: the pass kind, or
// for the fused z pass the boundary code of its (forward, inverse) kind pair
// (-1 if the pair is not one of a boundary code's)
int pass_kind_sel(int op, const PassArgs &p, const SolveArgs *sa);
// the compiled shapes: tables generated from the tuning tables at build time
// (cuhelmholtz_amd/kernelgen.py writes them and their index)
int gen_table_count();
const FixedEntry *gen_table(int q, int *n);
// fixed-shape kernel of a pass (arguments PassArgs, SolveArgs, const double *, double *), nullptr if
// none (CHZ_NO_FIXED in the environment disables them); given the pass arguments, the kernels
// with lean tile I/O are preferred where the pass windows allow them
const FixedEntry *fixed_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan,
                               const PassArgs *p = nullptr, int ks = -1);
// the lean-I/O single-problem kernel of a pass shape, nullptr if none is compiled
// (or CHZ_NO_LINE / CHZ_NO_FIXED is set)
const FixedEntry *line_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan, int ks = -1);
// fixed-shape batched kernel (arguments as k_batch), nullptr if none
const void *batch_fixed_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan);
// batched kernel with lean tile I/O (arguments as k_batch; the kind-specialised
// one if compiled, else the one for every kind), nullptr if none
// (CHZ_NO_LINE / CHZ_NO_FIXED in the environment: none); usable when every
// problem of the launch passes line_windows_ok
const FixedEntry *batch_line_kernel(int prec, int op, int N, int ept, int tpl, int lines, unsigned long long plan,
                                    int ks = -1);
// the windows of a pass contain slots 2..N-2 (precondition of the lean-I/O kernels)
bool line_windows_ok(const PassArgs &p, int N);
// rows of 10 ints per compiled fixed-shape kernel (form, prec, op, N, ept, tpl,
// lines, plan low 32 bits, plan high 32 bits, kind selector) into
// rows[0 .. 10 max_rows); returns the kernel count
int compiled_kernels(int *rows, int max_rows);
hipError_t launch_fixed(const FixedEntry *fx, const PassArgs &p, const SolveArgs *sa, const void *u_in, void *u_out,
                        int prec, int nb, hipStream_t s);

// Launch one axis pass (prec = 64 or 32; ept = elements per thread of the
// FFT; rows = x-pass row tiles, else column tiles) over nb lines of index b.
hipError_t launch_axis_pass(const PassArgs &p, int prec, int ept, bool rows, int nb, hipStream_t s);
hipError_t launch_solve_pass(const void *fn, const PassArgs &p, const SolveArgs &sa, int prec, int nb, hipStream_t s);
hipError_t launch_pass_ptrs(const void *fn, const PassArgs &p, const void *a0, const void *a1, int prec, int nb,
                            hipStream_t s);
hipError_t launch_pass_sa_ptr(const void *fn, const PassArgs &p, const SolveArgs &sa, const void *a0, int prec, int nb,
                              hipStream_t s);
size_t axis_pass_lds(int prec, int M, int L, int TPL);
// 1 / (sum of the 0-mode weights) over the unknown box
double mode0_inv_wsum(const SolveArgs &sa);
// subtract the weighted mean (the (0,0,0) mode) of u over the unknown box; acc: one device double
hipError_t launch_remove_mode0(double *u, const SolveArgs &sa, double *acc, hipStream_t s);
// acc = weighted sum of the folded rhs over the unknown box (PERTRB x weight sum)
hipError_t launch_fold_sum(const SolveArgs &sa, double *acc, hipStream_t s);

}  // namespace chz
