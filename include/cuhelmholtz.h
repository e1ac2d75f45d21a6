/* cuhelmholtz.h -- C ABI of the MI355X (gfx950) transform and Helmholtz library.
 *
 * Plain C: pointers, sizes and status codes only.  Every array argument is a
 * device pointer owned by the caller.  Calls are stream-ordered on `stream`
 * (a hipStream_t passed as void*, NULL = the default stream) and return
 * before the device work completes; synchronise the stream to read results.
 *
 * Streams: the single-problem calls share persistent per-device workspaces;
 * each call's launches are ordered after the previous call that used the
 * same workspace, on whatever stream (a stream wait on that call's end event,
 * no host wait), so calls on different streams serialise on the device where
 * they share a workspace and never overwrite each other's data.  A batched
 * plan is ordered the same way after its own previous solve or rebind.
 *
 * Status: 0 = OK, -1 = invalid argument, -2 = unsupported size,
 * -3 = device / launch error, -4 = device allocation failure.
 * chz_last_error() returns the message of the last failure on this thread.
 *
 * Supported sizes: any even N up to 4096 per axis (like the reference's cuFFT
 * plans, cudasymmfft/funcinterface.cu:54-60).  The FFT stages use radix 2, 3,
 * 4, 5, 6, 8, 10, 12 and 16 butterflies; other prime factors of N/2 run as
 * direct DFT stages (M * p multiply-adds per line for a factor p), which are
 * correct for every size but slow for large primes.
 */
#ifndef CUHELMHOLTZ_H
#define CUHELMHOLTZ_H

#ifdef __cplusplus
extern "C" {
#endif

/* Transform kind ordinals, identical to enum transform_kind
 * (reference cudasymmfft/global.h:6-15). */
#define CHZ_DCT_1 0
#define CHZ_DCT_2 1
#define CHZ_DCT_3 2
#define CHZ_DST_1 3
#define CHZ_DST_2 4
#define CHZ_DST_3 5
#define CHZ_DFT_R2C 6
#define CHZ_DFT_C2R 7

/* 3D separable transform of a padded (NX+2)(NY+2)(NZ+2) FP64 volume (x
 * fastest), X then Y then Z.  d_out == NULL or d_out == d_data: in place (any
 * shape); otherwise the result is written to d_out and d_data is unchanged.
 * Replaces do_transform / do_transform_cubic / do_transform_nocubic
 * (reference cudasymmfft/funcinterface.cu:25-40, :42-198, :200-334). */
int chz_do_transform(double *d_data, double *d_out, int NX, int NY, int NZ,
                     int x_kind, int y_kind, int z_kind, void *stream);

/* Same in FP32 (twiddles generated in FP64, rounded).  No reference
 * counterpart: the reference is FP64 only; used by the mixed-precision solve. */
int chz_do_transform_f32(float *d_data, float *d_out, int NX, int NY, int NZ,
                         int x_kind, int y_kind, int z_kind, void *stream);

/* FFT-diagonalised solve of laplace u + lambda u = f (7-point stencil) on the
 * node grid of [left,right]x[bottom,top]x[front,back] with NX, NY, NZ cells,
 * FP64 throughout.  Same arguments, layouts and boundary codes as
 * cuda_helmholtz_solver (reference cudahelmholtz/cuda_helmholtz_solver.h:5-11,
 * .cu:33-380): f holds (NX+1)(NY+1)(NZ+1) nodes (x fastest) with the rhs and
 * the Dirichlet values on entry and the solution on return; codes 0 periodic,
 * 1 D-D, 2 D-N, 3 N-N, 4 N-D; Neumann face data (derivative along +axis) x:
 * [k*(NY+1)+j], y: [k*(NX+1)+i], z: [j*(NX+1)+i], NULL where unused.  For
 * lambda == 0 with only periodic / Neumann axes the (0,0,0) mode is set to 0.
 * Uses a persistent per-device workspace of (NX+2)(NY+2)(NZ+2) doubles. */
int chz_helmholtz_solve(double left, double right, int NX, int xbc, const double *bcl, const double *bcr,
                        double bottom, double top, int NY, int ybc, const double *bcb, const double *bct,
                        double front, double back, int NZ, int zbc, const double *bcf, const double *bce,
                        double lambda, double *f, void *stream);

/* Scratch of chz_helmholtz_solve_noalloc: a fixed device array of this many
 * bytes, held by every device the library runs on until the process ends. */
#define CHZ_NOALLOC_SCRATCH_BYTES (64u << 20)

/* chz_helmholtz_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and result (to rounding: its twiddle and eigenvalue
 * tables are generated on the device) as chz_helmholtz_solve; the spectral
 * workspace and the tables live in a static device array of
 * CHZ_NOALLOC_SCRATCH_BYTES, so the call only launches kernels (no
 * allocation, copy or memset).  Needs
 *   8 (NX+2 rounded up to 16)(NY+2)(NZ+2) + 16 (3 N_a + 2) per axis + 8 (NX+NY+NZ+6)
 * bytes (each part rounded up to 256) within that size -- about 190^3 cells;
 * larger problems return -1 with a message.  Every call on one device shares
 * the array and leaves its contents behind: calls that overlap on one device
 * (two streams, two threads) race. */
int chz_helmholtz_solve_noalloc(double left, double right, int NX, int xbc, const double *bcl, const double *bcr,
                                double bottom, double top, int NY, int ybc, const double *bcb, const double *bct,
                                double front, double back, int NZ, int zbc, const double *bcf, const double *bce,
                                double lambda, double *f, void *stream);

/* The same solve with the 19-point compact operator: (L19 + lambda) u = B f,
 * L19 = sum_a D_a + sum_{a<b} (h_a^2 + h_b^2)/12 D_a D_b and B = I + sum_a
 * h_a^2/12 D_a, D_a the second difference along axis a with the boundary
 * closure of its code; the boundary data are folded into f as for the
 * 7-point operator.  Replaces the reference's _nonpoisson_devide_19points_
 * (cudahelmholtz/_cuda_helmholtz_.inl:255-274), which is defined but never
 * launched (cuda_helmholtz_solver.cu:325,338) and hard-codes cos(i pi/N) for
 * every axis (codes 1 and 3 only); here every code uses its own eigenvalues.
 * Fourth order for lambda = 0 on problems whose boundary extension is smooth
 * (periodic, homogeneous data); second order otherwise (lambda u is not
 * compact-corrected, as in the reference's kernel). */
int chz_helmholtz_solve19(double left, double right, int NX, int xbc, const double *bcl, const double *bcr,
                          double bottom, double top, int NY, int ybc, const double *bcb, const double *bct,
                          double front, double back, int NZ, int zbc, const double *bcf, const double *bce,
                          double lambda, double *f, void *stream);

/* Same solve with FP32 transforms and divide plus `refine` (0..16) steps of
 * FP64 iterative refinement: an FP32 solve of the folded rhs, then per step
 * the FP64 7-point residual r = f - (Delta_h u + lambda u), an FP32
 * correction solve and the FP64 update u += delta.  Twiddles and eigenvalues
 * are generated in FP64 and rounded; the residual and the update are the only
 * FP64 arithmetic.  No reference counterpart (the reference is FP64 only);
 * same arguments and layouts as chz_helmholtz_solve.  Workspaces:
 * (NX+2)(NY+2)(NZ+2) floats and, if refine > 0, (NX+1)(NY+1)(NZ+1) doubles. */
int chz_helmholtz_solve_mixed(double left, double right, int NX, int xbc, const double *bcl, const double *bcr,
                              double bottom, double top, int NY, int ybc, const double *bcb, const double *bct,
                              double front, double back, int NZ, int zbc, const double *bcf, const double *bce,
                              double lambda, double *f, int refine, void *stream);

/* Adjoint (vector-Jacobian product) of chz_helmholtz_solve for the same
 * geometry, codes and lambda.  The solve maps (f, Neumann face data) to the
 * node array u linearly; given gu = dL/du on all (NX+1)(NY+1)(NZ+1) nodes it
 * writes gf = dL/df (same layout; 0 at the periodic duplicate nodes, which
 * the solve overwrites), the face gradients gbcl..gbce (layouts of the face
 * arrays; NULL = not wanted, and not written where the codes read no face
 * data) and, if glambda is set, glambda[0] = dL/dlambda = -<A^-T gu, u> on
 * the unknown box, which needs u = the forward solution.  All pointers are
 * device pointers; gf must not alias gu.  Runs one FP64 solve (A^-T =
 * W A^-1 W^-1, W = 1/2 at Neumann end nodes) between two node-array
 * passes.  No reference counterpart (the reference has no derivatives). */
int chz_helmholtz_adjoint(double left, double right, int NX, int xbc,
                          double bottom, double top, int NY, int ybc,
                          double front, double back, int NZ, int zbc,
                          double lambda, const double *u, const double *gu, double *gf,
                          double *gbcl, double *gbcr, double *gbcb, double *gbct, double *gbcf, double *gbce,
                          double *glambda, void *stream);

/* chz_helmholtz_adjoint for `count` problems of one geometry and codes with
 * lambdas[k] (host array): u, gu, gf hold `count` consecutive node arrays,
 * each face gradient (NULL = not wanted) `count` consecutive face arrays, and
 * glambda (NULL = not wanted) `count` values.  The node-array steps run once
 * over the batch and the solves as one batched solve (chz_batch_solve_cached
 * plan).  No reference counterpart. */
int chz_helmholtz_adjoint_batch(double left, double right, int NX, int xbc,
                                double bottom, double top, int NY, int ybc,
                                double front, double back, int NZ, int zbc,
                                int count, const double *lambdas, const double *u, const double *gu, double *gf,
                                double *gbcl, double *gbcr, double *gbcb, double *gbct, double *gbcf, double *gbce,
                                double *glambda, void *stream);

/* One problem of a batched solve: the arguments of chz_helmholtz_solve. */
typedef struct chz_problem {
    double domain[6];        /* left, right, bottom, top, front, back */
    int N[3];                /* NX, NY, NZ cells */
    int bc[3];               /* boundary codes x, y, z (0..4) */
    double lambda;
    double *f;               /* (NX+1)(NY+1)(NZ+1) node array: rhs in, solution out */
    const double *faces[6];  /* bcl, bcr, bcb, bct, bcf, bce (NULL where unused) */
} chz_problem;

/* Plan a batched solve of `count` independent problems (any mix of sizes and
 * boundary codes): precision 64 = plain FP64 (refine must be 0), precision
 * 32 = FP32 transforms + `refine` FP64 refinement steps as in
 * chz_helmholtz_solve_mixed (singular problems then need refine == 0).
 * refine < 0 (precision 32): each problem gets its own count, at most
 * -refine: 1 where the spectral condition number max|e + lambda| /
 * min|e + lambda| of its operator is at most 1e5, 2 up to 1e7, 3 beyond
 * (chz_batch_refine_counts reports them).  The
 * plan owns one workspace per problem, device tables of per-problem pass
 * descriptors and its own copies of the twiddle and eigenvalue tables it
 * reads, and holds the whole solve as one hipGraph: one launch per pass
 * and size class (the size classes of a pass as parallel branches), then
 * the periodic copies.  The
 * problems' f arrays and face arrays are bound here: each chz_batch_solve
 * solves whatever they hold at that time.  Results equal the single-problem
 * solve's (same kernels and launch configurations).  No reference
 * counterpart (the reference solves one problem per call). */
int chz_batch_create(const chz_problem *problems, int count, int precision, int refine, void **plan);

/* Run the batched solve on `stream`: use_graph = 1 replays the plan's
 * graph (one hipGraphLaunch); 0 issues the same launches directly (and
 * records per-launch timing when chz_set_timing is on). */
int chz_batch_solve(void *plan, int use_graph, void *stream);

/* Point a plan at `count` problems of the same geometry as the ones it was
 * made for (sizes, codes, domains, which face arrays are given, lambda = 0 or
 * not): new f and face arrays and new lambdas.  The pass descriptors are
 * rewritten on the host and uploaded on `stream` (from a pinned staging
 * buffer), ordered after the plan's earlier solves on any stream; the plan's
 * workspaces, work lists, tables and graph are kept.  The host waits only for
 * the plan's previous rebind upload to have run.  No reference counterpart. */
int chz_batch_rebind(void *plan, const chz_problem *problems, int count, void *stream);

/* Batched solve through a plan kept by the library for repeated solves of
 * one geometry (a training loop, a time stepper): the first call makes the
 * plan as chz_batch_create does, later calls with the same geometry rebind it
 * (chz_batch_rebind) and replay its graph on `stream`.  Up to 4 plans are
 * kept per process (least recently used dropped first).  No reference
 * counterpart. */
int chz_batch_solve_cached(const chz_problem *problems, int count, int precision, int refine, void *stream);

/* Launch count of one batched solve and the plan's device memory. */
int chz_batch_info(void *plan, int *launches, long long *work_bytes, long long *table_bytes);

/* Refinement steps of each of the plan's `count` problems (all equal unless
 * the plan was made with refine < 0).  No reference counterpart. */
int chz_batch_refine_counts(void *plan, int *counts, int count);

/* Free the plan (waits for the plan's last solve or rebind to finish). */
int chz_batch_destroy(void *plan);

/* Message of the last failed call on this thread ("" if none). */
const char *chz_last_error(void);

/* Free cached twiddle tables and workspaces of the current device, and the
 * batched plans kept by chz_batch_solve_cached (after waiting for the device).
 * Plans made by chz_batch_create own their tables and stay usable; launch
 * configurations (tuning) are kept.
 * Replaces freeMemory_cubic (reference cudasymmfft/funcinterface.cu:337-342). */
void chz_release(void);

/* Launch configuration of a pass kernel for N cells: elements per thread,
 * threads per line, lines per workgroup and the radix plan (one byte per
 * stage).  prec = 64 or 32; role 0 = x-axis rows, 1 = y columns (and the
 * solve's y passes), 2 = first solve pass (folds + forward x), 3 = fused z
 * pass of the solve with periodic z, 4 = z columns of do_transform, 5 = fused
 * z pass of the solve with z codes 1-4 (a size without a role-5
 * configuration of its own uses role 3's).  No reference counterpart (the
 * reference's cuFFT plans are opaque). */
int chz_axis_config(int N, int prec, int role, int *ept, int *tpl, int *lines,
                    unsigned long long *plan);

/* Pin a launch configuration (used by the autotuner); -1 if it cannot run
 * (no radix plan for ept/tpl, more than 512 threads per tile, or LDS over the limit). */
int chz_set_axis_config(int N, int prec, int role, int ept, int tpl, int lines);

/* Forget pinned / cached configurations (back to the built-in choice). */
void chz_reset_axis_configs(void);

/* Pin every configuration of a tuning table (lines "role N prec ept tpl
 * lines"); returns how many were valid.  At start-up the library loads
 * $CHZ_TUNE_FILE or tuning/<gfx arch>.txt next to the shared library. */
int chz_load_tuning(const char *path);

/* The radix plan the library derives for an M-point FFT with at most ept
 * points per thread and tpl threads per line (one byte per stage, first stage
 * in the low byte); -1 if that shape cannot run.  Host only (no device call).
 * No reference counterpart (cuFFT plans are opaque). */
int chz_radix_plan(int M, int ept, int tpl, unsigned long long *plan);

/* The compiled fixed-shape pass kernels (generated at build time from the
 * tuning tables tuning/<arch>.txt and tuning/<arch>.alt.txt): for kernel i,
 * rows[10 i .. 10 i + 9] = form (0 single-problem lean tile I/O, 1 its batched
 * form, 2 single-problem fixed shape, 3 batched fixed shape), precision, pass
 * op (0 rows, 1 columns, 2 solve pass 1, 3 fused z pass, 4 residual, 5
 * update), N, points per thread, threads per line, lines per tile, radix plan
 * low and high 32 bits, kind selector (-1: every kind; else the kind of the
 * pass, for a fused z pass the boundary code 0-4).  Writes at most max_rows
 * rows (rows may be NULL) and returns the kernel count.  Host only.  No
 * reference counterpart. */
int chz_compiled_kernels(int *rows, int max_rows);

int chz_version(void);

/* Per-pass timing, the counterpart of the reference's TIME_TEST splits
 * (e.g. cudasymmfft/dst1funcinplace.cu:145-218): with timing on, every pass
 * of a transform / solve records an event at its start on the call's stream
 * (and each call one at its end).  chz_get_timing waits for the events
 * recorded since timing was switched on or last read, writes the pass
 * durations (ms) and their comma-separated names, clears them, and returns
 * the pass count. */
void chz_set_timing(int on);
int chz_get_timing(double *ms, int max_passes, char *names, int names_len);

#ifdef __cplusplus
}
#endif
#endif /* CUHELMHOLTZ_H */
