/* hw3crt_wrapper.h -- drop-in for the reference's CPU entry points,
 * cudahelmholtz/hw3crt_wrapper.h:1-19 (FISHPACK hw3crt and its C++ wrapper),
 * served by the GPU solve of this library.
 *
 * hw3crt_: FISHPACK's argument list (every argument by pointer).  Host arrays:
 *   F(LDIMF, MDIMF, N+1)     node (i,j,k) at i + LDIMF*(j + MDIMF*k): rhs and
 *                            Dirichlet values in, solution out
 *   BDXS, BDXF (MDIMF, N+1)  x-face derivatives, (j,k) at j + MDIMF*k
 *   BDYS, BDYF (LDIMF, N+1)  y-face derivatives, (i,k) at i + LDIMF*k
 *   BDZS, BDZF (LDIMF, M+1)  z-face derivatives, (i,j) at i + LDIMF*j
 * Codes per axis (LBDCND, MBDCND, NBDCND): 0 periodic, 1 Dirichlet-Dirichlet,
 * 2 Dirichlet-Neumann, 3 Neumann-Neumann, 4 Neumann-Dirichlet.
 * IERROR: 0 ok; 1 XS >= XF; 2 L < 3; 3 LBDCND outside 0..4; 4 YS >= YF;
 * 5 M < 3; 6 MBDCND; 7 ZS >= ZF; 8 N < 3; 9 NBDCND; 10 LDIMF < L+1;
 * 11 MDIMF < M+1; 12 ELMBDA > 0 (warning, solution computed); 13 odd cell
 * count (the GPU transforms take any even one; FISHPACK accepts odd ones
 * too); 20 device failure (chz_last_error() has the reason).
 * PERTRB: for ELMBDA = 0 with periodic / Neumann codes only, the weighted mean
 * of the folded rhs (the constant removed to make the problem consistent);
 * the solution returned then has a zero (0,0,0) mode.  W is not used.
 *
 * solver_3d: the reference wrapper (hw3crt_wrapper.cpp:5-33): packed node
 * array (LDIMF = NX+1, MDIMF = NY+1), prints "State code : <IERROR>".
 */
#ifndef CHZ_HW3CRT_WRAPPER_H
#define CHZ_HW3CRT_WRAPPER_H

extern "C" {
void hw3crt_(double *xs, double *xf, int *l, int *lbdcnd, double *bdxs, double *bdxf,
             double *ys, double *yf, int *m, int *mbdcnd, double *bdys, double *bdyf,
             double *zs, double *zf, int *n, int *nbdcnd, double *bdzs, double *bdzf,
             double *elmbda, int *ldimf, int *mdimf, double *f,
             double *pertrb, int *ierror, double *w);
}

void solver_3d(double xs, double xf, int NX, int xbc,
               double *bcxs, double *bcxf,
               double ys, double yf, int NY, int ybc,
               double *bcys, double *bcyf,
               double zs, double zf, int NZ, int zbc,
               double *bczs, double *bczf,
               double lambda, double *f);

#endif
