/* cuda_helmholtz_solver.h -- C++ drop-in for the reference's
 * cudahelmholtz/cuda_helmholtz_solver.h:5-11.
 *
 * Solves laplace u + lambda * u = f with the 7-point stencil on the node grid
 * of [left,right]x[bottom,top]x[front,back] with NX, NY, NZ cells (even).
 * f: device array of (NX+1)(NY+1)(NZ+1) FP64 nodes, x fastest; holds the rhs
 * and the Dirichlet boundary values on entry, the solution on return.
 * Boundary codes per axis (FISHPACK LBDCND): 0 periodic, 1 Dirichlet-Dirichlet,
 * 2 Dirichlet-Neumann, 3 Neumann-Neumann, 4 Neumann-Dirichlet.  Neumann data
 * (derivative along +axis) as device face arrays: x faces [k*(NY+1)+j],
 * y faces [k*(NX+1)+i], z faces [j*(NX+1)+i]; NULL where the code needs none.
 *
 * Returns 0 on success.  Unlike the reference (which exit()s on a bad code,
 * cuda_helmholtz_solver.cu:104-107) it prints the reason and returns a
 * negative status; chz_last_error() has the message.  The call is
 * synchronous, like the reference.
 */
#pragma once

int cuda_helmholtz_solver(double left, double right, int NX, int xbc,
                          double *bcl, double *bcr,
                          double bottom, double top, int NY, int ybc,
                          double *bcb, double *bct,
                          double front, double back, int NZ, int zbc,
                          double *bcf, double *bce,
                          double lambda, double *f);
