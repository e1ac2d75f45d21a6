// driver.h -- shared part of the command-line drivers that mirror the
// reference's test programs (cudahelmholtz/testgpu.cu, testgpurectangle.cu,
// testcpu.cpp, testcpurectangle.cpp): same arguments and printed lines.
//
// Problem (the reference's ufunc.h case 1): u = 1/3 sin(2 pi x) sin(2 pi y)
// sin(2 pi z) on [0,1]^3, lambda = -1, f = laplace u + lambda u at the nodes,
// Dirichlet values u and Neumann data du/dn (derivative along +axis) on the
// faces each code needs.  After the solve the program prints the time of the
// solve call and the largest |u_h - u| over all nodes.
#pragma once
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <vector>

namespace drv {

constexpr double PI = 3.14159265358979323846;
inline double u(double x, double y, double z) { return sin(2 * PI * x) * sin(2 * PI * y) * sin(2 * PI * z) / 3.0; }
inline double ux(double x, double y, double z) { return 2 * PI / 3.0 * cos(2 * PI * x) * sin(2 * PI * y) * sin(2 * PI * z); }
inline double uy(double x, double y, double z) { return 2 * PI / 3.0 * sin(2 * PI * x) * cos(2 * PI * y) * sin(2 * PI * z); }
inline double uz(double x, double y, double z) { return 2 * PI / 3.0 * sin(2 * PI * x) * sin(2 * PI * y) * cos(2 * PI * z); }
inline double lap(double x, double y, double z) { return -4 * PI * PI * sin(2 * PI * x) * sin(2 * PI * y) * sin(2 * PI * z); }

inline double now()
{
    timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec + t.tv_usec * 1e-6;
}

// host problem: node array f (x fastest) and the six Neumann face arrays
// (x faces [k(ny+1)+j], y faces [k(nx+1)+i], z faces [j(nx+1)+i]; empty where unused)
struct Problem {
    int nx, ny, nz, bc[3];
    double lambda = -1.0;
    std::vector<double> f, face[6];
    Problem(int nx_, int ny_, int nz_, int xbc, int ybc, int zbc) : nx(nx_), ny(ny_), nz(nz_), bc{xbc, ybc, zbc}
    {
        const double dx = 1.0 / nx, dy = 1.0 / ny, dz = 1.0 / nz;
        const long fx = nx + 1, fy = ny + 1;
        f.assign((size_t)fx * fy * (nz + 1), 0.0);
        for (int k = 0; k <= nz; k++)
            for (int j = 0; j <= ny; j++)
                for (int i = 0; i <= nx; i++)
                    f[i + fx * (j + fy * k)] = lap(i * dx, j * dy, k * dz) + lambda * u(i * dx, j * dy, k * dz);
        // Dirichlet values into the boundary nodes: x, then y, then z faces
        for (int k = 0; k <= nz; k++)
            for (int j = 0; j <= ny; j++) {
                if (xbc == 1 || xbc == 2) f[fx * (j + fy * k)] = u(0, j * dy, k * dz);
                if (xbc == 1 || xbc == 4) f[nx + fx * (j + fy * k)] = u(1, j * dy, k * dz);
            }
        for (int k = 0; k <= nz; k++)
            for (int i = 0; i <= nx; i++) {
                if (ybc == 1 || ybc == 2) f[i + fx * fy * k] = u(i * dx, 0, k * dz);
                if (ybc == 1 || ybc == 4) f[i + fx * (ny + fy * k)] = u(i * dx, 1, k * dz);
            }
        for (int j = 0; j <= ny; j++)
            for (int i = 0; i <= nx; i++) {
                if (zbc == 1 || zbc == 2) f[i + fx * j] = u(i * dx, j * dy, 0);
                if (zbc == 1 || zbc == 4) f[i + fx * (j + fy * nz)] = u(i * dx, j * dy, 1);
            }
        // Neumann data: low face for codes 3, 4; high face for codes 2, 3
        auto lo = [](int c) { return c == 3 || c == 4; };
        auto hi = [](int c) { return c == 2 || c == 3; };
        for (int s = 0; s < 2; s++)
            if (s == 0 ? lo(xbc) : hi(xbc)) {
                face[s].resize((size_t)(ny + 1) * (nz + 1));
                for (int k = 0; k <= nz; k++)
                    for (int j = 0; j <= ny; j++) face[s][k * (ny + 1) + j] = ux(s, j * dy, k * dz);
            }
        for (int s = 0; s < 2; s++)
            if (s == 0 ? lo(ybc) : hi(ybc)) {
                face[2 + s].resize((size_t)(nx + 1) * (nz + 1));
                for (int k = 0; k <= nz; k++)
                    for (int i = 0; i <= nx; i++) face[2 + s][k * (nx + 1) + i] = uy(i * dx, s, k * dz);
            }
        for (int s = 0; s < 2; s++)
            if (s == 0 ? lo(zbc) : hi(zbc)) {
                face[4 + s].resize((size_t)(nx + 1) * (ny + 1));
                for (int j = 0; j <= ny; j++)
                    for (int i = 0; i <= nx; i++) face[4 + s][j * (nx + 1) + i] = uz(i * dx, j * dy, s);
            }
    }
    double max_diff(const std::vector<double> &sol) const
    {
        const double dx = 1.0 / nx, dy = 1.0 / ny, dz = 1.0 / nz;
        const long fx = nx + 1, fy = ny + 1;
        double m = 0;
        for (int k = 0; k <= nz; k++)
            for (int j = 0; j <= ny; j++)
                for (int i = 0; i <= nx; i++) m = fmax(m, fabs(sol[i + fx * (j + fy * k)] - u(i * dx, j * dy, k * dz)));
        return m;
    }
};

// the reference's time_ macro line (time_.h): bar, location, seconds
inline void print_time(const char *where, double seconds) { printf("||||||||||||||||||| %s: \t%.6f\n", where, seconds); }

}  // namespace drv
