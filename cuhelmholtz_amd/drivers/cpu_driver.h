// cpu_driver.h -- the host-array test programs (testcpu.cpp, testcpurectangle.cpp
// of the reference): solver_3d on host arrays (here the hw3crt_ drop-in of this
// library, solved on the GPU), timed, then the max diff.
#pragma once
#include "driver.h"
#include "hw3crt_wrapper.h"

namespace drv {

inline int run_host(const char *prog, int nx, int ny, int nz, int xbc, int ybc, int zbc)
{
    Problem P(nx, ny, nz, xbc, ybc, zbc);
    double *fc[6];
    for (int s = 0; s < 6; s++) fc[s] = P.face[s].empty() ? nullptr : P.face[s].data();
    const double t0 = now();
    solver_3d(0, 1, nx, xbc, fc[0], fc[1], 0, 1, ny, ybc, fc[2], fc[3], 0, 1, nz, zbc, fc[4], fc[5], P.lambda, P.f.data());
    const double t1 = now();
    print_time(prog, t1 - t0);
    printf("max diff : %.14f\n", P.max_diff(P.f));
    return 0;
}

}  // namespace drv
