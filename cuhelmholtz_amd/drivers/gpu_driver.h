// gpu_driver.h -- the GPU test programs (testgpu.cu, testgpurectangle.cu of the
// reference): device copies of f and the faces, a warm-up solve, the f reset,
// the timed cuda_helmholtz_solver call, the copy back and the max diff.
#pragma once
#include <hip/hip_runtime.h>

#include "cuda_helmholtz_solver.h"
#include "driver.h"

namespace drv {

inline int run_gpu(const char *prog, int nx, int ny, int nz, int xbc, int ybc, int zbc)
{
    Problem P(nx, ny, nz, xbc, ybc, zbc);
    double *df = nullptr, *dface[6] = {};
    const size_t fb = sizeof(double) * P.f.size();
    if (hipMalloc(&df, fb) != hipSuccess) { printf("%s: device allocation failed\n", prog); return 1; }
    for (int s = 0; s < 6; s++)
        if (!P.face[s].empty()) {
            if (hipMalloc(&dface[s], sizeof(double) * P.face[s].size()) != hipSuccess ||
                hipMemcpy(dface[s], P.face[s].data(), sizeof(double) * P.face[s].size(), hipMemcpyHostToDevice) != hipSuccess) {
                printf("%s: device allocation failed\n", prog);
                return 1;
            }
        }
    auto solve = [&] {
        return cuda_helmholtz_solver(0, 1, nx, xbc, dface[0], dface[1], 0, 1, ny, ybc, dface[2], dface[3], 0, 1, nz, zbc,
                                     dface[4], dface[5], P.lambda, df);
    };
    int rc = hipMemcpy(df, P.f.data(), fb, hipMemcpyHostToDevice) != hipSuccess;
    rc = rc || solve();   // warm up
    rc = rc || hipMemcpy(df, P.f.data(), fb, hipMemcpyHostToDevice) != hipSuccess;
    const double t0 = now();
    rc = rc || solve();
    const double t1 = now();
    if (rc) { printf("%s: solve failed\n", prog); return 1; }
    print_time(prog, t1 - t0);
    std::vector<double> sol(P.f.size());
    if (hipMemcpy(sol.data(), df, fb, hipMemcpyDeviceToHost) != hipSuccess) { printf("%s: copy failed\n", prog); return 1; }
    (void)hipFree(df);
    for (double *q : dface) if (q) (void)hipFree(q);
    printf("max diff : %.14f\n", P.max_diff(sol));
    return 0;
}

}  // namespace drv
