// chz_dropin.cpp -- the reference's C++ API (include/funcinterface.h,
// include/cuda_helmholtz_solver.h) forwarding to the stream-ordered C ABI.
// Each call synchronises before returning, as the reference does.
#include <cstdio>

#include "../../include/funcinterface.h"
#include "chz_context.h"
#include "chz_transform.h"

static void report(int rc)
{
    if (rc != chz::OK) printf("cuhelmholtz_amd: %s\n", chz::last_error());
}

static void finish(int rc)
{
    if (rc == chz::OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = chz::fail(chz::ERR_DEVICE, std::string("CUDA ERROR: ") + hipGetErrorString(e));
    }
    report(rc);
}

// cudasymmfft/funcinterface.cu:25-40
void do_transform(double *d_data, double *d_out, int NX, int NY, int NZ, transform_kind X_TRANS,
                  transform_kind Y_TRANS, transform_kind Z_TRANS)
{
    if (d_out == nullptr || d_out == d_data) {
        if (!(NX == NY && NX == NZ)) {
            chz::set_error("the matrix for inplace transform must be cubic!");
            printf("the matrix for inplace transform must be cubic!\n");
            return;
        }
        finish(chz::transform3d<double>(d_data, d_data, NX, NY, NZ, X_TRANS, Y_TRANS, Z_TRANS, 0));
        return;
    }
    finish(chz::transform3d<double>(d_data, d_out, NX, NY, NZ, X_TRANS, Y_TRANS, Z_TRANS, 0));
}

// cudasymmfft/funcinterface.cu:200-334 (result in d_out)
void do_transform_nocubic(double *d_data, double *d_out, int NX, int NY, int NZ, transform_kind X_TRANS,
                          transform_kind Y_TRANS, transform_kind Z_TRANS)
{
    finish(chz::transform3d<double>(d_data, d_out ? d_out : d_data, NX, NY, NZ, X_TRANS, Y_TRANS, Z_TRANS, 0));
}

// cudasymmfft/funcinterface.cu:42-198 (in place)
void do_transform_cubic(double *d_data, int NX, int NY, int NZ, transform_kind X_TRANS, transform_kind Y_TRANS,
                        transform_kind Z_TRANS)
{
    finish(chz::transform3d<double>(d_data, d_data, NX, NY, NZ, X_TRANS, Y_TRANS, Z_TRANS, 0));
}

// cudasymmfft/funcinterface.cu:337-342: plans are cached per device, nothing to destroy per call
void freeMemory_cubic() {}
