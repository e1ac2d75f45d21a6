// FP32 refinement pass 1 (FP64 residual + forward x): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_solve_xres<float, e>
const void *solve_xres_kernel_f32(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
}  // namespace chz
