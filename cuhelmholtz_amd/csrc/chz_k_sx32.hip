// FP32 solve pass 1 (folds + forward x): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_solve_xfwd<float, e>
const void *solve_xfwd_kernel_f32(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
}  // namespace chz
