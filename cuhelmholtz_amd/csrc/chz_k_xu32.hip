// FP32 last pass of a mixed solve (inverse x + FP64 update): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_solve_xinv_upd<float, e>
const void *solve_xinv_upd_kernel_f32(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
}  // namespace chz
