// FP32 axis pass, row tiles (x): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_axis_pass<float, float, float, e, true>
const void *axis_kernel_f32_rows(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
}  // namespace chz
