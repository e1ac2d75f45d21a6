// FP32 solve pass 3 (forward z, divide, inverse z): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_solve_zmid<float, e>
const void *solve_zmid_kernel_f32(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
}  // namespace chz
