// FP64 solve pass 3 (forward z, divide, inverse z): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_solve_zmid<double, e>
const void *solve_zmid_kernel_f64(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
}  // namespace chz
