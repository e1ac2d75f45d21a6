// FP64 axis pass, column tiles (y, z): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_axis_pass<double, double, double, e, false>
const void *axis_kernel_f64_cols(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
}  // namespace chz
