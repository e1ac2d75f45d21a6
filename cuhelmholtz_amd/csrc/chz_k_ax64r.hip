// FP64 axis pass, row tiles (x): instantiations for every elements-per-thread value.
#include "chz_axis_kernels.h"

namespace chz {
#define KSEL(e) (const void *)k_axis_pass<double, double, double, e, true>
const void *axis_kernel_f64_rows(int ept) { CHZ_EPT_SWITCH(ept, KSEL) }
const void *periodic_copy_kernel() { return (const void *)k_periodic_copy<double>; }
}  // namespace chz
