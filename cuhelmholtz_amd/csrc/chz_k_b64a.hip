// FP64 batched passes: axis rows and columns, periodic copies.
#include "chz_axis_kernels.h"

namespace chz {
#define KROW(e) (const void *)k_batch<double, e, B_ROWS>
#define KCOL(e) (const void *)k_batch<double, e, B_COLS>
const void *batch_kernel_f64_axis(int op, int ept)
{
    if (op == B_ROWS) { CHZ_EPT_SWITCH(ept, KROW) }
    CHZ_EPT_SWITCH(ept, KCOL)
}
const void *batch_periodic_kernel() { return (const void *)k_batch_periodic<BatchCopy>; }
}  // namespace chz
