// FP32 batched passes: axis columns, residual pass.
#include "chz_axis_kernels.h"

namespace chz {
#define KCOL(e) (const void *)k_batch<float, e, B_COLS>
#define KR(e) (const void *)k_batch<float, e, B_XRES>
const void *batch_kernel_f32_a(int op, int ept)
{
    if (op == B_COLS) { CHZ_EPT_SWITCH(ept, KCOL) }
    if (op == B_XRES) { CHZ_EPT_SWITCH(ept, KR) }
    return nullptr;
}
}  // namespace chz
