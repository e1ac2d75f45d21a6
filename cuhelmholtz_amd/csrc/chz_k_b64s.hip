// FP64 batched solve passes: pass 1 with folds, fused z pass.
#include "chz_axis_kernels.h"

namespace chz {
const void *batch_kernel_f64_axis(int op, int ept);
#define KX(e) (const void *)k_batch<double, e, B_XFWD>
#define KZ(e) (const void *)k_batch<double, e, B_ZMID>
static const void *xfwd(int ept) { CHZ_EPT_SWITCH(ept, KX) }
static const void *zmid(int ept) { CHZ_EPT_SWITCH(ept, KZ) }
const void *batch_kernel_f64(int op, int ept)
{
    switch (op) {
        case B_ROWS:
        case B_COLS: return batch_kernel_f64_axis(op, ept);
        case B_XFWD: return xfwd(ept);
        case B_ZMID: return zmid(ept);
    }
    return nullptr;
}
}  // namespace chz
