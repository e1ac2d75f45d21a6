// FP32 batched solve passes: pass 1 with folds, fused z pass, update pass.
#include "chz_axis_kernels.h"

namespace chz {
const void *batch_kernel_f32_a(int op, int ept);
#define KX(e) (const void *)k_batch<float, e, B_XFWD>
#define KZ(e) (const void *)k_batch<float, e, B_ZMID>
#define KU(e) (const void *)k_batch<float, e, B_XUPD>
static const void *xfwd(int ept) { CHZ_EPT_SWITCH(ept, KX) }
static const void *zmid(int ept) { CHZ_EPT_SWITCH(ept, KZ) }
static const void *xupd(int ept) { CHZ_EPT_SWITCH(ept, KU) }
const void *batch_kernel_f32(int op, int ept)
{
    switch (op) {
        case B_COLS:
        case B_XRES: return batch_kernel_f32_a(op, ept);
        case B_XFWD: return xfwd(ept);
        case B_ZMID: return zmid(ept);
        case B_XUPD: return xupd(ept);
    }
    return nullptr;
}
}  // namespace chz
