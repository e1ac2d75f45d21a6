// chz_transform.h -- host orchestration of the 3D transform.
#pragma once
#include <hip/hip_runtime.h>

#include "chz_context.h"

namespace chz {

void in_window(int kind, int N, int *lo, int *hi);
void out_window(int kind, int N, int *lo, int *hi);
int check_axis(int N, int kind, const char *axis);
int fill_pass(PassArgs &p, const AxisCfg &c, int prec, int kind, const void *tw);
bool pass_in_bounds(const PassArgs &p, bool rows, int nb, long long src_elems, long long dst_elems);
int launch_checked(const PassArgs &p, int prec, int ept, bool rows, int nb, long long src_elems, long long dst_elems,
                   hipStream_t s, const char *what);

// Padded (NX+2)(NY+2)(NZ+2) volume, x fastest; src == dst is in place.
template <typename T>
int transform3d(const T *src, T *dst, int NX, int NY, int NZ, int kx, int ky, int kz, hipStream_t s);

}  // namespace chz
