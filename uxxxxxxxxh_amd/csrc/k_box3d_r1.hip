// box3d kernel instantiations for the radius-1 3D box benchmark stencil,
// FP64, FP32 and FP32 compensated, 1/2 points per thread along the
// unit-stride axis, one or three cells of LDS/coefficient prefetch.
#include "sweep_kernels.h"

namespace {
template <typename T, bool K, class S> LaunchFn pick_vec(int vec, int deep)
{
    if (vec == 0) return launch_box3d_rows2<T, S, K>; // two rows per thread
    switch (vec * 2 + deep) {
    case 2: return launch_box3d<T, S, 1, K, 1>;
    case 3: return launch_box3d<T, S, 1, K, 3>;
    case 4: return launch_box3d<T, S, 2, K, 1>;
    case 5: return launch_box3d<T, S, 2, K, 3>;
    }
    return nullptr;
}
} // namespace

LaunchFn uhs_find_box3d_r1(int dtype, int vec, int deep)
{
    using S = Prog_3d1r_box;
    switch (dtype) {
    case DT_F64: return pick_vec<double, false, S>(vec, deep);
    case DT_F32: return pick_vec<float, false, S>(vec, deep);
    case DT_F32K: return pick_vec<float, true, S>(vec, deep);
    }
    return nullptr;
}
