// win2d kernel instantiations: the six 2D benchmark stencils, FP64, FP32 and
// FP32 compensated, 1/2/4 points per thread along the unit-stride axis.
#include "sweep_kernels.h"

LaunchFn uhs_find_box3d_r1(int dtype, int vec, int deep);
LaunchFn uhs_find_box3d_r2(int dtype, int vec, int deep);
LaunchFn uhs_find_box3d_r3(int dtype, int vec, int deep);

namespace {
template <typename T, bool K, class S> LaunchFn pick_vec(int vec)
{
    switch (vec) {
    case 1: return launch_win2d<T, S, 1, K>;
    case 2: return launch_win2d<T, S, 2, K>;
    case 4:
        if constexpr (S::R <= 2) return launch_win2d<T, S, 4, K>;
        else return nullptr;
    }
    return nullptr;
}
template <class S> LaunchFn pick(int dtype, int vec)
{
    switch (dtype) {
    case DT_F64: return pick_vec<double, false, S>(vec);
    case DT_F32: return pick_vec<float, false, S>(vec);
    case DT_F32K: return pick_vec<float, true, S>(vec);
    }
    return nullptr;
}
} // namespace

LaunchFn uhs_find_win2d(int prog, int dtype, int vec)
{
    switch (prog) {
    case Prog_2d1r_star::ID: return pick<Prog_2d1r_star>(dtype, vec);
    case Prog_2d2r_star::ID: return pick<Prog_2d2r_star>(dtype, vec);
    case Prog_2d3r_star::ID: return pick<Prog_2d3r_star>(dtype, vec);
    case Prog_2d1r_box::ID: return pick<Prog_2d1r_box>(dtype, vec);
    case Prog_2d2r_box::ID: return pick<Prog_2d2r_box>(dtype, vec);
    case Prog_2d3r_box::ID: return pick<Prog_2d3r_box>(dtype, vec);
    }
    return nullptr;
}

LaunchFn uhs_find_box3d(int prog, int dtype, int vec, int deep)
{
    switch (prog) {
    case Prog_3d1r_box::ID: return uhs_find_box3d_r1(dtype, vec, deep);
    case Prog_3d2r_box::ID: return uhs_find_box3d_r2(dtype, vec, deep);
    case Prog_3d3r_box::ID: return uhs_find_box3d_r3(dtype, vec, deep);
    }
    return nullptr;
}
