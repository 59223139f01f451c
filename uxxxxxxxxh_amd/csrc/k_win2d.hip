// win2d kernel instantiations: the six 2D benchmark stencils, FP64, FP32 and
// FP32 compensated, 1/2/4 points per thread along the unit-stride axis, one,
// two or four rows of loads in flight or a ring of 2r+1 rows (prefetch 0).
#include "sweep_kernels.h"

LaunchFn uhs_find_box3d_r1(int dtype, int vec, int deep, int planes);
LaunchFn uhs_find_box3d_r2(int dtype, int vec, int deep, int planes);
LaunchFn uhs_find_box3d_r3(int dtype, int vec, int deep, int planes);

namespace {
template <typename T, bool K, class S, int PF> LaunchFn pick_vec(int vec)
{
    switch (vec) {
    case 1: return launch_win2d<T, S, 1, K, PF>;
    case 2: return launch_win2d<T, S, 2, K, PF>;
    case 4:
        if constexpr (S::R <= 2) return launch_win2d<T, S, 4, K, PF>;
        else return nullptr;
    }
    return nullptr;
}
template <typename T, bool K, class S> LaunchFn pick_pf(int vec, int prefetch)
{
    switch (prefetch) {
    case 0: return pick_vec<T, K, S, 0>(vec);
    case 1: return pick_vec<T, K, S, 1>(vec);
    case 2: return pick_vec<T, K, S, 2>(vec);
    case 4: return pick_vec<T, K, S, 4>(vec);
    }
    return nullptr;
}
template <class S> LaunchFn pick(int dtype, int vec, int prefetch)
{
    switch (dtype) {
    case DT_F64: return pick_pf<double, false, S>(vec, prefetch);
    case DT_F32: return pick_pf<float, false, S>(vec, prefetch);
    case DT_F32K: return pick_pf<float, true, S>(vec, prefetch);
    }
    return nullptr;
}
} // namespace

LaunchFn uhs_find_win2d(int prog, int dtype, int vec, int prefetch)
{
    switch (prog) {
    case Prog_2d1r_star::ID: return pick<Prog_2d1r_star>(dtype, vec, prefetch);
    case Prog_2d2r_star::ID: return pick<Prog_2d2r_star>(dtype, vec, prefetch);
    case Prog_2d3r_star::ID: return pick<Prog_2d3r_star>(dtype, vec, prefetch);
    case Prog_2d1r_box::ID: return pick<Prog_2d1r_box>(dtype, vec, prefetch);
    case Prog_2d2r_box::ID: return pick<Prog_2d2r_box>(dtype, vec, prefetch);
    case Prog_2d3r_box::ID: return pick<Prog_2d3r_box>(dtype, vec, prefetch);
    }
    return nullptr;
}

LaunchFn uhs_find_box3d(int prog, int dtype, int vec, int deep, int planes)
{
    switch (prog) {
    case Prog_3d1r_box::ID: return uhs_find_box3d_r1(dtype, vec, deep, planes);
    case Prog_3d2r_box::ID: return uhs_find_box3d_r2(dtype, vec, deep, planes);
    case Prog_3d3r_box::ID: return uhs_find_box3d_r3(dtype, vec, deep, planes);
    }
    return nullptr;
}
