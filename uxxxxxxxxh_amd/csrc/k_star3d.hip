// star3d kernel instantiations: the three 3D star benchmark stencils, FP64,
// FP32 and FP32 compensated, 1/2/4/8 points per thread along the unit-stride
// axis, one or two planes of prefetch.
#include "sweep_kernels.h"

namespace {
template <typename T, bool K, class S, int PF, int NH> LaunchFn pick_vec(int vec)
{
    switch (vec) {
    case 1: return launch_star3d<T, S, 1, K, PF, NH>;
    case 2: return launch_star3d<T, S, 2, K, PF, NH>;
    case 4: return launch_star3d<T, S, 4, K, PF, NH>;
    case 8: return launch_star3d<T, S, 8, K, PF, NH>;
    }
    return nullptr;
}
// Halo ring cells per thread: the radius-1 kernels come in 1, 2 and 4 (the
// ring loads of their full tiles are unconditional, so unused slots cost
// loads); the others always stage up to 4.
template <class S, int PF> LaunchFn pick_nh(int dtype, int vec, int nh)
{
    if constexpr (S::R == 1) {
        if (nh <= 1) nh = 1;
        else if (nh <= 2) nh = 2;
        else nh = 4;
    } else {
        nh = 4;
    }
    switch (dtype * 8 + nh) {
    case DT_F64 * 8 + 4: return pick_vec<double, false, S, PF, 4>(vec);
    case DT_F32 * 8 + 4: return pick_vec<float, false, S, PF, 4>(vec);
    case DT_F32K * 8 + 4: return pick_vec<float, true, S, PF, 4>(vec);
    }
    if constexpr (S::R == 1) {
        switch (dtype * 8 + nh) {
        case DT_F64 * 8 + 1: return pick_vec<double, false, S, PF, 1>(vec);
        case DT_F64 * 8 + 2: return pick_vec<double, false, S, PF, 2>(vec);
        case DT_F32 * 8 + 1: return pick_vec<float, false, S, PF, 1>(vec);
        case DT_F32 * 8 + 2: return pick_vec<float, false, S, PF, 2>(vec);
        case DT_F32K * 8 + 1: return pick_vec<float, true, S, PF, 1>(vec);
        case DT_F32K * 8 + 2: return pick_vec<float, true, S, PF, 2>(vec);
        }
    }
    return nullptr;
}
template <class S> LaunchFn pick(int dtype, int vec, int prefetch, int nh)
{
    switch (prefetch) {
    case 1: return pick_nh<S, 1>(dtype, vec, nh);
    case 2: return pick_nh<S, 2>(dtype, vec, nh);
    }
    return nullptr;
}
} // namespace

LaunchFn uhs_find_star3d(int prog, int dtype, int vec, int prefetch, int ring_cells)
{
    switch (prog) {
    case Prog_3d1r_star::ID: return pick<Prog_3d1r_star>(dtype, vec, prefetch, ring_cells);
    case Prog_3d2r_star::ID: return pick<Prog_3d2r_star>(dtype, vec, prefetch, ring_cells);
    case Prog_3d3r_star::ID: return pick<Prog_3d3r_star>(dtype, vec, prefetch, ring_cells);
    }
    return nullptr;
}
