// star3d kernel instantiations: the three 3D star benchmark stencils, FP64,
// FP32 and FP32 compensated, 1/2/4/8 points per thread along the unit-stride
// axis, one or two planes of prefetch.
#include "sweep_kernels.h"

namespace {
template <typename T, bool K, class S, int PF> LaunchFn pick_vec(int vec)
{
    switch (vec) {
    case 1: return launch_star3d<T, S, 1, K, PF>;
    case 2: return launch_star3d<T, S, 2, K, PF>;
    case 4: return launch_star3d<T, S, 4, K, PF>;
    case 8: return launch_star3d<T, S, 8, K, PF>;
    }
    return nullptr;
}
template <class S, int PF> LaunchFn pick_pf(int dtype, int vec)
{
    switch (dtype) {
    case DT_F64: return pick_vec<double, false, S, PF>(vec);
    case DT_F32: return pick_vec<float, false, S, PF>(vec);
    case DT_F32K: return pick_vec<float, true, S, PF>(vec);
    }
    return nullptr;
}
template <class S> LaunchFn pick(int dtype, int vec, int prefetch)
{
    switch (prefetch) {
    case 1: return pick_pf<S, 1>(dtype, vec);
    case 2: return pick_pf<S, 2>(dtype, vec);
    }
    return nullptr;
}
} // namespace

LaunchFn uhs_find_star3d(int prog, int dtype, int vec, int prefetch)
{
    switch (prog) {
    case Prog_3d1r_star::ID: return pick<Prog_3d1r_star>(dtype, vec, prefetch);
    case Prog_3d2r_star::ID: return pick<Prog_3d2r_star>(dtype, vec, prefetch);
    case Prog_3d3r_star::ID: return pick<Prog_3d3r_star>(dtype, vec, prefetch);
    }
    return nullptr;
}
