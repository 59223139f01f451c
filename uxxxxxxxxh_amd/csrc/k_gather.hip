// gather kernel: the fallback for any stencil of up to kGatherMaxPoints terms
// that has no specialised kernel (another term order, one-sided offsets, ...).
#include "sweep_kernels.h"

namespace {
template <typename T, bool K>
hipError_t launch_gather(const SweepArgs &a, hipStream_t st)
{
    GParams<T> p;
    const long long h = a.halo;
    const long long origin = (a.dims == 3 ? h * a.sz : 0) + h * a.sy + h;
    p.in = static_cast<const T *>(a.in) + origin;
    p.out = static_cast<T *>(a.out) + origin;
    // 2D grids run as one plane of n0 rows.
    p.n0 = a.dims == 3 ? a.n0 : 1;
    p.n1 = a.dims == 3 ? a.n1 : a.n0;
    p.n2 = a.n2;
    p.P = a.npoints;
    p.sy = (int)a.sy;
    p.sz = a.dims == 3 ? a.sz : 0;
    for (int t = 0; t < a.npoints; t++) {
        long long o;
        if (a.dims == 3)
            o = a.off[3 * t] * a.sz + a.off[3 * t + 1] * a.sy + a.off[3 * t + 2];
        else
            o = a.off[2 * t] * a.sy + a.off[2 * t + 1];
        p.loff[t] = (int)o;
        p.c[t] = (T)a.coef[t];
    }
    const long long n = (long long)a.n0 * a.n1 * a.n2;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL((k_gather<T, K>), dim3(blocks), dim3(256), 0, st, p);
    return hipGetLastError();
}
} // namespace

LaunchFn uhs_find_gather(int dtype)
{
    switch (dtype) {
    case DT_F64: return launch_gather<double, false>;
    case DT_F32: return launch_gather<float, false>;
    case DT_F32K: return launch_gather<float, true>;
    }
    return nullptr;
}
