// star3d_tb2 kernel: two timesteps per launch for the 3D star stencils
// (star3d_tb2_tile), FP64, FP32 and FP32 compensated, 1/2 columns per thread.
#include "sweep_kernels.h"

namespace {

template <typename T, class S, int VX, bool KAHAN, int PF>
__global__ void __launch_bounds__(kMaxThreads) k_star3d_tb2(TB2Params<T, S::P> tp)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const KParams<T, S::P> &p = tp.k;
    int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x;
    bid /= p.tiles_x;
    const int tile_y = bid % p.tiles_y, chunk = bid / p.tiles_y;
    const int zs = chunk * p.ZC;
    const TileGeom g{p.n0, p.n1, p.n2, p.sy, p.sz, p.TX, p.TY, p.BXT, (p.flags & UHS_CFG_NT_STORE) != 0};
    const T *h0 = (tp.k0 & 1) ? tp.hB : tp.hA;
    const T *h1 = (tp.k0 & 1) ? tp.hA : tp.hB;
    Coefs<T, S::P> cf(p);
    const int x0 = tile_x * p.TX, y0 = tile_y * p.TY, ze = min(zs + p.ZC, p.n0);
    T *lds = reinterpret_cast<T *>(smem_raw);
    if (star3d_tb2_full<S::R>(x0, y0, zs, ze, p.TX, p.TY, p.n0, p.n1, p.n2))
        star3d_tb2_tile<T, S, VX, KAHAN, PF, kTb2RingCells, true>(p.in, p.out, h0, h1, g, x0, y0, zs, ze, lds, cf);
    else
        star3d_tb2_tile<T, S, VX, KAHAN, PF, kTb2RingCells, false>(p.in, p.out, h0, h1, g, x0, y0, zs, ze, lds, cf);
}

template <typename T, class S, int VX, bool KAHAN, int PF>
hipError_t launch_star3d_tb2(const SweepArgs &a, hipStream_t st)
{
    TB2Params<T, S::P> tp;
    tp.k = make_params<T, S::P>(a);
    constexpr int R = S::R;
    const int bx = a.tx + 2 * R, by = a.ty + 2 * R;
    tp.k.BXT = bx / VX;
    const long long h = a.halo, origin = h * a.sz + h * a.sy + h;
    tp.hA = static_cast<const T *>(a.hA) + origin;
    tp.hB = static_cast<const T *>(a.hB) + origin;
    tp.k0 = a.k0;
    const int nt = (bx / VX) * by;
    const size_t lds = 2 * sizeof(T) * ((size_t)(bx + 2 * R) * (by + 2 * R) + (size_t)bx * by);
    hipLaunchKernelGGL((k_star3d_tb2<T, S, VX, KAHAN, PF>), dim3(tp.k.nblocks), dim3(nt), lds, st, tp);
    return hipGetLastError();
}

template <typename T, bool K, class S> LaunchFn pick_vec(int vec, int pf)
{
    if (pf != 1) return nullptr;
    switch (vec) {
    case 1: return launch_star3d_tb2<T, S, 1, K, 1>;
    case 2: return launch_star3d_tb2<T, S, 2, K, 1>;
    }
    return nullptr;
}
template <class S> LaunchFn pick(int dtype, int vec, int pf)
{
    switch (dtype) {
    case DT_F64: return pick_vec<double, false, S>(vec, pf);
    case DT_F32: return pick_vec<float, false, S>(vec, pf);
    case DT_F32K: return pick_vec<float, true, S>(vec, pf);
    }
    return nullptr;
}
} // namespace

LaunchFn uhs_find_star3d_tb2(int prog, int dtype, int vec, int prefetch)
{
    switch (prog) {
    case Prog_3d1r_star::ID: return pick<Prog_3d1r_star>(dtype, vec, prefetch);
    case Prog_3d2r_star::ID: return pick<Prog_3d2r_star>(dtype, vec, prefetch);
    case Prog_3d3r_star::ID: return pick<Prog_3d3r_star>(dtype, vec, prefetch);
    }
    return nullptr;
}
