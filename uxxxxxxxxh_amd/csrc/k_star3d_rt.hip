// star3d_rt: register-tiled 3D star sweep with no LDS and no barriers.
//
// Each wavefront owns a 64-column x TYR-row band of the grid and streams it
// along axis 0; waves never exchange data through LDS and never wait for one
// another.  A lane holds one x column of its band: TYR rows times the 2R+1
// planes z-R..z+R in registers.  Neighbours come from
//   axis 0: the lane's own plane queue,
//   axis 1: the lane's own rows, plus R rows above and below the band that
//           every lane loads for the current plane (L2-resident: they are
//           the neighbouring bands' rows),
//   axis 2: the neighbouring lanes, by DPP wavefront shifts.
// The R lanes at each end of the wave only load (they are the x halo of the
// 64 - 2R columns the wave computes), so bands overlap by 2R columns.
// Terms are evaluated in DSL order with the same product-then-fma chain as
// every other kernel (bitwise identical results).
#include "sweep_kernels.h"

namespace {

// Whole-wavefront shifts by one lane with DPP (a VALU move, no LDS trip):
// wave_shl:1 gives lane l the value of lane l+1, wave_shr:1 that of lane
// l-1.  The lane shifted in at the wave's end reads 0; those R lanes at each
// end only load and never store.
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;
template <int CTRL> __device__ __forceinline__ int dpp32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ __forceinline__ float dpp_shift(float v)
{
    return __builtin_bit_cast(float, dpp32<CTRL>(__builtin_bit_cast(int, v)));
}
template <int CTRL> __device__ __forceinline__ double dpp_shift(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp32<CTRL>((int)(unsigned)u);
    const unsigned hi = (unsigned)dpp32<CTRL>((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <typename T, class S, int TYR, bool KAHAN>
__global__ void __launch_bounds__(kMaxThreads) k_star3d_rt(KParams<T, S::P> p)
{
    constexpr int R = S::R, P = S::P, Q = 2 * R + 1, XW = 64 - 2 * R;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x;
    bid /= p.tiles_x;
    const int tile_y = bid % p.tiles_y, chunk = bid / p.tiles_y;
    const int zs = chunk * p.ZC, ze = min(zs + p.ZC, p.n0);
    const T *__restrict__ in = p.in;
    T *__restrict__ out = p.out;
    const long long sz = p.sz;
    const int sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;

    const int xg = tile_x * XW - R + lane;
    const bool xact = lane >= R && lane < 64 - R && xg < p.n2;
    const int xc = clampi(xg, -R, p.n2 + R - 1);
    const int y0 = tile_y * p.TY + wave * TYR;
    int roff[TYR], hoff[2 * R];
    bool yact[TYR];
#pragma unroll
    for (int i = 0; i < TYR; i++) {
        yact[i] = y0 + i < p.n1;
        roff[i] = clampi(y0 + i, -R, p.n1 + R - 1) * sy + xc;
    }
#pragma unroll
    for (int k = 0; k < R; k++) {
        hoff[k] = clampi(y0 - R + k, -R, p.n1 + R - 1) * sy + xc;
        hoff[R + k] = clampi(y0 + TYR + k, -R, p.n1 + R - 1) * sy + xc;
    }

    Coefs<T, P> cf(p);
    // q[j][i] = plane z-R+j, row y0+i; nq = plane z+R (in flight);
    // hy/nhy = the 2R halo rows of plane z / z+1.
    T q[Q][TYR], nq[TYR], hy[2 * R], nhy[2 * R];
#pragma unroll
    for (int j = 0; j < 2 * R; j++)
#pragma unroll
        for (int i = 0; i < TYR; i++) q[j][i] = in[(long long)(zs - R + j) * sz + roff[i]];
#pragma unroll
    for (int i = 0; i < TYR; i++) nq[i] = in[(long long)(zs + R) * sz + roff[i]];
#pragma unroll
    for (int k = 0; k < 2 * R; k++) nhy[k] = in[(long long)zs * sz + hoff[k]];

    for (int z = zs; z < ze; ++z) {
#pragma unroll
        for (int i = 0; i < TYR; i++) q[2 * R][i] = nq[i];
#pragma unroll
        for (int k = 0; k < 2 * R; k++) hy[k] = nhy[k];
        {
            const T *pn = in + (long long)min(z + 1 + R, p.n0 + R - 1) * sz;
            const T *ph = in + (long long)(z + 1) * sz;
#pragma unroll
            for (int i = 0; i < TYR; i++) nq[i] = pn[roff[i]];
#pragma unroll
            for (int k = 0; k < 2 * R; k++) nhy[k] = ph[hoff[k]];
        }
        cf.refresh();
        T *oz = out + (long long)z * sz;
        static_for<TYR>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            // The row's x neighbours at distance 1..R, one DPP shift each.
            T dn[R + 1], up[R + 1];
            dn[0] = up[0] = q[R][i];
            static_for<R>([&](auto dc) {
                constexpr int d = decltype(dc)::value + 1;
                dn[d] = dpp_shift<kDppWaveShl1>(dn[d - 1]);
                up[d] = dpp_shift<kDppWaveShr1>(up[d - 1]);
            });
            Acc<T, KAHAN> a;
            static_for<P>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr int dz = S::OFF[t][0], dy = S::OFF[t][1], dx = S::OFF[t][2];
                static_assert((dx != 0) + (dy != 0) + (dz != 0) <= 1, "star3d_rt needs axis-aligned terms");
                T val;
                if constexpr (dx > 0) val = dn[dx];
                else if constexpr (dx < 0) val = up[-dx];
                else if constexpr (dy != 0 && i + dy < 0) val = hy[R + i + dy];        // row above the band
                else if constexpr (dy != 0 && i + dy >= TYR) val = hy[R + i + dy - TYR]; // row below the band
                else if constexpr (dy != 0) val = q[R][i + dy];
                else val = q[R + dz][i];
                if constexpr (t == 0) a.first(cf.template get<t>(), val);
                else a.add(cf.template get<t>(), val);
            });
            if (xact && yact[i]) store_out(oz + roff[i], a.value(), nts);
        });
#pragma unroll
        for (int j = 0; j < 2 * R; j++)
#pragma unroll
            for (int i = 0; i < TYR; i++) q[j][i] = q[j + 1][i];
    }
}

template <typename T, class S, int TYR, bool KAHAN>
hipError_t launch_star3d_rt(const SweepArgs &a, hipStream_t st)
{
    auto p = make_params<T, S::P>(a);
    constexpr int XW = 64 - 2 * S::R;
    p.tiles_x = (a.n2 + XW - 1) / XW;
    const int chunks = (a.n0 + a.tz - 1) / a.tz;
    p.nblocks = p.tiles_x * p.tiles_y * chunks;
    const int nt = 64 * (a.ty / TYR);
    hipLaunchKernelGGL((k_star3d_rt<T, S, TYR, KAHAN>), dim3(p.nblocks), dim3(nt), 0, st, p);
    return hipGetLastError();
}

template <typename T, bool K, class S> LaunchFn pick_rows(int rows)
{
    switch (rows) {
    case 2: return launch_star3d_rt<T, S, 2, K>;
    case 4: return launch_star3d_rt<T, S, 4, K>;
    case 8: return launch_star3d_rt<T, S, 8, K>;
    }
    return nullptr;
}
template <class S> LaunchFn pick(int dtype, int rows)
{
    switch (dtype) {
    case DT_F64: return pick_rows<double, false, S>(rows);
    case DT_F32: return pick_rows<float, false, S>(rows);
    case DT_F32K: return pick_rows<float, true, S>(rows);
    }
    return nullptr;
}

} // namespace

// rows = points per lane along axis 1 (the config's vec).
LaunchFn uhs_find_star3d_rt(int prog, int dtype, int rows)
{
    switch (prog) {
    case Prog_3d1r_star::ID: return pick<Prog_3d1r_star>(dtype, rows);
    case Prog_3d2r_star::ID: return pick<Prog_3d2r_star>(dtype, rows);
    case Prog_3d3r_star::ID: return pick<Prog_3d3r_star>(dtype, rows);
    }
    return nullptr;
}
