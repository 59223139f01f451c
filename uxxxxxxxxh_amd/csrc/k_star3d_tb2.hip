// star3d_tb2: two timesteps per launch for the 3D star stencils (temporal
// blocking with an R-cell ghost zone), so each cell crosses HBM once per two
// updates instead of once per update.
//
// A workgroup owns an output tile of tile_x * tile_y columns and covers it
// extended by R on each side (the ghost zone): BX = tile_x + 2R by
// BY = tile_y + 2R columns, VX columns per thread along the unit-stride
// axis.  It streams planes along axis 0.  At iteration i (plane z1 of
// level 1):
//   level 0 (the input): each thread holds its columns' 2R+1 planes around z1
//     in a register queue (plus one plane in flight) and stages plane z1,
//     with an R-wide ring around the extended tile, in LDS buffer i&1;
//   level 1 (one step): the thread stages plane z1-1-R of level 1 (the
//     centre of its level-1 register queue) in LDS buffer i&1 too;
//   -- one barrier --
//   level 1 at plane z1 for every extended column: axis-0 neighbours from the
//     level-0 queue, in-plane neighbours from LDS;
//   level 2 at plane z1-1-R for the tile's own columns: axis-0 neighbours
//     from the level-1 queue, in-plane neighbours from LDS.
// Buffer i&1 is rewritten two iterations later, after the next barrier, so
// one barrier per plane suffices.
//
// Cells of a level outside the interior hold that level's halo: the halo of
// the buffer the one-step ping-pong would hold that state in (the sweep
// never writes halos), passed as hA/hB with the global step k0 of the
// input.  Every update is the product-then-fma chain in term order, so the
// result is bitwise identical to two single sweeps.
#include "sweep_kernels.h"

namespace {

template <typename T, int P> struct TB2Params {
    KParams<T, P> k;  // first member: the coefficient pack keeps its kernarg offset
    const T *hA, *hB; // interior origins of the grids the ping-pong calls A and B
    int k0;           // global step of the input state
};

template <typename T, class S, int VX, bool KAHAN, class CF>
__device__ __forceinline__ void star3d_tb2_tile(const T *__restrict__ in, T *__restrict__ out,
                                                const T *__restrict__ h0, const T *__restrict__ h1,
                                                const TileGeom &g, int x0, int y0, int zs, int ze,
                                                T *lds, CF &cf)
{
    constexpr int R = S::R, P = S::P, Q = 2 * R + 1, NH = kStarHaloPerThread;
    const int NT = blockDim.x, tid = threadIdx.x;
    const int BXT = g.BXT, BX = g.TX + 2 * R, BY = g.TY + 2 * R;
    const int LW = BX + 2 * R, L0SZ = LW * (BY + 2 * R), L1SZ = BX * BY;
    T *L0 = lds;            // 2 x L0SZ: level-0 plane + ring
    T *L1 = lds + 2 * L0SZ; // 2 x L1SZ: level-1 plane over the extended tile
    const int ox = x0 - R, oy = y0 - R; // interior coordinates of extended cell (0, 0)
    const int tx = tid % BXT, ty = tid / BXT;
    const long long sz = g.sz;
    const int sy = g.sy, n0 = g.n0, n1 = g.n1, n2 = g.n2;

    int coff[VX], l0c[VX], l1c[VX];
    bool xyin[VX], core[VX];
    const int yg = oy + ty;
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tx + v * BXT, xg = ox + ex;
        xyin[v] = xg >= 0 && xg < n2 && yg >= 0 && yg < n1;
        core[v] = xyin[v] && ex >= R && ex < BX - R && ty >= R && ty < BY - R;
        coff[v] = clampi(yg, -R, n1 + R - 1) * sy + clampi(xg, -R, n2 + R - 1);
        l0c[v] = (ty + R) * LW + ex + R;
        l1c[v] = ty * BX + ex;
    }
    // Level-0 ring around the extended tile: R rows above and below, R
    // columns left and right (a star needs no corners).
    const int H = 2 * R * (BX + BY);
    int hoff[NH], hl[NH];
    bool hin[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int c = tid + k * NT;
        hl[k] = -1;
        hoff[k] = 0;
        hin[k] = false;
        if (c < H) {
            int ly, lx;
            if (c < 2 * R * BX) {
                const int row = c / BX;
                ly = row < R ? row : BY + row;
                lx = R + c % BX;
            } else {
                const int cc = c - 2 * R * BX, col = cc % (2 * R);
                ly = R + cc / (2 * R);
                lx = col < R ? col : BX + col;
            }
            const int gy = oy - R + ly, gx = ox - R + lx;
            hl[k] = ly * LW + lx;
            hin[k] = gy >= 0 && gy < n1 && gx >= 0 && gx < n2;
            hoff[k] = clampi(gy, -R, n1 + R - 1) * sy + clampi(gx, -R, n2 + R - 1);
        }
    }
    // Level-0 value of a cell at plane z: the input inside the interior, the
    // level-0 halo outside it.  Planes beyond the allocation are clamped;
    // their values feed only level-1 cells outside the interior, which take
    // the level-1 halo instead.
    auto lev0 = [&](bool cin, int z, int off) -> T {
        const long long a = (long long)clampi(z, -R, n0 + R - 1) * sz + off;
        return (cin && z >= 0 && z < n0) ? in[a] : h0[a];
    };

    const int z1s = zs - R;             // first level-1 plane
    const int N = (ze - zs) + 2 * R;    // level-1 planes
    T q0[VX][Q], q1[VX][Q], nq0[VX], nh[NH];
#pragma unroll
    for (int v = 0; v < VX; v++) {
#pragma unroll
        for (int j = 0; j < 2 * R; j++) q0[v][j] = lev0(xyin[v], z1s - R + j, coff[v]);
        nq0[v] = lev0(xyin[v], z1s + R, coff[v]);
#pragma unroll
        for (int j = 0; j < Q; j++) q1[v][j] = T(0);
    }
#pragma unroll
    for (int k = 0; k < NH; k++)
        if (hl[k] >= 0) nh[k] = lev0(hin[k], z1s, hoff[k]);

    for (int i = 0; i <= N; ++i) {
        const int z1 = z1s + i;
        T *S0 = L0 + (i & 1) * L0SZ;
        T *S1 = L1 + (i & 1) * L1SZ;
        if (i < N) {
#pragma unroll
            for (int v = 0; v < VX; v++) {
                q0[v][2 * R] = nq0[v];
                S0[l0c[v]] = q0[v][R];
            }
#pragma unroll
            for (int k = 0; k < NH; k++)
                if (hl[k] >= 0) S0[hl[k]] = nh[k];
            if (i + 1 < N) {
#pragma unroll
                for (int v = 0; v < VX; v++) nq0[v] = lev0(xyin[v], z1 + R + 1, coff[v]);
#pragma unroll
                for (int k = 0; k < NH; k++)
                    if (hl[k] >= 0) nh[k] = lev0(hin[k], z1 + 1, hoff[k]);
            }
        }
#pragma unroll
        for (int v = 0; v < VX; v++) S1[l1c[v]] = q1[v][R];
        __syncthreads();
        cf.refresh();
        // Level 2 at plane z2 = z1 - 1 - R: the level-1 queue holds planes
        // z1-1-2R .. z1-1, the LDS buffer plane z2.
        const int z2 = z1 - 1 - R;
        if (z2 >= zs) {
            T *oz = out + (long long)z2 * sz;
#pragma unroll
            for (int v = 0; v < VX; v++) {
                if (!core[v]) continue; // ghost columns: their neighbours leave the tile
                Acc<T, KAHAN> a;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int dz = S::OFF[t][0], dy = S::OFF[t][1], dx = S::OFF[t][2];
                    T val;
                    if constexpr (dy == 0 && dx == 0) val = q1[v][R + dz];
                    else val = S1[l1c[v] + dy * BX + dx];
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                store_out(oz + coff[v], a.value(), g.nts);
            }
        }
        if (i < N) {
            // Level 1 at plane z1 for every extended column.
            const bool zin = z1 >= 0 && z1 < n0;
#pragma unroll
            for (int v = 0; v < VX; v++) {
                T v1;
                if (zin && xyin[v]) {
                    Acc<T, KAHAN> a;
                    static_for<P>([&](auto tc) {
                        constexpr int t = decltype(tc)::value;
                        constexpr int dz = S::OFF[t][0], dy = S::OFF[t][1], dx = S::OFF[t][2];
                        T val;
                        if constexpr (dy == 0 && dx == 0) val = q0[v][R + dz];
                        else val = S0[l0c[v] + dy * LW + dx];
                        if constexpr (t == 0) a.first(cf.template get<t>(), val);
                        else a.add(cf.template get<t>(), val);
                    });
                    v1 = a.value();
                } else {
                    v1 = h1[(long long)z1 * sz + coff[v]];
                }
#pragma unroll
                for (int j = 0; j < 2 * R; j++) q1[v][j] = q1[v][j + 1];
                q1[v][2 * R] = v1;
#pragma unroll
                for (int j = 0; j < 2 * R; j++) q0[v][j] = q0[v][j + 1];
            }
        }
    }
}

template <typename T, class S, int VX, bool KAHAN>
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
    star3d_tb2_tile<T, S, VX, KAHAN>(p.in, p.out, h0, h1, g, tile_x * p.TX, tile_y * p.TY, zs,
                                     min(zs + p.ZC, p.n0), reinterpret_cast<T *>(smem_raw), cf);
}

template <typename T, class S, int VX, bool KAHAN>
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
    hipLaunchKernelGGL((k_star3d_tb2<T, S, VX, KAHAN>), dim3(tp.k.nblocks), dim3(nt), lds, st, tp);
    return hipGetLastError();
}

template <typename T, bool K, class S> LaunchFn pick_vec(int vec)
{
    switch (vec) {
    case 1: return launch_star3d_tb2<T, S, 1, K>;
    case 2: return launch_star3d_tb2<T, S, 2, K>;
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

LaunchFn uhs_find_star3d_tb2(int prog, int dtype, int vec)
{
    switch (prog) {
    case Prog_3d1r_star::ID: return pick<Prog_3d1r_star>(dtype, vec);
    case Prog_3d2r_star::ID: return pick<Prog_3d2r_star>(dtype, vec);
    case Prog_3d3r_star::ID: return pick<Prog_3d3r_star>(dtype, vec);
    }
    return nullptr;
}
