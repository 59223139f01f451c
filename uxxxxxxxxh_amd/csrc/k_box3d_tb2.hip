// box3d_tb2 kernel: two timesteps per launch for the 3D boxes whose terms
// come plane by plane in ascending axis-0 offset, every plane listing the
// same in-plane cells in the same order (the benchmark boxes).  Compiled for
// the radius-1 box, the one HBM bounds: the radius-2 and radius-3 boxes are
// bound by FP64 FMAs, which the ghost zone would add to.  FP64, FP32 and FP32
// compensated, 1 or 2 columns per thread (and 4 in FP32).
//
// A workgroup owns a tile_x * tile_y output tile and covers it extended by R
// on each side (BX = tile_x + 2R by BY = tile_y + 2R cells, VX columns per
// thread along the unit-stride axis), streaming planes along axis 0.  At
// step s, level-0 plane p = zs - 2R + s (the input) arrives: it is staged in
// LDS with an R-wide ring (corners included), and level-1 plane p - R - 1
// (computed at step s - 1) is staged beside it.
//   -- one barrier --
//   level 2: each thread reads the (2R+1)^2 in-plane cells of level-1 plane
//     p - R - 1 around its own cells once, in term order, and applies each
//     to the 2R+1 pending output planes it is a term of (a running sum per
//     pending plane); output plane p - 2R - 1 completes and is stored;
//   level 1: the same with level-0 plane p over the extended tile; plane
//     p - R completes and is kept in a register until the next step stages
//     it.
// An output plane q receives its plane groups in ascending axis-0 offset as
// the planes arrive, and each group's cells in term order, so every sum is
// the product-then-fma chain in term order: bitwise that of two single
// sweeps.  The running sums sit in a ring indexed by the step at which their
// plane completes; the plane loop is unrolled by 2R+1 so the ring slots are
// constants, and 2R+1 planes of loads are in flight in a register ring.
//
// Cells of a level outside the interior hold that level's halo: the halo of
// the grid the one-step ping-pong holds that state in (h0 for the input, h1
// for level 1).
#include "sweep_kernels.h"

namespace {

template <int R>
__device__ __forceinline__ bool box3d_tb2_full(int x0, int y0, int zs, int ze, int TX, int TY, int n0, int n1, int n2)
{
    return x0 >= 2 * R && x0 + TX + 2 * R <= n2 && y0 >= 2 * R && y0 + TY + 2 * R <= n1 && zs >= 2 * R &&
           ze + 2 * R <= n0;
}

// FULL: every cell whose value is used (the extended tile, its ring and the
// slab's 2R planes either side) lies in the interior: no halo selects and
// no clamps but the prefetch's past the last plane.
template <typename T, class S, int VX, bool KAHAN, bool FULL, class CF>
__device__ __forceinline__ void box3d_tb2_tile(const T *__restrict__ in, T *__restrict__ out,
                                               const T *__restrict__ h0, const T *__restrict__ h1,
                                               const KParams<T, S::P> &p, int x0, int y0, int zs, int ze, T *lds,
                                               CF &cf)
{
    constexpr int R = S::R, P = S::P, W = 2 * R + 1, G = P / W, NH = kBoxTb2RingCells;
    static_assert(uniform_planes<S>() && S::OFF[0][0] == -R && S::OFF[P - 1][0] == R,
                  "box3d_tb2 needs 2R+1 uniform plane groups");
    const int NT = blockDim.x, tid = threadIdx.x;
    const int BXT = p.BXT, BX = p.TX + 2 * R, BY = p.TY + 2 * R;
    const int LW = BX + 2 * R, L0SZ = LW * (BY + 2 * R), L1SZ = BX * BY;
    T *L0 = lds;            // 2 x L0SZ: level-0 plane over the extended tile + ring
    T *L1 = lds + 2 * L0SZ; // 2 x L1SZ: level-1 plane over the extended tile
    lds_poison(lds, 2 * (L0SZ + L1SZ));
    const int tx = tid % BXT, ty = tid / BXT;
    const long long sz = p.sz;
    const int sy = p.sy, n0 = p.n0, n1 = p.n1, n2 = p.n2;
    const bool nts = p.flags & UHS_CFG_NT_STORE;
    const int ox = x0 - R, oy = y0 - R; // interior coordinates of extended cell (0, 0)

    // Own extended cells (tx + v*BXT, ty).
    const int yg = oy + ty;
    int coff[VX], l0c[VX], l1c[VX];
    bool xyin[VX], core[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tx + v * BXT, xg = ox + ex;
        xyin[v] = FULL || (xg >= 0 && xg < n2 && yg >= 0 && yg < n1);
        core[v] = ex >= R && ex < BX - R && ty >= R && ty < BY - R && (FULL || (xg < n2 && yg < n1));
        coff[v] = FULL ? yg * sy + xg : clampi(yg, -R, n1 + R - 1) * sy + clampi(xg, -R, n2 + R - 1);
        l0c[v] = (ty + R) * LW + ex + R;
        l1c[v] = ty * BX + ex;
    }
    // Level-0 ring around the extended tile: R rows above and below (full
    // width, corners included), R columns left and right.  Threads past the
    // ring load a valid cell they never stage, so the loads stay
    // unconditional.
    const int H = 2 * R * LW + 2 * R * BY;
    int hoff[NH], hl[NH];
    bool hin[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int c = tid + k * NT;
        hl[k] = -1;
        hoff[k] = coff[0];
        hin[k] = xyin[0];
        if (c < H) {
            int ly, lx;
            if (c < 2 * R * LW) {
                const int row = c / LW;
                ly = row < R ? row : BY + row;
                lx = c % LW;
            } else {
                const int cc = c - 2 * R * LW, col = cc % (2 * R);
                ly = R + cc / (2 * R);
                lx = col < R ? col : BX + col;
            }
            const int gy = oy - R + ly, gx = ox - R + lx;
            hl[k] = ly * LW + lx;
            hin[k] = FULL || (gy >= 0 && gy < n1 && gx >= 0 && gx < n2);
            hoff[k] = FULL ? gy * sy + gx : clampi(gy, -R, n1 + R - 1) * sy + clampi(gx, -R, n2 + R - 1);
        }
    }
    // Level-0 value at plane z: the input inside the interior, h0 outside it.
    // Planes beyond the allocation are clamped; their values are never used.
    auto lev0 = [&](bool cin, int z, int off) -> T {
        if constexpr (FULL) {
            return in[(long long)min(z, n0 + R - 1) * sz + off];
        } else {
            const T *base = (cin && z >= 0 && z < n0) ? in : h0;
            return base[(long long)clampi(z, -R, n0 + R - 1) * sz + off];
        }
    };

    // Prefetch ring: slot f holds the level-0 plane staged at the next step
    // of phase f, loaded 2R+1 planes ahead.
    T nrow[W][VX], nh[W][NH];
#pragma unroll
    for (int f = 0; f < W; f++) {
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[f][v] = lev0(xyin[v], zs - 2 * R + f, coff[v]);
#pragma unroll
        for (int k = 0; k < NH; k++) nh[f][k] = lev0(hin[k], zs - 2 * R + f, hoff[k]);
    }
    // Running sums per pending plane and level: slot = phase of the step at
    // which the plane completes.
    Acc<T, KAHAN> s0[VX][W], s1[VX][W];
    T v1[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        v1[v] = T(0);
#pragma unroll
        for (int a = 0; a < W; a++) {
            s0[v][a].zero();
            s1[v][a].zero();
        }
    }
    const int N = (ze - zs) + 4 * R + 1; // steps

    // Plane arriving at a step of phase PH, read around LDS cell c of a plane
    // of row width lw: output plane q = (arriving) - R + k takes the group of
    // axis-0 offset R - k and completes k steps later, in slot (PH + k) mod W.
    auto feed = [&](auto phc, Acc<T, KAHAN>(&sums)[W], const T *Lp, int c, int lw) {
        constexpr int PH = decltype(phc)::value;
        static_for<G>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const T val = Lp[c + S::OFF[i][1] * lw + S::OFF[i][2]];
            static_for<W>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int t = (2 * R - k) * G + i, SL = (PH + k) % W;
                if constexpr (t == 0) sums[SL].first(cf.template get<t>(), val);
                else sums[SL].add(cf.template get<t>(), val);
            });
            // After each row of cells, pin the sums: the compiler then
            // finishes that row's FMAs before it issues later LDS reads, so
            // at most about two rows of cells are live (FP64, 2 columns per
            // thread: 113 VGPRs instead of 164, 4 waves per SIMD instead of
            // 3; 373-376 against 368 GSt/s at 512^3).
            if constexpr ((i + 1) % W == 0 && i + 1 < G) {
#pragma unroll
                for (int j = 0; j < W; j++) sums[j].pin();
            }
        });
    };

    // One step.  L1ON: the completing level-1 plane is one the tile needs
    // (s >= 2R); W1ON: a level-1 plane arrives (s >= 2R+1); OUT: the
    // completing level-2 plane is stored (s >= 4R+1).
    auto step = [&](auto phc, auto l1c_, auto w1c, auto outc, int s) {
        constexpr int PH = decltype(phc)::value;
        constexpr bool L1ON = decltype(l1c_)::value, W1ON = decltype(w1c)::value, OUT = decltype(outc)::value;
        const int z = zs - 2 * R + s; // arriving level-0 plane
        T *S0 = L0 + (s & 1) * L0SZ;
        T *S1 = L1 + (s & 1) * L1SZ;
#pragma unroll
        for (int v = 0; v < VX; v++) S0[l0c[v]] = nrow[PH][v];
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (hl[k] >= 0) S0[hl[k]] = nh[PH][k];
        if constexpr (W1ON) {
#pragma unroll
            for (int v = 0; v < VX; v++) S1[l1c[v]] = v1[v];
        }
        // Refill the slot with the plane 2R+1 ahead (clamped past the end).
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[PH][v] = lev0(xyin[v], z + W, coff[v]);
#pragma unroll
        for (int k = 0; k < NH; k++) nh[PH][k] = lev0(hin[k], z + W, hoff[k]);
        __syncthreads();
        cf.refresh();
        if constexpr (W1ON) {
            // Level 2: level-1 plane z - R - 1 arrives, plane z - 2R - 1
            // completes.  Ghost cells' neighbours leave the tile: skipped.
            T *oz = out + (long long)(z - 2 * R - 1) * sz;
#pragma unroll
            for (int v = 0; v < VX; v++) {
                if (!core[v]) continue;
                feed(phc, s1[v], S1, l1c[v], BX);
                if constexpr (OUT) store_out(oz + coff[v], s1[v][PH].value(), nts);
            }
        }
        // Level 1: level-0 plane z arrives, plane z - R completes.
        const int q = z - R;
        const bool qin = FULL || (q >= 0 && q < n0);
#pragma unroll
        for (int v = 0; v < VX; v++) {
            feed(phc, s0[v], S0, l0c[v], LW);
            if constexpr (L1ON) {
                v1[v] = s0[v][PH].value();
                if (!FULL && !(qin && xyin[v])) v1[v] = h1[(long long)clampi(q, -R, n0 + R - 1) * sz + coff[v]];
            }
        }
    };
    using B0 = std::false_type;
    using B1 = std::true_type;
    // Steps 0 .. 4R: no level-2 output yet (2W - 1 steps).
    static_for<2 * W - 1>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        step(std::integral_constant<int, i % W>{}, std::integral_constant<bool, (i >= 2 * R)>{},
             std::integral_constant<bool, (i >= 2 * R + 1)>{}, B0{}, i);
    });
    // Whole ring turns from step 2W - 1 (phase W - 1), then the rest.
    int s0i = 2 * W - 1;
    for (; s0i + W <= N; s0i += W)
        static_for<W>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            step(std::integral_constant<int, (W - 1 + i) % W>{}, B1{}, B1{}, B1{}, s0i + i);
        });
    static_for<W - 1>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if (s0i + i < N) step(std::integral_constant<int, (W - 1 + i) % W>{}, B1{}, B1{}, B1{}, s0i + i);
    });
}

template <typename T, class S, int VX, bool KAHAN>
__global__ void __launch_bounds__(kMaxThreads) k_box3d_tb2(TB2Params<T, S::P> tp)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const KParams<T, S::P> &p = tp.k;
    int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x;
    bid /= p.tiles_x;
    const int tile_y = bid % p.tiles_y, chunk = bid / p.tiles_y;
    const int x0 = tile_x * p.TX, y0 = tile_y * p.TY, zs = chunk * p.ZC, ze = min(zs + p.ZC, p.n0);
    const T *h0 = (tp.k0 & 1) ? tp.hB : tp.hA;
    const T *h1 = (tp.k0 & 1) ? tp.hA : tp.hB;
    Coefs<T, S::P> cf(p);
    T *lds = reinterpret_cast<T *>(smem_raw);
    if (box3d_tb2_full<S::R>(x0, y0, zs, ze, p.TX, p.TY, p.n0, p.n1, p.n2))
        box3d_tb2_tile<T, S, VX, KAHAN, true>(p.in, p.out, h0, h1, p, x0, y0, zs, ze, lds, cf);
    else
        box3d_tb2_tile<T, S, VX, KAHAN, false>(p.in, p.out, h0, h1, p, x0, y0, zs, ze, lds, cf);
}

template <typename T, class S, int VX, bool KAHAN>
hipError_t launch_box3d_tb2(const SweepArgs &a, hipStream_t st)
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
    const size_t lds = 2 * sizeof(T) * ((size_t)(bx + 2 * R) * (by + 2 * R) + (size_t)bx * by);
    hipLaunchKernelGGL((k_box3d_tb2<T, S, VX, KAHAN>), dim3(tp.k.nblocks), dim3((bx / VX) * by), lds, st, tp);
    return hipGetLastError();
}

template <typename T, bool K, class S> LaunchFn pick_vec(int vec)
{
    switch (vec) {
    case 1: return launch_box3d_tb2<T, S, 1, K>;
    case 2: return launch_box3d_tb2<T, S, 2, K>;
    case 4: // FP64 spills at 4 columns per thread
        if constexpr (sizeof(T) == 4) return launch_box3d_tb2<T, S, 4, K>;
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

LaunchFn uhs_find_box3d_tb2(int prog, int dtype, int vec)
{
    switch (prog) {
    case Prog_3d1r_box::ID: return pick<Prog_3d1r_box>(dtype, vec);
    }
    return nullptr;
}
