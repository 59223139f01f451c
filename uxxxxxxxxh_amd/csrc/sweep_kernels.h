// The sweep kernel family for gfx950.  One template per data-flow shape; the
// stencil program S (term offsets in DSL order) is a compile-time parameter
// so every offset and register index folds to a constant, while the tiling
// (tile_x/tile_y/tile_z) and the coefficients are runtime arguments.
//
// What one kernel launch computes is one sweep of the reference's generated
// kernel stencil_<name>_kernel (ConvertStencilToStandard.cpp:91-149 loop nest
// over the interior, :228-272 loads at the term offsets, :339-381 store):
//     out[p] = c0*in[p+o0] + c1*in[p+o1] + ...      (term order)
// with halo cells of `out` untouched.  How it is computed is MI355X-specific:
//
//   star3d  3D star stencils.  A workgroup owns a tile_x * tile_y column of
//           the grid and streams a slab of tile_z planes along axis 0.  Each
//           thread keeps its own column of 2R+1 values in registers
//           (the axis-0 neighbours), stages the current plane plus an R-wide
//           ring of halo cells in LDS (the axis-1/axis-2 neighbours), and
//           prefetches the next plane while computing this one.
//   box3d   3D stencils whose terms come plane by plane in ascending axis-0
//           offset (every 3D box in the benchmark set).  The whole plane tile
//           with halo is staged in LDS; each thread keeps one running sum per
//           pending output plane and adds the terms of each incoming plane in
//           term order, so the summation order is exactly the term order.
//   win2d   2D stencils of any term order.  A workgroup streams a strip of
//           rows; each thread keeps the (2R+1)x(2R+1) window around its
//           point(s) in registers, fed one row per step through LDS.
//   gather  any stencil (fallback): one thread per point, terms read from
//           global memory through the caches.
//
// All loads of out-of-tile cells are clamped into the allocation
// ([-R, n+R-1] per axis in interior coordinates), so partial edge tiles stay
// in bounds; their results are simply not stored.
#pragma once
#include "uhs_internal.h"

// --------------------------------------------------------------------------
// star3d
// --------------------------------------------------------------------------
// Geometry of one tile's grid: interior extents, strides, tile shape.
struct TileGeom {
    int n0, n1, n2;
    int sy;
    long long sz;
    int TX, TY, BXT;
    bool nts; // non-temporal output stores
};

template <typename T> __device__ __forceinline__ void store_out(T *p, T v, bool nt)
{
    if (nt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// One star3d tile: the tile_x * tile_y column at (x0, y0), output planes
// [zs, ze).  `in`/`out` point at the interior origin.  Shared by the
// per-grid kernel and the batched work-queue kernel.
// FULL: the tile lies inside the grid (every point stored).  Then the
// stores and the halo-ring loads are issued on every path (idle ring slots
// load a harmless in-bounds cell), which lets the compiler count them when it
// waits for older loads instead of waiting for all loads in flight.  NH: halo
// ring cells staged per thread (the launch picks the smallest that covers the
// ring).
template <typename T, class S, int VX, bool KAHAN, int PF, bool FULL = false, int NH = kStarHaloPerThread,
          class CF>
__device__ __forceinline__ void star3d_tile(const T *__restrict__ in, T *__restrict__ out,
                                            const TileGeom &p, int x0, int y0, int zs, int ze,
                                            T *lds, CF &cf)
{
    constexpr int R = S::R, P = S::P;
    const int NT = blockDim.x, tid = threadIdx.x;
    const int BXT = p.BXT, TX = p.TX, TY = p.TY;
    const int LW = TX + 2 * R, LSZ = LW * (TY + 2 * R);
    const int tx = tid % BXT, ty = tid / BXT;
    const long long sz = p.sz;
    const int sy = p.sy;
    lds_poison(lds, 2 * LSZ);

    // This thread's points (x0 + tx + v*BXT, y0 + ty).
    int coff[VX], lc[VX];
    bool act[VX];
    const int yg = y0 + ty, ycl = clampi(yg, -R, p.n1 + R - 1);
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int xg = x0 + tx + v * BXT;
        act[v] = xg < p.n2 && yg < p.n1;
        coff[v] = ycl * sy + clampi(xg, -R, p.n2 + R - 1);
        lc[v] = (ty + R) * LW + tx + v * BXT + R;
    }
    // Halo ring cells this thread stages: R rows above/below the tile, R
    // columns left/right of it (corners are not needed by a star).
    const int H = 2 * R * (TX + TY);
    int hoff[NH], hl[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int c = tid + k * NT;
        hl[k] = -1;
        hoff[k] = 0;
        if (c < H) {
            int ly, lx;
            if (c < 2 * R * TX) {
                const int row = c / TX;
                ly = row < R ? row : TY + row;
                lx = R + c % TX;
            } else {
                const int cc = c - 2 * R * TX, col = cc % (2 * R);
                ly = R + cc / (2 * R);
                lx = col < R ? col : TX + col;
            }
            hl[k] = ly * LW + lx;
            hoff[k] = clampi(y0 - R + ly, -R, p.n1 + R - 1) * sy +
                      clampi(x0 - R + lx, -R, p.n2 + R - 1);
        }
    }

    // Register ring: plane p of this thread's point v lives in slot
    // (p - zs + R) mod M -- 2R+1 planes in use, PF in flight.  The halo ring
    // of plane p lives in slot (p - zs) mod 2.  The z loop is unrolled by hand
    // by U (a multiple of M and of 2; the compiler does not unroll a loop
    // with a barrier), so every slot index is a constant and no value still
    // in flight is ever copied (a copy would make the loop wait for it).
    constexpr int M = 2 * R + 1 + PF, U = (M % 2 == 0) ? M : 2 * M;
    T rq[VX][M], rh[2][NH];
    const bool nts = p.nts;
    // Halo first, then the own columns plane by plane: the order the loop
    // consumes them in, so the waits on entering the loop count the loads
    // still in flight instead of draining them.
#pragma unroll
    for (int k = 0; k < NH; k++)
        if (FULL || hl[k] >= 0) rh[0][k] = in[(long long)zs * sz + hoff[k]];
#pragma unroll
    for (int j = 0; j < 2 * R + PF; j++)
#pragma unroll
        for (int v = 0; v < VX; v++)
            rq[v][j] = in[(long long)min(zs - R + j, p.n0 + R - 1) * sz + coff[v]];

    auto step = [&](auto phc, int z) {
        constexpr int PH = decltype(phc)::value; // (z - zs) mod U
        constexpr int C = (PH + R) % M;           // slot of plane z
        T *L = lds + (PH & 1) * LSZ;
#pragma unroll
        for (int v = 0; v < VX; v++) L[lc[v]] = rq[v][C];
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (hl[k] >= 0) L[hl[k]] = rh[PH & 1][k];
        // Halo of plane z+1 first (exec-masked), then plane z+R+PF of the own
        // points into the slot of plane z-R-1: loads issued on every path
        // come last, so later waits can count them.
        {
            const T *ph = in + (long long)(z + 1) * sz;
#pragma unroll
            for (int k = 0; k < NH; k++)
                if (FULL || hl[k] >= 0) rh[(PH + 1) & 1][k] = ph[hoff[k]];
            const T *pn = in + (long long)min(z + R + PF, p.n0 + R - 1) * sz;
#pragma unroll
            for (int v = 0; v < VX; v++) rq[v][(PH + 2 * R + PF) % M] = pn[coff[v]];
        }
        __syncthreads();
        cf.refresh();
        T *oz = out + (long long)z * sz;
#pragma unroll
        for (int v = 0; v < VX; v++) {
            Acc<T, KAHAN> a;
            static_for<P>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr int dz = S::OFF[t][0], dy = S::OFF[t][1], dx = S::OFF[t][2];
                T val;
                if constexpr (dy == 0 && dx == 0) {
                    val = rq[v][(PH + R + dz) % M];
                } else {
                    static_assert(dz == 0, "star3d needs axis-aligned terms");
                    val = L[lc[v] + dy * LW + dx];
                }
                if constexpr (t == 0) a.first(cf.template get<t>(), val);
                else a.add(cf.template get<t>(), val);
            });
            if (FULL || act[v]) store_out(oz + coff[v], a.value(), nts);
        }
    };
    // Whole turns of the ring in the loop, with no branch inside it (a
    // branch makes the compiler's wait counts conservative), the rest after.
    int z0 = zs;
    for (; z0 + U <= ze; z0 += U)
        static_for<U>([&](auto phc) { step(phc, z0 + decltype(phc)::value); });
    static_for<U - 1>([&](auto phc) {
        const int z = z0 + decltype(phc)::value;
        if (z < ze) step(phc, z);
    });
}

// The same tile with shifting register queues (radius 3, where the ring
// form's larger unrolled body measured slower: 287 against 295 GSt/s at
// 512^3; at radius 2 the ring form is the faster, 322 against 310).
template <typename T, class S, int VX, bool KAHAN, int PF, class CF>
__device__ __forceinline__ void star3d_tile_shift(const T *__restrict__ in, T *__restrict__ out,
                                            const TileGeom &p, int x0, int y0, int zs, int ze,
                                            T *lds, CF &cf)
{
    constexpr int R = S::R, P = S::P, Q = 2 * R + 1, NH = kStarHaloPerThread;
    const int NT = blockDim.x, tid = threadIdx.x;
    const int BXT = p.BXT, TX = p.TX, TY = p.TY;
    const int LW = TX + 2 * R, LSZ = LW * (TY + 2 * R);
    const int tx = tid % BXT, ty = tid / BXT;
    const long long sz = p.sz;
    const int sy = p.sy;
    lds_poison(lds, 2 * LSZ);

    // This thread's points (x0 + tx + v*BXT, y0 + ty).
    int coff[VX], lc[VX];
    bool act[VX];
    const int yg = y0 + ty, ycl = clampi(yg, -R, p.n1 + R - 1);
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int xg = x0 + tx + v * BXT;
        act[v] = xg < p.n2 && yg < p.n1;
        coff[v] = ycl * sy + clampi(xg, -R, p.n2 + R - 1);
        lc[v] = (ty + R) * LW + tx + v * BXT + R;
    }
    // Halo ring cells this thread stages: R rows above/below the tile, R
    // columns left/right of it (corners are not needed by a star).
    const int H = 2 * R * (TX + TY);
    int hoff[NH], hl[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int c = tid + k * NT;
        hl[k] = -1;
        hoff[k] = 0;
        if (c < H) {
            int ly, lx;
            if (c < 2 * R * TX) {
                const int row = c / TX;
                ly = row < R ? row : TY + row;
                lx = R + c % TX;
            } else {
                const int cc = c - 2 * R * TX, col = cc % (2 * R);
                ly = R + cc / (2 * R);
                lx = col < R ? col : TX + col;
            }
            hl[k] = ly * LW + lx;
            hoff[k] = clampi(y0 - R + ly, -R, p.n1 + R - 1) * sy +
                      clampi(x0 - R + lx, -R, p.n2 + R - 1);
        }
    }

    // Register queue: q[v][j] = plane z-R+j at this thread's point v; nq[f]
    // holds plane z+R+f, in flight (PF planes of prefetch).
    T q[VX][Q], nq[PF][VX], nh[NH];
    const bool nts = p.nts;
#pragma unroll
    for (int v = 0; v < VX; v++) {
#pragma unroll
        for (int j = 0; j < 2 * R; j++) q[v][j] = in[(long long)(zs - R + j) * sz + coff[v]];
#pragma unroll
        for (int f = 0; f < PF; f++)
            nq[f][v] = in[(long long)min(zs + R + f, p.n0 + R - 1) * sz + coff[v]];
    }
#pragma unroll
    for (int k = 0; k < NH; k++)
        if (hl[k] >= 0) nh[k] = in[(long long)zs * sz + hoff[k]];

    for (int z = zs; z < ze; ++z) {
        T *L = lds + (z & 1) * LSZ;
#pragma unroll
        for (int v = 0; v < VX; v++) {
            q[v][2 * R] = nq[0][v];
            L[lc[v]] = q[v][R];
        }
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (hl[k] >= 0) L[hl[k]] = nh[k];
        // Prefetch plane z+R+PF (own column) and the halo of plane z+1.
        {
            const T *pn = in + (long long)min(z + R + PF, p.n0 + R - 1) * sz;
            const T *ph = in + (long long)(z + 1) * sz;
#pragma unroll
            for (int f = 0; f + 1 < PF; f++)
#pragma unroll
                for (int v = 0; v < VX; v++) nq[f][v] = nq[f + 1][v];
#pragma unroll
            for (int v = 0; v < VX; v++) nq[PF - 1][v] = pn[coff[v]];
#pragma unroll
            for (int k = 0; k < NH; k++)
                if (hl[k] >= 0) nh[k] = ph[hoff[k]];
        }
        __syncthreads();
        cf.refresh();
        T *oz = out + (long long)z * sz;
#pragma unroll
        for (int v = 0; v < VX; v++) {
            Acc<T, KAHAN> a;
            static_for<P>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr int dz = S::OFF[t][0], dy = S::OFF[t][1], dx = S::OFF[t][2];
                T val;
                if constexpr (dy == 0 && dx == 0) {
                    val = q[v][R + dz];
                } else {
                    static_assert(dz == 0, "star3d needs axis-aligned terms");
                    val = L[lc[v] + dy * LW + dx];
                }
                if constexpr (t == 0) a.first(cf.template get<t>(), val);
                else a.add(cf.template get<t>(), val);
            });
            if (act[v]) store_out(oz + coff[v], a.value(), nts);
        }
#pragma unroll
        for (int v = 0; v < VX; v++)
#pragma unroll
            for (int j = 0; j < 2 * R; j++) q[v][j] = q[v][j + 1];
    }
}

template <typename T, class S, int VX, bool KAHAN, int PF, int NH>
__global__ void __launch_bounds__(kMaxThreads) k_star3d(KParams<T, S::P> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x;
    bid /= p.tiles_x;
    const int tile_y = bid % p.tiles_y, chunk = bid / p.tiles_y;
    const int zs = chunk * p.ZC;
    const TileGeom g{p.n0, p.n1, p.n2, p.sy, p.sz, p.TX, p.TY, p.BXT, (p.flags & UHS_CFG_NT_STORE) != 0};
    Coefs<T, S::P> cf(p);
    if constexpr (S::R <= 2) {
        const int x0 = tile_x * p.TX, y0 = tile_y * p.TY, ze = min(zs + p.ZC, p.n0);
        T *lds = reinterpret_cast<T *>(smem_raw);
        if (x0 + p.TX <= p.n2 && y0 + p.TY <= p.n1)
            star3d_tile<T, S, VX, KAHAN, PF, true, NH>(p.in, p.out, g, x0, y0, zs, ze, lds, cf);
        else
            star3d_tile<T, S, VX, KAHAN, PF, false, NH>(p.in, p.out, g, x0, y0, zs, ze, lds, cf);
    } else
        star3d_tile_shift<T, S, VX, KAHAN, PF>(p.in, p.out, g, tile_x * p.TX, tile_y * p.TY, zs,
                                               min(zs + p.ZC, p.n0), reinterpret_cast<T *>(smem_raw), cf);
}

// --------------------------------------------------------------------------
// star3d_tb2
// --------------------------------------------------------------------------
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
//     in a register ring (plus one plane in flight) and stages plane z1,
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
// FULL: every level-0 cell the tile reads (the extended tile, its ring and
// the slab's 2R planes either side) lies in the interior, so no value comes
// from a halo: the loads take the plane's base address plus a per-lane
// offset, with no clamp or halo select, and no level-1 halo is loaded
// (star3d_tb2_full tells whether a tile qualifies).
// Arguments of a two-step launch (3D stars and 2D stencils).
template <typename T, int P> struct TB2Params {
    KParams<T, P> k;  // first member: the coefficient pack keeps its kernarg offset
    const T *hA, *hB; // interior origins of the grids the ping-pong calls A and B
    int k0;           // global step of the input state
};

template <int R>
__device__ __forceinline__ bool star3d_tb2_full(int x0, int y0, int zs, int ze, int TX, int TY, int n0, int n1,
                                                int n2)
{
    return x0 >= 2 * R && x0 + TX + 2 * R <= n2 && y0 >= 2 * R && y0 + TY + 2 * R <= n1 && zs >= 2 * R &&
           ze + 2 * R <= n0;
}

template <typename T, class S, int VX, bool KAHAN, int PF, int NH = kStarHaloPerThread, bool FULL = false,
          int AHEAD = 2, class CF>
__device__ __forceinline__ void star3d_tb2_tile(const T *__restrict__ in, T *__restrict__ out,
                                                const T *__restrict__ h0, const T *__restrict__ h1,
                                                const TileGeom &g, int x0, int y0, int zs, int ze,
                                                T *lds, CF &cf)
{
    constexpr int R = S::R, P = S::P, Q = 2 * R + 1;
    const int NT = blockDim.x, tid = threadIdx.x;
    const int BXT = g.BXT, BX = g.TX + 2 * R, BY = g.TY + 2 * R;
    const int LW = BX + 2 * R, L0SZ = LW * (BY + 2 * R), L1SZ = BX * BY;
    T *L0 = lds;            // 2 x L0SZ: level-0 plane + ring
    T *L1 = lds + 2 * L0SZ; // 2 x L1SZ: level-1 plane over the extended tile
    lds_poison(lds, 2 * (L0SZ + L1SZ));
    const int ox = x0 - R, oy = y0 - R; // interior coordinates of extended cell (0, 0)
    const int tx = tid % BXT, ty = tid / BXT;
    // Threads beyond the extended tile (a workgroup larger than it, as in
    // the batched kernel) only stage ring cells and join the barriers.
    const bool col = ty < BY;
    const long long sz = g.sz;
    const int sy = g.sy, n0 = g.n0, n1 = g.n1, n2 = g.n2;

    int coff[VX], l0c[VX], l1c[VX];
    bool xyin[VX], core[VX];
    const int yg = oy + ty;
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tx + v * BXT, xg = ox + ex;
        xyin[v] = col && (FULL || (xg >= 0 && xg < n2 && yg >= 0 && yg < n1));
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
            hin[k] = FULL || (gy >= 0 && gy < n1 && gx >= 0 && gx < n2);
            hoff[k] = FULL ? gy * sy + gx : clampi(gy, -R, n1 + R - 1) * sy + clampi(gx, -R, n2 + R - 1);
        }
    }
    // Level-0 value of a cell at plane z: the input inside the interior, the
    // level-0 halo outside it.  Planes beyond the allocation are clamped;
    // their values feed only level-1 cells outside the interior, which take
    // the level-1 halo instead.  In a FULL tile every plane the values are
    // used from is interior; only the prefetch past the slab's last plane is
    // clamped (its value is never used).
    auto lev0 = [&](bool cin, int z, int off) -> T {
        if constexpr (FULL) {
            const T *pz = in + (long long)min(z, n0 + R - 1) * sz;
            return pz[off];
        } else {
            const long long a = (long long)clampi(z, -R, n0 + R - 1) * sz + off;
            return (cin && z >= 0 && z < n0) ? in[a] : h0[a];
        }
    };

    const int z1s = zs - R;             // first level-1 plane
    const int N = (ze - zs) + 2 * R;    // level-1 planes
    // Register rings instead of shifting queues: level-0 plane p of the own
    // columns lives in slot (p - z1s + R) mod M0, the ring cells of plane p
    // in slot (p - z1s) mod 2.  AHEAD = 2: every load is issued two
    // iterations before the iteration that first reads it, at the end of an
    // iteration, into the slot whose value that iteration read last:
    // iteration i reads level-0 planes z1-R .. z1+R, plane z1+R+1 is in
    // flight, and plane z1+R+2 is loaded into the slot of plane z1-R.  So the
    // waits of an iteration count the previous iteration's loads as still in
    // flight (the memory counter is in order) and wait only for older ones.
    // AHEAD = 1: each iteration loads, before its barrier, plane z1+R+1 into
    // the slot of plane z1-R-1 and the ring (and level-1 halo) of plane z1+1
    // -- fewer loads in flight, which the batched kernel, sharing the CUs
    // with the box tiles, runs faster with (measured on the benchmark batch).
    // The loop below is unrolled by M0 by hand (the compiler does not unroll a
    // loop with a barrier), so every slot index is a constant: a loaded
    // value stays in its register until its last use, and no copy makes the
    // loop wait for a load still in flight.  Level-1 plane p of the own
    // columns lives in slot (p - z1s) mod M0 (it may be a halo value loaded
    // from memory, so it must not be copied either).
    static_assert(PF == 1, "one plane in flight between iterations");
    static_assert(AHEAD == 1 || AHEAD == 2, "loads one or two iterations ahead");
    constexpr int M0 = Q + 1;
    // rh1[v][i & 1]: the level-1 halo value of plane z1 for a column that
    // takes it (outside the interior, or z1 outside it), loaded AHEAD
    // iterations ahead into a register of its own.
    T r0[VX][M0], rh[2][NH], r1[VX][M0], rh1[VX][2];
    auto takes_halo1 = [&](int v, int z) { return !FULL && !(xyin[v] && z >= 0 && z < n0); };
    auto halo1 = [&](int v, int z) { return h1[(long long)clampi(z, -R, n0 + R - 1) * sz + coff[v]]; };
    // Planes z1s-R .. z1s+R-1 first, then for g < AHEAD plane z1s+R+g (slot
    // 2R+g) with the level-1 halo and the ring of plane z1s+g, in the order
    // an iteration issues its loads (level-1 halo, own columns, ring), so the
    // waits at the loop's entry, which must hold for the first turn too,
    // still count the last group as in flight.
#pragma unroll
    for (int j = 0; j < 2 * R; j++)
#pragma unroll
        for (int v = 0; v < VX; v++) r0[v][j] = lev0(xyin[v], z1s - R + j, coff[v]);
#pragma unroll
    for (int v = 0; v < VX; v++) {
#pragma unroll
        for (int j = 0; j < M0; j++) r1[v][j] = T(0);
        rh1[v][0] = rh1[v][1] = T(0);
    }
#pragma unroll
    for (int g = 0; g < AHEAD; g++) {
#pragma unroll
        for (int v = 0; v < VX; v++)
            if (col && takes_halo1(v, z1s + g)) rh1[v][g] = halo1(v, z1s + g);
#pragma unroll
        for (int v = 0; v < VX; v++) r0[v][2 * R + g] = lev0(xyin[v], z1s + R + g, coff[v]);
#pragma unroll
        for (int k = 0; k < NH; k++) rh[g][k] = lev0(hin[k], z1s + g, hoff[k]);
    }

    // One iteration, i = z1 - z1s with i % M0 == PH (and i % 2 == PH % 2,
    // M0 being even).
    auto step = [&](auto phc, int i) {
        constexpr int PH = decltype(phc)::value;
        const int z1 = z1s + i;
        T *S0 = L0 + (PH & 1) * L0SZ;
        T *S1 = L1 + (PH & 1) * L1SZ;
        // Staging and prefetch run on every iteration, also the last (i == N,
        // whose staged plane nobody reads): loads issued on every path let
        // the compiler count them when it waits for older loads, instead of
        // waiting for all of them.
#pragma unroll
        for (int v = 0; v < VX; v++)
            if (col) S0[l0c[v]] = r0[v][(PH + R) % M0];
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (hl[k] >= 0) S0[hl[k]] = rh[PH & 1][k];
        if constexpr (AHEAD == 1) {
            // Plane z1+R+1 into the slot of plane z1-R-1 (no longer used),
            // the ring of plane z1+1 into the other ring slot.  Addresses are
            // clamped, so the loads past the slab end are harmless.
#pragma unroll
            for (int v = 0; v < VX; v++) r0[v][(PH + M0 - 1) % M0] = lev0(xyin[v], z1 + R + 1, coff[v]);
#pragma unroll
            for (int k = 0; k < NH; k++) rh[(PH + 1) & 1][k] = lev0(hin[k], z1 + 1, hoff[k]);
        }
        // Level-1 plane z1-1+d (d <= 0) is in slot (PH - 1 + d) mod M0.
        constexpr int L1S = PH - 1 - R + 2 * M0; // slot of plane z2 = z1-1-R, plus 2*M0
        if (col) {
#pragma unroll
            for (int v = 0; v < VX; v++) S1[l1c[v]] = r1[v][L1S % M0];
        }
        __syncthreads();
        cf.refresh();
        // Level 2 at plane z2 = z1 - 1 - R: the level-1 ring holds planes
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
                    if constexpr (dy == 0 && dx == 0) val = r1[v][(L1S + dz) % M0];
                    else val = S1[l1c[v] + dy * BX + dx];
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                store_out(oz + coff[v], a.value(), g.nts);
            }
        }
        if (i < N && col) {
            // Level 1 at plane z1 for every extended column (the columns
            // outside the interior take the level-1 halo); the halo values
            // they take at plane z1+AHEAD are loaded before (AHEAD = 1) or
            // after (AHEAD = 2) it.
            if constexpr (AHEAD == 1) {
#pragma unroll
                for (int v = 0; v < VX; v++)
                    if (takes_halo1(v, z1 + 1)) rh1[v][(PH + 1) & 1] = halo1(v, z1 + 1);
            }
            const bool zin = FULL || (z1 >= 0 && z1 < n0);
#pragma unroll
            for (int v = 0; v < VX; v++) {
                T v1;
                if (FULL || (zin && xyin[v])) {
                    Acc<T, KAHAN> a;
                    static_for<P>([&](auto tc) {
                        constexpr int t = decltype(tc)::value;
                        constexpr int dz = S::OFF[t][0], dy = S::OFF[t][1], dx = S::OFF[t][2];
                        T val;
                        if constexpr (dy == 0 && dx == 0) val = r0[v][(PH + R + dz) % M0];
                        else val = S0[l0c[v] + dy * LW + dx];
                        if constexpr (t == 0) a.first(cf.template get<t>(), val);
                        else a.add(cf.template get<t>(), val);
                    });
                    v1 = a.value();
                } else {
                    v1 = rh1[v][PH & 1];
                }
                r1[v][PH] = v1; // the slot of plane z1 - M0, last used above
            }
            if constexpr (AHEAD == 2) {
#pragma unroll
                for (int v = 0; v < VX; v++)
                    if (takes_halo1(v, z1 + 2)) rh1[v][PH & 1] = halo1(v, z1 + 2);
            }
        }
        if constexpr (AHEAD == 2) {
            // Plane z1+R+2 into the slot of plane z1-R (last read by level 1
            // above), the ring of plane z1+2 into the slot of plane z1 (staged
            // above).  Addresses are clamped, so the loads past the slab end
            // are harmless.
#pragma unroll
            for (int v = 0; v < VX; v++) r0[v][PH] = lev0(xyin[v], z1 + R + 2, coff[v]);
#pragma unroll
            for (int k = 0; k < NH; k++) rh[PH & 1][k] = lev0(hin[k], z1 + 2, hoff[k]);
        }
    };
    // Whole ring turns without a branch around any step, so the waits at the
    // loop's entry count the loads still in flight instead of draining them;
    // then the remaining iterations i0 .. N (fewer than M0).
    int i0 = 0;
    for (; i0 + M0 - 1 <= N; i0 += M0)
        static_for<M0>([&](auto phc) { step(phc, i0 + decltype(phc)::value); });
    static_for<M0>([&](auto phc) {
        const int i = i0 + decltype(phc)::value;
        if (i <= N) step(phc, i);
    });
}


// --------------------------------------------------------------------------
// box3d (terms grouped by ascending axis-0 offset)
// --------------------------------------------------------------------------
template <class S> constexpr bool plane_major()
{
    for (int t = 1; t < S::P; t++)
        if (S::OFF[t][0] < S::OFF[t - 1][0]) return false;
    return true;
}

// True when the terms form NJ equal plane groups (one per axis-0 offset,
// ascending) that list the same in-plane cells in the same order -- every
// 3D box of the benchmark set.  Then each staged LDS value can be applied to
// all NJ pending sums at once (one LDS read per cell instead of one per
// term) while every sum still receives its terms in term order.
template <class S> constexpr bool uniform_planes()
{
    if (!plane_major<S>()) return false;
    constexpr int NJ = S::OFF[S::P - 1][0] - S::OFF[0][0] + 1;
    if (S::P % NJ) return false;
    constexpr int G = S::P / NJ;
    for (int g = 0; g < NJ; g++)
        for (int i = 0; i < G; i++) {
            if (S::OFF[g * G + i][0] != S::OFF[0][0] + g) return false;
            if (S::OFF[g * G + i][1] != S::OFF[i][1]) return false;
            if (S::OFF[g * G + i][2] != S::OFF[i][2]) return false;
        }
    return true;
}

// Position of term t in the coefficient pack the box3d kernel reads: cell
// major (all NJ coefficients of one in-plane cell adjacent, so one scalar
// load fetches them) when uniform_planes<S>(), term order otherwise.
template <class S> constexpr int box_pack_index(int t)
{
    if constexpr (uniform_planes<S>()) {
        constexpr int NJ = S::OFF[S::P - 1][0] - S::OFF[0][0] + 1;
        constexpr int G = S::P / NJ;
        return (t % G) * NJ + t / G;
    } else {
        return t;
    }
}

// One box3d tile: the tile_x * tile_y column at (x0, y0), output planes
// [zs, ze).  `in`/`out` point at the interior origin; CF supplies the
// coefficients in box_pack_index order.  Shared by the per-grid kernel and
// the batched work-queue kernel.

template <typename T, class S, int VX, bool KAHAN, int PDEEP, int GPF = 1, int NL = box_cells_per_thread(S::R),
          class CF>
__device__ __forceinline__ void box3d_tile(const T *__restrict__ in, T *__restrict__ out,
                                           const TileGeom &p, int x0, int y0, int zs, int ze,
                                           T *lds, CF &cf)
{
    static_assert(plane_major<S>(), "box3d needs terms in ascending axis-0 offset");
    constexpr int R = S::R, P = S::P;
    constexpr int DZMIN = S::OFF[0][0], DZMAX = S::OFF[P - 1][0];
    constexpr int NJ = DZMAX - DZMIN + 1; // pending output planes per point
    const int NT = blockDim.x, tid = threadIdx.x;
    const int BXT = p.BXT, TX = p.TX, TY = p.TY;
    const int LW = TX + 2 * R, LSZ = LW * (TY + 2 * R);
    const int tx = tid % BXT, ty = tid / BXT;
    const long long sz = p.sz;
    const int sy = p.sy;
    lds_poison(lds, 2 * LSZ);

    int coff[VX], lc[VX];
    bool act[VX];
    const int yg = y0 + ty;
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int xg = x0 + tx + v * BXT;
        act[v] = xg < p.n2 && yg < p.n1;
        coff[v] = act[v] ? yg * sy + xg : 0;
        lc[v] = (ty + R) * LW + tx + v * BXT + R;
    }
    // Plane-tile cells (with R-wide halo, corners included) staged by this thread.
    int goff[NL];
    bool gv[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int c = tid + k * NT;
        gv[k] = c < LSZ;
        const int ly = gv[k] ? c / LW : 0, lx = gv[k] ? c % LW : 0;
        goff[k] = clampi(y0 - R + ly, -R, p.n1 + R - 1) * sy +
                  clampi(x0 - R + lx, -R, p.n2 + R - 1);
    }

    // pl[f] holds plane zp + f of the staged tile (GPF planes in flight).
    T pl[GPF][NL];
#pragma unroll
    for (int f = 0; f < GPF; f++) {
        const T *pp = in + (long long)min(zs + DZMIN + f, p.n0 + R - 1) * sz;
#pragma unroll
        for (int k = 0; k < NL; k++)
            if (gv[k]) pl[f][k] = pp[goff[k]];
    }
    // acc[v][j] is the running sum of output plane zp - DZMAX + j.
    Acc<T, KAHAN> acc[VX][NJ];
#pragma unroll
    for (int v = 0; v < VX; v++)
#pragma unroll
        for (int j = 0; j < NJ; j++) acc[v][j].zero();
    const int steps = (ze - zs) + NJ - 1;
    for (int s = 0; s < steps; ++s) {
        const int zp = zs + DZMIN + s;
        T *L = lds + (s & 1) * LSZ;
#pragma unroll
        for (int k = 0; k < NL; k++)
            if (gv[k]) L[tid + k * NT] = pl[0][k];
        {
            const T *pn = in + (long long)min(zp + GPF, p.n0 + R - 1) * sz;
#pragma unroll
            for (int f = 0; f + 1 < GPF; f++)
#pragma unroll
                for (int k = 0; k < NL; k++) pl[f][k] = pl[f + 1][k];
#pragma unroll
            for (int k = 0; k < NL; k++)
                if (gv[k]) pl[GPF - 1][k] = pn[goff[k]];
        }
        __syncthreads();
        cf.refresh();
        const int zo = zp - DZMAX;
        T *oz = out + (long long)zo * sz;
#pragma unroll
        for (int v = 0; v < VX; v++) {
            if constexpr (uniform_planes<S>()) {
                // Cell by cell: the value and the NJ coefficients of cell
                // i+PD are fetched (LDS read, scalar loads) while cell i is
                // applied to the NJ pending sums, so their latency is covered
                // by PD cells of FMAs; pin() keeps the compiler from hoisting
                // every LDS read to the top of the step (which would hold all
                // G cells live).
                constexpr int G = P / NJ;
                constexpr int PD = G < PDEEP ? G : PDEEP;
                T vb[PD], cb[PD][NJ];
                static_for<PD>([&](auto dc) {
                    constexpr int d = decltype(dc)::value;
                    vb[d] = L[lc[v] + S::OFF[d][1] * LW + S::OFF[d][2]];
                    static_for<NJ>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        cb[d][g] = cf.template get<box_pack_index<S>(g * G + d)>();
                    });
                });
                static_for<G>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    constexpr int slot = i % PD;
                    const T cur = vb[slot];
                    T cc[NJ];
#pragma unroll
                    for (int g = 0; g < NJ; g++) cc[g] = cb[slot][g];
                    if constexpr (i + PD < G) {
                        vb[slot] = L[lc[v] + S::OFF[i + PD][1] * LW + S::OFF[i + PD][2]];
                        static_for<NJ>([&](auto gc) {
                            constexpr int g = decltype(gc)::value;
                            cb[slot][g] = cf.template get<box_pack_index<S>(g * G + i + PD)>();
                        });
                    }
                    static_for<NJ>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        constexpr int t = g * G + i, j = NJ - 1 - g;
                        if constexpr (t == 0) acc[v][j].first(cc[g], cur);
                        else acc[v][j].add(cc[g], cur);
                    });
                    if constexpr (i + 1 < G) {
#pragma unroll
                        for (int j = 0; j < NJ; j++) acc[v][j].pin();
                    }
                });
            } else {
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int dz = S::OFF[t][0], dy = S::OFF[t][1], dx = S::OFF[t][2];
                    constexpr int j = DZMAX - dz;
                    const T val = L[lc[v] + dy * LW + dx];
                    if constexpr (t == 0) acc[v][j].first(cf.template get<t>(), val);
                    else acc[v][j].add(cf.template get<t>(), val);
                });
            }
            if (zo >= zs && act[v]) store_out(oz + coff[v], acc[v][0].value(), p.nts);
#pragma unroll
            for (int j = 0; j + 1 < NJ; j++) acc[v][j] = acc[v][j + 1];
        }
    }
}

// True when every plane group lists its cells row by row (ascending
// axis-1 offset), every row with the same sequence of axis-2 offsets -- the
// benchmark boxes.  Then a value staged at row y+dy is cell (dy, dx) of
// output row y and cell (dy-1, dx) of output row y+1, and visiting staged
// rows in ascending order feeds both outputs their cells in term order.
template <class S> constexpr bool row_grouped()
{
    if (!uniform_planes<S>()) return false;
    constexpr int NJ = S::OFF[S::P - 1][0] - S::OFF[0][0] + 1;
    constexpr int G = S::P / NJ, W = 2 * S::R + 1;
    if (G != W * W) return false;
    for (int i = 0; i < G; i++) {
        if (S::OFF[i][1] != -S::R + i / W) return false;
        if (S::OFF[i][2] != S::OFF[i % W][2]) return false;
    }
    return true;
}

// box3d with two output rows per thread (vec 1): every staged value is read
// from LDS once and applied to the pending sums of both rows, which halves
// the LDS reads per output -- the boxes' other bottleneck next to FP64 FMA.
// Whether box3d_tile_rows2 keeps its plane loads in a register ring (see
// there).  Measured on MI355X at 512^3: the radius-2 box gains 10% and the
// radius-3 box 4%; the radius-1 box loses 20% (its 8 cells per thread become
// 8 unconditional loads where a typical tile needs 3).  The compensated
// radius-3 form is left out: its unrolled loop moves the running sums to
// scratch memory and runs 60x slower.
template <class S, bool KAHAN> constexpr bool box_plane_ring() { return S::R == 2 || (S::R == 3 && !KAHAN); }

template <typename T, class S, bool KAHAN, int GPF, int NL = box_cells_per_thread(S::R), class CF>
__device__ __forceinline__ void box3d_tile_rows2(const T *__restrict__ in, T *__restrict__ out,
                                                 const TileGeom &p, int x0, int y0, int zs, int ze,
                                                 T *lds, CF &cf)
{
    static_assert(row_grouped<S>(), "two-row box3d needs row-grouped plane cells");
    constexpr int R = S::R, P = S::P, W = 2 * R + 1;
    constexpr int DZMIN = S::OFF[0][0], DZMAX = S::OFF[P - 1][0];
    constexpr int NJ = DZMAX - DZMIN + 1, G = P / NJ;
    const int NT = blockDim.x, tid = threadIdx.x;
    const int TX = p.TX, TY = p.TY;
    const int LW = TX + 2 * R, LSZ = LW * (TY + 2 * R);
    const int tx = tid % TX, ty = 2 * (tid / TX);
    const long long sz = p.sz;
    const int sy = p.sy;
    lds_poison(lds, 2 * LSZ);

    const int xg = x0 + tx, yg = y0 + ty;
    bool act[2];
    int coff[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        act[r] = xg < p.n2 && yg + r < p.n1;
        coff[r] = act[r] ? (yg + r) * sy + xg : 0;
    }
    const int lc = (ty + R) * LW + tx + R;
    int goff[NL];
    bool gv[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int c = tid + k * NT;
        gv[k] = c < LSZ;
        const int ly = gv[k] ? c / LW : 0, lx = gv[k] ? c % LW : 0;
        goff[k] = clampi(y0 - R + ly, -R, p.n1 + R - 1) * sy + clampi(x0 - R + lx, -R, p.n2 + R - 1);
    }
    constexpr bool RING = GPF > 1 && box_plane_ring<S, KAHAN>();
    T pl[GPF][NL];
#pragma unroll
    for (int f = 0; f < GPF; f++) {
        const T *pp = in + (long long)min(zs + DZMIN + f, p.n0 + R - 1) * sz;
#pragma unroll
        for (int k = 0; k < NL; k++)
            if (RING || gv[k]) pl[f][k] = pp[goff[k]];
    }
    Acc<T, KAHAN> acc[2][NJ];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int j = 0; j < NJ; j++) acc[r][j].zero();
    const int steps = (ze - zs) + NJ - 1;
    // With RING the plane loads in flight form a register ring: step s
    // stages slot s mod GPF and refills that same slot with the plane GPF
    // steps ahead.  The step loop is then unrolled by hand by GPF so slots
    // are constants and no load in flight is ever copied (a copy waits for
    // it at once), and the loads are unconditional (goff of an unused cell
    // is a valid cell; a load under a branch makes the compiler drain every
    // load before the next wait).  Without RING slot 0 is staged and the
    // slots shift down by one per step.
    auto step = [&](auto phc, int s) {
        constexpr int PH = decltype(phc)::value;
        const int zp = zs + DZMIN + s;
        T *L = lds + (s & 1) * LSZ;
#pragma unroll
        for (int k = 0; k < NL; k++)
            if (gv[k]) L[tid + k * NT] = pl[PH][k];
        {
            const T *pn = in + (long long)min(zp + GPF, p.n0 + R - 1) * sz;
            if constexpr (RING) {
#pragma unroll
                for (int k = 0; k < NL; k++) pl[PH][k] = pn[goff[k]];
            } else {
#pragma unroll
                for (int f = 0; f + 1 < GPF; f++)
#pragma unroll
                    for (int k = 0; k < NL; k++) pl[f][k] = pl[f + 1][k];
#pragma unroll
                for (int k = 0; k < NL; k++)
                    if (gv[k]) pl[GPF - 1][k] = pn[goff[k]];
            }
        }
        __syncthreads();
        cf.refresh();
        const int zo = zp - DZMAX;
        // Staged rows j = 0..W (offset j - R from row 0), columns k in the
        // rows' axis-2 order; item u = j * W + k.  The value and coefficients
        // of item u+1 are fetched while item u is applied.
        constexpr int NU = (W + 1) * W;
        auto value = [&](auto uc) {
            constexpr int u = decltype(uc)::value;
            return L[lc + (u / W - R) * LW + S::OFF[u % W][2]];
        };
        auto coefs = [&](auto uc, T (&c)[2][NJ]) {
            constexpr int u = decltype(uc)::value, j = u / W, k = u % W;
            static_for<NJ>([&](auto gc) {
                constexpr int g = decltype(gc)::value;
                if constexpr (j < W) c[0][g] = cf.template get<box_pack_index<S>(g * G + j * W + k)>();
                if constexpr (j >= 1) c[1][g] = cf.template get<box_pack_index<S>(g * G + (j - 1) * W + k)>();
            });
        };
        T cur = value(std::integral_constant<int, 0>{});
        T cc[2][NJ];
        coefs(std::integral_constant<int, 0>{}, cc);
        static_for<NU>([&](auto uc) {
            constexpr int u = decltype(uc)::value, j = u / W, k = u % W;
            T nxt = cur, nc[2][NJ];
            if constexpr (u + 1 < NU) {
                nxt = value(std::integral_constant<int, u + 1>{});
                coefs(std::integral_constant<int, u + 1>{}, nc);
            }
            static_for<NJ>([&](auto gc) {
                constexpr int g = decltype(gc)::value;
                if constexpr (j < W) { // cell (j - R, k) of row 0
                    constexpr int t = g * G + j * W + k;
                    if constexpr (t == 0) acc[0][NJ - 1 - g].first(cc[0][g], cur);
                    else acc[0][NJ - 1 - g].add(cc[0][g], cur);
                }
                if constexpr (j >= 1) { // cell (j - 1 - R, k) of row 1
                    constexpr int t = g * G + (j - 1) * W + k;
                    if constexpr (t == 0) acc[1][NJ - 1 - g].first(cc[1][g], cur);
                    else acc[1][NJ - 1 - g].add(cc[1][g], cur);
                }
            });
            if constexpr (u + 1 < NU) {
#pragma unroll
                for (int r = 0; r < 2; r++)
#pragma unroll
                    for (int q = 0; q < NJ; q++) acc[r][q].pin();
                cur = nxt;
#pragma unroll
                for (int r = 0; r < 2; r++)
#pragma unroll
                    for (int g = 0; g < NJ; g++) cc[r][g] = nc[r][g];
            }
        });
        T *oz = out + (long long)zo * sz;
#pragma unroll
        for (int r = 0; r < 2; r++) {
            if (zo >= zs && act[r]) store_out(oz + coff[r], acc[r][0].value(), p.nts);
#pragma unroll
            for (int j = 0; j + 1 < NJ; j++) acc[r][j] = acc[r][j + 1];
        }
    };
    if constexpr (RING) {
        int s0 = 0;
        for (; s0 + GPF <= steps; s0 += GPF)
            static_for<GPF>([&](auto phc) { step(phc, s0 + decltype(phc)::value); });
        static_for<GPF - 1>([&](auto phc) {
            if (s0 + decltype(phc)::value < steps) step(phc, s0 + decltype(phc)::value);
        });
    } else {
        for (int s = 0; s < steps; ++s) step(std::integral_constant<int, 0>{}, s);
    }
}

template <typename T, class S, bool KAHAN, int GPF>
__global__ void __launch_bounds__(kMaxThreads) k_box3d_rows2(KParams<T, S::P> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x;
    bid /= p.tiles_x;
    const int tile_y = bid % p.tiles_y, chunk = bid / p.tiles_y;
    const int zs = chunk * p.ZC;
    const TileGeom g{p.n0, p.n1, p.n2, p.sy, p.sz, p.TX, p.TY, p.TX, (p.flags & UHS_CFG_NT_STORE) != 0};
    Coefs<T, S::P> cf(p);
    box3d_tile_rows2<T, S, KAHAN, GPF>(p.in, p.out, g, tile_x * p.TX, tile_y * p.TY, zs, min(zs + p.ZC, p.n0),
                                  reinterpret_cast<T *>(smem_raw), cf);
}

template <typename T, class S, bool KAHAN, int GPF>
hipError_t launch_box3d_rows2(const SweepArgs &a, hipStream_t st)
{
    auto p = make_params<T, S::P>(a);
    for (int t = 0; t < S::P; t++) p.c[box_pack_index<S>(t)] = (T)a.coef[t];
    const int nt = a.tx * (a.ty / 2);
    const size_t lds = 2 * sizeof(T) * (a.tx + 2 * S::R) * (a.ty + 2 * S::R);
    hipLaunchKernelGGL((k_box3d_rows2<T, S, KAHAN, GPF>), dim3(p.nblocks), dim3(nt), lds, st, p);
    return hipGetLastError();
}

template <typename T, class S, int VX, bool KAHAN, int PDEEP, int GPF>
__global__ void __launch_bounds__(kMaxThreads) k_box3d(KParams<T, S::P> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x;
    bid /= p.tiles_x;
    const int tile_y = bid % p.tiles_y, chunk = bid / p.tiles_y;
    const int zs = chunk * p.ZC;
    const TileGeom g{p.n0, p.n1, p.n2, p.sy, p.sz, p.TX, p.TY, p.BXT, (p.flags & UHS_CFG_NT_STORE) != 0};
    Coefs<T, S::P> cf(p);
    box3d_tile<T, S, VX, KAHAN, PDEEP, GPF>(p.in, p.out, g, tile_x * p.TX, tile_y * p.TY, zs,
                                min(zs + p.ZC, p.n0), reinterpret_cast<T *>(smem_raw), cf);
}

// --------------------------------------------------------------------------
// win2d
// --------------------------------------------------------------------------
template <typename T, class S, int VX, bool KAHAN, int PF>
__global__ void __launch_bounds__(kMaxThreads) k_win2d(KParams<T, S::P> p)
{
    constexpr int R = S::R, P = S::P, W = 2 * R + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);

    const int tid = threadIdx.x;
    const int BXT = p.BXT, TX = p.TX, LW = TX + 2 * R;
    lds_poison(lds, 2 * LW);
    const int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x, chunk = bid / p.tiles_x;
    const int x0 = tile_x * TX;
    const int ys = chunk * p.ZC, ye = min(ys + p.ZC, p.n0);
    const T *__restrict__ in = p.in;
    T *__restrict__ out = p.out;
    const long long sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;

    int xcl[VX], xg[VX];
    bool act[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        xg[v] = x0 + tid + v * BXT;
        act[v] = xg[v] < p.n2;
        xcl[v] = clampi(xg[v], -R, p.n2 + R - 1);
    }
    // Threads 0..2R-1 also stage the R halo columns on each side.
    const bool hthr = tid < 2 * R;
    const int hlx = tid < R ? tid : TX + tid;
    const int hoff = clampi(x0 - R + hlx, -R, p.n2 + R - 1);

    // PF rows of loads in flight, shifted down one slot per row; or, with
    // PF == 0, a ring of W = 2R+1 rows: step s stages slot s mod W (the
    // step loop is already unrolled by W) and refills it with the row W
    // ahead, so no load in flight is ever copied (a copy waits for it at
    // once) and W rows stay in flight.  The ring's loads are unconditional
    // (hoff is a valid cell for every thread): a load under a branch makes
    // the compiler drain every load before the next wait.
    constexpr bool RING = PF == 0;
    constexpr int NS = RING ? W : PF;
    // nrow[f] / nh[f]: row ys - R + s + f.
    T nrow[NS][VX], nh[NS];
#pragma unroll
    for (int f = 0; f < NS; f++) {
        const T *pr = in + (long long)min(ys - R + f, p.n0 + R - 1) * sy;
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[f][v] = pr[xcl[v]];
        if constexpr (RING) nh[f] = pr[hoff];
        else nh[f] = hthr ? pr[hoff] : T(0);
    }
    Coefs<T, P> cf(p);
    // Window ring: w[v][r][b] holds row y of the window around point v in
    // slot r = (y - ys + R) mod W, column x_v - R + b.  The row loop is
    // unrolled by hand by W (the compiler does not unroll a loop with a
    // barrier), so slots are constants and the window never shifts through
    // register copies (2R*W moves per row otherwise, next to (2R+1)^2 FMAs
    // for a box).
    T w[VX][W][W];
#pragma unroll
    for (int v = 0; v < VX; v++)
#pragma unroll
        for (int a = 0; a < W; a++)
#pragma unroll
            for (int b = 0; b < W; b++) w[v][a][b] = T(0);
    const int steps = (ye - ys) + 2 * R;
    auto step = [&](auto phc, int s) {
        constexpr int PH = decltype(phc)::value; // s mod W
        const int yr = ys - R + s;               // row staged at this step (slot PH)
        T *L = lds + (s & 1) * LW;
        constexpr int SL = RING ? PH : 0;
#pragma unroll
        for (int v = 0; v < VX; v++) L[R + tid + v * BXT] = nrow[SL][v];
        if (hthr) L[hlx] = nh[SL];
        if constexpr (RING) {
            const T *pr = in + (long long)min(yr + W, p.n0 + R - 1) * sy;
#pragma unroll
            for (int v = 0; v < VX; v++) nrow[PH][v] = pr[xcl[v]];
            nh[PH] = pr[hoff];
        } else {
#pragma unroll
            for (int f = 0; f + 1 < PF; f++) {
#pragma unroll
                for (int v = 0; v < VX; v++) nrow[f][v] = nrow[f + 1][v];
                nh[f] = nh[f + 1];
            }
            const T *pr = in + (long long)min(yr + PF, p.n0 + R - 1) * sy;
#pragma unroll
            for (int v = 0; v < VX; v++) nrow[PF - 1][v] = pr[xcl[v]];
            if (hthr) nh[PF - 1] = pr[hoff];
        }
        __syncthreads();
        cf.refresh();
        const int yo = yr - R;
        T *orow = out + (long long)yo * sy;
#pragma unroll
        for (int v = 0; v < VX; v++) {
#pragma unroll
            for (int b = 0; b < W; b++) w[v][PH][b] = L[tid + v * BXT + b];
            if (s >= 2 * R) {
                // Output row yo: its row yo + dy is in slot (PH - R + dy) mod W.
                Acc<T, KAHAN> acc;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int dy = S::OFF[t][0], dx = S::OFF[t][1];
                    const T val = w[v][(PH + W - R + dy) % W][R + dx];
                    if constexpr (t == 0) acc.first(cf.template get<t>(), val);
                    else acc.add(cf.template get<t>(), val);
                });
                if (act[v]) store_out(orow + xg[v], acc.value(), nts);
            }
        }
    };
    if constexpr (RING) {
        // Whole ring turns in the loop (no branch inside it), the rest after.
        int s0 = 0;
        for (; s0 + W <= steps; s0 += W)
            static_for<W>([&](auto phc) { step(phc, s0 + decltype(phc)::value); });
        static_for<W - 1>([&](auto phc) {
            const int s = s0 + decltype(phc)::value;
            if (s < steps) step(phc, s);
        });
    } else {
        for (int s0 = 0; s0 < steps; s0 += W)
            static_for<W>([&](auto phc) {
                const int s = s0 + decltype(phc)::value;
                if (s < steps) step(phc, s);
            });
    }
}

// --------------------------------------------------------------------------
// gather (any stencil)
// --------------------------------------------------------------------------
template <typename T> struct GParams {
    const T *in;
    T *out;
    int n0, n1, n2, P;
    int sy;
    long long sz;
    int loff[kGatherMaxPoints];
    T c[kGatherMaxPoints];
};

template <typename T, bool KAHAN>
__global__ void __launch_bounds__(256) k_gather(GParams<T> p)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long plane = (long long)p.n1 * p.n2;
    if (i >= plane * p.n0) return;
    const int z = (int)(i / plane);
    const int rem = (int)(i - (long long)z * plane);
    const int y = rem / p.n2, x = rem - y * p.n2;
    const long long pt = (long long)z * p.sz + (long long)y * p.sy + x;
    const T *__restrict__ in = p.in + pt;
    Acc<T, KAHAN> a;
    a.first(p.c[0], in[p.loff[0]]);
    for (int t = 1; t < p.P; t++) a.add(p.c[t], in[p.loff[t]]);
    p.out[pt] = a.value();
}

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
template <typename T> constexpr int dtype_of(bool kahan)
{
    return sizeof(T) == 8 ? DT_F64 : (kahan ? DT_F32K : DT_F32);
}

template <typename T, class S, int VX, bool KAHAN, int PF, int NH = kStarHaloPerThread>
hipError_t launch_star3d(const SweepArgs &a, hipStream_t st)
{
    auto p = make_params<T, S::P>(a);
    const int nt = (a.tx / VX) * a.ty;
    const size_t lds = 2 * sizeof(T) * (a.tx + 2 * S::R) * (a.ty + 2 * S::R);
    hipLaunchKernelGGL((k_star3d<T, S, VX, KAHAN, PF, NH>), dim3(p.nblocks), dim3(nt), lds, st, p);
    return hipGetLastError();
}

template <typename T, class S, int VX, bool KAHAN, int PDEEP, int GPF>
hipError_t launch_box3d(const SweepArgs &a, hipStream_t st)
{
    auto p = make_params<T, S::P>(a);
    for (int t = 0; t < S::P; t++) p.c[box_pack_index<S>(t)] = (T)a.coef[t];
    const int nt = (a.tx / VX) * a.ty;
    const size_t lds = 2 * sizeof(T) * (a.tx + 2 * S::R) * (a.ty + 2 * S::R);
    hipLaunchKernelGGL((k_box3d<T, S, VX, KAHAN, PDEEP, GPF>), dim3(p.nblocks), dim3(nt), lds, st, p);
    return hipGetLastError();
}

template <typename T, class S, int VX, bool KAHAN, int PF>
hipError_t launch_win2d(const SweepArgs &a, hipStream_t st)
{
    auto p = make_params<T, S::P>(a);
    const int nt = a.tx / VX;
    const size_t lds = 2 * sizeof(T) * (a.tx + 2 * S::R);
    hipLaunchKernelGGL((k_win2d<T, S, VX, KAHAN, PF>), dim3(p.nblocks), dim3(nt), lds, st, p);
    return hipGetLastError();
}
