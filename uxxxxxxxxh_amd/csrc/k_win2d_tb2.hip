// win2d_tb2 kernel: two timesteps per launch for the 2D stencils (temporal
// blocking with an R-column ghost zone), so each cell crosses HBM once per
// two updates instead of once per update.  FP64, FP32 and FP32 compensated.
// Three forms: the window form below (any 2D stencil; UHS_CFG_ROWRING), the
// ring form of win2d_tb2_star_tile (stars) and the sum form of
// win2d_tb2_sum_tile (boxes).
//
// Window form.  A workgroup owns output columns [x0, x0 + tile_x) of rows
// [ys, ye) and covers the columns extended by R on each side (BX = tile_x +
// 2R, VX columns per thread along the unit-stride axis).  It streams rows
// along axis 0.  At step s it stages level-0 row r0 = ys - 2R + s (the
// input) in LDS, and level-1 row r0 - R - 1 (one step, computed at step
// s - 1):
//   -- one barrier --
//   each thread moves the 2R+1 cells of row r0 around each of its columns
//     into its level-0 window (a register ring of 2R+1 rows), and the 2R+1
//     level-1 cells of row r0 - R - 1 into its level-1 window;
//   level 2 at row r0 - 2R - 1 for the tile's own columns, from the level-1
//     window, is stored;
//   level 1 at row r0 - R for every extended column, from the level-0
//     window, is kept in a register until the next step stages it.
// Buffer s&1 is rewritten two steps later, after the next barrier, so one
// barrier per row suffices.  The windows are rings indexed by the step's
// phase, with the row loop unrolled by hand by 2R+1, so no register copies
// shift them; a star's window cells off its centre row and column are dead
// and take no registers.  2R+1 rows of loads are in flight in a register
// ring.
//
// Cells of a level outside the interior hold that level's halo: the halo of
// the grid the one-step ping-pong holds that state in (h0 for the input, h1
// for level 1).  Every update is the product-then-fma chain in term order,
// so the result is bitwise that of two single sweeps.
#include "sweep_kernels.h"

namespace {

template <int R> __device__ __forceinline__ bool win2d_tb2_full(int x0, int TX, int ys, int ye, int n0, int n2)
{
    return x0 >= 2 * R && x0 + TX + 2 * R <= n2 && ys >= 2 * R && ye + 2 * R <= n0;
}

// FULL: every cell whose value is used (level-0 columns [x0-2R, x0+TX+2R)
// of rows [ys-2R, ye+2R)) lies in the interior: no halo selects, no clamps
// but the prefetch's past the last row.
template <typename T, class S, int VX, bool KAHAN, bool FULL, class CF>
__device__ __forceinline__ void win2d_tb2_tile(const T *__restrict__ in, T *__restrict__ out,
                                               const T *__restrict__ h0, const T *__restrict__ h1,
                                               const KParams<T, S::P> &p, int x0, int ys, int ye, T *lds, CF &cf)
{
    constexpr int R = S::R, P = S::P, W = 2 * R + 1;
    const int tid = threadIdx.x, BXT = p.BXT;
    const int BX = p.TX + 2 * R, LW = BX + 2 * R;
    T *L0 = lds;              // 2 x LW: level-0 row, columns x0-2R ..
    T *L1 = lds + 2 * LW + R; // 2 x BX: level-1 row, columns x0-R .. (R spare cells either side)
    lds_poison(lds, 2 * LW + 2 * BX + 2 * R);
    const int n0 = p.n0, n2 = p.n2;
    const long long sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;

    // Extended column ex = tid + v*BXT is interior column x0 - R + ex.
    int xcl[VX];
    bool xin[VX], core[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tid + v * BXT, xg = x0 - R + ex;
        xin[v] = FULL || (xg >= 0 && xg < n2);
        core[v] = ex >= R && ex < BX - R && (FULL || xg < n2);
        xcl[v] = FULL ? xg : clampi(xg, -R, n2 + R - 1);
    }
    // Threads 0 .. 2R-1 also stage the R level-0 cells beyond the extended
    // columns on each side (LDS index hlx, interior column x0 - 2R + hlx);
    // the others load a valid cell they never stage, so the loads stay
    // unconditional.
    const bool hthr = tid < 2 * R;
    const int hlx = tid < R ? tid : BX + tid;
    const int hxg = x0 - 2 * R + hlx;
    const bool hin = FULL || (hxg >= 0 && hxg < n2);
    const int hcl = FULL && hthr ? hxg : clampi(hxg, -R, n2 + R - 1);

    // Level-0 value at row r: the input inside the interior, h0 outside it.
    // Rows beyond the allocation are clamped; their values are never used.
    auto lev0 = [&](bool cin, int r, int c) -> T {
        if constexpr (FULL) {
            return in[(long long)min(r, n0 + R - 1) * sy + c];
        } else {
            const T *base = (cin && r >= 0 && r < n0) ? in : h0;
            return base[(long long)clampi(r, -R, n0 + R - 1) * sy + c];
        }
    };

    // Prefetch ring: slot f holds the level-0 row staged at the next step
    // with phase f, loaded 2R+1 rows ahead.
    T nrow[W][VX], nh[W];
#pragma unroll
    for (int f = 0; f < W; f++) {
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[f][v] = lev0(xin[v], ys - 2 * R + f, xcl[v]);
        nh[f] = lev0(hin, ys - 2 * R + f, hcl);
    }
    // Windows: w0[v][k][b] is level-0 row r (staged at a step of phase k)
    // at column x - R + b; w1 the same for level 1.
    T w0[VX][W][W], w1[VX][W][W], v1[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        v1[v] = T(0);
#pragma unroll
        for (int a = 0; a < W; a++)
#pragma unroll
            for (int b = 0; b < W; b++) w0[v][a][b] = w1[v][a][b] = T(0);
    }
    const int N = (ye - ys) + 4 * R + 1; // steps

    // One step s with phase PH = s mod W.  L1ON: level 1 is needed (s >= 2R);
    // W1ON: the level-1 row read is one of them (s >= 2R+1); OUT: level 2 is
    // stored (s >= 4R+1).
    auto step = [&](auto phc, auto l1c, auto w1c, auto outc, int s) {
        constexpr int PH = decltype(phc)::value;
        constexpr bool L1ON = decltype(l1c)::value, W1ON = decltype(w1c)::value, OUT = decltype(outc)::value;
        const int r0 = ys - 2 * R + s;
        T *S0 = L0 + (s & 1) * LW;
        T *S1 = L1 + (s & 1) * BX;
#pragma unroll
        for (int v = 0; v < VX; v++) S0[R + tid + v * BXT] = nrow[PH][v];
        if (hthr) S0[hlx] = nh[PH];
        if constexpr (W1ON) {
#pragma unroll
            for (int v = 0; v < VX; v++) S1[tid + v * BXT] = v1[v];
        }
        // Refill the slot with the row 2R+1 ahead (clamped past the end).
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[PH][v] = lev0(xin[v], r0 + W, xcl[v]);
        nh[PH] = lev0(hin, r0 + W, hcl);
        __syncthreads();
        cf.refresh();
#pragma unroll
        for (int v = 0; v < VX; v++)
#pragma unroll
            for (int b = 0; b < W; b++) w0[v][PH][b] = S0[tid + v * BXT + b];
        if constexpr (W1ON) {
            // Columns near the tile edges read the spare cells or the other
            // buffer; their level-2 sums are not stored.
#pragma unroll
            for (int v = 0; v < VX; v++)
#pragma unroll
                for (int b = 0; b < W; b++) w1[v][PH][b] = S1[tid + v * BXT - R + b];
        }
        // Row r of a window sits in the slot of the step that staged it: the
        // level-0 row r0 - R + dy in slot (PH + R + 1 + dy) mod W, and the
        // level-1 row r0 - 2R - 1 + dy the same.
        if constexpr (OUT) {
            const int y2 = r0 - 2 * R - 1;
            T *orow = out + (long long)y2 * sy;
#pragma unroll
            for (int v = 0; v < VX; v++) {
                Acc<T, KAHAN> a;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int dy = S::OFF[t][0], dx = S::OFF[t][1];
                    const T val = w1[v][(PH + R + 1 + dy) % W][R + dx];
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                if (core[v]) store_out(orow + xcl[v], a.value(), nts);
            }
        }
        if constexpr (L1ON) {
            const int y1 = r0 - R;
            const bool yin = FULL || (y1 >= 0 && y1 < n0);
#pragma unroll
            for (int v = 0; v < VX; v++) {
                Acc<T, KAHAN> a;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int dy = S::OFF[t][0], dx = S::OFF[t][1];
                    const T val = w0[v][(PH + R + 1 + dy) % W][R + dx];
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                v1[v] = a.value();
                if (!FULL && !(yin && xin[v])) v1[v] = h1[(long long)clampi(y1, -R, n0 + R - 1) * sy + xcl[v]];
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
    int s0 = 2 * W - 1;
    for (; s0 + W <= N; s0 += W)
        static_for<W>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            step(std::integral_constant<int, (W - 1 + i) % W>{}, B1{}, B1{}, B1{}, s0 + i);
        });
    static_for<W - 1>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if (s0 + i < N) step(std::integral_constant<int, (W - 1 + i) % W>{}, B1{}, B1{}, B1{}, s0 + i);
    });
}

// The star form (every term on the centre row or the centre column): the
// 3D two-step star tile without its y axis.  Each thread keeps its columns'
// rows z1-R .. z1+R of level 0 in a register ring, loaded from memory two
// iterations before their first use, and the rows of level 1 likewise;
// only the centre rows go through LDS (level 0 at row z1, level 1 at row
// z2 = z1-1-R), for the in-row neighbours.  Iteration i = z1 - (ys - R):
//   stage both centre rows -- one barrier --
//   level 2 at row z2 (own columns): column terms from the level-1 ring,
//     row terms from LDS; stored;
//   level 1 at row z1 (extended columns): column terms from the level-0
//     ring, row terms from LDS; into the level-1 ring;
//   load level-0 row z1+R+2 into the slot row z1-R held.
// Registers per column: two rings of 2R+2 cells, instead of two windows.
template <typename T, class S, int VX, bool KAHAN, bool FULL, class CF>
__device__ __forceinline__ void win2d_tb2_star_tile(const T *__restrict__ in, T *__restrict__ out,
                                                    const T *__restrict__ h0, const T *__restrict__ h1,
                                                    const KParams<T, S::P> &p, int x0, int ys, int ye, T *lds,
                                                    CF &cf)
{
    constexpr int R = S::R, P = S::P, M0 = 2 * R + 2;
    const int tid = threadIdx.x, BXT = p.BXT;
    const int BX = p.TX + 2 * R, LW = BX + 2 * R;
    T *L0 = lds;              // 2 x LW: level-0 row z1, columns x0-2R ..
    T *L1 = lds + 2 * LW + R; // 2 x BX: level-1 row z2, columns x0-R .. (R spare cells either side)
    lds_poison(lds, 2 * LW + 2 * BX + 2 * R);
    const int n0 = p.n0, n2 = p.n2;
    const long long sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;

    int xcl[VX];
    bool xin[VX], core[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tid + v * BXT, xg = x0 - R + ex;
        xin[v] = FULL || (xg >= 0 && xg < n2);
        core[v] = ex >= R && ex < BX - R && (FULL || xg < n2);
        xcl[v] = FULL ? xg : clampi(xg, -R, n2 + R - 1);
    }
    const bool hthr = tid < 2 * R;
    const int hlx = tid < R ? tid : BX + tid;
    const int hxg = x0 - 2 * R + hlx;
    const bool hin = FULL || (hxg >= 0 && hxg < n2);
    const int hcl = FULL && hthr ? hxg : clampi(hxg, -R, n2 + R - 1);
    auto lev0 = [&](bool cin, int r, int c) -> T {
        if constexpr (FULL) {
            return in[(long long)min(r, n0 + R - 1) * sy + c];
        } else {
            const T *base = (cin && r >= 0 && r < n0) ? in : h0;
            return base[(long long)clampi(r, -R, n0 + R - 1) * sy + c];
        }
    };
    auto takes_halo1 = [&](int v, int r) { return !FULL && !(xin[v] && r >= 0 && r < n0); };
    auto halo1 = [&](int v, int r) { return h1[(long long)clampi(r, -R, n0 + R - 1) * sy + xcl[v]]; };

    const int z1s = ys - R;          // first level-1 row
    const int N = (ye - ys) + 2 * R; // level-1 rows; iterations 0 .. N
    // r0[v][(p - z1s + R) mod M0]: level-0 row p; r1[v][(p - z1s) mod M0]:
    // level-1 row p; rh[(p - z1s) mod 2]: the ring cells of level-0 row p;
    // rh1[v][(p - z1s) mod 2]: the level-1 halo value of row p.
    T r0[VX][M0], r1[VX][M0], rh[2], rh1[VX][2];
#pragma unroll
    for (int v = 0; v < VX; v++) {
#pragma unroll
        for (int j = 0; j < M0; j++) r1[v][j] = T(0);
        rh1[v][0] = rh1[v][1] = T(0);
    }
#pragma unroll
    for (int j = 0; j < 2 * R; j++)
#pragma unroll
        for (int v = 0; v < VX; v++) r0[v][j] = lev0(xin[v], z1s - R + j, xcl[v]);
#pragma unroll
    for (int g = 0; g < 2; g++) {
#pragma unroll
        for (int v = 0; v < VX; v++)
            if (takes_halo1(v, z1s + g)) rh1[v][g] = halo1(v, z1s + g);
#pragma unroll
        for (int v = 0; v < VX; v++) r0[v][2 * R + g] = lev0(xin[v], z1s + R + g, xcl[v]);
        rh[g] = lev0(hin, z1s + g, hcl);
    }

    auto step = [&](auto phc, int i) {
        constexpr int PH = decltype(phc)::value; // i mod M0 (M0 even: PH & 1 == i & 1)
        const int z1 = z1s + i;
        T *S0 = L0 + (PH & 1) * LW;
        T *S1 = L1 + (PH & 1) * BX;
        constexpr int L1S = PH - 1 - R + 2 * M0; // slot of level-1 row z2 = z1-1-R, plus 2*M0
#pragma unroll
        for (int v = 0; v < VX; v++) {
            S0[R + tid + v * BXT] = r0[v][(PH + R) % M0];
            S1[tid + v * BXT] = r1[v][L1S % M0];
        }
        if (hthr) S0[hlx] = rh[PH & 1];
        __syncthreads();
        cf.refresh();
        const int z2 = z1 - 1 - R;
        if (z2 >= ys) {
            // Columns near the tile edges read the spare cells or the other
            // buffer; their sums are not stored.
            T *orow = out + (long long)z2 * sy;
#pragma unroll
            for (int v = 0; v < VX; v++) {
                Acc<T, KAHAN> a;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int d0 = S::OFF[t][0], d1 = S::OFF[t][1];
                    T val;
                    if constexpr (d1 == 0) val = r1[v][(L1S + d0) % M0];
                    else val = S1[tid + v * BXT + d1];
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                if (core[v]) store_out(orow + xcl[v], a.value(), nts);
            }
        }
        if (i < N) {
            const bool zin = FULL || (z1 >= 0 && z1 < n0);
#pragma unroll
            for (int v = 0; v < VX; v++) {
                Acc<T, KAHAN> a;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int d0 = S::OFF[t][0], d1 = S::OFF[t][1];
                    T val;
                    if constexpr (d1 == 0) val = r0[v][(PH + R + d0) % M0];
                    else val = S0[R + tid + v * BXT + d1];
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                r1[v][PH] = (FULL || (zin && xin[v])) ? a.value() : rh1[v][PH & 1];
            }
#pragma unroll
            for (int v = 0; v < VX; v++)
                if (takes_halo1(v, z1 + 2)) rh1[v][PH & 1] = halo1(v, z1 + 2);
        }
        // Row z1+R+2 into the slot of row z1-R (last read above), the ring
        // cells of row z1+2 into the slot of row z1 (staged above).
        // Addresses are clamped, so the loads past the slab end are harmless.
#pragma unroll
        for (int v = 0; v < VX; v++) r0[v][PH] = lev0(xin[v], z1 + R + 2, xcl[v]);
        rh[PH & 1] = lev0(hin, z1 + 2, hcl);
    };
    int i0 = 0;
    for (; i0 + M0 - 1 <= N; i0 += M0)
        static_for<M0>([&](auto phc) { step(phc, i0 + decltype(phc)::value); });
    static_for<M0>([&](auto phc) {
        const int i = i0 + decltype(phc)::value;
        if (i <= N) step(phc, i);
    });
}

// Sum form (2D boxes; UHS_CFG_TEMPORAL2 without UHS_CFG_ROWRING).  A box's
// terms split into a row part, the longest prefix whose row offsets never
// decrease, and a tail (the rest; the radius-3 box's two outer columns).
// For each level and column a thread keeps one running sum per pending
// output row (2R+1 of them) and, for each column the tail uses, the last
// 2R+1 rows' values in a register ring, instead of a (2R+1)^2 window.  When
// a row arrives through LDS, its cells are applied, term by term in term
// order, to every pending sum they are a row-part term of (row offsets
// arrive in increasing order, so each sum takes its terms in term order);
// the sum whose last row this was then takes its tail terms, in term order,
// from the ring, and is complete.  Schedule, staging and halo rules are
// those of the window form (steps, level-0 row r0 = ys - 2R + s, level-1 row
// r0 - R - 1, one barrier per row); the sum of an output row sits in the
// ring slot of the phase of the step at which its first row (offset -R)
// arrived, and completes 2R steps later.
template <class S> constexpr int sum_row_part()
{
    int t = 1;
    while (t < S::P && S::OFF[t][0] >= S::OFF[t - 1][0]) t++;
    return t;
}
template <class S> constexpr bool sum_first_of_col(int t) // first tail term of its column
{
    for (int u = sum_row_part<S>(); u < t; u++)
        if (S::OFF[u][1] == S::OFF[t][1]) return false;
    return true;
}
template <class S> constexpr int sum_tail_cols()
{
    int n = 0;
    for (int t = sum_row_part<S>(); t < S::P; t++) n += sum_first_of_col<S>(t);
    return n;
}
template <class S> constexpr int sum_tail_col_of(int dx) // ring index of tail column dx
{
    int n = 0;
    for (int t = sum_row_part<S>(); t < S::P; t++)
        if (sum_first_of_col<S>(t)) {
            if (S::OFF[t][1] == dx) return n;
            n++;
        }
    return -1;
}
template <class S> constexpr int sum_tail_col_dx(int i)
{
    int n = 0;
    for (int t = sum_row_part<S>(); t < S::P; t++)
        if (sum_first_of_col<S>(t)) {
            if (n == i) return S::OFF[t][1];
            n++;
        }
    return 0;
}

constexpr int kTermGroup = 8; // sum form: coefficients fetched per group of terms

template <typename T, class S, int VX, bool KAHAN, bool FULL, class CF>
__device__ __forceinline__ void win2d_tb2_sum_tile(const T *__restrict__ in, T *__restrict__ out,
                                                   const T *__restrict__ h0, const T *__restrict__ h1,
                                                   const KParams<T, S::P> &p, int x0, int ys, int ye, T *lds, CF &cf)
{
    constexpr int R = S::R, P = S::P, W = 2 * R + 1;
    constexpr int TR = sum_row_part<S>(), NC = sum_tail_cols<S>(), NCA = NC > 0 ? NC : 1;
    static_assert(S::OFF[0][0] == -R, "the first term lies on the first row of the box");
    const int tid = threadIdx.x, BXT = p.BXT;
    const int BX = p.TX + 2 * R, LW = BX + 2 * R;
    T *L0 = lds;              // 2 x LW: level-0 row, columns x0-2R ..
    T *L1 = lds + 2 * LW + R; // 2 x BX: level-1 row, columns x0-R .. (R spare cells either side)
    lds_poison(lds, 2 * LW + 2 * BX + 2 * R);
    const int n0 = p.n0, n2 = p.n2;
    const long long sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;

    int xcl[VX];
    bool xin[VX], core[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tid + v * BXT, xg = x0 - R + ex;
        xin[v] = FULL || (xg >= 0 && xg < n2);
        core[v] = ex >= R && ex < BX - R && (FULL || xg < n2);
        xcl[v] = FULL ? xg : clampi(xg, -R, n2 + R - 1);
    }
    const bool hthr = tid < 2 * R;
    const int hlx = tid < R ? tid : BX + tid;
    const int hxg = x0 - 2 * R + hlx;
    const bool hin = FULL || (hxg >= 0 && hxg < n2);
    const int hcl = FULL && hthr ? hxg : clampi(hxg, -R, n2 + R - 1);
    auto lev0 = [&](bool cin, int r, int c) -> T {
        if constexpr (FULL) {
            return in[(long long)min(r, n0 + R - 1) * sy + c];
        } else {
            const T *base = (cin && r >= 0 && r < n0) ? in : h0;
            return base[(long long)clampi(r, -R, n0 + R - 1) * sy + c];
        }
    };
    T nrow[W][VX], nh[W];
#pragma unroll
    for (int f = 0; f < W; f++) {
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[f][v] = lev0(xin[v], ys - 2 * R + f, xcl[v]);
        nh[f] = lev0(hin, ys - 2 * R + f, hcl);
    }
    // a0 / c0: level-1 sums and tail ring (fed by level-0 rows); a1 / c1 the
    // same one level up.
    Acc<T, KAHAN> a0[VX][W], a1[VX][W];
    T c0[VX][NCA][W], c1[VX][NCA][W], v1[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        v1[v] = T(0);
#pragma unroll
        for (int k = 0; k < W; k++) {
            a0[v][k].zero();
            a1[v][k].zero();
#pragma unroll
            for (int i = 0; i < NCA; i++) c0[v][i][k] = c1[v][i][k] = T(0);
        }
    }
    const int N = (ye - ys) + 4 * R + 1; // steps

    // Rows arrive at a step of phase PH: row0 (level 0, a pointer to its
    // cell at column x - R) into the sums a0 and the ring c0, and with L2,
    // row1 (level 1) into a1 and c1.  The sums in slot (PH + 1) mod W take
    // their last row and their tail terms; they are returned in o0 / o1
    // (meaningful once all their rows have arrived).  Both levels take each
    // coefficient in turn, kTermGroup terms at a time: the sums are pinned
    // after each group and the coefficient pointer refreshed before the
    // next, so a large stencil's coefficients are fetched a group at a time
    // instead of all at once (they do not fit the scalar registers).
    auto arrive = [&](auto phc, auto l2c, int v, const T *row0, const T *row1, T &o0, T &o1) {
        constexpr int PH = decltype(phc)::value;
        constexpr bool L2 = decltype(l2c)::value;
        constexpr int G = kTermGroup;
        T e0[W], e1[W];
#pragma unroll
        for (int b = 0; b < W; b++) {
            e0[b] = row0[b];
            e1[b] = L2 ? row1[b] : T(0);
        }
        static_for<NC>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            c0[v][i][PH] = e0[R + sum_tail_col_dx<S>(i)];
            if constexpr (L2) c1[v][i][PH] = e1[R + sum_tail_col_dx<S>(i)];
        });
        constexpr int fs = (PH + 1) % W;
        static_for<P>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            constexpr int dy = S::OFF[t][0], dx = S::OFF[t][1];
            if constexpr (t % G == 0 && t > 0) {
#pragma unroll
                for (int k = 0; k < W; k++) {
                    a0[v][k].pin();
                    if constexpr (L2) a1[v][k].pin();
                }
                cf.refresh();
            }
            const T c = cf.template get<t>();
            if constexpr (t < TR) {
                constexpr int slot = ((PH - dy - R) % W + W) % W;
                if constexpr (t == 0) {
                    a0[v][slot].first(c, e0[R + dx]);
                    if constexpr (L2) a1[v][slot].first(c, e1[R + dx]);
                } else {
                    a0[v][slot].add(c, e0[R + dx]);
                    if constexpr (L2) a1[v][slot].add(c, e1[R + dx]);
                }
            } else {
                constexpr int rs = ((PH - R + dy) % W + W) % W; // the ring slot of row y + dy
                constexpr int col = sum_tail_col_of<S>(dx);
                a0[v][fs].add(c, c0[v][col][rs]);
                if constexpr (L2) a1[v][fs].add(c, c1[v][col][rs]);
            }
        });
        o0 = a0[v][fs].value();
        o1 = a1[v][fs].value();
    };

    auto step = [&](auto phc, auto l1c, auto w1c, auto outc, int s) {
        constexpr int PH = decltype(phc)::value;
        constexpr bool L1ON = decltype(l1c)::value, W1ON = decltype(w1c)::value, OUT = decltype(outc)::value;
        const int r0 = ys - 2 * R + s;
        T *S0 = L0 + (s & 1) * LW;
        T *S1 = L1 + (s & 1) * BX;
#pragma unroll
        for (int v = 0; v < VX; v++) S0[R + tid + v * BXT] = nrow[PH][v];
        if (hthr) S0[hlx] = nh[PH];
        if constexpr (W1ON) {
#pragma unroll
            for (int v = 0; v < VX; v++) S1[tid + v * BXT] = v1[v];
        }
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[PH][v] = lev0(xin[v], r0 + W, xcl[v]);
        nh[PH] = lev0(hin, r0 + W, hcl);
        __syncthreads();
        cf.refresh();
        // Level-0 row r0 completes level-1 row y1 = r0 - R; level-1 row r0 -
        // R - 1 (from W1ON on) completes level-2 row y2 = r0 - 2R - 1.
        // Columns near the tile edges read the spare cells or the other
        // buffer; their level-2 sums are not stored.
        const int y1 = r0 - R, y2 = r0 - 2 * R - 1;
        const bool yin = FULL || (y1 >= 0 && y1 < n0);
        T *orow = out + (long long)y2 * sy;
#pragma unroll
        for (int v = 0; v < VX; v++) {
            T o0, o1;
            arrive(phc, std::integral_constant<bool, W1ON>{}, v, S0 + tid + v * BXT, S1 + tid + v * BXT - R, o0, o1);
            if constexpr (OUT)
                if (core[v]) store_out(orow + xcl[v], o1, nts);
            if constexpr (L1ON) {
                v1[v] = o0;
                if (!FULL && !(yin && xin[v])) v1[v] = h1[(long long)clampi(y1, -R, n0 + R - 1) * sy + xcl[v]];
            }
        }
    };
    using B0 = std::false_type;
    using B1 = std::true_type;
    static_for<2 * W - 1>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        step(std::integral_constant<int, i % W>{}, std::integral_constant<bool, (i >= 2 * R)>{},
             std::integral_constant<bool, (i >= 2 * R + 1)>{}, B0{}, i);
    });
    int s0 = 2 * W - 1;
    for (; s0 + W <= N; s0 += W)
        static_for<W>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            step(std::integral_constant<int, (W - 1 + i) % W>{}, B1{}, B1{}, B1{}, s0 + i);
        });
    static_for<W - 1>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if (s0 + i < N) step(std::integral_constant<int, (W - 1 + i) % W>{}, B1{}, B1{}, B1{}, s0 + i);
    });
}

// FORM: 1 the window form (any 2D stencil), 0 the star ring form, 2 the sum
// form (boxes).
template <typename T, class S, int VX, bool KAHAN, int FORM>
__global__ void __launch_bounds__(kMaxThreads) k_win2d_tb2(TB2Params<T, S::P> tp)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const KParams<T, S::P> &p = tp.k;
    const int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x, chunk = bid / p.tiles_x;
    const int x0 = tile_x * p.TX, ys = chunk * p.ZC, ye = min(ys + p.ZC, p.n0);
    const T *h0 = (tp.k0 & 1) ? tp.hB : tp.hA;
    const T *h1 = (tp.k0 & 1) ? tp.hA : tp.hB;
    Coefs<T, S::P> cf(p);
    T *lds = reinterpret_cast<T *>(smem_raw);
    const bool full = win2d_tb2_full<S::R>(x0, p.TX, ys, ye, p.n0, p.n2);
    if constexpr (FORM == 1) {
        if (full) win2d_tb2_tile<T, S, VX, KAHAN, true>(p.in, p.out, h0, h1, p, x0, ys, ye, lds, cf);
        else win2d_tb2_tile<T, S, VX, KAHAN, false>(p.in, p.out, h0, h1, p, x0, ys, ye, lds, cf);
    } else if constexpr (FORM == 2) {
        if (full) win2d_tb2_sum_tile<T, S, VX, KAHAN, true>(p.in, p.out, h0, h1, p, x0, ys, ye, lds, cf);
        else win2d_tb2_sum_tile<T, S, VX, KAHAN, false>(p.in, p.out, h0, h1, p, x0, ys, ye, lds, cf);
    } else {
        static_assert(S::STAR, "the ring form needs a star");
        if (full) win2d_tb2_star_tile<T, S, VX, KAHAN, true>(p.in, p.out, h0, h1, p, x0, ys, ye, lds, cf);
        else win2d_tb2_star_tile<T, S, VX, KAHAN, false>(p.in, p.out, h0, h1, p, x0, ys, ye, lds, cf);
    }
}

template <typename T, class S, int VX, bool KAHAN, int FORM>
hipError_t launch_win2d_tb2(const SweepArgs &a, hipStream_t st)
{
    TB2Params<T, S::P> tp;
    tp.k = make_params<T, S::P>(a);
    constexpr int R = S::R;
    const int bx = a.tx + 2 * R;
    tp.k.BXT = bx / VX;
    const long long h = a.halo, origin = h * a.sy + h;
    tp.hA = static_cast<const T *>(a.hA) + origin;
    tp.hB = static_cast<const T *>(a.hB) + origin;
    tp.k0 = a.k0;
    const size_t lds = sizeof(T) * (2 * (size_t)(bx + 2 * R) + 2 * (size_t)bx + 2 * R);
    hipLaunchKernelGGL((k_win2d_tb2<T, S, VX, KAHAN, FORM>), dim3(tp.k.nblocks), dim3(bx / VX), lds, st, tp);
    return hipGetLastError();
}

// Compiled where the registers hold the tile without spilling.  The ring
// form (stars) keeps two rings of 2R+2 cells per column: 1, 2 or 4 columns
// per thread.  The window form keeps two windows of (2R+1)^2 cells per
// column (a star's dead cells excepted): 4 columns at radius 1; at radius 2
// one (FP64) or two (FP32); at radius 3 one, and none for the FP64 box.  The
// sum form (boxes) keeps 2R+1 sums and the tail rings per column and level:
// one or two columns.
template <typename T, class S, int FORM> constexpr int tb2_max_vec()
{
    constexpr bool f64 = sizeof(T) == 8;
    if constexpr (FORM == 0) return S::STAR ? 4 : 0;
    else if constexpr (FORM == 2) return S::STAR ? 0 : 2;
    else if constexpr (S::R == 1) return 4;
    else if constexpr (S::R == 2) return f64 ? 1 : 2;
    else return (f64 && !S::STAR) ? 0 : 1;
}
template <typename T, bool K, class S, int FORM> LaunchFn pick_vec(int vec)
{
    constexpr int M = tb2_max_vec<T, S, FORM>();
    switch (vec) {
    case 1:
        if constexpr (M >= 1) return launch_win2d_tb2<T, S, 1, K, FORM>;
        else return nullptr;
    case 2:
        if constexpr (M >= 2) return launch_win2d_tb2<T, S, 2, K, FORM>;
        else return nullptr;
    case 4:
        if constexpr (M >= 4) return launch_win2d_tb2<T, S, 4, K, FORM>;
        else return nullptr;
    }
    return nullptr;
}
// window: the window form; otherwise a star's ring form or a box's sum form.
template <typename T, bool K, class S> LaunchFn pick_form(int vec, bool window)
{
    if (window) return pick_vec<T, K, S, 1>(vec);
    return S::STAR ? pick_vec<T, K, S, 0>(vec) : pick_vec<T, K, S, 2>(vec);
}
template <class S> LaunchFn pick(int dtype, int vec, bool window)
{
    switch (dtype) {
    case DT_F64: return pick_form<double, false, S>(vec, window);
    case DT_F32: return pick_form<float, false, S>(vec, window);
    case DT_F32K: return pick_form<float, true, S>(vec, window);
    }
    return nullptr;
}
} // namespace

LaunchFn uhs_find_win2d_tb2(int prog, int dtype, int vec, int window)
{
    switch (prog) {
    case Prog_2d1r_star::ID: return pick<Prog_2d1r_star>(dtype, vec, window);
    case Prog_2d2r_star::ID: return pick<Prog_2d2r_star>(dtype, vec, window);
    case Prog_2d3r_star::ID: return pick<Prog_2d3r_star>(dtype, vec, window);
    case Prog_2d1r_box::ID: return pick<Prog_2d1r_box>(dtype, vec, window);
    case Prog_2d2r_box::ID: return pick<Prog_2d2r_box>(dtype, vec, window);
    case Prog_2d3r_box::ID: return pick<Prog_2d3r_box>(dtype, vec, window);
    }
    return nullptr;
}
