// win2d_tbk: K timesteps per launch (UHS_CFG_TEMPORALK, K = tile_y) for the
// 2D stars and the radius-1 and radius-2 2D boxes (temporal blocking with a
// (K-1)R-column ghost zone), so each cell crosses HBM once per K updates.
// FP64, FP32 and FP32 compensated.  Three forms: the ring form (stars; 1 or
// 2 columns per thread, and 4 for the radius-1 star), the sum form (boxes
// whose terms come row by row: the radius-2 box) and the window form (the
// radius-1 box); the boxes take 1 or 2 columns per thread.  Each k_win2d_tbK.hip instantiates one K (a
// translation unit per K keeps the build parallel).
#ifndef UHS_K_WIN2D_TBK_H
#define UHS_K_WIN2D_TBK_H
#include "sweep_kernels.h"

namespace {

// K steps per launch, ring form (2D stars; UHS_CFG_TEMPORALK with K =
// tile_y).
// The two-step ring form with K-1 computed levels: at iteration i level k
// (1 <= k <= K) is computed at row z_k = z1 - (k-1)(R+1), z1 = z1s + i, from
// the 2R+1 rows of level k-1 around it (its column terms, from level k-1's
// register ring) and the centre row of level k-1 staged in LDS (its row
// terms).  Every level keeps 2R+2 rows of each column in a register ring;
// the row of level k computed at iteration i sits in slot i mod (2R+2), the
// level-0 row p in slot (p - z1s + R) mod (2R+2), loaded two iterations
// before its first use.  Level K is the output, stored for rows [ys, ye).
// All levels share one barrier per row.  The workgroup covers tile_x output
// columns plus (K-1)R on each side (the columns level 1 needs); level k is
// only used (K-k)R columns into that margin, the rest of its columns are
// computed and never read.  A level-k cell outside the interior takes the
// halo of the grid that holds state k0 + k in the one-step ping-pong (hA for
// even states, hB for odd).
template <typename T, class S, int VX, bool KAHAN, int K, bool FULL, class CF>
__device__ __forceinline__ void win2d_tbk_star_tile(const T *__restrict__ in, T *__restrict__ out,
                                                    const T *__restrict__ hA, const T *__restrict__ hB, int k0,
                                                    const KParams<T, S::P> &p, int x0, int ys, int ye, T *lds,
                                                    CF &cf)
{
    constexpr int R = S::R, P = S::P, M0 = 2 * R + 2;
    static_assert(K >= 2, "at least two steps");
    const int tid = threadIdx.x, BXT = p.BXT;
    const int BX = p.TX + 2 * (K - 1) * R, LW = BX + 2 * R;
    // LDS: level 0 (2 x LW, columns x0-KR ..), then per level k = 1 .. K-1
    // 2 x BX cells (columns x0-(K-1)R ..) with R spare cells either side.
    T *L0 = lds;
    auto Lk = [&](int k) { return lds + 2 * LW + (k - 1) * (2 * BX + 2 * R) + R; };
    lds_poison(lds, 2 * LW + (K - 1) * (2 * BX + 2 * R));
    const int n0 = p.n0, n2 = p.n2;
    const long long sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;
    const T *h0 = (k0 & 1) ? hB : hA;

    // Column ex = tid + v*BXT is interior column x0 - (K-1)R + ex.
    int xcl[VX];
    bool xin[VX], core[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tid + v * BXT, xg = x0 - (K - 1) * R + ex;
        xin[v] = FULL || (xg >= 0 && xg < n2);
        core[v] = ex >= (K - 1) * R && ex < BX - (K - 1) * R && (FULL || xg < n2);
        xcl[v] = FULL ? xg : clampi(xg, -R, n2 + R - 1);
    }
    const bool hthr = tid < 2 * R;
    const int hlx = tid < R ? tid : BX + tid;
    const int hxg = x0 - K * R + hlx;
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

    const int z1s = ys - (K - 1) * R;                            // first level-1 row
    const int NI = (ye - ys) - 1 + (K - 1) * (2 * R + 1);        // last iteration
    // rr[0]: level-0 rows; rr[k]: level-k rows (k = 1 .. K-1).
    T rr[K][VX][M0], rh[2];
#pragma unroll
    for (int k = 1; k < K; k++)
#pragma unroll
        for (int v = 0; v < VX; v++)
#pragma unroll
            for (int j = 0; j < M0; j++) rr[k][v][j] = T(0);
#pragma unroll
    for (int j = 0; j < 2 * R; j++)
#pragma unroll
        for (int v = 0; v < VX; v++) rr[0][v][j] = lev0(xin[v], z1s - R + j, xcl[v]);
#pragma unroll
    for (int g = 0; g < 2; g++) {
#pragma unroll
        for (int v = 0; v < VX; v++) rr[0][v][2 * R + g] = lev0(xin[v], z1s + R + g, xcl[v]);
        rh[g] = lev0(hin, z1s + g, hcl);
    }

    auto step = [&](auto phc, int i) {
        constexpr int PH = decltype(phc)::value; // i mod M0 (M0 even: PH & 1 == i & 1)
        constexpr int CS = (PH - 1 - R + 2 * M0) % M0; // slot of the centre row a level above reads
        const int z1 = z1s + i;
        T *S0 = L0 + (PH & 1) * LW;
#pragma unroll
        for (int v = 0; v < VX; v++) S0[R + tid + v * BXT] = rr[0][v][(PH + R) % M0];
        if (hthr) S0[hlx] = rh[PH & 1];
        static_for<K - 1>([&](auto kc) {
            constexpr int k = decltype(kc)::value + 1;
            T *Sk = Lk(k) + (PH & 1) * BX;
#pragma unroll
            for (int v = 0; v < VX; v++) Sk[tid + v * BXT] = rr[k][v][CS];
        });
        __syncthreads();
        cf.refresh();
        // Levels K .. 1.  Level k's input is level k-1: its row z_k + d0 is in
        // slot (PH - R - 1 + d0) of ring k-1 (level 0: slot PH + R + d0), its
        // centre row in LDS.
        static_for<K>([&](auto kc) {
            constexpr int k = K - decltype(kc)::value;
            const int zk = z1 - (k - 1) * (R + 1);
            const int first = ys - (K - k) * R; // first row of level k that is used
            if (zk < first) return;
            const T *Sin;
            if constexpr (k == 1) Sin = S0 + R;
            else Sin = Lk(k - 1) + (PH & 1) * BX;
            T *orow = out + (long long)zk * sy;
            const bool zin = FULL || (zk >= 0 && zk < n0);
            const T *hk = ((k0 + k) & 1) ? hB : hA;
#pragma unroll
            for (int v = 0; v < VX; v++) {
                Acc<T, KAHAN> a;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int d0 = S::OFF[t][0], d1 = S::OFF[t][1];
                    T val;
                    if constexpr (d1 == 0) {
                        if constexpr (k == 1) val = rr[0][v][(PH + R + d0 + M0) % M0];
                        else val = rr[k - 1][v][(PH - R - 1 + d0 + 2 * M0) % M0];
                    } else {
                        val = Sin[tid + v * BXT + d1];
                    }
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                if constexpr (k == K) {
                    if (core[v]) store_out(orow + xcl[v], a.value(), nts);
                } else {
                    T val = a.value();
                    if (!FULL && !(zin && xin[v])) val = hk[(long long)clampi(zk, -R, n0 + R - 1) * sy + xcl[v]];
                    rr[k][v][PH] = val;
                }
            }
        });
        // Level-0 row z1+R+2 into the slot of row z1-R (last read above), the
        // ring cells of row z1+2 into the slot of row z1 (staged above).
#pragma unroll
        for (int v = 0; v < VX; v++) rr[0][v][PH] = lev0(xin[v], z1 + R + 2, xcl[v]);
        rh[PH & 1] = lev0(hin, z1 + 2, hcl);
    };
    int i0 = 0;
    for (; i0 + M0 - 1 <= NI; i0 += M0)
        static_for<M0>([&](auto phc) { step(phc, i0 + decltype(phc)::value); });
    static_for<M0>([&](auto phc) {
        const int i = i0 + decltype(phc)::value;
        if (i <= NI) step(phc, i);
    });
}

// Every term of S in non-decreasing row offset (the sum form's condition).
template <class S> constexpr bool win2d_rows_in_order()
{
    for (int t = 1; t < S::P; t++)
        if (S::OFF[t][0] < S::OFF[t - 1][0]) return false;
    return true;
}

// K steps per launch, sum form (2D boxes whose terms come row by row in
// non-decreasing row offset: of the benchmark boxes, the radius-2 box; the
// radius-1 box lists its centre row last).  The two-step
// sum form with K levels: every level keeps, per column, one running sum per
// pending output row (2R+1 of them, in a ring indexed by the step's phase).
// At step s level-0 row r0 = ys - KR + s arrives through LDS, and so does
// each intermediate level's row computed at step s-1; each arriving row's
// cells are applied, term by term in term order, to the pending sums of the
// level above, and the sum whose last row this was completes: level k row
// r0 - k(R+1) + 1.  Level K is the output.  The levels of a step read only
// rows staged before its barrier, so one barrier per row serves all.  Rows
// and columns a level computes outside the range the level above uses are
// never read.  Same columns, ghost zone and halo rules as the ring form.
template <typename T, class S, int VX, bool KAHAN, int K, bool FULL, class CF>
__device__ __forceinline__ void win2d_tbk_sum_tile(const T *__restrict__ in, T *__restrict__ out,
                                                   const T *__restrict__ hA, const T *__restrict__ hB, int k0,
                                                   const KParams<T, S::P> &p, int x0, int ys, int ye, T *lds,
                                                   CF &cf)
{
    constexpr int R = S::R, P = S::P, W = 2 * R + 1;
    static_assert(K >= 2, "at least two steps");
    static_assert(S::OFF[0][0] == -R, "the first term lies on the first row of the box");
    static_assert(win2d_rows_in_order<S>(), "every term in non-decreasing row offset");
    const int tid = threadIdx.x, BXT = p.BXT;
    const int BX = p.TX + 2 * (K - 1) * R, LW = BX + 2 * R;
    T *L0 = lds;
    auto Lk = [&](int k) { return lds + 2 * LW + (k - 1) * (2 * BX + 2 * R) + R; };
    lds_poison(lds, 2 * LW + (K - 1) * (2 * BX + 2 * R));
    const int n0 = p.n0, n2 = p.n2;
    const long long sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;
    const T *h0 = (k0 & 1) ? hB : hA;

    int xcl[VX];
    bool xin[VX], core[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tid + v * BXT, xg = x0 - (K - 1) * R + ex;
        xin[v] = FULL || (xg >= 0 && xg < n2);
        core[v] = ex >= (K - 1) * R && ex < BX - (K - 1) * R && (FULL || xg < n2);
        xcl[v] = FULL ? xg : clampi(xg, -R, n2 + R - 1);
    }
    const bool hthr = tid < 2 * R;
    const int hlx = tid < R ? tid : BX + tid;
    const int hxg = x0 - K * R + hlx;
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
    const int rs0 = ys - K * R; // level-0 row of step 0
    // Prefetch ring: slot f holds the level-0 row staged at the next step of
    // phase f, loaded 2R+1 rows ahead.
    T nrow[W][VX], nh[W];
#pragma unroll
    for (int f = 0; f < W; f++) {
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[f][v] = lev0(xin[v], rs0 + f, xcl[v]);
        nh[f] = lev0(hin, rs0 + f, hcl);
    }
    // a[k-1]: the level-k sums; vk[k-1]: level k's row completed at the
    // previous step (staged at this one), k = 1 .. K-1.
    Acc<T, KAHAN> a[K][VX][W];
    T vk[K - 1][VX];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int v = 0; v < VX; v++)
#pragma unroll
            for (int j = 0; j < W; j++) a[k][v][j].zero();
#pragma unroll
    for (int k = 0; k < K - 1; k++)
#pragma unroll
        for (int v = 0; v < VX; v++) vk[k][v] = T(0);
    const int N = (ye - ys) + K * (2 * R + 1) - 1; // steps

    auto step = [&](auto phc, int s) {
        constexpr int PH = decltype(phc)::value;
        constexpr int fs = (PH + 1) % W; // the slot of the sums that complete
        const int r0 = rs0 + s;
        T *S0 = L0 + (s & 1) * LW;
#pragma unroll
        for (int v = 0; v < VX; v++) S0[R + tid + v * BXT] = nrow[PH][v];
        if (hthr) S0[hlx] = nh[PH];
        static_for<K - 1>([&](auto kc) {
            constexpr int k = decltype(kc)::value + 1;
            T *Sk = Lk(k) + (s & 1) * BX;
#pragma unroll
            for (int v = 0; v < VX; v++) Sk[tid + v * BXT] = vk[k - 1][v];
        });
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[PH][v] = lev0(xin[v], r0 + W, xcl[v]);
        nh[PH] = lev0(hin, r0 + W, hcl);
        __syncthreads();
        cf.refresh();
        static_for<K>([&](auto kc) {
            constexpr int k = K - decltype(kc)::value; // level K first
            const int yk = r0 - k * (R + 1) + 1;       // the level-k row that completes
            const T *row;                               // arriving row, cell of column x - R
            if constexpr (k == 1) row = S0;
            else row = Lk(k - 1) + (s & 1) * BX - R;
            const bool yin = FULL || (yk >= 0 && yk < n0);
            const T *hk = ((k0 + k) & 1) ? hB : hA;
#pragma unroll
            for (int v = 0; v < VX; v++) {
                T e[W];
#pragma unroll
                for (int b = 0; b < W; b++) e[b] = row[tid + v * BXT + b];
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int dy = S::OFF[t][0], dx = S::OFF[t][1];
                    constexpr int slot = ((PH - dy - R) % W + W) % W;
                    if constexpr (t == 0) a[k - 1][v][slot].first(cf.template get<t>(), e[R + dx]);
                    else a[k - 1][v][slot].add(cf.template get<t>(), e[R + dx]);
                });
                // Pin the level's sums: its FMAs finish before the next
                // level's LDS reads, so one level's cells are live at a time.
#pragma unroll
                for (int j = 0; j < W; j++) a[k - 1][v][j].pin();
                const T o = a[k - 1][v][fs].value();
                if constexpr (k == K) {
                    if (yk >= ys && core[v]) store_out(out + (long long)yk * sy + xcl[v], o, nts);
                } else {
                    vk[k - 1][v] = (FULL || (yin && xin[v])) ? o
                                                            : hk[(long long)clampi(yk, -R, n0 + R - 1) * sy + xcl[v]];
                }
            }
        });
    };
    int s0 = 0;
    for (; s0 + W <= N; s0 += W)
        static_for<W>([&](auto ic) { step(ic, s0 + decltype(ic)::value); });
    static_for<W - 1>([&](auto ic) {
        const int s = s0 + decltype(ic)::value;
        if (s < N) step(ic, s);
    });
}

// K steps per launch, window form (2D boxes in any term order; compiled for
// the radius-1 box, whose centre row comes last).  The two-step window form
// with K levels: every level keeps, per column, the (2R+1)^2 window of the
// level below in a register ring of 2R+1 rows (slot: the phase of the step
// that staged the row).  At step s level-0 row r0 = ys - KR + s and each
// intermediate level's row computed at step s-1 arrive through LDS; level k
// then computes row r0 - k(R+1) + 1 from its window, in term order.  Same
// schedule, columns, ghost zone and halo rules as the sum form.
template <typename T, class S, int VX, bool KAHAN, int K, bool FULL, class CF>
__device__ __forceinline__ void win2d_tbk_win_tile(const T *__restrict__ in, T *__restrict__ out,
                                                   const T *__restrict__ hA, const T *__restrict__ hB, int k0,
                                                   const KParams<T, S::P> &p, int x0, int ys, int ye, T *lds,
                                                   CF &cf)
{
    constexpr int R = S::R, P = S::P, W = 2 * R + 1;
    static_assert(K >= 2, "at least two steps");
    const int tid = threadIdx.x, BXT = p.BXT;
    const int BX = p.TX + 2 * (K - 1) * R, LW = BX + 2 * R;
    T *L0 = lds;
    auto Lk = [&](int k) { return lds + 2 * LW + (k - 1) * (2 * BX + 2 * R) + R; };
    lds_poison(lds, 2 * LW + (K - 1) * (2 * BX + 2 * R));
    const int n0 = p.n0, n2 = p.n2;
    const long long sy = p.sy;
    const bool nts = p.flags & UHS_CFG_NT_STORE;
    const T *h0 = (k0 & 1) ? hB : hA;

    int xcl[VX];
    bool xin[VX], core[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        const int ex = tid + v * BXT, xg = x0 - (K - 1) * R + ex;
        xin[v] = FULL || (xg >= 0 && xg < n2);
        core[v] = ex >= (K - 1) * R && ex < BX - (K - 1) * R && (FULL || xg < n2);
        xcl[v] = FULL ? xg : clampi(xg, -R, n2 + R - 1);
    }
    const bool hthr = tid < 2 * R;
    const int hlx = tid < R ? tid : BX + tid;
    const int hxg = x0 - K * R + hlx;
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
    const int rs0 = ys - K * R; // level-0 row of step 0
    T nrow[W][VX], nh[W];
#pragma unroll
    for (int f = 0; f < W; f++) {
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[f][v] = lev0(xin[v], rs0 + f, xcl[v]);
        nh[f] = lev0(hin, rs0 + f, hcl);
    }
    // w[k-1][v][slot][b]: the window level k computes from (level k-1 rows,
    // column x - R + b); vk[k-1]: level k's row computed at the previous
    // step (staged at this one), k = 1 .. K-1.
    T w[K][VX][W][W], vk[K - 1][VX];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int v = 0; v < VX; v++)
#pragma unroll
            for (int j = 0; j < W; j++)
#pragma unroll
                for (int b = 0; b < W; b++) w[k][v][j][b] = T(0);
#pragma unroll
    for (int k = 0; k < K - 1; k++)
#pragma unroll
        for (int v = 0; v < VX; v++) vk[k][v] = T(0);
    const int N = (ye - ys) + K * (2 * R + 1) - 1; // steps

    auto step = [&](auto phc, int s) {
        constexpr int PH = decltype(phc)::value;
        const int r0 = rs0 + s;
        T *S0 = L0 + (s & 1) * LW;
#pragma unroll
        for (int v = 0; v < VX; v++) S0[R + tid + v * BXT] = nrow[PH][v];
        if (hthr) S0[hlx] = nh[PH];
        static_for<K - 1>([&](auto kc) {
            constexpr int k = decltype(kc)::value + 1;
            T *Sk = Lk(k) + (s & 1) * BX;
#pragma unroll
            for (int v = 0; v < VX; v++) Sk[tid + v * BXT] = vk[k - 1][v];
        });
#pragma unroll
        for (int v = 0; v < VX; v++) nrow[PH][v] = lev0(xin[v], r0 + W, xcl[v]);
        nh[PH] = lev0(hin, r0 + W, hcl);
        __syncthreads();
        cf.refresh();
        static_for<K>([&](auto kc) {
            constexpr int k = K - decltype(kc)::value; // level K first
            const int yk = r0 - k * (R + 1) + 1;       // the level-k row computed
            const T *row;                               // arriving row, cell of column x - R
            if constexpr (k == 1) row = S0;
            else row = Lk(k - 1) + (s & 1) * BX - R;
            const bool yin = FULL || (yk >= 0 && yk < n0);
            const T *hk = ((k0 + k) & 1) ? hB : hA;
#pragma unroll
            for (int v = 0; v < VX; v++) {
#pragma unroll
                for (int b = 0; b < W; b++) w[k - 1][v][PH][b] = row[tid + v * BXT + b];
                Acc<T, KAHAN> a;
                static_for<P>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr int dy = S::OFF[t][0], dx = S::OFF[t][1];
                    const T val = w[k - 1][v][(PH + R + 1 + dy) % W][R + dx];
                    if constexpr (t == 0) a.first(cf.template get<t>(), val);
                    else a.add(cf.template get<t>(), val);
                });
                a.pin();
                if constexpr (k == K) {
                    if (yk >= ys && core[v]) store_out(out + (long long)yk * sy + xcl[v], a.value(), nts);
                } else {
                    vk[k - 1][v] = (FULL || (yin && xin[v]))
                                       ? a.value()
                                       : hk[(long long)clampi(yk, -R, n0 + R - 1) * sy + xcl[v]];
                }
            }
        });
    };
    int s0 = 0;
    for (; s0 + W <= N; s0 += W)
        static_for<W>([&](auto ic) { step(ic, s0 + decltype(ic)::value); });
    static_for<W - 1>([&](auto ic) {
        const int s = s0 + decltype(ic)::value;
        if (s < N) step(ic, s);
    });
}

template <typename T, class S, int VX, bool KAHAN, int K>
__global__ void __launch_bounds__(kMaxThreads) k_win2d_tbk(TB2Params<T, S::P> tp)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const KParams<T, S::P> &p = tp.k;
    const int bid = xcd_remap(blockIdx.x, p.nblocks);
    const int tile_x = bid % p.tiles_x, chunk = bid / p.tiles_x;
    const int x0 = tile_x * p.TX, ys = chunk * p.ZC, ye = min(ys + p.ZC, p.n0);
    Coefs<T, S::P> cf(p);
    T *lds = reinterpret_cast<T *>(smem_raw);
    constexpr int KR = K * S::R;
    const bool full = x0 >= KR && x0 + p.TX + KR <= p.n2 && ys >= KR && ye + KR <= p.n0;
    if constexpr (S::STAR) {
        if (full) win2d_tbk_star_tile<T, S, VX, KAHAN, K, true>(p.in, p.out, tp.hA, tp.hB, tp.k0, p, x0, ys, ye, lds, cf);
        else win2d_tbk_star_tile<T, S, VX, KAHAN, K, false>(p.in, p.out, tp.hA, tp.hB, tp.k0, p, x0, ys, ye, lds, cf);
    } else if constexpr (win2d_rows_in_order<S>()) {
        if (full) win2d_tbk_sum_tile<T, S, VX, KAHAN, K, true>(p.in, p.out, tp.hA, tp.hB, tp.k0, p, x0, ys, ye, lds, cf);
        else win2d_tbk_sum_tile<T, S, VX, KAHAN, K, false>(p.in, p.out, tp.hA, tp.hB, tp.k0, p, x0, ys, ye, lds, cf);
    } else {
        if (full) win2d_tbk_win_tile<T, S, VX, KAHAN, K, true>(p.in, p.out, tp.hA, tp.hB, tp.k0, p, x0, ys, ye, lds, cf);
        else win2d_tbk_win_tile<T, S, VX, KAHAN, K, false>(p.in, p.out, tp.hA, tp.hB, tp.k0, p, x0, ys, ye, lds, cf);
    }
}

template <typename T, class S, int VX, bool KAHAN, int K>
hipError_t launch_win2d_tbk(const SweepArgs &a, hipStream_t st)
{
    TB2Params<T, S::P> tp;
    tp.k = make_params<T, S::P>(a);
    constexpr int R = S::R;
    const int bx = a.tx + 2 * (K - 1) * R;
    tp.k.BXT = bx / VX;
    const long long h = a.halo, origin = h * a.sy + h;
    tp.hA = static_cast<const T *>(a.hA) + origin;
    tp.hB = static_cast<const T *>(a.hB) + origin;
    tp.k0 = a.k0;
    const size_t lds = sizeof(T) * (2 * (size_t)(bx + 2 * R) + (K - 1) * (2 * (size_t)bx + 2 * R));
    hipLaunchKernelGGL((k_win2d_tbk<T, S, VX, KAHAN, K>), dim3(tp.k.nblocks), dim3(bx / VX), lds, st, tp);
    return hipGetLastError();
}

// Stars: 1 or 2 columns per thread, and 4 for the radius-1 star.  Boxes whose terms come row by row
// (sum form; the radius-2 box): 1 or 2 (1 in FP64, where two columns would
// spill).  Other radius-1 boxes (window form): 1 or 2.
template <typename T, bool KAHAN, class S, int K> LaunchFn win2d_tbk_pick_vec(int vec)
{
    if constexpr (!S::STAR && !win2d_rows_in_order<S>() && S::R > 1) return nullptr;
    else switch (vec) {
        case 1: return launch_win2d_tbk<T, S, 1, KAHAN, K>;
        case 2:
            if constexpr (!S::STAR && S::R == 2 && sizeof(T) == 8) return nullptr;
            else return launch_win2d_tbk<T, S, 2, KAHAN, K>;
        case 4: // the radius-1 star only: wider stars spill at four columns
            if constexpr (S::STAR && S::R == 1) return launch_win2d_tbk<T, S, 4, KAHAN, K>;
            else return nullptr;
        }
    return nullptr;
}
template <class S, int K> LaunchFn win2d_tbk_pick(int dtype, int vec)
{
    switch (dtype) {
    case DT_F64: return win2d_tbk_pick_vec<double, false, S, K>(vec);
    case DT_F32: return win2d_tbk_pick_vec<float, false, S, K>(vec);
    case DT_F32K: return win2d_tbk_pick_vec<float, true, S, K>(vec);
    }
    return nullptr;
}
template <int K> LaunchFn win2d_tbk_find(int prog, int dtype, int vec)
{
    switch (prog) {
    case Prog_2d1r_star::ID: return win2d_tbk_pick<Prog_2d1r_star, K>(dtype, vec);
    case Prog_2d2r_star::ID: return win2d_tbk_pick<Prog_2d2r_star, K>(dtype, vec);
    case Prog_2d3r_star::ID: return win2d_tbk_pick<Prog_2d3r_star, K>(dtype, vec);
    case Prog_2d1r_box::ID: return win2d_tbk_pick<Prog_2d1r_box, K>(dtype, vec);
    case Prog_2d2r_box::ID: return win2d_tbk_pick<Prog_2d2r_box, K>(dtype, vec);
    }
    return nullptr;
}

} // namespace
#endif
