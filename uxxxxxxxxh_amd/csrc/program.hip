// Fused multi-kernel programs (uhs_program_sweep_f64): the output kernel's
// terms read the input grid or a producer kernel evaluated at the shifted
// point, the form the reference's kernel fusion produces
// (UHStencil/lib/Dialect/Stencil/StencilKernelFusion.cpp:66-180: a producer
// with a single consumer is cloned into the consumer at every access offset,
// with no intermediate grid).
//
// Two kernels, bitwise the same result (a producer's value at a cell is the
// same sum in the same order whichever access asks for it):
//   k_program_stream  a tile of the plane per workgroup, streamed plane by
//                     plane along a slab: each input plane is staged in LDS
//                     once (the next one prefetched into registers), each
//                     producer is evaluated once per cell of the planes it is
//                     read at, into LDS rings, and the output kernel reads
//                     only LDS.
//   k_program         one thread per written cell, producers re-evaluated at
//                     every access, grid read through the caches: the general
//                     form for frames beyond the LDS.
// The term tables travel in the kernel arguments (uniform: scalar loads).
#include "uhs.h"
#include "uhs_host.h"
#include "uhs_tune.h"
#include "uhs_internal.h"

#include <algorithm>
#include <array>
#include <mutex>
#include <unordered_map>
#include <vector>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <string>

namespace {

constexpr int kMaxTerms = UHS_MAX_PROGRAM_TERMS, kMaxProd = UHS_MAX_PRODUCERS;
constexpr int kTileThreads = 256;

struct ProgParams {
    const void *in;                // grids of T (double, or float for the FP32 forms)
    void *out;
    long long lo0, lo1, lo2;       // first written cell (halo-inclusive coordinates)
    long long n0, n1, n2;          // written extents
    long long sy, sz;              // row and plane strides
    int nout;                      // output terms
    int nprod;
    int pstart[kMaxProd + 1];      // producer k's terms: [pstart[k], pstart[k+1])
    int osrc[kMaxTerms];           // -1 input, else producer index
    int ooff[kMaxTerms];           // linear offsets of the output terms
    double ocoef[kMaxTerms];
    int poff[kMaxTerms * kMaxProd];
    double pcoef[kMaxTerms * kMaxProd];
    signed char psrc[kMaxTerms * kMaxProd]; // producer terms: -1 input, else an earlier producer
};

// Producer k at linear cell q, its sources evaluated recursively.  D bounds
// the depth at compile time: producer k reads only producers below k, so a
// chain from producer kMaxProd - 1 ends at producer 0, which reads only the
// input.
template <typename T, bool KAHAN, int D>
__device__ T eval_producer(const ProgParams &p, int k, long long q)
{
    const T *in = static_cast<const T *>(p.in);
    const int b = p.pstart[k], e = p.pstart[k + 1];
    Acc<T, KAHAN> s;
    for (int t = b; t < e; t++) {
        const long long r = q + p.poff[t];
        T v;
        if constexpr (D > 0) {
            const int src = p.psrc[t];
            v = src < 0 ? in[r] : eval_producer<T, KAHAN, D - 1>(p, src, r);
        } else {
            v = in[r];
        }
        if (t == b) s.first((T)p.pcoef[t], v);
        else s.add((T)p.pcoef[t], v);
    }
    return s.value();
}

// DEPTH: producer nesting the launch evaluates (0: producers read only the
// input grid).
template <typename T, bool KAHAN, int DEPTH>
__global__ void __launch_bounds__(256) k_program(ProgParams p)
{
    const T *in = static_cast<const T *>(p.in);
    T *out = static_cast<T *>(p.out);
    const long long total = p.n0 * p.n1 * p.n2;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long plane = p.n1 * p.n2;
        const long long z = i / plane, r = i - z * plane;
        const long long y = r / p.n2, x = r - y * p.n2;
        const long long c = (z + p.lo0) * p.sz + (y + p.lo1) * p.sy + (x + p.lo2);
        Acc<T, KAHAN> acc;
        for (int j = 0; j < p.nout; j++) {
            const long long q = c + p.ooff[j];
            const int k = p.osrc[j];
            const T v = k < 0 ? in[q] : eval_producer<T, KAHAN, DEPTH>(p, k, q);
            if (j == 0) acc.first((T)p.ocoef[0], v);
            else acc.add((T)p.ocoef[j], v);
        }
        out[c] = acc.value();
    }
}

// The streaming kernel.  The grid is seen as (a, b, c): a the streamed axis
// (axis 0 in 3D, the rows in 2D), b the other outer axis (axis 1 in 3D, a
// single index in 2D), c the contiguous axis.  A workgroup owns a TB x TC
// tile of (b, c) and a slab of TA planes along a, which it walks one plane
// per step.  In-plane, each plane of its frame (tile plus reach) is Hb x Hc
// cells, Hc a multiple of 64; every staged plane has that layout, so a
// term's in-plane LDS offset is the constant db * Hc + dc.  Planes live in
// LDS rings, exactly as long as they need to be:
//   input ring    the planes one step reads, span_a + 1 (each producer
//                 reads its newest plane's inputs, the output kernel its own
//                 input terms: offsets [nlo_a, nlo_a + span_a] from the
//                 written plane), plus the one the next step stores before
//                 its first barrier;
//   producer k    cr_k + 1 planes of its values.
// Ring index of a plane = its distance from the slab's first plane of that
// ring; a term at distance d reads ring index s + d at step s.  Each lane
// keeps its term's slot and advances it by one per step.
struct StreamParams {
    const void *in;                // grids of T (double, or float for the FP32 forms)
    void *out;
    long long base;                // grid index of the domain's first cell
    long long sa, sb;              // strides of axes a and b
    int na, nb, nc;                // domain extents
    int TA, TB, TC;                // slab planes, tile rows and columns
    int tiles_c, tiles_b, ntiles;
    int Hc, lnch;                  // frame row length = 64 << lnch
    int plane;                     // cells per staged plane (Hb * Hc)
    int flo_b, flo_c;              // frame in-plane origin relative to the tile (<= 0)
    int ilo_a, ilo_b, ilo_c;       // input reach lower offsets
    int ir_a, ir_b, ir_c;          // input reach widths
    int nlo_a, span_a;             // input planes one step reads: [nlo_a, nlo_a + span_a]
    int NI;                        // input ring length
    int lead;                      // planes before the slab's first written plane
    int lds_cells;                 // LDS cells the rings span
    int nprod, nout;
    int pring[kMaxProd];           // producer k's ring: first LDS cell, length
    int pN[kMaxProd];
    int clo_b[kMaxProd], clo_c[kMaxProd];   // producer k is read at [clo, clo + cr]
    int cr_a[kMaxProd], cr_b[kMaxProd], cr_c[kMaxProd];
    int pstart[kMaxProd + 1];
    int pbar;                      // bit k: a barrier before producer k (it reads a producer
                                   // computed since the last barrier)
    int pd[kMaxTerms * kMaxProd];  // producer term: distance in its source's ring,
    int psN[kMaxTerms * kMaxProd]; // that ring's length, its first cell plus the
    int poff[kMaxTerms * kMaxProd]; // in-plane offset, and the coefficient
    double pcoef[kMaxTerms * kMaxProd];
    int oring[kMaxTerms];          // output term: its ring's first cell, length, distance, offset
    int oN[kMaxTerms];
    int od[kMaxTerms];
    int ooff[kMaxTerms];
    double ocoef[kMaxTerms];
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Value of lane l (uniform l) of a per-lane table.
__device__ __forceinline__ int lane_int(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double lane_dbl(double v, int l)
{
    const long long b = __builtin_bit_cast(long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Left-to-right sum of terms [b, e) at the cells li[] (s already holds the
// sum up to term b - 1): fetch(q, off, cf) gives term q's LDS offset and
// coefficient.  Terms go four at a time, their LDS reads issued before the
// first fused multiply-add; the order of the additions is unchanged.  (The
// term loop reads its tables with readlane, which the compiler does not
// unroll.)
template <typename T, bool KAHAN, int U, class Fetch>
__device__ __forceinline__ void accumulate(const T *lds, const int (&li)[U], int b, int e,
                                           Acc<T, KAHAN> (&s)[U], Fetch &&fetch)
{
    int q = b;
    for (; q + 4 <= e; q += 4) {
        int off[4];
        T cf[4], v[4][U];
#pragma unroll
        for (int g = 0; g < 4; g++) fetch(q + g, off[g], cf[g]);
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int u = 0; u < U; u++) v[g][u] = lds[li[u] + off[g]];
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int u = 0; u < U; u++) s[u].add(cf[g], v[g][u]);
    }
    for (; q < e; q++) {
        int off;
        T cf;
        fetch(q, off, cf);
#pragma unroll
        for (int u = 0; u < U; u++) s[u].add(cf, lds[li[u] + off]);
    }
}

// Items: the staged plane's Hb rows of 64-cell chunks, item i = row
// i >> lnch, chunk i & (2^lnch - 1), dealt to the 4 waves (i = wave + 4u).
// US: staged items per wave (input plane); UC: items per wave of the
// compute phases, whose rows are the tile's plus the producers' reach
// (fewer than the staged rows).  PF: planes of loads in flight.  Items
// outside a phase's region are clamped onto its last cell (loads stay in
// bounds) and not written.
template <typename T, bool KAHAN, int US, int UC, int PF>
__global__ void __launch_bounds__(kTileThreads) k_program_stream(StreamParams p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    lds_poison(lds, p.lds_cells);
    constexpr int waves = kTileThreads / 64;
    const int lane = threadIdx.x & 63, wave = uniform(threadIdx.x >> 6);
    const int t = xcd_remap(blockIdx.x, p.ntiles);
    const int tc = t % p.tiles_c, tb = (t / p.tiles_c) % p.tiles_b, ta = t / (p.tiles_c * p.tiles_b);
    const int c0 = tc * p.TC, b0 = tb * p.TB, a0 = ta * p.TA;
    const int ec = min(p.TC, p.nc - c0), eb = min(p.TB, p.nb - b0), a_end = min(a0 + p.TA, p.na);
    const int cmask = (1 << p.lnch) - 1;
    // Item u of this lane clamped onto a region of rows x cols; true for a
    // real cell.
    auto cell = [&](int u, int rows, int cols, int &r, int &c) {
        const int i = wave + u * waves;
        const int rr = i >> p.lnch, cc = ((i & cmask) << 6) + lane;
        r = min(rr, rows - 1);
        c = min(cc, cols - 1);
        return rr < rows && cc < cols;
    };

    // Term tables in lanes: lane l holds output term l, producer term l of
    // each of two tables, and producer l's ring slot and region.  Each step
    // advances the slots by one with one vector operation per table; the
    // term loops read them with readlane (no scalar arithmetic or scalar
    // loads per term).
    const int ol = min(lane, p.nout - 1);
    const int o_N = p.oN[ol], o_fix = p.oring[ol] + p.ooff[ol];
    int o_slot = p.od[ol];
    const double o_c = p.ocoef[ol];
    // Producer tables: table 0 holds producers 0 and 1 (at most 64 terms),
    // table 1 producers 2 and 3, each from its first producer's first term.
    const int nptot = p.pstart[p.nprod];
    const int t1 = p.pstart[min(2, p.nprod)];
    const int pl0 = min(lane, max(t1 - 1, 0)), pl1 = min(t1 + lane, max(nptot - 1, 0));
    const int p_off0 = p.poff[pl0], p_off1 = p.poff[pl1];
    const int p_N0 = p.psN[pl0], p_N1 = p.psN[pl1];
    int p_slot0 = p.pd[pl0], p_slot1 = p.pd[pl1];
    const double p_c0 = p.pcoef[pl0], p_c1 = p.pcoef[pl1];
    // Lane k: producer k's ring slot and length, and its region.
    const int kl = min(lane, kMaxProd - 1);
    const int k_N = lane < p.nprod ? p.pN[kl] : 1;
    int k_slot = lane < p.nprod ? p.cr_a[kl] : 0;
    const int k_rows = p.cr_b[kl], k_cols = p.cr_c[kl], k_ring = p.pring[kl];
    const int k_li = (p.clo_b[kl] - p.flo_b) * p.Hc + (p.clo_c[kl] - p.flo_c);
    const int k_first = p.pstart[kl] - (kl < 2 ? 0 : t1), k_end = p.pstart[kl + 1] - (kl < 2 ? 0 : t1);
    // Wait for the tables here, once: otherwise the first use inside the
    // plane loop waits on a count that also covers the planes in flight.
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_sched_barrier(0);

    // Input items: fixed for the whole slab.
    const int rin = eb + p.ir_b, cin = ec + p.ir_c;
    long long gin[US];
    int lin[US];
    bool okin[US];
#pragma unroll
    for (int u = 0; u < US; u++) {
        int r, c;
        okin[u] = cell(u, rin, cin, r, c);
        gin[u] = (long long)(b0 + p.ilo_b + r) * p.sb + (c0 + p.ilo_c + c);
        lin[u] = (r + p.ilo_b - p.flo_b) * p.Hc + (c + p.ilo_c - p.flo_c);
    }
    const T *gbase = static_cast<const T *>(p.in) + p.base;
    // Nominal plane q (domain-relative) is read from plane clamp(q): planes
    // outside [qmin, qmax] are staged but never read.
    const int qmin = a0 + p.ilo_a, qmax = a_end - 1 + p.ilo_a + p.ir_a;
    auto load_plane = [&](int q, T (&v)[US]) {
        const T *g = gbase + (long long)min(max(q, qmin), qmax) * p.sa;
#pragma unroll
        for (int u = 0; u < US; u++) v[u] = g[gin[u]];
    };
    auto store_plane = [&](int slot, const T (&v)[US]) {
        T *s = lds + slot * p.plane;
#pragma unroll
        for (int u = 0; u < US; u++)
            if (okin[u]) s[lin[u]] = v[u];
    };

    const int A = a0 - p.lead;     // first step: planes [A + nlo_a, ...] staged
    const int nhi = p.nlo_a + p.span_a;
    for (int i = 0; i < p.span_a; i++) {
        T v[US];
        load_plane(A + p.nlo_a + i, v);
        store_plane(i, v);
    }
    // PF planes in flight in a register ring: step s stores ring slot s % PF
    // (the plane of ring index s + span_a) and refills it with the plane PF
    // steps ahead.  The loop is unrolled by PF so slots are registers; the
    // refill loads are unconditional (clamped planes) and never copied.
    T pf[PF][US];
    // Issued in slot order (no reordering across slots), so the loop entry
    // holds the same loads in flight as the loop's back edge.
    static_for<PF>([&](auto f) {
        load_plane(A + nhi + f, pf[f]);
        __builtin_amdgcn_sched_barrier(0);
    });
    int in_slot = p.span_a;        // slot of the plane stored at the current step
    auto step = [](int v, int n) { return v + 1 == n ? 0 : v + 1; };
    for (int i0 = 0; A + i0 < a_end; i0 += PF)
        static_for<PF>([&](auto ph) {
        const int a = A + i0 + ph;
        // This step's slots; then advance every ring by one plane.
        const int sin = in_slot, so = o_slot, sp0 = p_slot0, sp1 = p_slot1, sk = k_slot;
        in_slot = step(in_slot, p.NI);
        o_slot = step(o_slot, o_N);
        p_slot0 = step(p_slot0, p_N0);
        p_slot1 = step(p_slot1, p_N1);
        k_slot = step(k_slot, k_N);
        if (a < a_end) store_plane(sin, pf[ph]);
        load_plane(a + PF + nhi, pf[ph]);
        if (a >= a_end) return;
        __syncthreads();
        // Producer k's plane a + chi_a, in dependency order; its terms read the
        // input ring or an earlier producer's ring.
        const int pb0 = sp0 * p.plane + p_off0;
        const int pb1 = sp1 * p.plane + p_off1;
        for (int k = 0; k < p.nprod; k++) {
            if ((p.pbar >> k) & 1) __syncthreads();
            const int rows = eb + lane_int(k_rows, k), cols = ec + lane_int(k_cols, k);
            if (rows <= 0) continue;   // a producer the output kernel never needs
            T *dst = lds + lane_int(k_ring, k) + lane_int(sk, k) * p.plane;
            const int b = lane_int(k_first, k), e = lane_int(k_end, k), lk = lane_int(k_li, k);
            // This producer's table (a per-lane select once per producer).
            const int pb = k < 2 ? pb0 : pb1;
            const double pc = k < 2 ? p_c0 : p_c1;
            int li[UC];
            bool ok[UC];
            Acc<T, KAHAN> s[UC];
#pragma unroll
            for (int u = 0; u < UC; u++) {
                int r, c;
                ok[u] = cell(u, rows, cols, r, c);
                li[u] = r * p.Hc + c + lk;
            }
            auto term = [&](int q, int &off, T &cf) {
                off = lane_int(pb, q);
                cf = (T)lane_dbl(pc, q);
            };
            {
                int off;
                T cf;
                term(b, off, cf);
#pragma unroll
                for (int u = 0; u < UC; u++) s[u].first(cf, lds[li[u] + off]);
            }
            accumulate<T, KAHAN, UC>(lds, li, b + 1, e, s, term);
#pragma unroll
            for (int u = 0; u < UC; u++)
                if (ok[u]) dst[li[u]] = s[u].value();
        }
        if (p.nprod) __syncthreads();
        if (a < a0) return;
        // The output kernel on plane a.
        {
            int li[UC];
            long long go[UC];
            bool ok[UC];
            Acc<T, KAHAN> acc[UC];
#pragma unroll
            for (int u = 0; u < UC; u++) {
                int r, c;
                ok[u] = cell(u, eb, ec, r, c);
                li[u] = (r - p.flo_b) * p.Hc + (c - p.flo_c);
                go[u] = (long long)a * p.sa + (long long)(b0 + r) * p.sb + (c0 + c);
            }
            const int ob = o_fix + so * p.plane;
            auto term = [&](int j, int &off, T &cf) {
                off = lane_int(ob, j);
                cf = (T)lane_dbl(o_c, j);
            };
            {
                int off;
                T cf;
                term(0, off, cf);
#pragma unroll
                for (int u = 0; u < UC; u++) acc[u].first(cf, lds[li[u] + off]);
            }
            accumulate<T, KAHAN, UC>(lds, li, 1, p.nout, acc, term);
            T *out = static_cast<T *>(p.out) + p.base;
#pragma unroll
            for (int u = 0; u < UC; u++)
                if (ok[u]) out[go[u]] = acc[u].value();
        }
        });
}

int perr(int code, const std::string &m) { return uhs_set_error(code, m.c_str()); }

// Offsets as (axis-0, axis-1, axis-2) triples, 2D padded with a zero axis 0.
void off3(const int *offs, int d, int t, long long v[3])
{
    v[0] = d == 3 ? offs[t * 3] : 0;
    v[1] = offs[t * d + d - 2];
    v[2] = offs[t * d + d - 1];
}

struct Reach {
    long long lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool any = false;
    void add(const long long v[3])
    {
        for (int a = 0; a < 3; a++) {
            lo[a] = any ? std::min(lo[a], v[a]) : v[a];
            hi[a] = any ? std::max(hi[a], v[a]) : v[a];
        }
        any = true;
    }
};

// Source of producer k's term t: -1 the input grid, else an earlier producer.
int psrc(const uhs_program *pg, int k, int t)
{
    return pg->producer_sources ? pg->producer_sources[k * kMaxTerms + t] : -1;
}

// Everything both kernels need, validated.
struct Checked {
    long long shp[3], lo[3], hi[3], sy, sz;
    Reach in;                      // every input read, relative to the written cell
    Reach cons[kMaxProd];          // offsets (from the written cell) at which producer k is needed
    Reach pin[kMaxProd];           // input offsets producer k reads, relative to its own cell
};

int check_program(const int64_t *shape, const uhs_program *pg, Checked &c)
{
    const int d = pg->dims;
    if (d != 2 && d != 3) return perr(UHS_E_INVAL, "dims must be 2 or 3");
    if (pg->nproducers < 0 || pg->nproducers > kMaxProd || (pg->nproducers && !pg->producers))
        return perr(UHS_E_INVAL, "0 to 4 producers");
    const uhs_stencil &o = pg->output;
    if (o.dims != d || o.npoints < 1 || o.npoints > kMaxTerms || !o.offsets || !o.coeffs || !pg->sources)
        return perr(UHS_E_INVAL, "the output kernel needs 1 to 32 terms, offsets, coefficients and sources");
    for (int k = 0; k < 3; k++) { c.shp[k] = 1; c.lo[k] = 0; c.hi[k] = 1; }
    for (int k = 0; k < d; k++) {
        const int a = 3 - d + k;
        if (shape[k] < 1 || shape[k] > (1LL << 30)) return perr(UHS_E_INVAL, "grid shape out of range");
        c.shp[a] = shape[k];
        c.lo[a] = pg->lo[k];
        c.hi[a] = pg->hi[k];
        if (c.lo[a] < 0 || c.hi[a] > c.shp[a] || c.lo[a] >= c.hi[a])
            return perr(UHS_E_INVAL, "domain outside the grid or empty");
    }
    c.sy = c.shp[2];
    c.sz = c.shp[1] * c.shp[2];
    if (c.sz * c.shp[0] > (1LL << 40)) return perr(UHS_E_INVAL, "grid too large");
    // Producers in dependency order; each one's input reach.
    for (int k = 0; k < pg->nproducers; k++) {
        const uhs_stencil &s = pg->producers[k];
        if (s.dims != d || s.npoints < 1 || s.npoints > kMaxTerms || !s.offsets || !s.coeffs)
            return perr(UHS_E_INVAL, "each producer needs 1 to 32 terms, offsets and coefficients");
        for (int t = 0; t < s.npoints; t++) {
            const int src = psrc(pg, k, t);
            if (src < -1 || src >= k)
                return perr(UHS_E_INVAL, "a producer term reads itself or a later producer");
            long long w[3];
            off3(s.offsets, d, t, w);
            if (src < 0) {
                c.pin[k].add(w);
            } else {
                long long lo[3], hi[3];
                for (int a = 0; a < 3; a++) { lo[a] = w[a] + c.pin[src].lo[a]; hi[a] = w[a] + c.pin[src].hi[a]; }
                c.pin[k].add(lo);
                c.pin[k].add(hi);
            }
        }
    }
    for (int j = 0; j < o.npoints; j++) {
        const int k = pg->sources[j];
        if (k < -1 || k >= pg->nproducers) return perr(UHS_E_INVAL, "source index out of range");
        long long v[3];
        off3(o.offsets, d, j, v);
        Reach r;
        if (k < 0) {
            r.add(v);
        } else {
            c.cons[k].add(v);
            long long lo[3], hi[3];
            for (int a = 0; a < 3; a++) { lo[a] = v[a] + c.pin[k].lo[a]; hi[a] = v[a] + c.pin[k].hi[a]; }
            r.add(lo);
            r.add(hi);
        }
        for (int a = 0; a < 3; a++)
            if (c.lo[a] + r.lo[a] < 0 || c.hi[a] - 1 + r.hi[a] >= c.shp[a]) {
                char b[160];
                snprintf(b, sizeof b, "output term %d reads outside the grid for cells of the domain", j);
                return perr(UHS_E_INVAL, b);
            }
        c.in.add(r.lo);
        c.in.add(r.hi);
    }
    // Where each producer is needed: by the output kernel, and by later
    // producers where they are needed (consumers first).
    for (int k = pg->nproducers - 1; k >= 0; k--) {
        if (!c.cons[k].any) continue;
        const uhs_stencil &s = pg->producers[k];
        for (int t = 0; t < s.npoints; t++) {
            const int src = psrc(pg, k, t);
            if (src < 0) continue;
            long long w[3], lo[3], hi[3];
            off3(s.offsets, d, t, w);
            for (int a = 0; a < 3; a++) { lo[a] = c.cons[k].lo[a] + w[a]; hi[a] = c.cons[k].hi[a] + w[a]; }
            c.cons[src].add(lo);
            c.cons[src].add(hi);
        }
    }
    return UHS_OK;
}

int build_pointwise(const uhs_program *pg, const Checked &c, ProgParams &p)
{
    const int d = pg->dims;
    const uhs_stencil &o = pg->output;
    p = ProgParams{};
    p.lo0 = c.lo[0]; p.lo1 = c.lo[1]; p.lo2 = c.lo[2];
    p.n0 = c.hi[0] - c.lo[0]; p.n1 = c.hi[1] - c.lo[1]; p.n2 = c.hi[2] - c.lo[2];
    p.sy = c.sy; p.sz = c.sz;
    p.nout = o.npoints;
    p.nprod = pg->nproducers;
    auto lin = [&](const long long v[3], int *out) {
        const long long l = v[0] * c.sz + v[1] * c.sy + v[2];
        if (l > INT32_MAX || l < -INT32_MAX) return false;
        *out = (int)l;
        return true;
    };
    int nterms = 0;
    for (int k = 0; k < p.nprod; k++) {
        const uhs_stencil &s = pg->producers[k];
        p.pstart[k] = nterms;
        for (int t = 0; t < s.npoints; t++, nterms++) {
            long long v[3];
            off3(s.offsets, d, t, v);
            if (!lin(v, &p.poff[nterms])) return perr(UHS_E_INVAL, "term offset exceeds 2^31 cells");
            p.pcoef[nterms] = s.coeffs[t];
            p.psrc[nterms] = (signed char)psrc(pg, k, t);
        }
    }
    p.pstart[p.nprod] = nterms;
    for (int j = 0; j < o.npoints; j++) {
        long long v[3];
        off3(o.offsets, d, j, v);
        p.osrc[j] = pg->sources[j];
        if (!lin(v, &p.ooff[j])) return perr(UHS_E_INVAL, "term offset exceeds 2^31 cells");
        p.ocoef[j] = o.coeffs[j];
    }
    return UHS_OK;
}

// Largest staged-plane rows x chunks per wave a launch may need.
constexpr int kMaxItemsPerWave = 8;

// Builds the streaming kernel's parameters for tile {TC, TB, TA} (3D) or
// {TC, TA, 1} (2D).  Returns the LDS bytes of one workgroup, SIZE_MAX when
// the frame cannot fit; *us, *uc = staged and compute items per wave.
size_t build_stream(const uhs_program *pg, const Checked &c, const int tile[3], size_t esz, StreamParams &p, int *us,
                    int *uc)
{
    const int d = pg->dims;
    const uhs_stencil &o = pg->output;
    // Axis roles in the padded 3D view (axis 0 is a unit axis in 2D).
    const int ax_a = d == 3 ? 0 : 1, ax_b = d == 3 ? 1 : 0, ax_c = 2;
    const long long stride[3] = {c.sz, c.sy, 1};
    p = StreamParams{};
    p.base = c.lo[0] * c.sz + c.lo[1] * c.sy + c.lo[2];
    p.sa = stride[ax_a];
    p.sb = stride[ax_b];
    p.na = (int)(c.hi[ax_a] - c.lo[ax_a]);
    p.nb = (int)(c.hi[ax_b] - c.lo[ax_b]);
    p.nc = (int)(c.hi[ax_c] - c.lo[ax_c]);
    p.TC = tile[0];
    p.TB = d == 3 ? tile[1] : 1;
    p.TA = d == 3 ? tile[2] : tile[1];
    p.tiles_c = (p.nc + p.TC - 1) / p.TC;
    p.tiles_b = (p.nb + p.TB - 1) / p.TB;
    const long long ntiles = (long long)p.tiles_c * p.tiles_b * ((p.na + p.TA - 1) / p.TA);
    if (ntiles > INT32_MAX) return SIZE_MAX;
    p.ntiles = (int)ntiles;
    // In-plane frame: the tile (offset 0), every input read and every
    // producer read.
    long long flo[3] = {0, 0, 0}, fhi[3] = {0, 0, 0};
    auto widen = [&](const long long *lo, const long long *hi) {
        for (int a = 0; a < 3; a++) {
            flo[a] = std::min(flo[a], lo[a]);
            fhi[a] = std::max(fhi[a], hi[a]);
        }
    };
    widen(c.in.lo, c.in.hi);
    int maxcr = 0;
    for (int k = 0; k < pg->nproducers; k++)
        if (c.cons[k].any) {
            widen(c.cons[k].lo, c.cons[k].hi);
            maxcr = (int)std::max<long long>(maxcr, c.cons[k].hi[ax_a] - c.cons[k].lo[ax_a]);
        }
    const long long kCells = (long long)(kMaxLdsBytes / esz);
    const long long hb = p.TB + fhi[ax_b] - flo[ax_b], wc = p.TC + fhi[ax_c] - flo[ax_c];
    const long long ira = c.in.hi[ax_a] - c.in.lo[ax_a];
    if (hb > kCells || wc > kCells || ira > kCells) return SIZE_MAX;
    int lnch = 0;
    while ((64LL << lnch) < wc) lnch++;
    const long long hc = 64LL << lnch;
    if (hb * hc > kCells) return SIZE_MAX;
    *us = (int)(((hb << lnch) + 3) / 4);
    if (*us > kMaxItemsPerWave) return SIZE_MAX;
    // Compute rows: the tile's, plus the widest producer reach along b.
    long long crb = 0;
    for (int k = 0; k < pg->nproducers; k++)
        if (c.cons[k].any) crb = std::max(crb, c.cons[k].hi[ax_b] - c.cons[k].lo[ax_b]);
    *uc = (int)((((p.TB + crb) << lnch) + 3) / 4);
    p.Hc = (int)hc;
    p.lnch = lnch;
    p.plane = (int)(hb * hc);
    p.flo_b = (int)flo[ax_b];
    p.flo_c = (int)flo[ax_c];
    p.ilo_a = (int)c.in.lo[ax_a];
    p.ilo_b = (int)c.in.lo[ax_b];
    p.ilo_c = (int)c.in.lo[ax_c];
    p.ir_a = (int)ira;
    p.ir_b = (int)(c.in.hi[ax_b] - c.in.lo[ax_b]);
    p.ir_c = (int)(c.in.hi[ax_c] - c.in.lo[ax_c]);
    // Input planes one step reads (offsets from the written plane a): the
    // producers' newest planes a + chi_k read a + chi_k + w; the output
    // kernel's input terms a + v.
    long long nlo = LLONG_MAX, nhi = LLONG_MIN;
    for (int k = 0; k < pg->nproducers; k++) {
        if (!c.cons[k].any) continue;
        const uhs_stencil &s = pg->producers[k];
        for (int t = 0; t < s.npoints; t++) {
            if (psrc(pg, k, t) >= 0) continue;
            long long w[3];
            off3(s.offsets, d, t, w);
            nlo = std::min(nlo, c.cons[k].hi[ax_a] + w[ax_a]);
            nhi = std::max(nhi, c.cons[k].hi[ax_a] + w[ax_a]);
        }
    }
    for (int j = 0; j < o.npoints; j++)
        if (pg->sources[j] < 0) {
            long long v[3];
            off3(o.offsets, d, j, v);
            nlo = std::min(nlo, v[ax_a]);
            nhi = std::max(nhi, v[ax_a]);
        }
    p.nlo_a = (int)nlo;
    p.span_a = (int)(nhi - nlo);
    const int ni = p.span_a + 2;
    p.NI = ni;
    p.lead = maxcr;
    p.nprod = pg->nproducers;
    p.nout = o.npoints;
    long long cells = (long long)ni * p.plane;
    for (int k = 0; k < p.nprod; k++) {
        const Reach &r = c.cons[k];
        const int cra = r.any ? (int)(r.hi[ax_a] - r.lo[ax_a]) : 0;
        const int np = cra + 1;
        if (cells + (long long)np * p.plane > kCells) return SIZE_MAX;
        p.pring[k] = (int)cells;
        p.pN[k] = np;
        cells += (long long)np * p.plane;
        p.cr_a[k] = cra;
        // A producer nothing reads: an empty region (no rows).
        p.clo_b[k] = r.any ? (int)r.lo[ax_b] : 0;
        p.clo_c[k] = r.any ? (int)r.lo[ax_c] : 0;
        p.cr_b[k] = r.any ? (int)(r.hi[ax_b] - r.lo[ax_b]) : -p.TB;
        p.cr_c[k] = r.any ? (int)(r.hi[ax_c] - r.lo[ax_c]) : 0;
    }
    int nterms = 0;
    for (int k = 0; k < p.nprod; k++) {
        const uhs_stencil &s = pg->producers[k];
        const Reach &r = c.cons[k];
        const long long chi_a = r.any ? r.hi[ax_a] : 0;
        p.pstart[k] = nterms;
        for (int t = 0; t < s.npoints; t++, nterms++) {
            long long w[3];
            off3(s.offsets, d, t, w);
            // Plane a + chi_a + w_a of its source: ring distance from the step
            // (the input ring starts at plane A + nlo_a, producer j's at
            // A + clo_j).
            const int src = psrc(pg, k, t);
            if (r.any) {
                // A needed producer's sources are needed too (their reach covers it).
                const long long base = src < 0 ? nlo : c.cons[src].lo[ax_a];
                p.pd[nterms] = (int)(chi_a + w[ax_a] - base);
            } else {
                p.pd[nterms] = 0;
            }
            p.psN[nterms] = src < 0 ? p.NI : p.pN[src];
            p.poff[nterms] = (src < 0 ? 0 : p.pring[src]) + (int)(w[ax_b] * p.Hc + w[ax_c]);
            p.pcoef[nterms] = s.coeffs[t];
        }
    }
    p.pstart[p.nprod] = nterms;
    p.lds_cells = (int)cells;
    // A barrier before a producer that reads one computed since the last one.
    p.pbar = 0;
    for (int k = 0, first = 0; k < p.nprod; k++) {
        bool reads_recent = false;
        for (int t = 0; t < pg->producers[k].npoints; t++) reads_recent |= psrc(pg, k, t) >= first;
        if (reads_recent) {
            p.pbar |= 1 << k;
            first = k;
        }
    }
    for (int j = 0; j < o.npoints; j++) {
        long long v[3];
        off3(o.offsets, d, j, v);
        const int k = pg->sources[j];
        if (k < 0) {
            p.oring[j] = 0;
            p.oN[j] = p.NI;
            p.od[j] = (int)(v[ax_a] - nlo);
        } else {
            p.oring[j] = p.pring[k];
            p.oN[j] = p.pN[k];
            p.od[j] = (int)(v[ax_a] - c.cons[k].lo[ax_a]);
        }
        p.ooff[j] = (int)(v[ax_b] * p.Hc + v[ax_c]);
        p.ocoef[j] = o.coeffs[j];
    }
    return (size_t)cells * esz;
}

// Default tiles, first fit: rows of 64 << n cells once the reach is added;
// 2D 32-row slabs of 1024-cell rows, 3D 6 rows of 128 cells x 8-plane slabs
// (the fastest of a sweep over both reference example programs), then
// smaller tiles for programs whose rings need more LDS.
int default_tiles(const Checked &c, int d, int tiles[][3])
{
    const long long rc = c.in.hi[2] - c.in.lo[2];
    int n = 0;
    if (d == 2) {
        // 512-cell rows first: the reference's nested example times the same
        // with 512 x 16 as with 1024 x 32 tiles, its one-sided example 11%
        // faster.
        for (int w : {512, 1024, 256})
            if (w - rc >= 64) { tiles[n][0] = (int)(w - rc); tiles[n][1] = w == 1024 ? 32 : 16; tiles[n][2] = 1; n++; }
    } else {
        // Narrower rows only when the wider ones leave no room in the LDS.
        for (int w : {128, 64})
            if (w - rc >= 32)
                for (int tb : {6, 4, 2, 1}) { tiles[n][0] = (int)(w - rc); tiles[n][1] = tb; tiles[n][2] = 8; n++; }
    }
    return n;
}

typedef void (*LaunchProgram)(StreamParams);

// Instantiations: staged items per wave 1, 2, 4 or 8, compute items 1 to 4
// or 8 (never more than staged); planes in flight more where a plane is few
// registers.
template <typename T, bool KAHAN>
LaunchProgram pick_stream(int us, int uc)
{
    if (us <= 1) return k_program_stream<T, KAHAN, 1, 1, 4>;
    if (us <= 2) return uc <= 1 ? k_program_stream<T, KAHAN, 2, 1, 4> : k_program_stream<T, KAHAN, 2, 2, 4>;
    if (us <= 4) {
        switch (uc) {
        case 1: return k_program_stream<T, KAHAN, 4, 1, 2>;
        case 2: return k_program_stream<T, KAHAN, 4, 2, 2>;
        case 3: return k_program_stream<T, KAHAN, 4, 3, 2>;
        default: return k_program_stream<T, KAHAN, 4, 4, 2>;
        }
    }
    switch (uc) {
    case 1: return k_program_stream<T, KAHAN, 8, 1, 1>;
    case 2: return k_program_stream<T, KAHAN, 8, 2, 1>;
    case 3: return k_program_stream<T, KAHAN, 8, 3, 1>;
    case 4: return k_program_stream<T, KAHAN, 8, 4, 1>;
    default: return k_program_stream<T, KAHAN, 8, 8, 1>;
    }
}

// Tiles uhs_tune_program_f64 picked in this process, by program, grid shape
// and element size: the tile an untiled call uses.
std::mutex g_tuned_mu;
std::unordered_map<std::string, std::array<int, 3>> g_tuned;

std::string program_key(const int64_t *shape, const uhs_program *pg, size_t esz)
{
    std::string k;
    auto put = [&](const void *p, size_t n) { k.append(static_cast<const char *>(p), n); };
    const int d = pg->dims;
    put(&esz, sizeof esz);
    put(&d, sizeof d);
    put(shape, sizeof(int64_t) * d);
    put(pg->lo, sizeof(int64_t) * d);
    put(pg->hi, sizeof(int64_t) * d);
    put(&pg->nproducers, sizeof(int));
    auto stencil = [&](const uhs_stencil &s) {
        put(&s.npoints, sizeof(int));
        put(s.offsets, sizeof(int) * s.npoints * d);
        put(s.coeffs, sizeof(double) * s.npoints);
    };
    for (int k2 = 0; k2 < pg->nproducers; k2++) {
        stencil(pg->producers[k2]);
        for (int t = 0; t < pg->producers[k2].npoints; t++) {
            const int s = psrc(pg, k2, t);
            put(&s, sizeof s);
        }
    }
    stencil(pg->output);
    put(pg->sources, sizeof(int) * pg->output.npoints);
    return k;
}

// T sweeps of a program on grids of T (KAHAN: compensated sums).
template <typename T, bool KAHAN>
int program_sweep(void *A, void *B, const int64_t *shape, const uhs_program *pg, int timesteps, int kernel,
                  const int *tile, uhs_stream_t stream)
{
    if (!A || !B || !shape || !pg) return perr(UHS_E_INVAL, "null argument");
    if (A == B) return perr(UHS_E_INVAL, "A and B must be distinct grids");
    if (timesteps < 0) return perr(UHS_E_INVAL, "negative timestep count");
    if (kernel < UHS_PROGRAM_AUTO || kernel > UHS_PROGRAM_POINTWISE) return perr(UHS_E_INVAL, "unknown program kernel");
    Checked c;
    int rc = check_program(shape, pg, c);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;

    bool tiled = kernel != UHS_PROGRAM_POINTWISE;
    StreamParams sp;
    size_t lds = SIZE_MAX;
    int us = 0, uc = 0;
    if (tiled) {
        int cand[12][3], n = 0;   // a tuned tile, then at most 8 defaults
        if (tile) {
            for (int a = 0; a < 3; a++) cand[0][a] = tile[a];
            if (tile[0] < 1 || tile[1] < 1 || tile[2] < 1 || tile[0] > 4096 || tile[1] > 4096 || tile[2] > 4096 ||
                (pg->dims == 2 && tile[2] != 1))
                return perr(UHS_E_CONFIG, "tile extents must be 1..4096 (tile_z 1 for 2D)");
            n = 1;
        } else {
            n = 0;
            {
                std::lock_guard<std::mutex> lk(g_tuned_mu);
                auto it = g_tuned.find(program_key(shape, pg, sizeof(T)));
                if (it != g_tuned.end()) {
                    for (int a = 0; a < 3; a++) cand[0][a] = it->second[a];
                    n = 1;
                }
            }
            n += default_tiles(c, pg->dims, cand + n);
        }
        for (int i = 0; i < n; i++) {
            lds = build_stream(pg, c, cand[i], sizeof(T), sp, &us, &uc);
            if (lds <= kMaxLdsBytes) break;
        }
        if (lds > kMaxLdsBytes) {
            if (kernel == UHS_PROGRAM_TILED || tile) {
                char b[200];
                if (lds == SIZE_MAX)
                    snprintf(b, sizeof b, "the tiled kernel's frame does not fit in %zu bytes of LDS "
                             "(or a staged plane exceeds %d rows of 64 cells per wave)", kMaxLdsBytes,
                             kMaxItemsPerWave);
                else
                    snprintf(b, sizeof b, "the tiled kernel's frame needs %zu bytes of LDS (at most %zu)", lds,
                             kMaxLdsBytes);
                return perr(UHS_E_CONFIG, b);
            }
            tiled = false;
        }
    }
    if (tiled) {
        const LaunchProgram kern = pick_stream<T, KAHAN>(us, uc);
        for (int st = 1; st <= timesteps; st++) {
            sp.in = (st & 1) ? A : B;
            sp.out = (st & 1) ? B : A;
            hipLaunchKernelGGL(kern, dim3(sp.ntiles), dim3(kTileThreads), lds, s, sp);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return perr(UHS_E_HIP, std::string("program launch: ") + hipGetErrorString(e));
        }
        return UHS_OK;
    }
    ProgParams p;
    rc = build_pointwise(pg, c, p);
    if (rc) return rc;
    const long long total = p.n0 * p.n1 * p.n2;
    const unsigned blocks = (unsigned)std::min<long long>((total + 255) / 256, 65536);
    for (int st = 1; st <= timesteps; st++) {
        p.in = (st & 1) ? A : B;
        p.out = (st & 1) ? B : A;
        if (pg->producer_sources)
            hipLaunchKernelGGL((k_program<T, KAHAN, kMaxProd - 1>), dim3(blocks), dim3(256), 0, s, p);
        else
            hipLaunchKernelGGL((k_program<T, KAHAN, 0>), dim3(blocks), dim3(256), 0, s, p);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return perr(UHS_E_HIP, std::string("program launch: ") + hipGetErrorString(e));
    }
    return UHS_OK;
}

} // namespace

extern "C" int uhs_program_sweep_ex_f64(double *A, double *B, const int64_t *shape, const uhs_program *pg,
                                        int timesteps, int kernel, const int *tile, uhs_stream_t stream)
{
    return program_sweep<double, false>(A, B, shape, pg, timesteps, kernel, tile, stream);
}

extern "C" int uhs_program_sweep_ex_f32(float *A, float *B, const int64_t *shape, const uhs_program *pg,
                                        int timesteps, int compensated, int kernel, const int *tile,
                                        uhs_stream_t stream)
{
    return compensated ? program_sweep<float, true>(A, B, shape, pg, timesteps, kernel, tile, stream)
                       : program_sweep<float, false>(A, B, shape, pg, timesteps, kernel, tile, stream);
}

extern "C" int uhs_program_sweep_f32(float *A, float *B, const int64_t *shape, const uhs_program *pg,
                                     int timesteps, int compensated, uhs_stream_t stream)
{
    return uhs_program_sweep_ex_f32(A, B, shape, pg, timesteps, compensated, UHS_PROGRAM_AUTO, nullptr, stream);
}

extern "C" int uhs_program_sweep_f64(double *A, double *B, const int64_t *shape, const uhs_program *pg,
                                     int timesteps, uhs_stream_t stream)
{
    return uhs_program_sweep_ex_f64(A, B, shape, pg, timesteps, UHS_PROGRAM_AUTO, nullptr, stream);
}

extern "C" int uhs_tune_program_f64(double *A, double *B, const int64_t *shape, const uhs_program *pg,
                                    int reps, int *tile_out, double *best_ms, int *evaluated,
                                    uhs_stream_t stream)
{
    if (!A || !B || !shape || !pg || !tile_out) return perr(UHS_E_INVAL, "null argument");
    if (reps < 1) return perr(UHS_E_INVAL, "reps must be at least 1");
    Checked c;
    int rc = check_program(shape, pg, c);
    if (rc) return rc;
    // Candidates: rows of 64 << n cells once the reach is added, a range of
    // tile heights and slab lengths; only those whose rings fit are timed.
    const long long rc2 = c.in.hi[2] - c.in.lo[2];
    std::vector<std::array<int, 3>> cand;
    if (pg->dims == 2) {
        for (int w : {256, 512, 1024})
            if (w - rc2 >= 32)
                for (int ta : {8, 16, 32, 64}) cand.push_back({(int)(w - rc2), ta, 1});
    } else {
        for (int w : {64, 128, 256})
            if (w - rc2 >= 16)
                for (int tb : {1, 2, 4, 6, 8, 12})
                    for (int ta : {8, 16, 32}) cand.push_back({(int)(w - rc2), tb, ta});
    }
    const hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return perr(UHS_E_HIP, "event creation failed");
    double best = 0.0;
    int n = 0, found = 0;
    for (const auto &t : cand) {
        StreamParams sp;
        int us, uc;
        if (build_stream(pg, c, t.data(), sizeof(double), sp, &us, &uc) > kMaxLdsBytes) continue;
        rc = uhs_program_sweep_ex_f64(A, B, shape, pg, 1, UHS_PROGRAM_TILED, t.data(), stream);
        if (rc) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = perr(UHS_E_HIP, "event record failed"); break; }
        rc = uhs_program_sweep_ex_f64(A, B, shape, pg, reps, UHS_PROGRAM_TILED, t.data(), stream);
        if (rc) break;
        if (hipEventRecord(e1, s) != hipSuccess) { rc = perr(UHS_E_HIP, "event record failed"); break; }
        if (hipEventSynchronize(e1) != hipSuccess) { rc = perr(UHS_E_HIP, "event wait failed"); break; }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = perr(UHS_E_HIP, "event timing failed"); break; }
        const double per = ms / reps;
        n++;
        if (!found || per < best) {
            best = per;
            for (int a = 0; a < 3; a++) tile_out[a] = t[a];
            found = 1;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (evaluated) *evaluated = n;
    if (!found) return perr(UHS_E_CONFIG, "no tile of the streaming kernel fits this program in the LDS");
    if (best_ms) *best_ms = best;
    {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        g_tuned[program_key(shape, pg, sizeof(double))] = {tile_out[0], tile_out[1], tile_out[2]};
    }
    return UHS_OK;
}
