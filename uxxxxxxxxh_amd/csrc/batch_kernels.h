// Device side of the batched path: many independent 3D FP64 grids, mixed
// stencils, one persistent launch per time block, workgroups pulling tiles
// from a device-side queue.
//
// star_tb_tile advances a tile of a star stencil by K timesteps at once
// (temporal blocking with ghost zones).  The workgroup streams planes along
// axis 0; level 0 is the input plane read from HBM, level j (1..K) is the
// state after j steps, computed one plane behind level j-1 (R planes) on a
// region that shrinks by R per level.  Levels 0..K-1 live in LDS rings of
// 2R+1 planes; level K is written to the destination grid.  Every update is
// the same product-then-fma chain in term order as the one-step kernels, so
// the result is bitwise identical to K single sweeps.  Cells of a level that
// lie outside the interior hold that level's halo, i.e. the halo of the
// buffer the one-step ping-pong would hold it in (halos never change).
#pragma once
#include "uhs.h"
#include "sweep_kernels.h"

constexpr int kBatchThreads = 256;
// Workgroup size of the star launches of a split batch (UHS_BATCH_SPLIT):
// the per-grid two-step kernel's tile shapes need twice the threads.
constexpr int kBatchStarThreads = 512;
// Staging registers per thread of the batch's tiles (box: plane-tile cells,
// two-step star: ring cells); the host rejects tiles that need more.
constexpr int kBatchBoxCells = 4, kBatchRingCells = 2;

// Per-grid record, in device memory.
struct BGrid {
    const double *A;
    const double *B;      // halo-inclusive grids (interior origins below)
    int n0, n1, n2, halo;
    int sy;
    long long sz;
    int kind;             // built-in stencil id (kBuiltins index)
    int TX, TY;           // tile
    int mode;             // kMode* bits: which tile code runs this grid's tasks
    int TX1, TY1;         // stars: the tile of single-step tasks (kModeRq1)
    double coef[UHS_MAX_POINTS]; // term order (stars) or box_pack_index order (boxes)
};

// One unit of work: a tile column and slab of one grid, advanced by `depth`
// steps from global step k0 (src/dst are the interior origins).
struct BTask {
    const double *src;
    double *dst;
    int grid;
    int x0, y0, zs, ze;
    int depth;
    int k0;
    int pad;
};

// BGrid::mode bits.  Stars: kModeRq1 = single steps use the register-queue
// tile (star3d_tile, 2 points per thread, tile TX1 x TY1 = 512 points: the
// batch's star tile when it has 512 points, 64 x 8 for the leading single
// steps of a temporally blocked batch otherwise); kModeRq2 = two-
// step tasks use the register-queue two-step tile (star3d_tb2_tile, 2 columns
// per thread over the tile plus its ghost zone).  Boxes: kModeRows2 = two
// rows per thread (tile = 512 points), kModePlanes2 = two planes of tile
// loads in flight.
constexpr int kModeRq1 = 1, kModeRq2 = 2, kModeRows2 = 4, kModePlanes2 = 8;

// Tile forms a k_batch instantiation compiles in (template argument F): the
// forms of kMode* plus the LDS-ring star tile, the one-row box tile and one
// plane of box loads in flight.  A launch uses an instantiation with just
// the forms its tasks need, so its register footprint (and occupancy) is
// that of those forms, not of every form.
constexpr int kFormRq1 = 1, kFormRq2 = 2, kFormRows2 = 4, kFormPlanes2 = 8, kFormStarLds = 16, kFormBoxStd = 32,
              kFormPlanes1 = 64;
constexpr int kFormAll = 127;

struct BLaunch {
    const BTask *tasks;
    const BGrid *grids;
    unsigned *counter;
    int ntasks;
};

__device__ __forceinline__ int ring_slot(int p, int q)
{
    const int r = p % q;
    return r < 0 ? r + q : r;
}

template <class S, int K, class CF>
__device__ __forceinline__ void star_tb_tile(const BGrid &g, const BTask &t, double *lds, CF &cf)
{
    constexpr int R = S::R, Q = 2 * R + 1, P = S::P, M = K * R;
    const int NT = blockDim.x, tid = threadIdx.x;
    const int TX = g.TX, TY = g.TY;
    const int WX = TX + 2 * M, WY = TY + 2 * M, WC = WX * WY;
    lds_poison(lds, K * Q * WC);
    const int ox = t.x0 - M, oy = t.y0 - M; // interior coordinates of LDS cell (0, 0)
    const int h = g.halo, n0 = g.n0, n1 = g.n1, n2 = g.n2, sy = g.sy;
    const long long sz = g.sz;
    const double *__restrict__ src = t.src;
    double *__restrict__ dst = t.dst;
    const long long origin = (long long)h * sz + (long long)h * sy + h;
    const double *hA = g.A + origin, *hB = g.B + origin;
    const int zs = t.zs, ze = t.ze;
    const int steps = (ze - zs) + 2 * M;
    for (int s = 0; s < steps; ++s) {
        const int z0 = zs - M + s;
        {
            // Level 0 is the state after global step k0: interior cells come
            // from src, which holds it; halo cells from the buffer the
            // one-step ping-pong would hold it in (src itself unless earlier
            // launches advanced an even number of steps per buffer swap).
            double *ring = lds + ring_slot(z0, Q) * WC;
            const long long zoff = (long long)clampi(z0, -h, n0 + h - 1) * sz;
            const double *pz = src + zoff;
            const double *hz = ((t.k0 & 1) ? hB : hA) + zoff;
            const bool zin = z0 >= 0 && z0 < n0;
            for (int c = tid; c < WC; c += NT) {
                const int ly = c / WX, lx = c - ly * WX;
                const int gy = oy + ly, gx = ox + lx;
                const int off = clampi(gy, -h, n1 + h - 1) * sy + clampi(gx, -h, n2 + h - 1);
                const bool inside = zin && gy >= 0 && gy < n1 && gx >= 0 && gx < n2;
                ring[c] = inside ? pz[off] : hz[off];
            }
        }
        __syncthreads();
        cf.refresh();
        static_for<K>([&](auto jc) {
            constexpr int j = decltype(jc)::value + 1;
            const int zj = z0 - j * R;
            const int m = (K - j) * R;
            if (zj >= zs - m && zj <= ze - 1 + m) {
                const double *prev = lds + (j - 1) * Q * WC;
                const double *pp[Q];
#pragma unroll
                for (int d = 0; d < Q; d++) pp[d] = prev + ring_slot(zj - R + d, Q) * WC;
                double *cur = lds + (j < K ? j : 0) * Q * WC + ring_slot(zj, Q) * WC;
                // Level j is the state after global step k0 + j: its halo is
                // the halo of B when that step is odd, of A when even.
                const double *hbuf = ((t.k0 + j) & 1) ? hB : hA;
                const bool zin = zj >= 0 && zj < n0;
                const int RW = TX + 2 * m, RH = TY + 2 * m;
                double *oz = dst + (long long)zj * sz;
                for (int c = tid; c < RW * RH; c += NT) {
                    const int ry = c / RW, rx = c - ry * RW;
                    const int ly = M - m + ry, lx = M - m + rx;
                    const int gy = oy + ly, gx = ox + lx;
                    const bool inside = zin && gy >= 0 && gy < n1 && gx >= 0 && gx < n2;
                    const int li = ly * WX + lx;
                    if (inside) {
                        Acc<double, false> a;
                        static_for<P>([&](auto tc) {
                            constexpr int tt = decltype(tc)::value;
                            constexpr int dz = S::OFF[tt][0], dy = S::OFF[tt][1], dx = S::OFF[tt][2];
                            const double v = pp[R + dz][li + dy * WX + dx];
                            if constexpr (tt == 0) a.first(cf.template get<tt>(), v);
                            else a.add(cf.template get<tt>(), v);
                        });
                        if constexpr (j < K) cur[li] = a.value();
                        else oz[gy * sy + gx] = a.value();
                    } else if constexpr (j < K) {
                        cur[li] = hbuf[(long long)clampi(zj, -h, n0 + h - 1) * sz +
                                       clampi(gy, -h, n1 + h - 1) * sy + clampi(gx, -h, n2 + h - 1)];
                    }
                }
            }
            __syncthreads();
        });
    }
}

// Runs task `t` of grid `g` with program S (star: depth t.depth <= K; box:
// one step), with the tile forms F allows (the host picks F to cover every
// task of the launch).
template <class S, int K, int F>
__device__ __forceinline__ void run_task(const BGrid &g, const BTask &t, double *lds)
{
    typedef const __attribute__((address_space(4))) double *kptr;
    Coefs<double, S::P> cf((kptr)g.coef);
    if constexpr (S::STAR) {
        if constexpr (K == 1) {
            // Single-step launches come first in a grid's schedule, so src
            // always holds the halo of its own state (src is the ping-pong
            // buffer) and the one-step tiles may read halos from it.
            if constexpr ((F & kFormRq1) != 0) {
                if (g.mode & kModeRq1) {
                    const TileGeom geo{g.n0, g.n1, g.n2, g.sy, g.sz, g.TX1, g.TY1, g.TX1 / 2, false};
                    star3d_tile<double, S, 2, false, 1>(t.src, t.dst, geo, t.x0, t.y0, t.zs, t.ze, lds, cf);
                    return;
                }
            }
            if constexpr ((F & kFormStarLds) != 0) star_tb_tile<S, 1>(g, t, lds, cf);
        } else if constexpr (K == 2) {
            if constexpr ((F & kFormRq2) != 0) {
                if (g.mode & kModeRq2) {
                    constexpr int R = S::R;
                    const long long origin = (long long)g.halo * g.sz + (long long)g.halo * g.sy + g.halo;
                    const double *hA = g.A + origin, *hB = g.B + origin;
                    const TileGeom geo{g.n0, g.n1, g.n2, g.sy, g.sz, g.TX, g.TY, (g.TX + 2 * R) / 2, false};
                    const double *h0 = (t.k0 & 1) ? hB : hA, *h1 = (t.k0 & 1) ? hA : hB;
                    // Loads one iteration ahead (fewer in flight): beside
                    // the box tiles this runs the benchmark batch in 120 ms
                    // against 129 with the per-grid kernel's two.
                    if (star3d_tb2_full<R>(t.x0, t.y0, t.zs, t.ze, g.TX, g.TY, g.n0, g.n1, g.n2))
                        star3d_tb2_tile<double, S, 2, false, 1, kBatchRingCells, true, 1>(
                            t.src, t.dst, h0, h1, geo, t.x0, t.y0, t.zs, t.ze, lds, cf);
                    else
                        star3d_tb2_tile<double, S, 2, false, 1, kBatchRingCells, false, 1>(
                            t.src, t.dst, h0, h1, geo, t.x0, t.y0, t.zs, t.ze, lds, cf);
                    return;
                }
            }
            if constexpr ((F & kFormStarLds) != 0) star_tb_tile<S, 2>(g, t, lds, cf);
        } else {
            if constexpr ((F & kFormStarLds) != 0) star_tb_tile<S, K>(g, t, lds, cf);
        }
    } else {
        const TileGeom geo{g.n0, g.n1, g.n2, g.sy, g.sz, g.TX, g.TY, g.TX, false};
        const bool pl2 = g.mode & kModePlanes2;
        if (g.mode & kModeRows2) {
            if constexpr ((F & kFormRows2) != 0) {
                if constexpr ((F & kFormPlanes2) != 0) {
                    if (pl2) {
                        box3d_tile_rows2<double, S, false, 2, kBatchBoxCells>(t.src, t.dst, geo, t.x0, t.y0, t.zs, t.ze, lds, cf);
                        return;
                    }
                }
                if constexpr ((F & kFormPlanes1) != 0)
                    box3d_tile_rows2<double, S, false, 1, kBatchBoxCells>(t.src, t.dst, geo, t.x0, t.y0, t.zs, t.ze, lds, cf);
            }
        } else {
            if constexpr ((F & kFormBoxStd) != 0) {
                if constexpr ((F & kFormPlanes2) != 0) {
                    if (pl2) {
                        box3d_tile<double, S, 1, false, 1, 2, kBatchBoxCells>(t.src, t.dst, geo, t.x0, t.y0, t.zs, t.ze, lds, cf);
                        return;
                    }
                }
                if constexpr ((F & kFormPlanes1) != 0)
                    box3d_tile<double, S, 1, false, 1, 1, kBatchBoxCells>(t.src, t.dst, geo, t.x0, t.y0, t.zs, t.ze, lds, cf);
            }
        }
    }
}

// The forms task `t` of grid `g` needs (host and device agree on this).
__host__ __device__ inline int task_forms(const BGrid &g, const BTask &t, bool star)
{
    if (star) {
        if (t.depth == 1) return (g.mode & kModeRq1) ? kFormRq1 : kFormStarLds;
        if (t.depth == 2) return (g.mode & kModeRq2) ? kFormRq2 : kFormStarLds;
        return kFormStarLds;
    }
    return ((g.mode & kModeRows2) ? kFormRows2 : kFormBoxStd) | ((g.mode & kModePlanes2) ? kFormPlanes2 : kFormPlanes1);
}

// The persistent kernel: each workgroup takes the next task index from the
// launch's counter until the list is exhausted (every workgroup reaches the
// exit).  Progs lists the stencils this instantiation can run; a grid's kind
// selects one.  NTB: threads per workgroup (the launch uses exactly that).
template <int K, int F, int NTB, class... Progs>
__global__ void __launch_bounds__(NTB) k_batch(BLaunch L)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double *lds = reinterpret_cast<double *>(smem_raw);
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(L.counter, 1u);
        __syncthreads();
        // Every lane read the same value: say so, so the task and grid
        // records (and the coefficient pointer) live in scalar registers.
        const int i = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (i >= L.ntasks) break;
        const BTask t = L.tasks[i];
        const BGrid &g = L.grids[t.grid];
        const int kind = g.kind;
        ((kind == Progs::ID ? (run_task<Progs, K, F>(g, t, lds), 0) : 0), ...);
        __syncthreads();
    }
}

__global__ void k_zero_u32(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
