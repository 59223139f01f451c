// Batched sweep of many independent 3D FP64 grids (uneven extents, mixed
// stencils) as one hipGraph: a counter-reset node, then one persistent
// work-queue launch per time block.  Star-stencil grids advance `depth`
// steps per launch (temporal blocking); box-stencil grids one step.
#include "batch_kernels.h"
#include "uhs_host.h"
#include "uhs_tune.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

using MixFn = void (*)(BLaunch);

// The benchmark mix (7-point star + 125-point box) gets its own kernels,
// one per combination of the tile forms a launch can hold (star: none, the
// register-queue one- or two-step tile, or the LDS-ring tile; box: one or
// two rows per thread, one or two planes in flight), so each launch carries
// the register footprint of its own forms only.  Other stencil sets use one
// kernel with every form per depth.
template <int K, int SF> MixFn mix_box(int bf)
{
    switch (bf) {
    case kFormBoxStd | kFormPlanes1: return k_batch<K, SF | kFormBoxStd | kFormPlanes1, kBatchThreads, Prog_3d1r_star, Prog_3d2r_box>;
    case kFormBoxStd | kFormPlanes2: return k_batch<K, SF | kFormBoxStd | kFormPlanes2, kBatchThreads, Prog_3d1r_star, Prog_3d2r_box>;
    case kFormRows2 | kFormPlanes1: return k_batch<K, SF | kFormRows2 | kFormPlanes1, kBatchThreads, Prog_3d1r_star, Prog_3d2r_box>;
    case kFormRows2 | kFormPlanes2: return k_batch<K, SF | kFormRows2 | kFormPlanes2, kBatchThreads, Prog_3d1r_star, Prog_3d2r_box>;
    }
    return nullptr;
}
template <int K> MixFn all_kernel()
{
    return k_batch<K, kFormAll, kBatchThreads, Prog_3d1r_star, Prog_3d2r_star, Prog_3d3r_star, Prog_3d1r_box,
                   Prog_3d2r_box, Prog_3d3r_box>;
}

MixFn pick_kernel(bool mix, int K, int forms)
{
    if (mix) {
        const int sf = forms & (kFormRq1 | kFormRq2 | kFormStarLds);
        const int bf = forms & (kFormRows2 | kFormBoxStd | kFormPlanes1 | kFormPlanes2);
        MixFn f = nullptr;
        // Box-only launches use the kernel that also has the (light)
        // register-queue star tile: an instantiation without any star form
        // does not compile (the backend rejects it).
        if (K == 1 && (sf == 0 || sf == kFormRq1)) f = mix_box<1, kFormRq1>(bf);
        if (K == 1 && sf == kFormStarLds) f = mix_box<1, kFormStarLds>(bf);
        if (K == 2 && sf == kFormRq2) f = mix_box<2, kFormRq2>(bf);
        if (K == 2 && sf == kFormStarLds) f = mix_box<2, kFormStarLds>(bf);
        if (K == 4 && sf == kFormStarLds) f = mix_box<4, kFormStarLds>(bf);
        if (f) return f;
        if (K == 1) return k_batch<1, kFormAll, kBatchThreads, Prog_3d1r_star, Prog_3d2r_box>;
        if (K == 2) return k_batch<2, kFormAll, kBatchThreads, Prog_3d1r_star, Prog_3d2r_box>;
        if (K == 4) return k_batch<4, kFormAll, kBatchThreads, Prog_3d1r_star, Prog_3d2r_box>;
        return nullptr;
    }
    switch (K) {
    case 1: return all_kernel<1>();
    case 2: return all_kernel<2>();
    case 4: return all_kernel<4>();
    }
    return nullptr;
}

// Star launches of a split batch of the benchmark mix: star tile forms only,
// kBatchStarThreads threads per workgroup.
MixFn pick_star_kernel(int K, int forms)
{
    const int sf = forms & (kFormRq1 | kFormRq2 | kFormStarLds);
    using S = Prog_3d1r_star;
    using B = Prog_3d2r_box;
    constexpr int N = kBatchStarThreads;
    if (K == 1 && sf == kFormRq1) return k_batch<1, kFormRq1, N, S, B>;
    if (K == 1 && sf == kFormStarLds) return k_batch<1, kFormStarLds, N, S, B>;
    if (K == 2 && sf == kFormRq2) return k_batch<2, kFormRq2, N, S, B>;
    if (K == 2 && sf == kFormStarLds) return k_batch<2, kFormStarLds, N, S, B>;
    if (K == 4 && sf == kFormStarLds) return k_batch<4, kFormStarLds, N, S, B>;
    return nullptr;
}

struct LaunchRec {
    int depth;
    size_t first, count; // slice of the task array
    size_t lds;
    int workgroups;
    int threads;         // per workgroup
    int chain;           // split batches: 0 = the star grids' launches, 1 = the box grids'
    MixFn fn;
};

std::string g_batch_err;

int berr(int code, const std::string &m)
{
    return uhs_set_error(code, m.c_str());
}

} // namespace

// A batch keeps one captured graph per buffer parity.  Run r (r = 0, 1, ...)
// advances every grid from global step r*T; when T is odd, the state after an
// odd number of runs sits in B, so that run reads the grids with the roles of
// A and B swapped (grid records n..2n-1, task list ntasks..2*ntasks-1), which
// is exactly a fresh run on the swapped pair.  Even T needs graph 0 only.
struct uhs_batch {
    BGrid *d_grids = nullptr;   // [2n]: as given, then with A and B swapped
    BTask *d_tasks = nullptr;   // [2 * ntasks] (the second half only for odd T)
    unsigned *d_counters = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<LaunchRec> launches;
    long long ntasks = 0;
    int timesteps = 0;
    int phase = 0;              // graph the next run launches
};

static std::vector<int> star_schedule(int T, int K)
{
    // Single steps first, then the K-step blocks.  The result must end in A
    // for even T and in B for odd T: the number of launches (each reads one
    // buffer and writes the other) must have T's parity; turning one K-block
    // into K single steps flips it (K even).  With the single steps first, a
    // single-step launch always reads the buffer the one-step ping-pong
    // would hold its state in.
    if (K <= 1) return std::vector<int>(T, 1);
    int ones = T % K, blocks = T / K;
    if ((ones + blocks) % 2 != T % 2 && blocks > 0) {
        blocks--;
        ones += K;
    }
    std::vector<int> d(ones, 1);
    d.insert(d.end(), blocks, K);
    return d;
}

extern "C" {

int uhs_batch_partition(const uhs_grid_desc *grids, int n, int ngpus, const double *cost, int *gpu_of,
                        double *load)
{
    if (!grids || n < 1 || !gpu_of) return berr(UHS_E_INVAL, "partition needs grids and an output array");
    if (ngpus < 1) return berr(UHS_E_INVAL, "partition needs ngpus >= 1");
    std::vector<double> c(n);
    for (int i = 0; i < n; i++) {
        if (cost) {
            if (!(cost[i] >= 0.0) || !std::isfinite(cost[i])) return berr(UHS_E_INVAL, "costs must be finite and >= 0");
            c[i] = cost[i];
            continue;
        }
        const uhs_grid_desc &d = grids[i];
        const int prog = d.dims == 3 && d.stencil.dims == 3 && d.stencil.offsets
                             ? uhs_match_builtin(3, d.stencil.npoints, d.stencil.offsets)
                             : -1;
        if (prog < 0) return berr(UHS_E_UNSUPPORTED, "cost estimates exist for the built-in 3D stencils only");
        double pts = 1.0;
        for (int k = 0; k < 3; k++) {
            if (d.ext[k] < 1) return berr(UHS_E_INVAL, "extent out of range");
            pts *= (double)d.ext[k];
        }
        // Time per update ~ max(DRAM time, FMA time): memory-bound up to a
        // few tens of terms, then proportional to the term count.  The
        // measured FP64 rates at 512^3 (stars 300-340 GSt/s, the 125-point
        // box 183, the 343-point box 78) fit max(1, terms / 64).
        c[i] = pts * std::max(1.0, kBuiltins[prog].npoints / 64.0);
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return c[a] > c[b]; });
    std::vector<double> l(ngpus, 0.0);
    for (int i : ord) {
        int best = 0;
        for (int g = 1; g < ngpus; g++)
            if (l[g] < l[best]) best = g;
        gpu_of[i] = best;
        l[best] += c[i];
    }
    if (load)
        for (int g = 0; g < ngpus; g++) load[g] = l[g];
    return UHS_OK;
}

int uhs_batch_default_config(uhs_batch_config *c)
{
    if (!c) return berr(UHS_E_INVAL, "null config");
    // The benchmark mix's best measured configuration (64 uneven grids,
    // 7-point stars and 125-point boxes): stars two steps per launch in
    // 62 x 14 tiles (512-thread workgroups), boxes two rows per thread with
    // two planes in flight, the two kinds as concurrent chains.
    *c = uhs_batch_config{2, 62, 14, 32, 32, 16, 128,
                          UHS_BATCH_BOX_ROWS2 | UHS_BATCH_BOX_PLANES2 | UHS_BATCH_SPLIT | UHS_BATCH_CONCURRENT};
    return UHS_OK;
}

int uhs_batch_create(const uhs_grid_desc *grids, int n, int timesteps, const uhs_batch_config *cfg_in,
                     uhs_batch_t *out)
{
    if (!grids || n < 1 || !out) return berr(UHS_E_INVAL, "batch needs grids and an output handle");
    if (timesteps < 1) return berr(UHS_E_INVAL, "batch needs timesteps >= 1");
    uhs_batch_config cfg;
    if (cfg_in) {
        cfg = *cfg_in;
    } else {
        // Other stencil sets than the benchmark mix: shared launches with
        // 32 x 12 star tiles in 256-thread workgroups.
        bool mix = true;
        for (int i = 0; i < n; i++) {
            const int prog = grids[i].stencil.offsets && grids[i].dims == 3
                                 ? uhs_match_builtin(3, grids[i].stencil.npoints, grids[i].stencil.offsets)
                                 : -1;
            if (prog != Prog_3d1r_star::ID && prog != Prog_3d2r_box::ID) mix = false;
        }
        uhs_batch_default_config(&cfg);
        if (!mix) cfg = uhs_batch_config{2, 32, 12, 64, 32, 16, 128, UHS_BATCH_BOX_ROWS2 | UHS_BATCH_BOX_PLANES2};
    }
    const int K = cfg.star_depth;
    if (K != 1 && K != 2 && K != 4) return berr(UHS_E_CONFIG, "star_depth must be 1, 2 or 4");
    if (cfg.flags & ~(UHS_BATCH_BOX_ROWS2 | UHS_BATCH_BOX_PLANES2 | UHS_BATCH_SPLIT | UHS_BATCH_CONCURRENT))
        return berr(UHS_E_CONFIG, "unknown batch flags");
    const bool rows2 = cfg.flags & UHS_BATCH_BOX_ROWS2;
    const bool split = cfg.flags & UHS_BATCH_SPLIT, concurrent = cfg.flags & UHS_BATCH_CONCURRENT;
    if (concurrent && !split) return berr(UHS_E_CONFIG, "UHS_BATCH_CONCURRENT needs UHS_BATCH_SPLIT");
    if (cfg.box_tile_x < 1 || cfg.box_tile_y < 1 ||
        cfg.box_tile_x * cfg.box_tile_y != (rows2 ? 2 : 1) * kBatchThreads || (rows2 && cfg.box_tile_y % 2))
        return berr(UHS_E_CONFIG, "box tiles must hold exactly 256 points (tile_x * tile_y), or 512 with an even "
                                  "tile_y for two rows per thread");
    if (cfg.star_tile_x < 1 || cfg.star_tile_y < 1 || cfg.star_tile_z < 1 || cfg.box_tile_z < 1)
        return berr(UHS_E_CONFIG, "tile sizes must be positive");

    int dev = 0, cus = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return berr(UHS_E_HIP, "cannot query the device");

    // The benchmark mix (only 7-point stars and 125-point boxes) has its own
    // kernels, and a split batch of it runs its star launches with
    // kBatchStarThreads-thread workgroups.
    bool mix = true;
    for (int i = 0; i < n; i++) {
        const int prog = grids[i].stencil.offsets && grids[i].dims == 3
                             ? uhs_match_builtin(3, grids[i].stencil.npoints, grids[i].stencil.offsets)
                             : -1;
        if (prog != Prog_3d1r_star::ID && prog != Prog_3d2r_box::ID) mix = false;
    }
    // Star launches of a split batch take 512-thread workgroups when the star
    // tile needs them (two steps: its ghost-zone extension at two columns per
    // thread exceeds 256 threads; single steps: a 1024-point tile), else 256.
    int NS = kBatchThreads; // threads of star tasks
    if (split && mix) {
        const int bx = cfg.star_tile_x + 2, by = cfg.star_tile_y + 2; // radius 1
        const bool big = K == 2 ? bx / 2 * by > kBatchThreads
                                : K == 1 && cfg.star_tile_x * cfg.star_tile_y == 2 * kBatchStarThreads;
        if (big) NS = kBatchStarThreads;
    }

    // Per-grid records and validation.
    std::vector<BGrid> hg(n);
    for (int i = 0; i < n; i++) {
        const uhs_grid_desc &d = grids[i];
        char msg[160];
        if (d.dims != 3 || d.stencil.dims != 3 || !d.A || !d.B || d.A == d.B || !d.stencil.offsets) {
            snprintf(msg, sizeof msg, "grid %d: the batch takes 3D FP64 grids with distinct A and B", i);
            return berr(UHS_E_INVAL, msg);
        }
        const int prog = uhs_match_builtin(3, d.stencil.npoints, d.stencil.offsets);
        if (prog < 0 || !d.stencil.coeffs) {
            snprintf(msg, sizeof msg, "grid %d: the batch runs the six built-in 3D stencils only", i);
            return berr(UHS_E_UNSUPPORTED, msg);
        }
        const BuiltinStencil &b = kBuiltins[prog];
        if (d.halo < b.radius) {
            snprintf(msg, sizeof msg, "grid %d: halo %d smaller than radius %d", i, d.halo, b.radius);
            return berr(UHS_E_INVAL, msg);
        }
        for (int k = 0; k < 3; k++)
            if (d.ext[k] < 1 || d.ext[k] + 2LL * d.halo > (1 << 20)) {
                snprintf(msg, sizeof msg, "grid %d: extent out of range", i);
                return berr(UHS_E_INVAL, msg);
            }
        BGrid &g = hg[i];
        memset(&g, 0, sizeof g);
        g.n0 = (int)d.ext[0]; g.n1 = (int)d.ext[1]; g.n2 = (int)d.ext[2];
        g.halo = d.halo;
        g.sy = (int)(d.ext[2] + 2 * d.halo);
        g.sz = (long long)g.sy * (d.ext[1] + 2 * d.halo);
        if (g.sz > (1LL << 31)) return berr(UHS_E_INVAL, "plane too large");
        g.A = d.A; g.B = d.B;
        g.kind = prog;
        if (b.star) {
            g.TX = cfg.star_tile_x; g.TY = cfg.star_tile_y;
            const int R = b.radius, bx = g.TX + 2 * R, by = g.TY + 2 * R;
            // Single steps: the star tile when it is a register-queue tile
            // (two points per thread); in a temporally blocked batch, whose
            // only single steps are the few leading ones that keep the
            // result in the right grid, a 64-column register-queue tile
            // otherwise (the LDS-ring form is several times slower per step).
            auto rq1 = [&](int tx, int ty) {
                return tx * ty == 2 * NS && tx % 2 == 0 && 2 * R * (tx + ty) <= kStarHaloPerThread * NS;
            };
            if (rq1(g.TX, g.TY)) {
                g.mode |= kModeRq1;
                g.TX1 = g.TX; g.TY1 = g.TY;
            } else if (K > 1 && rq1(64, 2 * NS / 64)) {
                g.mode |= kModeRq1;
                g.TX1 = 64; g.TY1 = 2 * NS / 64;
            }
            if (K == 2 && bx % 2 == 0 && bx / 2 * by <= NS && 2 * R * (bx + by) <= kBatchRingCells * NS)
                g.mode |= kModeRq2;
            for (int t = 0; t < b.npoints; t++) g.coef[t] = d.stencil.coeffs[t];
        } else {
            g.TX = cfg.box_tile_x; g.TY = cfg.box_tile_y;
            if (rows2) g.mode |= kModeRows2;
            if (cfg.flags & UHS_BATCH_BOX_PLANES2) g.mode |= kModePlanes2;
            // box3d reads the coefficients cell-major (box_pack_index).
            for (int t = 0; t < b.npoints; t++) {
                int idx = t;
                switch (prog) {
                case Prog_3d1r_box::ID: idx = box_pack_index<Prog_3d1r_box>(t); break;
                case Prog_3d2r_box::ID: idx = box_pack_index<Prog_3d2r_box>(t); break;
                case Prog_3d3r_box::ID: idx = box_pack_index<Prog_3d3r_box>(t); break;
                }
                g.coef[idx] = d.stencil.coeffs[t];
            }
            if ((long long)(g.TX + 2 * b.radius) * (g.TY + 2 * b.radius) > kBatchBoxCells * kBatchThreads)
                return berr(UHS_E_CONFIG, "box tile plus halo exceeds the per-thread staging");
        }
    }

    // Schedules and tasks.
    const std::vector<int> ssched = star_schedule(timesteps, K);
    const int nls = (int)ssched.size();
    // Star launch j runs in box launch slot[j]'s time: spread evenly over the
    // box grids' one-step launches, so every part of the run has HBM-bound
    // star tiles beside the VALU-bound box tiles (each grid's own launches
    // stay in order).  Unsplit, both share that launch; split, the star
    // launches are launches of their own (indices 0 .. nls-1, box launch l
    // is nls + l), captured right after box launch slot[j] in one chain, or
    // as a second chain beside the box chain with UHS_BATCH_CONCURRENT.
    const int Lm = std::max<int>(nls, timesteps);
    const int L = split ? nls + timesteps : Lm;
    std::vector<int> slot(nls);
    for (int j = 0; j < nls; j++) slot[j] = (int)((long long)j * Lm / nls);
    std::vector<int> ldepth(L, 1), lchain(L, 0);
    for (int j = 0; j < nls; j++) ldepth[split ? j : slot[j]] = ssched[j];
    if (split)
        for (int l = 0; l < timesteps; l++) lchain[nls + l] = 1;
    std::vector<std::vector<BTask>> per(L);
    std::vector<std::vector<double>> cost(L);
    for (int i = 0; i < n; i++) {
        const BGrid &g = hg[i];
        const BuiltinStencil &b = kBuiltins[g.kind];
        const int tz = b.star ? cfg.star_tile_z : cfg.box_tile_z;
        const long long origin = (long long)g.halo * g.sz + (long long)g.halo * g.sy + g.halo;
        int k0 = 0;
        const int nl = b.star ? (int)ssched.size() : timesteps;
        for (int l = 0; l < nl; l++) {
            const int depth = b.star ? ssched[l] : 1;
            const double *src = (l & 1) ? g.B : g.A;
            const double *dst = (l & 1) ? g.A : g.B;
            // The tile this launch's tasks of the grid cover.
            const bool t1 = b.star && depth == 1 && (g.mode & kModeRq1);
            const int TX = t1 ? g.TX1 : g.TX, TY = t1 ? g.TY1 : g.TY;
            for (int z = 0; z < g.n0; z += tz)
                for (int y = 0; y < g.n1; y += TY)
                    for (int x = 0; x < g.n2; x += TX) {
                        BTask t{};
                        t.src = src + origin;
                        t.dst = const_cast<double *>(dst) + origin;
                        t.grid = i;
                        t.x0 = x; t.y0 = y; t.zs = z; t.ze = std::min(z + tz, g.n0);
                        t.depth = depth;
                        t.k0 = k0;
                        const int gl = b.star ? (split ? l : slot[l]) : (split ? nls + l : l);
                        per[gl].push_back(t);
                        const double pts = (double)std::min(TX, g.n2 - x) * std::min(TY, g.n1 - y) *
                                           (t.ze - t.zs);
                        cost[gl].push_back(pts * (b.star ? 1.0 : b.npoints / 8.0) * depth);
                    }
            k0 += depth;
        }
    }

    auto *B = new uhs_batch();
    B->timesteps = timesteps;
    std::vector<BTask> all;
    for (int l = 0; l < L; l++) {
        // Order: within each kind (star / box) longest first, so a launch
        // ends on small tasks; across kinds, a proportional interleave, so
        // HBM-bound star tiles and VALU-bound box tiles run side by side on
        // every CU instead of one phase after the other.
        std::vector<size_t> kind_ord[2];
        double kind_cost[2] = {0, 0};
        for (size_t k = 0; k < per[l].size(); k++) {
            const int kd = kBuiltins[hg[per[l][k].grid].kind].star ? 0 : 1;
            kind_ord[kd].push_back(k);
            kind_cost[kd] += cost[l][k];
        }
        for (auto &v : kind_ord)
            std::stable_sort(v.begin(), v.end(), [&](size_t a, size_t b) { return cost[l][a] > cost[l][b]; });
        std::vector<size_t> ord;
        size_t pos[2] = {0, 0};
        double taken[2] = {0, 0};
        while (pos[0] < kind_ord[0].size() || pos[1] < kind_ord[1].size()) {
            int kd;
            if (pos[0] >= kind_ord[0].size()) kd = 1;
            else if (pos[1] >= kind_ord[1].size()) kd = 0;
            else kd = taken[0] / kind_cost[0] <= taken[1] / kind_cost[1] ? 0 : 1;
            const size_t k = kind_ord[kd][pos[kd]++];
            taken[kd] += cost[l][k];
            ord.push_back(k);
        }
        LaunchRec r{};
        r.depth = ldepth[l];
        r.first = all.size();
        r.count = ord.size();
        size_t lds = 0;
        int forms = 0;
        for (size_t k : ord) {
            const BTask &t = per[l][k];
            const BGrid &g = hg[t.grid];
            const BuiltinStencil &b = kBuiltins[g.kind];
            forms |= task_forms(g, t, b.star);
            const int R = b.radius;
            size_t need;
            if (b.star) {
                const size_t wc = (size_t)(g.TX + 2 * t.depth * R) * (g.TY + 2 * t.depth * R);
                need = sizeof(double) * t.depth * (2 * R + 1) * wc;
                if (t.depth == 1 && (g.mode & kModeRq1)) // star3d_tile: two staged planes
                    need = 2 * sizeof(double) * (size_t)(g.TX1 + 2 * R) * (g.TY1 + 2 * R);
                if (t.depth == 2 && (g.mode & kModeRq2)) { // star3d_tb2_tile
                    const size_t bx = g.TX + 2 * R, by = g.TY + 2 * R;
                    need = 2 * sizeof(double) * ((bx + 2 * R) * (by + 2 * R) + bx * by);
                }
            } else {
                need = 2 * sizeof(double) * (size_t)(g.TX + 2 * R) * (g.TY + 2 * R);
            }
            lds = std::max(lds, need);
            all.push_back(t);
        }
        if (lds > (size_t)max_lds) {
            delete B;
            return berr(UHS_E_CONFIG, "tile and depth need more LDS than a workgroup may use");
        }
        r.lds = lds;
        r.chain = lchain[l];
        const bool star_launch = split && lchain[l] == 0;
        r.threads = star_launch ? NS : kBatchThreads;
        r.fn = star_launch && NS != kBatchThreads ? pick_star_kernel(r.depth, forms) : pick_kernel(mix, r.depth, forms);
        if (!r.fn && r.count) {
            delete B;
            return berr(UHS_E_CONFIG, "no batch kernel for this combination of tile forms");
        }
        B->launches.push_back(r);
    }
    B->ntasks = (long long)all.size();
    const bool odd = timesteps % 2 != 0;
    const size_t nt = all.size();
    // The swapped-role copies for the odd-numbered runs of an odd T.
    for (int i = 0; i < n; i++) {
        BGrid g = hg[i];
        std::swap(g.A, g.B);
        hg.push_back(g);
    }
    if (odd) {
        all.reserve(2 * nt);
        for (size_t k = 0; k < nt; k++) {
            BTask t = all[k];
            const BGrid &g = hg[t.grid];
            const long long origin = (long long)g.halo * g.sz + (long long)g.halo * g.sy + g.halo;
            const double *a = g.A + origin, *b = g.B + origin;
            t.src = t.src == a ? b : a;
            t.dst = const_cast<double *>(t.dst == b ? a : b);
            t.grid += n;
            all.push_back(t);
        }
    }

    auto fail_hip = [&](hipError_t e, const char *what) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        uhs_batch_destroy(B);
        return berr(UHS_E_HIP, m);
    };
    hipError_t e;
    if ((e = hipMalloc(&B->d_grids, sizeof(BGrid) * hg.size())) != hipSuccess) return fail_hip(e, "grid table");
    if ((e = hipMalloc(&B->d_tasks, sizeof(BTask) * std::max<size_t>(1, all.size()))) != hipSuccess)
        return fail_hip(e, "task table");
    if ((e = hipMalloc(&B->d_counters, sizeof(unsigned) * L)) != hipSuccess) return fail_hip(e, "counters");
    if ((e = hipMemcpy(B->d_grids, hg.data(), sizeof(BGrid) * hg.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return fail_hip(e, "grid upload");
    if (!all.empty() &&
        (e = hipMemcpy(B->d_tasks, all.data(), sizeof(BTask) * all.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return fail_hip(e, "task upload");

    for (auto &r : B->launches) {
        if (!r.count) continue;
        if (r.lds > 64 * 1024 &&
            (e = hipFuncSetAttribute((const void *)r.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)r.lds)) != hipSuccess)
            return fail_hip(e, "LDS limit");
        int per_cu = 0;
        if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)r.fn, r.threads,
                                                              r.lds)) != hipSuccess)
            return fail_hip(e, "occupancy");
        const long long wg = (long long)std::max(1, per_cu) * cus;
        r.workgroups = (int)std::max<long long>(1, std::min<long long>(wg, (long long)r.count));
    }

    // Capture order: unsplit, the launches in index order; split, star
    // launch j right after box launch slot[j] in one chain, or the star
    // launches as a second chain forked off after the counter reset and
    // joined before the end (UHS_BATCH_CONCURRENT).
    std::vector<int> order;
    if (!split) {
        for (int l = 0; l < L; l++) order.push_back(l);
    } else {
        int j = 0;
        for (int l = 0; l < timesteps; l++) {
            order.push_back(nls + l);
            for (; j < nls && slot[j] <= l; j++) order.push_back(j);
        }
        for (; j < nls; j++) order.push_back(j);
    }
    // Capture: zero the counters with a kernel (not a memset node), then the
    // launches; one graph per buffer parity that a run can start from.
    for (int ph = 0; ph < (odd ? 2 : 1); ph++) {
        hipStream_t cs = nullptr, cs2 = nullptr;
        hipEvent_t fork = nullptr, join = nullptr;
        if ((e = hipStreamCreate(&cs)) != hipSuccess) return fail_hip(e, "capture stream");
        if (concurrent && ((e = hipStreamCreate(&cs2)) != hipSuccess ||
                           (e = hipEventCreateWithFlags(&fork, hipEventDisableTiming)) != hipSuccess ||
                           (e = hipEventCreateWithFlags(&join, hipEventDisableTiming)) != hipSuccess)) {
            (void)hipStreamDestroy(cs);
            if (cs2) (void)hipStreamDestroy(cs2);
            if (fork) (void)hipEventDestroy(fork);
            return fail_hip(e, "capture stream");
        }
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_zero_u32, dim3((L + 255) / 256), dim3(256), 0, cs, B->d_counters, L);
            if (concurrent) {
                if (e == hipSuccess) e = hipEventRecord(fork, cs);
                if (e == hipSuccess) e = hipStreamWaitEvent(cs2, fork, 0);
            }
            for (int l : order) {
                const LaunchRec &r = B->launches[l];
                if (!r.count) continue;
                hipStream_t ls = concurrent && r.chain == 0 ? cs2 : cs;
                BLaunch bl{B->d_tasks + ph * nt + r.first, B->d_grids, B->d_counters + l, (int)r.count};
                hipLaunchKernelGGL(r.fn, dim3(r.workgroups), dim3(r.threads), r.lds, ls, bl);
            }
            if (concurrent) {
                if (e == hipSuccess) e = hipEventRecord(join, cs2);
                if (e == hipSuccess) e = hipStreamWaitEvent(cs, join, 0);
            }
            hipError_t e2 = hipStreamEndCapture(cs, &B->graph[ph]);
            if (e == hipSuccess) e = e2;
        }
        (void)hipStreamDestroy(cs);
        if (cs2) (void)hipStreamDestroy(cs2);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (e != hipSuccess) return fail_hip(e, "graph capture");
        if ((e = hipGraphInstantiate(&B->exec[ph], B->graph[ph], nullptr, nullptr, 0)) != hipSuccess)
            return fail_hip(e, "graph instantiate");
    }
    *out = B;
    return UHS_OK;
}

int uhs_batch_run(uhs_batch_t b, uhs_stream_t stream)
{
    if (!b || !b->exec[0]) return berr(UHS_E_INVAL, "null batch");
    hipError_t e = hipGraphLaunch(b->exec[b->phase], (hipStream_t)stream);
    if (e != hipSuccess) return berr(UHS_E_HIP, std::string("graph launch: ") + hipGetErrorString(e));
    if (b->timesteps % 2) b->phase ^= 1;
    return UHS_OK;
}

int uhs_batch_destroy(uhs_batch_t b)
{
    if (!b) return UHS_OK;
    for (int ph = 0; ph < 2; ph++) {
        if (b->exec[ph]) (void)hipGraphExecDestroy(b->exec[ph]);
        if (b->graph[ph]) (void)hipGraphDestroy(b->graph[ph]);
    }
    if (b->d_grids) (void)hipFree(b->d_grids);
    if (b->d_tasks) (void)hipFree(b->d_tasks);
    if (b->d_counters) (void)hipFree(b->d_counters);
    delete b;
    return UHS_OK;
}

int uhs_batch_info(uhs_batch_t b, int *launches, long long *tasks, int *max_workgroups,
                   size_t *max_lds)
{
    if (!b) return berr(UHS_E_INVAL, "null batch");
    int wg = 0;
    size_t lds = 0, nl = 0;
    for (auto &r : b->launches) {
        if (!r.count) continue;
        nl++;
        wg = std::max(wg, r.workgroups);
        lds = std::max(lds, r.lds);
    }
    if (launches) *launches = (int)nl;
    if (tasks) *tasks = b->ntasks;
    if (max_workgroups) *max_workgroups = wg;
    if (max_lds) *max_lds = lds;
    return UHS_OK;
}

int uhs_sweep_each_f64(const uhs_grid_desc *grids, int n, int timesteps, uhs_stream_t stream)
{
    if (!grids || n < 1) return berr(UHS_E_INVAL, "no grids");
    // One launch per timestep per grid: the one-step kernel with its
    // default tile, whatever the per-grid default (which may advance two
    // steps per launch) or a tuned entry would choose.
    for (int i = 0; i < n; i++) {
        const uhs_grid_desc &d = grids[i];
        Plan pl;
        int rc = uhs_make_plan(d.dims, d.ext, d.halo, &d.stencil, DT_F64, nullptr, &pl);
        if (rc) return rc;
        const uhs_config c = uhs_heuristic_config(pl, true);
        rc = uhs_sweep_f64(d.A, d.B, d.dims, d.ext, d.halo, &d.stencil, timesteps, &c, stream);
        if (rc) return rc;
    }
    return UHS_OK;
}

} // extern "C"
