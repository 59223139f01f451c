// Tuning of the batched graph: the candidate space of batch configurations,
// a timed search over any list of them, and the learned cost model layered
// on top (features of a batch and a configuration, model-pruned search).
//
// The reference tunes by compiling and timing sampled points of its tiling
// space, one per stencil (scripts/evaluation_scripts_GPU/test_driver.py:
// 198-256, collect.py:30-69).  Here a batch configuration is a runtime
// argument, so candidates are planned, captured and run directly; the model
// predicts every candidate and only the few it ranks fastest are timed.
#include "uhs.h"
#include "uhs_host.h"
#include "uhs_internal.h"
#include "uhs_tune.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The candidate space: temporal depth {1, 2, 4} with its star tiles (512-
// point register-queue tiles for single steps; tiles whose ghost-zone
// extension fits the workgroup for two steps; the LDS-ring form for four),
// 64- or 512-plane star slabs, box tiles of 256 points (one row per thread)
// or 512 (two rows), 64- or 128-plane box slabs, one or two planes of box
// loads in flight.
std::vector<uhs_batch_config> batch_space()
{
    static const int s1[][2] = {{64, 8}, {128, 4}, {32, 16}};
    static const int s2[][2] = {{62, 6}, {30, 14}, {32, 12}, {32, 8}, {64, 4}};
    static const int s4[][2] = {{32, 8}};
    static const int stz[] = {64, 512};
    static const int btiles[][3] = {{32, 8, 0},
                                    {64, 4, 0},
                                    {32, 16, UHS_BATCH_BOX_ROWS2},
                                    {64, 8, UHS_BATCH_BOX_ROWS2},
                                    {16, 32, UHS_BATCH_BOX_ROWS2}};
    static const int btz[] = {64, 128};
    static const int bpl[] = {0, UHS_BATCH_BOX_PLANES2};
    std::vector<uhs_batch_config> cand;
    auto add = [&](int d, const int (*tiles)[2], int nt) {
        for (int i = 0; i < nt; i++)
            for (int sz : stz)
                for (auto &b : btiles)
                    for (int bz : btz)
                        for (int pl : bpl)
                            cand.push_back(uhs_batch_config{d, tiles[i][0], tiles[i][1], sz, b[0], b[1], bz, b[2] | pl});
    };
    add(1, s1, 3);
    add(2, s2, 5);
    add(4, s4, 1);
    // Split batches, star and box launches as two concurrent chains: two-step
    // star tiles of 256 threads and of 512 (whose ghost-zone extension needs
    // 512-thread workgroups).  The split without concurrency (one chain) is
    // left out: it measured slower than shared launches for every tile.
    static const int s2w[][2] = {{32, 12}, {30, 14}, {62, 6}, {62, 14}, {30, 30}, {126, 6}};
    static const int stzw[] = {32, 64};
    static const int btw[][3] = {{32, 16, UHS_BATCH_BOX_ROWS2 | UHS_BATCH_BOX_PLANES2},
                                 {64, 8, UHS_BATCH_BOX_ROWS2 | UHS_BATCH_BOX_PLANES2},
                                 {32, 8, UHS_BATCH_BOX_PLANES2}};
    for (auto &t : s2w)
        for (int sz : stzw)
            for (auto &b : btw)
                for (int bz : btz)
                    cand.push_back(uhs_batch_config{2, t[0], t[1], sz, b[0], b[1], bz,
                                                    b[2] | UHS_BATCH_SPLIT | UHS_BATCH_CONCURRENT});
    return cand;
}

// Times each candidate: plan, capture, one warm-up run and one timed run of
// `timesteps` steps (grid contents are overwritten).  ms[i] < 0 for
// candidates that cannot be built.
int time_list(const uhs_grid_desc *grids, int n, int timesteps, const uhs_batch_config *cand, int ncand,
              double *ms, hipStream_t stream)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return uhs_set_error(UHS_E_HIP, "events");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return uhs_set_error(UHS_E_HIP, "events");
    }
    int rc = UHS_OK;
    for (int i = 0; i < ncand && rc == UHS_OK; i++) {
        ms[i] = -1.0;
        uhs_batch_t b = nullptr;
        if (uhs_batch_create(grids, n, timesteps, &cand[i], &b) != UHS_OK) continue;
        float t = 0.f;
        bool ok = uhs_batch_run(b, stream) == UHS_OK; // warm-up
        ok = ok && hipEventRecord(e0, stream) == hipSuccess;
        ok = ok && uhs_batch_run(b, stream) == UHS_OK;
        ok = ok && hipEventRecord(e1, stream) == hipSuccess;
        ok = ok && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess;
        uhs_batch_destroy(b);
        if (!ok) rc = uhs_set_error(UHS_E_HIP, "batch run failed during tuning");
        else ms[i] = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int pick_best(const uhs_batch_config *cand, const double *ms, int n, uhs_batch_config *best, double *best_ms)
{
    int bi = -1;
    for (int i = 0; i < n; i++)
        if (ms[i] >= 0 && (bi < 0 || ms[i] < ms[bi])) bi = i;
    if (bi < 0) return uhs_set_error(UHS_E_CONFIG, "no batch configuration could be built");
    *best = cand[bi];
    if (best_ms) *best_ms = ms[bi];
    return UHS_OK;
}

} // namespace

extern "C" {

int uhs_batch_space(uhs_batch_config *out, int cap)
{
    const auto sp = batch_space();
    if (out)
        for (int i = 0; i < (int)sp.size() && i < cap; i++) out[i] = sp[i];
    return (int)sp.size();
}

int uhs_batch_tune_list(const uhs_grid_desc *grids, int n, int timesteps, const uhs_batch_config *cands,
                        int ncand, uhs_batch_config *best, double *best_ms, double *ms_out, uhs_stream_t stream)
{
    if (!grids || n < 1 || !cands || ncand < 1 || !best || timesteps < 1)
        return uhs_set_error(UHS_E_INVAL, "batch tuning needs grids, candidates, steps and an output");
    std::vector<double> ms(ncand, -1.0);
    if (int rc = time_list(grids, n, timesteps, cands, ncand, ms.data(), (hipStream_t)stream)) return rc;
    if (ms_out) std::copy(ms.begin(), ms.end(), ms_out);
    return pick_best(cands, ms.data(), ncand, best, best_ms);
}

int uhs_batch_tune(const uhs_grid_desc *grids, int n, int timesteps, uhs_batch_config *best, double *best_ms,
                   uhs_batch_config *cands_out, double *ms_out, int ncand_cap, int *ncand, uhs_stream_t stream)
{
    if (!grids || n < 1 || !best) return uhs_set_error(UHS_E_INVAL, "batch tuning needs grids and an output");
    const auto cand = batch_space();
    std::vector<double> ms(cand.size(), -1.0);
    if (int rc = time_list(grids, n, timesteps, cand.data(), (int)cand.size(), ms.data(), (hipStream_t)stream))
        return rc;
    for (int i = 0; i < (int)cand.size() && i < ncand_cap; i++) {
        if (cands_out) cands_out[i] = cand[i];
        if (ms_out) ms_out[i] = ms[i];
    }
    if (ncand) *ncand = (int)cand.size();
    return pick_best(cand.data(), ms.data(), (int)cand.size(), best, best_ms);
}

int uhs_batch_features(const uhs_grid_desc *grids, int n, const uhs_batch_config *c, float *f)
{
    if (!grids || n < 1 || !c || !f) return uhs_set_error(UHS_E_INVAL, "bad batch feature request");
    std::fill(f, f + UHS_MLP_INPUTS, 0.f);
    f[c->star_depth >= 4 ? 2 : c->star_depth >= 2 ? 1 : 0] = 1.f;
    f[3] = log2f((float)std::max(1, c->star_tile_x));
    f[4] = log2f((float)std::max(1, c->star_tile_y));
    f[5] = log2f((float)std::max(1, c->star_tile_z));
    f[6] = log2f((float)std::max(1, c->box_tile_x));
    f[7] = log2f((float)std::max(1, c->box_tile_y));
    f[8] = log2f((float)std::max(1, c->box_tile_z));
    f[9] = (float)((c->flags & UHS_BATCH_BOX_ROWS2 ? 1 : 0) + (c->flags & UHS_BATCH_BOX_PLANES2 ? 2 : 0));
    f[10] = (float)((c->flags & UHS_BATCH_SPLIT ? 1 : 0) + (c->flags & UHS_BATCH_CONCURRENT ? 1 : 0));
    double star_pts = 0, box_pts = 0, star_work = 0, box_work = 0;
    double le[3] = {0, 0, 0}, le2 = 0;
    int radius_star = 1;
    for (int i = 0; i < n; i++) {
        const uhs_grid_desc &d = grids[i];
        if (d.dims != 3 || !d.stencil.offsets) return uhs_set_error(UHS_E_INVAL, "batch grids are 3D");
        const int prog = uhs_match_builtin(3, d.stencil.npoints, d.stencil.offsets);
        if (prog < 0) return uhs_set_error(UHS_E_UNSUPPORTED, "the batch runs the built-in 3D stencils only");
        double pts = 1;
        for (int k = 0; k < 3; k++) {
            if (d.ext[k] < 1) return uhs_set_error(UHS_E_INVAL, "extent out of range");
            pts *= (double)d.ext[k];
            const double l = std::log2((double)d.ext[k]);
            le[k] += l;
            le2 += l * l;
        }
        if (kBuiltins[prog].star) {
            star_pts += pts;
            star_work += pts * kBuiltins[prog].npoints;
            radius_star = std::max(radius_star, kBuiltins[prog].radius);
        } else {
            box_pts += pts;
            box_work += pts * kBuiltins[prog].npoints;
        }
    }
    f[11] = star_pts > 0 ? (float)std::log2(star_pts) : 0.f;
    f[12] = box_pts > 0 ? (float)std::log2(box_pts) : 0.f;
    f[13] = log2f((float)n);
    double mean_all = 0;
    for (int k = 0; k < 3; k++) {
        f[14 + k] = (float)(le[k] / n);
        mean_all += le[k] / (3.0 * n);
    }
    f[17] = (float)std::sqrt(std::max(0.0, le2 / (3.0 * n) - mean_all * mean_all)); // spread of log extents
    // Mean terms per point of each kind (log2), and whether the star tile
    // takes the register-queue two-step form at depth 2 (its ghost-zone
    // extension fits the 256-thread workgroup, two columns per thread).
    f[18] = (float)std::log2(1.0 + (star_pts > 0 ? star_work / star_pts : 0.0) +
                             (box_pts > 0 ? box_work / box_pts : 0.0));
    const int bx = c->star_tile_x + 2 * radius_star, by = c->star_tile_y + 2 * radius_star;
    const int nts = (c->flags & UHS_BATCH_SPLIT) ? 512 : 256;
    f[19] = (c->star_depth == 2 && bx % 2 == 0 && bx / 2 * by <= nts) ? 1.f : 0.f;
    return UHS_OK;
}

int uhs_batch_tune_model(const uhs_grid_desc *grids, int n, int timesteps, uhs_mlp_t model, int topk,
                         uhs_batch_config *best, double *best_ms, double *seconds, uhs_stream_t stream)
{
    if (!grids || n < 1 || !model || topk < 1 || !best || timesteps < 1)
        return uhs_set_error(UHS_E_INVAL, "model-pruned batch tuning needs grids, a model, k and an output");
    const double t0 = now_s();
    const auto cand = batch_space();
    std::vector<float> feat(cand.size() * UHS_MLP_INPUTS), pred(cand.size());
    for (size_t i = 0; i < cand.size(); i++)
        if (int rc = uhs_batch_features(grids, n, &cand[i], &feat[i * UHS_MLP_INPUTS])) return rc;
    if (int rc = uhs_mlp_predict(model, feat.data(), (int)cand.size(), pred.data())) return rc;
    std::vector<size_t> ord(cand.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return pred[a] < pred[b]; });
    std::vector<uhs_batch_config> top;
    for (size_t i = 0; i < ord.size() && (int)top.size() < topk; i++) top.push_back(cand[ord[i]]);
    std::vector<double> ms(top.size(), -1.0);
    if (int rc = time_list(grids, n, timesteps, top.data(), (int)top.size(), ms.data(), (hipStream_t)stream))
        return rc;
    const int rc = pick_best(top.data(), ms.data(), (int)top.size(), best, best_ms);
    if (seconds) *seconds = now_s() - t0;
    return rc;
}

} // extern "C"
