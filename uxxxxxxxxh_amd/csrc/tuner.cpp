// Autotuner: timed search over the runtime tiling space of a sweep, with an
// append-only tuning database that makes a search resumable and doubles as
// the cost model's training set.
//
// Replaces the reference's Python random search (searchSpace_2d/3d and
// FilterParams2d/3d in scripts/evaluation_scripts_GPU/test_driver.py:36-99,
// 169-196,258-285; collect.py:30-69), where every configuration was a
// recompile plus an nsys run.  Here a configuration is a runtime argument, so
// a point of the space costs a few sweeps.  The reference's legality filters
// (threads per block, LDS size, exact divisibility) become uhs_check_config
// (register staging, 64 KiB LDS, 512 threads; partial tiles are allowed).
//
// Database line (CSV, one per timed configuration):
//   device,stencil,dtype,dims,e0,e1,e2,halo,tile_x,tile_y,tile_z,vec,flags,ms,status
// `stencil` is the built-in case name or "custom"; status 0 = timed, <0 = the
// UHS_E_* code the configuration failed with (skipped, the search goes on).
#include "uhs.h"
#include "uhs_host.h"
#include "uhs_tune.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace {

std::string device_tag()
{
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
    char b[96];
    snprintf(b, sizeof b, "%s-%dcu", p.gcnArchName, p.multiProcessorCount);
    for (char *c = b; *c; c++)
        if (*c == ',' || *c == ':') *c = '_';
    return b;
}

struct Key {
    std::string dev, stencil;
    int dtype, dims;
    long long e0, e1, e2;
    int halo;
    bool operator<(const Key &o) const
    {
        return std::tie(dev, stencil, dtype, dims, e0, e1, e2, halo) <
               std::tie(o.dev, o.stencil, o.dtype, o.dims, o.e0, o.e1, o.e2, o.halo);
    }
};

typedef std::tuple<int, int, int, int, int> CfgT;

CfgT ct(const uhs_config &c) { return CfgT(c.tile_x, c.tile_y, c.tile_z, c.vec, c.flags); }

// Timings read back from a database file, per problem.
std::map<Key, std::map<CfgT, std::pair<double, int>>> load_db(const char *path)
{
    std::map<Key, std::map<CfgT, std::pair<double, int>>> db;
    if (!path) return db;
    FILE *f = fopen(path, "r");
    if (!f) return db;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char dev[96], st[64];
        int dt, dims, halo, tx, ty, tz, vec, flags, status;
        long long e0, e1, e2;
        double ms;
        if (sscanf(line, "%95[^,],%63[^,],%d,%d,%lld,%lld,%lld,%d,%d,%d,%d,%d,%d,%lf,%d", dev, st, &dt,
                   &dims, &e0, &e1, &e2, &halo, &tx, &ty, &tz, &vec, &flags, &ms, &status) != 15)
            continue; // header or a torn last line
        db[Key{dev, st, dt, dims, e0, e1, e2, halo}][CfgT(tx, ty, tz, vec, flags)] = {ms, status};
    }
    fclose(f);
    return db;
}

std::mutex g_db_mu;

void append_db(const char *path, const Key &k, const uhs_config &c, double ms, int status)
{
    if (!path) return;
    std::lock_guard<std::mutex> lk(g_db_mu);
    FILE *f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s,%s,%d,%d,%lld,%lld,%lld,%d,%d,%d,%d,%d,%d,%.6f,%d\n", k.dev.c_str(), k.stencil.c_str(),
            k.dtype, k.dims, k.e0, k.e1, k.e2, k.halo, c.tile_x, c.tile_y, c.tile_z, c.vec, c.flags, ms,
            status);
    fclose(f);
}

Key problem_key(const Plan &pl, int dims, const int64_t *ext, int halo, int dtype)
{
    Key k;
    k.dev = device_tag();
    k.stencil = pl.prog >= 0 ? kBuiltins[pl.prog].case_name : "custom";
    k.dtype = dtype;
    k.dims = dims;
    k.e0 = ext[0];
    k.e1 = ext[1];
    k.e2 = dims == 3 ? ext[2] : 1;
    k.halo = halo;
    return k;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The search's timer: the GPU (uhs_time_config), or a caller's function set
// by uhs_tune_set_timer (a mock backend for tests of the search logic).
uhs_tune_timer_fn g_timer = nullptr;
void *g_timer_user = nullptr;

} // namespace

// The candidate space of a plan, filtered by uhs_check_config.
std::vector<uhs_config> uhs_config_space(const Plan &pl)
{
    std::vector<uhs_config> out;
    if (pl.family == FAM_GATHER) return out;
    static const int txs3[] = {32, 64, 128, 256, 512};
    static const int tys3[] = {1, 2, 4, 8, 16};
    static const int tzs[] = {16, 32, 64, 128, 256, 512};
    static const int txs2[] = {64, 128, 256, 512, 1024};
    const std::vector<int> vecs = pl.family == FAM_STAR3D ? std::vector<int>{1, 2, 4, 8} : std::vector<int>{1, 2, 4};
    const bool d3 = pl.args.dims == 3;
    std::vector<int> flagset{0, UHS_CFG_NT_STORE, UHS_CFG_PREFETCH2, UHS_CFG_NT_STORE | UHS_CFG_PREFETCH2,
                             UHS_CFG_PREFETCH4, UHS_CFG_NT_STORE | UHS_CFG_PREFETCH4,
                             UHS_CFG_ROWRING, UHS_CFG_NT_STORE | UHS_CFG_ROWRING};
    if (pl.family == FAM_BOX3D) {
        flagset.clear();
        for (int base : {0, UHS_CFG_PREFETCH2, UHS_CFG_ROWS2, UHS_CFG_PLANES2, UHS_CFG_ROWS2 | UHS_CFG_PLANES2})
            for (int nt : {0, UHS_CFG_NT_STORE}) flagset.push_back(base | nt);
    }
    if (pl.family == FAM_STAR3D)
        flagset = {0, UHS_CFG_NT_STORE, UHS_CFG_PREFETCH2, UHS_CFG_NT_STORE | UHS_CFG_PREFETCH2};
    char why[160];
    for (int tx : (d3 ? std::vector<int>(txs3, txs3 + 5) : std::vector<int>(txs2, txs2 + 5)))
        for (int ty : (d3 ? std::vector<int>(tys3, tys3 + 5) : std::vector<int>{1}))
            for (int tz : tzs)
                for (int v : vecs)
                    for (int fl : flagset) {
                        uhs_config c{tx, ty, tz, v, fl};
                        if (uhs_check_config(pl, c, why, sizeof why) != UHS_OK) continue;
                        // Skip slabs much longer than the streamed axis: they
                        // time the same as the axis length.
                        if (tz > 2 * pl.args.n0 && tz > 16) continue;
                        out.push_back(c);
                    }
    if (pl.family == FAM_STAR3D) { // two steps per launch: output tile + ghost zone of 32/64/128 columns
        const int R = pl.radius;
        for (int bx : {32, 64, 128})
            for (int by : {6, 8, 10, 12, 16})
                for (int tz : {64, 128, 256})
                    for (int v : {1, 2})
                        for (int fl : {UHS_CFG_TEMPORAL2, UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE}) {
                            if (by <= 2 * R) continue;
                            uhs_config c{bx - 2 * R, by - 2 * R, tz, v, fl};
                            if (uhs_check_config(pl, c, why, sizeof why) != UHS_OK) continue;
                            if (tz > 2 * pl.args.n0 && tz > 16) continue;
                            out.push_back(c);
                        }
    }
    if (pl.family == FAM_BOX3D && uhs_find_box3d_tb2(pl.prog, pl.args.dtype, 1)) { // two steps per launch
        const int R = pl.radius;
        for (int bx : {32, 64, 128})
            for (int by : {4, 6, 8, 10, 12, 16})
                for (int tz : {16, 32, 64, 128})
                    for (int v : {1, 2, 4})
                        for (int fl : {UHS_CFG_TEMPORAL2, UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE}) {
                            if (by <= 2 * R) continue;
                            uhs_config c{bx - 2 * R, by - 2 * R, tz, v, fl};
                            if (uhs_check_config(pl, c, why, sizeof why) != UHS_OK) continue;
                            if (tz > 2 * pl.args.n0 && tz > 16) continue;
                            out.push_back(c);
                        }
    }
    if (pl.family == FAM_WIN2D) { // two steps per launch: output columns + ghost zone of 64 .. 1024 columns
        const int R = pl.radius;
        for (int bx : {64, 128, 256, 512, 1024})
            for (int tz : {16, 32, 64, 128})
                for (int v : {1, 2, 4})
                    for (int fl : {UHS_CFG_TEMPORAL2, UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE,
                                   UHS_CFG_TEMPORAL2 | UHS_CFG_ROWRING,
                                   UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE | UHS_CFG_ROWRING}) {
                        uhs_config c{bx - 2 * R, 1, tz, v, fl};
                        if (uhs_check_config(pl, c, why, sizeof why) != UHS_OK) continue;
                        if (tz > 2 * pl.args.n0 && tz > 16) continue;
                        out.push_back(c);
                    }
        // K = tile_y steps per launch (stars, radius-1 and radius-2 boxes):
        // output columns + (K-1)r each side
        for (int K : {3, 4, 5, 6, 8})
            for (int bx : {64, 128, 256, 512, 1024})
                for (int tz : {16, 32, 64, 128})
                    for (int v : {1, 2, 4})
                        for (int fl : {UHS_CFG_TEMPORALK, UHS_CFG_TEMPORALK | UHS_CFG_NT_STORE}) {
                            uhs_config c{bx - 2 * (K - 1) * R, K, tz, v, fl};
                            if (c.tile_x < 1 || uhs_check_config(pl, c, why, sizeof why) != UHS_OK) continue;
                            if (tz > 2 * pl.args.n0 && tz > 16) continue;
                            out.push_back(c);
                        }
    }
    if (pl.family == FAM_STAR3D) // register-tiled variants
        for (int rows : {2, 4, 8})
            for (int waves : {1, 2, 4, 8})
                for (int tz : tzs)
                    for (int fl : {UHS_CFG_REGTILE, UHS_CFG_REGTILE | UHS_CFG_NT_STORE}) {
                        uhs_config c{64, rows * waves, tz, rows, fl};
                        if (uhs_check_config(pl, c, why, sizeof why) != UHS_OK) continue;
                        if (tz > 2 * pl.args.n0 && tz > 16) continue;
                        out.push_back(c);
                    }
    // The untuned default is always a candidate (its slab depth is shortened
    // for small grids, outside the grid of tile sizes above), so a search
    // never picks something slower than not searching, timing noise aside.
    const uhs_config d = uhs_heuristic_config(pl);
    if (pl.family != FAM_GATHER &&
        std::none_of(out.begin(), out.end(), [&](const uhs_config &c) { return ct(c) == ct(d); }))
        out.push_back(d);
    return out;
}

// Times `reps` sweeps of one configuration (after one warm-up sweep).  A
// two-steps-per-launch configuration is timed over a multiple of 4 steps
// (an even number of two-step launches, no leading single steps), a
// K-steps-per-launch one over a multiple of K (of 2K for even K).
int uhs_time_config(Plan &pl, void *A, void *B, const uhs_config &c, int reps, hipStream_t st,
                    double *ms)
{
    Plan p = pl;
    char why[160];
    if (uhs_check_config(p, c, why, sizeof why) != UHS_OK) return UHS_E_CONFIG;
    if (uhs_plan_apply(p, c) != UHS_OK) return UHS_E_CONFIG;
    // K = tile_y steps per launch: 2K when K is even (an even number of
    // launches), K when odd.
    const int per = (c.flags & UHS_CFG_TEMPORALK)   ? (c.tile_y % 2 ? c.tile_y : 2 * c.tile_y)
                    : (c.flags & UHS_CFG_TEMPORAL2) ? 4
                                                    : 1;
    reps = (reps + per - 1) / per * per;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return UHS_E_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return UHS_E_HIP;
    }
    int rc = uhs_run_plan(p, A, B, per, st);
    if (rc == UHS_OK) {
        (void)hipEventRecord(e0, st);
        rc = uhs_run_plan(p, A, B, reps, st);
        (void)hipEventRecord(e1, st);
    }
    float t = 0.f;
    if (rc == UHS_OK) {
        hipError_t e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) rc = UHS_E_HIP;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return rc;
}

// Search over `cands` (all of them, or a pruned list), resuming from db.
int uhs_search(Plan &pl, void *A, void *B, int dims, const int64_t *ext, int halo, int dtype,
               const std::vector<uhs_config> &cands, int reps, const char *db_path, hipStream_t st,
               uhs_tune_result *res)
{
    const double t0 = now_s();
    const Key key = problem_key(pl, dims, ext, halo, dtype);
    auto db = load_db(db_path);
    auto &known = db[key];
    uhs_tune_result r{};
    r.best_ms = 1e30;
    for (const uhs_config &c : cands) {
        double ms = 0;
        int status;
        auto it = known.find(ct(c));
        if (it != known.end()) {
            ms = it->second.first;
            status = it->second.second;
            r.reused++;
        } else {
            if (g_timer) {
                Plan p = pl;
                char why[160];
                status = uhs_check_config(p, c, why, sizeof why) != UHS_OK ? UHS_E_CONFIG
                                                                          : g_timer(&c, reps, &ms, g_timer_user);
            } else {
                status = uhs_time_config(pl, A, B, c, reps, st, &ms);
            }
            if (status == UHS_E_HIP) {
                // A launch failure leaves the stream unusable: stop here.
                (void)hipGetLastError();
                append_db(db_path, key, c, 0.0, status);
                return uhs_set_error(UHS_E_HIP, "a configuration failed to launch; search stopped");
            }
            append_db(db_path, key, c, ms, status);
        }
        if (status != UHS_OK) {
            r.failed++;
            continue;
        }
        r.evaluated++;
        if (ms < r.best_ms) {
            r.best_ms = ms;
            r.best = c;
        }
    }
    r.seconds = now_s() - t0;
    if (r.evaluated == 0) return uhs_set_error(UHS_E_CONFIG, "no configuration could be timed");
    uhs_store_tuned(uhs_config_key(pl.family, pl.prog, dtype, dims, ext, halo), r.best);
    if (res) *res = r;
    return UHS_OK;
}

extern "C" {

int uhs_tune_set_timer(uhs_tune_timer_fn fn, void *user)
{
    g_timer = fn;
    g_timer_user = fn ? user : nullptr;
    return UHS_OK;
}

int uhs_tune_sweep(void *A, void *B, int dims, const int64_t *ext, int halo, const uhs_stencil *st,
                   int dtype, int reps, const char *db_path, uhs_tune_result *out, uhs_stream_t stream)
{
    if (!A || !B || A == B) return uhs_set_error(UHS_E_INVAL, "tuning needs two distinct grids");
    if (reps < 1) reps = 1;
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, nullptr, &pl);
    if (rc) return rc;
    if (pl.family == FAM_GATHER) return uhs_set_error(UHS_E_UNSUPPORTED, "the general kernel has no tiling to tune");
    return uhs_search(pl, A, B, dims, ext, halo, dtype, uhs_config_space(pl), reps, db_path,
                      (hipStream_t)stream, out);
}

int uhs_tune_sweep_list(void *A, void *B, int dims, const int64_t *ext, int halo, const uhs_stencil *st,
                        int dtype, const uhs_config *cands, int ncand, int reps, const char *db_path,
                        uhs_tune_result *out, uhs_stream_t stream)
{
    if (!A || !B || A == B || !cands || ncand < 1)
        return uhs_set_error(UHS_E_INVAL, "tuning needs two distinct grids and candidates");
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, nullptr, &pl);
    if (rc) return rc;
    if (pl.family == FAM_GATHER) return uhs_set_error(UHS_E_UNSUPPORTED, "the general kernel has no tiling to tune");
    std::vector<uhs_config> c(cands, cands + ncand);
    return uhs_search(pl, A, B, dims, ext, halo, dtype, c, reps < 1 ? 1 : reps, db_path, (hipStream_t)stream, out);
}

int uhs_tune_space(int dims, const int64_t *ext, int halo, const uhs_stencil *st, int dtype,
                   uhs_config *out, int cap)
{
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, nullptr, &pl);
    if (rc) return rc;
    const auto sp = uhs_config_space(pl);
    if (out)
        for (int i = 0; i < (int)sp.size() && i < cap; i++) out[i] = sp[i];
    return (int)sp.size();
}

int uhs_tune_time(void *A, void *B, int dims, const int64_t *ext, int halo, const uhs_stencil *st,
                  int dtype, const uhs_config *cfg, int reps, double *ms, uhs_stream_t stream)
{
    if (!A || !B || !cfg || !ms) return uhs_set_error(UHS_E_INVAL, "null argument");
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, cfg, &pl);
    if (rc) return rc;
    return uhs_time_config(pl, A, B, *cfg, reps < 1 ? 1 : reps, (hipStream_t)stream, ms);
}

int uhs_tune_load_db(const char *db_path)
{
    if (!db_path) return uhs_set_error(UHS_E_INVAL, "null path");
    FILE *f = fopen(db_path, "r");
    if (!f) return uhs_set_error(UHS_E_IO, "cannot open the tuning database");
    fclose(f);
    const std::string dev = device_tag();
    int loaded = 0;
    for (auto &kv : load_db(db_path)) {
        const Key &k = kv.first;
        if (k.dev != dev || k.stencil == "custom") continue;
        double best = 1e30;
        CfgT bc{};
        for (auto &c : kv.second)
            if (c.second.second == UHS_OK && c.second.first < best) {
                best = c.second.first;
                bc = c.first;
            }
        if (best >= 1e30) continue;
        const int prog = uhs_builtin_find(k.stencil.c_str());
        if (prog < 0) continue;
        const BuiltinStencil &b = kBuiltins[prog];
        const int64_t ext[3] = {k.e0, k.e1, k.e2};
        uhs_config c{std::get<0>(bc), std::get<1>(bc), std::get<2>(bc), std::get<3>(bc), std::get<4>(bc)};
        // Only entries this build can still run: the database outlives
        // builds whose kernels and legality rules differ.
        const uhs_stencil st{b.dims, b.npoints, b.off, b.coef};
        Plan pl;
        if (k.dims != b.dims || uhs_make_plan(k.dims, ext, k.halo, &st, k.dtype, &c, &pl) != UHS_OK) continue;
        uhs_store_tuned(uhs_config_key(pl.family, pl.prog, k.dtype, k.dims, ext, k.halo), c);
        loaded++;
    }
    return loaded;
}

int uhs_tune_clear(void)
{
    uhs_clear_tuned();
    return UHS_OK;
}

} // extern "C"
