// C ABI of the sweep library (include/uhs.h): argument checks, stencil ->
// kernel dispatch, configuration choice, the ping-pong time loop, and the
// device plumbing the tests and the benchmark use.
#include "uhs.h"
#include "uhs_internal.h"
#include "uhs_host.h"
#include "uhs_util.h"

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <mutex>
#include <string>
#include <unordered_map>

namespace {
thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(UHS_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

const char *family_name(int f)
{
    static const char *names[] = {"star3d", "box3d", "win2d", "gather"};
    return f >= 0 && f < 4 ? names[f] : "?";
}

std::mutex g_tuned_mu;
std::unordered_map<std::string, uhs_config> g_tuned;
} // namespace

int uhs_set_error(int code, const char *msg)
{
    g_err = msg ? msg : "";
    return code;
}

int uhs_match_builtin(int dims, int npoints, const int *off)
{
    for (int i = 0; i < (int)(sizeof(kBuiltins) / sizeof(kBuiltins[0])); i++) {
        const BuiltinStencil &b = kBuiltins[i];
        if (b.dims != dims || b.npoints != npoints) continue;
        if (memcmp(b.off, off, sizeof(int) * npoints * dims) == 0) return i;
    }
    return -1;
}

std::string uhs_config_key(int family, int prog, int dtype, int dims, const int64_t *ext, int halo)
{
    char buf[128];
    snprintf(buf, sizeof buf, "%d/%d/%d/%lldx%lldx%lld/h%d", family, prog, dtype, (long long)ext[0],
             (long long)ext[1], dims == 3 ? (long long)ext[2] : 1LL, halo);
    return buf;
}

void uhs_store_tuned(const std::string &key, const uhs_config &c)
{
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned[key] = c;
}

bool uhs_lookup_tuned(const std::string &key, uhs_config *c)
{
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned.find(key);
    if (it == g_tuned.end()) return false;
    *c = it->second;
    return true;
}

void uhs_clear_tuned()
{
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned.clear();
}

static long long ceil_div(long long a, long long b) { return (a + b - 1) / b; }

// Workgroups a configuration launches for the given problem.
long long uhs_config_blocks(const SweepArgs &a, const uhs_config &c, int radius)
{
    const long long xw = (c.flags & UHS_CFG_REGTILE) ? 64 - 2 * radius : c.tile_x;
    const long long tiles = ceil_div(a.n2, xw) * (a.dims == 3 ? ceil_div(a.n1, c.tile_y) : 1);
    return tiles * ceil_div(a.n0, c.tile_z);
}

LaunchFn uhs_pick_fn(const Plan &pl, const uhs_config &c)
{
    const int dt = pl.args.dtype;
    switch (pl.family) {
    case FAM_STAR3D:
        if (c.flags & UHS_CFG_TEMPORAL2)
            return uhs_find_star3d_tb2(pl.prog, dt, c.vec, (c.flags & UHS_CFG_PREFETCH2) ? 2 : 1);
        if (c.flags & UHS_CFG_REGTILE) return uhs_find_star3d_rt(pl.prog, dt, c.vec);
    {
        const long long nt = (long long)std::max(1, c.tile_x / std::max(1, c.vec)) * c.tile_y;
        const long long ring = 2LL * pl.radius * (c.tile_x + c.tile_y);
        return uhs_find_star3d(pl.prog, dt, c.vec,
                               (c.flags & UHS_CFG_PREFETCH4) ? 4 : (c.flags & UHS_CFG_PREFETCH2) ? 2 : 1,
                               (int)ceil_div(ring, nt));
    }
    case FAM_BOX3D:
    {
        if (c.flags & UHS_CFG_TEMPORAL2) return uhs_find_box3d_tb2(pl.prog, dt, c.vec);
        const int planes = (c.flags & UHS_CFG_PLANES2) ? 2 : 1;
        if (c.flags & UHS_CFG_ROWS2) return c.vec == 1 ? uhs_find_box3d(pl.prog, dt, 0, 0, planes) : nullptr;
        return uhs_find_box3d(pl.prog, dt, c.vec, (c.flags & UHS_CFG_PREFETCH2) ? 1 : 0, planes);
    }
    case FAM_WIN2D:
        if (c.flags & UHS_CFG_TEMPORALK) return uhs_find_win2d_tbk(pl.prog, dt, c.vec, c.tile_y);
        if (c.flags & UHS_CFG_TEMPORAL2)
            return uhs_find_win2d_tb2(pl.prog, dt, c.vec, (c.flags & UHS_CFG_ROWRING) ? 1 : 0);
        return uhs_find_win2d(pl.prog, dt, c.vec,
                              (c.flags & UHS_CFG_ROWRING)     ? 0
                              : (c.flags & UHS_CFG_PREFETCH4) ? 4
                              : (c.flags & UHS_CFG_PREFETCH2) ? 2
                                                              : 1);
    default: return uhs_find_gather(dt);
    }
}

int uhs_check_config(const Plan &pl, const uhs_config &c, char *why, size_t whylen)
{
    const SweepArgs &a = pl.args;
    const int R = pl.radius;
    auto bad = [&](const std::string &msg) {
        snprintf(why, whylen, "%s", msg.c_str());
        return UHS_E_CONFIG;
    };
    auto str = [](const char *fmt, long long a = 0, long long b = 0) {
        char buf[160];
        snprintf(buf, sizeof buf, fmt, a, b);
        return std::string(buf);
    };
    if (pl.family == FAM_GATHER) return UHS_OK;
    if (c.flags & ~(UHS_CFG_NT_STORE | UHS_CFG_PREFETCH2 | UHS_CFG_REGTILE | UHS_CFG_PREFETCH4 | UHS_CFG_ROWS2 |
                    UHS_CFG_TEMPORAL2 | UHS_CFG_PLANES2 | UHS_CFG_ROWRING | UHS_CFG_TEMPORALK))
        return bad("unknown config flags");
    if ((c.flags & UHS_CFG_ROWRING) &&
        (pl.family != FAM_WIN2D || (c.flags & (UHS_CFG_PREFETCH2 | UHS_CFG_PREFETCH4))))
        return bad("the row ring exists for 2D stencils, without UHS_CFG_PREFETCH2/4");
    if ((c.flags & UHS_CFG_PLANES2) &&
        (pl.family != FAM_BOX3D || (c.flags & UHS_CFG_PREFETCH2)))
        return bad("two planes of tile loads in flight exist for 3D boxes without deeper cell prefetch");
    if (c.tile_x < 1 || c.tile_z < 1 || c.tile_y < 1) return bad("tile sizes must be positive");
    if (c.flags & UHS_CFG_TEMPORALK) {
        // K = tile_y steps per launch, 2D stars (ring form), the radius-2 2D
        // box, whose terms come row by row (sum form), and the radius-1 box
        // (window form): output columns tile_x plus (K-1)r on each side, vec
        // columns per thread, one row per step.
        const long long K = c.tile_y;
        if (pl.family != FAM_WIN2D || (!kBuiltins[pl.prog].star && R > 2))
            return bad("K steps per launch exist for 2D stars and the radius-1 and radius-2 2D boxes only");
        if (c.flags & ~(UHS_CFG_TEMPORALK | UHS_CFG_NT_STORE))
            return bad("K steps per launch combine with UHS_CFG_NT_STORE only");
        if (K != 3 && K != 4 && K != 5 && K != 6 && K != 8)
            return bad(str("K steps per launch: tile_y = K must be 3, 4, 5, 6 or 8 (got %lld)", K));
        const long long bx = c.tile_x + 2 * (K - 1) * R;
        if (c.vec != 1 && c.vec != 2 && c.vec != 4)
            return bad(str("%lld steps per launch need vec 1, 2 or 4 (got %lld)", K, c.vec));
        if (bx % c.vec) return bad(str("extended tile of %lld columns not a multiple of vec %lld", bx, c.vec));
        const long long nt = bx / c.vec;
        if (nt > kMaxThreads) return bad(str("%lld threads per workgroup exceeds %lld", nt, kMaxThreads));
        if (nt < 2 * R) return bad(str("%lld threads cannot stage a radius-%lld ring", nt, R));
        const size_t esz = a.dtype == DT_F64 ? 8 : 4;
        const size_t lds = esz * (size_t)(2 * (bx + 2 * R) + (K - 1) * (2 * bx + 2 * R));
        if (lds > kMaxLdsBytes) return bad(str("%lld bytes of LDS exceeds %lld", (long long)lds, (long long)kMaxLdsBytes));
        if (uhs_config_blocks(a, c, R) > INT_MAX) return bad("too many workgroups");
        if (!uhs_pick_fn(pl, c)) return bad(str("multi-step 2D kernel not compiled for vec %lld", c.vec));
        return UHS_OK;
    }
    if (c.flags & UHS_CFG_TEMPORAL2) {
        if (pl.family == FAM_BOX3D) {
            // Output tile plus an R-wide ghost zone: (tile_x + 2r) / vec *
            // (tile_y + 2r) threads; the level-0 plane with its R-wide ring
            // (corners included) and the level-1 plane double-buffered in LDS.
            if (c.flags & ~(UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE))
                return bad("two steps per launch combines with UHS_CFG_NT_STORE only");
            const long long bx = c.tile_x + 2LL * R, by = c.tile_y + 2LL * R;
            if (c.vec != 1 && c.vec != 2 && c.vec != 4)
                return bad(str("two steps per launch needs vec 1, 2 or 4 (got %lld)", c.vec));
            if (bx % c.vec) return bad(str("tile_x + 2r = %lld not a multiple of vec %lld", bx, c.vec));
            const long long nt = bx / c.vec * by;
            if (nt > kMaxThreads) return bad(str("%lld threads per workgroup exceeds %lld", nt, kMaxThreads));
            const long long ring = 2LL * R * (bx + 2 * R) + 2LL * R * by;
            if (ceil_div(ring, nt) > kBoxTb2RingCells)
                return bad(str("ring of %lld cells needs more than %lld per thread", ring, kBoxTb2RingCells));
            const size_t esz = a.dtype == DT_F64 ? 8 : 4;
            const size_t lds = 2 * esz * (size_t)((bx + 2 * R) * (by + 2 * R) + bx * by);
            if (lds > kMaxLdsBytes) return bad(str("%lld bytes of LDS exceeds %lld", (long long)lds, (long long)kMaxLdsBytes));
            if (uhs_config_blocks(a, c, R) > INT_MAX) return bad("too many workgroups");
            if (!uhs_pick_fn(pl, c)) return bad("two steps per launch exists for the radius-1 3D box only");
            return UHS_OK;
        }
        if (pl.family != FAM_STAR3D && pl.family != FAM_WIN2D)
            return bad("two steps per launch exists for 3D stars, the radius-1 3D box and 2D stencils only");
        if (pl.family == FAM_WIN2D) {
            // Output columns tile_x plus an R-wide ghost zone each side, vec
            // columns per thread; one row per step.  UHS_CFG_ROWRING selects
            // the window form (2r+1 rows of loads in flight in a register
            // ring, a (2r+1)^2 window per column and level); without it a
            // star takes the ring form (2r+2 rows of each column in
            // registers, loads two rows ahead) and a box the sum form (2r+1
            // running sums per column and level).
            if (c.flags & ~(UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE | UHS_CFG_ROWRING))
                return bad("2D two steps per launch combines with UHS_CFG_NT_STORE and UHS_CFG_ROWRING only");
            const long long bx = c.tile_x + 2LL * R;
            if (c.tile_y != 1) return bad("2D stencils need tile_y = 1");
            if (c.vec != 1 && c.vec != 2 && c.vec != 4)
                return bad(str("two steps per launch needs vec 1, 2 or 4 (got %lld)", c.vec));
            if (bx % c.vec) return bad(str("tile_x + 2r = %lld not a multiple of vec %lld", bx, c.vec));
            const long long nt = bx / c.vec;
            if (nt > kMaxThreads) return bad(str("%lld threads per workgroup exceeds %lld", nt, kMaxThreads));
            if (nt < 2 * R) return bad(str("%lld threads cannot stage a radius-%lld ring", nt, R));
            const size_t esz = a.dtype == DT_F64 ? 8 : 4;
            const size_t lds = esz * (size_t)(2 * (bx + 2 * R) + 2 * bx + 2 * R);
            if (lds > kMaxLdsBytes) return bad(str("%lld bytes of LDS exceeds %lld", (long long)lds, (long long)kMaxLdsBytes));
            if (uhs_config_blocks(a, c, R) > INT_MAX) return bad("too many workgroups");
            if (!uhs_pick_fn(pl, c)) return bad(str("two-step 2D kernel not compiled for vec %lld", c.vec));
            return UHS_OK;
        }
        if (c.flags & ~(UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE))
            return bad("two steps per launch combines with UHS_CFG_NT_STORE only");
        const long long bx = c.tile_x + 2LL * R, by = c.tile_y + 2LL * R;
        if (c.vec != 1 && c.vec != 2) return bad(str("two steps per launch needs vec 1 or 2 (got %lld)", c.vec));
        if (bx % c.vec) return bad(str("tile_x + 2r = %lld not a multiple of vec %lld", bx, c.vec));
        const long long nt = bx / c.vec * by;
        if (nt > kMaxThreads) return bad(str("%lld threads per workgroup exceeds %lld", nt, kMaxThreads));
        if (ceil_div(2LL * R * (bx + by), nt) > kTb2RingCells)
            return bad(str("ring of %lld cells needs more than %lld per thread", 2 * R * (bx + by), kTb2RingCells));
        const size_t esz = a.dtype == DT_F64 ? 8 : 4;
        const size_t lds = 2 * esz * (size_t)((bx + 2 * R) * (by + 2 * R) + bx * by);
        if (lds > kMaxLdsBytes) return bad(str("%lld bytes of LDS exceeds %lld", (long long)lds, (long long)kMaxLdsBytes));
        if (uhs_config_blocks(a, c, R) > INT_MAX) return bad("too many workgroups");
        if (!uhs_pick_fn(pl, c)) return bad("two-step kernel not compiled for this stencil");
        return UHS_OK;
    }
    if (c.flags & UHS_CFG_ROWS2) {
        if (pl.family != FAM_BOX3D) return bad("two rows per thread exists for 3D boxes only");
        if (c.vec != 1 || (c.tile_y & 1) || (c.flags & UHS_CFG_PREFETCH2))
            return bad("two rows per thread needs vec 1, an even tile_y and no deeper prefetch");
    }
    if ((c.flags & UHS_CFG_PREFETCH4) && ((c.flags & UHS_CFG_PREFETCH2) || pl.family != FAM_WIN2D))
        return bad("four rows of prefetch is a 2D variant of its own");
    if (c.flags & UHS_CFG_REGTILE) {
        if (pl.family != FAM_STAR3D) return bad("the register-tiled kernel exists for 3D star stencils only");
        if (c.flags & UHS_CFG_PREFETCH2) return bad("the register-tiled kernel has no two-plane prefetch variant");
        if (c.tile_x != 64) return bad("the register-tiled kernel needs tile_x = 64 (one wavefront)");
        if (c.vec != 2 && c.vec != 4 && c.vec != 8) return bad(str("rows per lane must be 2, 4 or 8 (got %lld)", c.vec));
        if (c.tile_y % c.vec) return bad(str("tile_y %lld not a multiple of the rows per lane %lld", c.tile_y, c.vec));
        const long long nt = 64LL * (c.tile_y / c.vec);
        if (nt > kMaxThreads) return bad(str("%lld threads per workgroup exceeds %lld", nt, kMaxThreads));
        if (uhs_config_blocks(a, c, R) > INT_MAX) return bad("too many workgroups");
        if (!uhs_pick_fn(pl, c)) return bad("register-tiled kernel not compiled for this stencil");
        return UHS_OK;
    }
    if (c.vec != 1 && c.vec != 2 && c.vec != 4 && !(c.vec == 8 && pl.family == FAM_STAR3D))
        return bad(str("vec must be 1, 2 or 4, or 8 for 3D stars (got %lld)", c.vec));
    if (c.tile_x % c.vec) return bad(str("tile_x %lld not a multiple of vec %lld", c.tile_x, c.vec));
    if (a.dims == 2 && c.tile_y != 1) return bad("2D stencils need tile_y = 1");
    const long long nt = (long long)(c.tile_x / c.vec) * c.tile_y / ((c.flags & UHS_CFG_ROWS2) ? 2 : 1);
    if (nt > kMaxThreads) return bad(str("%lld threads per workgroup exceeds %lld", nt, kMaxThreads));
    if (nt < 2 * R) return bad(str("%lld threads cannot stage a radius-%lld halo", nt, R));
    const size_t esz = a.dtype == DT_F64 ? 8 : 4;
    size_t lds;
    if (pl.family == FAM_WIN2D) {
        lds = 2 * esz * (c.tile_x + 2 * R);
    } else {
        lds = 2 * esz * (size_t)(c.tile_x + 2 * R) * (c.tile_y + 2 * R);
        if (pl.family == FAM_STAR3D &&
            ceil_div(2LL * R * (c.tile_x + c.tile_y), nt) > kStarHaloPerThread)
            return bad(str("halo ring of %lld cells needs more than %lld per thread",
                           2 * R * (c.tile_x + c.tile_y), kStarHaloPerThread));
        if (pl.family == FAM_BOX3D &&
            ceil_div((long long)(c.tile_x + 2 * R) * (c.tile_y + 2 * R), nt) > box_cells_per_thread(R))
            return bad(str("plane tile of %lld cells needs more than %lld per thread",
                           (c.tile_x + 2 * R) * (c.tile_y + 2 * R), box_cells_per_thread(R)));
    }
    if (lds > kMaxLdsBytes) return bad(str("%lld bytes of LDS exceeds %lld", (long long)lds, (long long)kMaxLdsBytes));
    if (uhs_config_blocks(a, c, R) > INT_MAX) return bad("too many workgroups");
    if ((c.flags & (UHS_CFG_ROWS2 | UHS_CFG_PLANES2 | UHS_CFG_REGTILE | UHS_CFG_TEMPORAL2)) && pl.family == FAM_WIN2D)
        return bad("2D stencils take UHS_CFG_NT_STORE, UHS_CFG_PREFETCH2/4 and UHS_CFG_ROWRING only");
    if (!uhs_pick_fn(pl, c)) return bad(std::string("no ") + family_name(pl.family) + str(" kernel compiled for vec %lld", c.vec));
    return UHS_OK;
}

// Defaults per built-in stencil: the fastest configuration of a timed sweep
// over the configuration space on MI355X at 512^3 / 8192^2 (FP64).
static const uhs_config kBuiltinDefaults[12] = {
    {256, 1, 256, 4, UHS_CFG_ROWRING},                        // 2d1r_star
    {256, 1, 256, 4, UHS_CFG_ROWRING | UHS_CFG_NT_STORE},     // 2d2r_star
    {512, 1, 128, 1, UHS_CFG_ROWRING | UHS_CFG_NT_STORE},     // 2d3r_star
    {1024, 1, 256, 4, UHS_CFG_ROWRING | UHS_CFG_NT_STORE},    // 2d1r_box
    {1024, 1, 128, 4, UHS_CFG_NT_STORE},                      // 2d2r_box
    {256, 1, 64, 1, UHS_CFG_ROWRING},                         // 2d3r_box
    {128, 16, 256, 8, UHS_CFG_NT_STORE},                     // 3d1r_star
    {128, 16, 256, 8, UHS_CFG_NT_STORE},                     // 3d2r_star
    {128, 16, 256, 4, UHS_CFG_NT_STORE | UHS_CFG_PREFETCH2}, // 3d3r_star
    {64, 16, 128, 1, UHS_CFG_ROWS2 | UHS_CFG_PLANES2},       // 3d1r_box
    {32, 16, 64, 1, UHS_CFG_ROWS2 | UHS_CFG_PLANES2 | UHS_CFG_NT_STORE}, // 3d2r_box
    {64, 16, 128, 1, UHS_CFG_ROWS2 | UHS_CFG_PLANES2 | UHS_CFG_NT_STORE}, // 3d3r_box
};

// FP32 defaults (plain and compensated): the fastest configuration of a
// timed sweep of the compensated form at 512^3 / 8192^2.  Four-byte cells
// favour wider tiles and more points per thread than FP64.
static const uhs_config kBuiltinDefaultsF32[12] = {
    {256, 1, 64, 4, 0},                                      // 2d1r_star
    {512, 1, 256, 2, UHS_CFG_ROWRING | UHS_CFG_NT_STORE},    // 2d2r_star
    {256, 1, 256, 2, UHS_CFG_ROWRING},                       // 2d3r_star
    {1024, 1, 64, 4, 0},                                     // 2d1r_box
    {512, 1, 64, 1, UHS_CFG_ROWRING},                        // 2d2r_box
    {256, 1, 64, 1, UHS_CFG_ROWRING | UHS_CFG_NT_STORE},     // 2d3r_box
    {256, 16, 128, 8, 0},                                    // 3d1r_star
    {256, 16, 128, 8, UHS_CFG_NT_STORE},                     // 3d2r_star
    {128, 8, 128, 2, UHS_CFG_NT_STORE | UHS_CFG_PREFETCH2},  // 3d3r_star
    {128, 8, 64, 1, UHS_CFG_ROWS2},                          // 3d1r_box
    {128, 8, 128, 1, UHS_CFG_ROWS2 | UHS_CFG_PLANES2},       // 3d2r_box
    {32, 32, 128, 1, UHS_CFG_ROWS2 | UHS_CFG_NT_STORE},      // 3d3r_box
};

// Defaults that advance two steps per launch (UHS_CFG_TEMPORAL2) or K
// (UHS_CFG_TEMPORALK, K = tile_y, the 2D stars), where a timed sweep at 512^3 / 8192^2
// (T = 20 to 100) found them faster than the one-step default.  The results
// are bitwise those of the one-step kernels.
//   K steps per launch for the 2D stars (T = 240, medians of 3, K of 3, 4,
//   5, 6 and 8 and 108-120 tiles per star and precision): FP64 2d1r_star
//   1318 GSt/s at K = 8 (583 for the two-step default), 2d2r_star 956 at
//   K = 6 (542), 2d3r_star 696 at K = 4 (509); FP32 2143 (K = 8), 1542
//   (6), 1177 (5) against 1071, 1023, 897; FP32 compensated 1811 (8), 1131
//   (6), 791 (5) against 1025, 891, 739.  The radius-2 box (sum form, its
//   25 terms row by row) in FP64: 604 at K = 4 against 510 for two steps;
//   in FP32 (854 against 860) and compensated (453 against 496) it stays at
//   two steps.  The radius-1 box (window form, its centre row last): 1103
//   at K = 6 against 590 (FP64), 1700 at K = 6 against 1066 (FP32), 1096 at
//   K = 4 against 997 (compensated).
//   FP64: 3d1r_star 446 against 337 GSt/s one step (62x14 output tile, 64x16
//   with the ghost zone, 2 columns per thread, 32-plane slabs); 3d2r_star 327
//   against 320 (28x12 output tile, 64-plane slabs); 3d1r_box 373-376
//   against 310-322 (62x14 output tile); the 2D stencils 1.4-1.8x
//   (2d1r_star 588 against 335, 2d2r_star 541 / 323, 2d3r_star 508 / 313,
//   2d1r_box 602 / 335, 2d2r_box 431 / 311).  The radius-3 star stays one
//   step per launch (209 against 293 GSt/s for its best two-step tile), and
//   so does the 2D radius-3 box: its two windows do not fit the registers,
//   and its sum form (170 VGPRs, 3 waves per SIMD) times the same as one
//   step (289 against 286).
//   With two steps per launch 2d1r_box ran faster in the sum form (585 against
//   570 for the window form); it now takes K steps per launch.
//   FP32 plain: 3d1r_star 727 against 627; 3d1r_box 564 against 451-483;
//   2D 1.3-1.8x (2d1r_star 1080
//   against 595, 2d2r_star 1029 / 619, 2d3r_star 909 / 601, 2d1r_box 1099 /
//   605, 2d2r_box 818 / 559).
//   FP32 compensated: 2D 1.1-1.7x (2d1r_star 1058 against 634, 2d2r_star 897
//   / 582, 2d3r_star 747 / 532, 2d1r_box 982 / 590, 2d2r_box 464 / 418).
// 2D stars take the window form (UHS_CFG_ROWRING) or the ring form, whichever
// timed faster; the two are within 1-2% at radius 1 and 2, the ring form is
// 4-17% faster at radius 3.  2D boxes likewise take the window form or the
// sum form (no UHS_CFG_ROWRING).
// Slower with two steps per launch, so absent: the 3D radius-2 and radius-3
// stars in FP32 (537 against 586, 327 against 440 GSt/s), the compensated
// 3d1r_star (630 against 636) and 3d1r_box (347 against 351-361), and the FP32
// 2D radius-3 box (299 against 424).  The 3D radius-2 and radius-3 boxes,
// bound by FP64 FMAs, have no two-step kernel.
static const uhs_config kMultiStepNone = {0, 0, 0, 0, 0};
constexpr int T2 = UHS_CFG_TEMPORAL2, T2NT = UHS_CFG_TEMPORAL2 | UHS_CFG_NT_STORE, RR = UHS_CFG_ROWRING;
constexpr int TK = UHS_CFG_TEMPORALK, TKNT = UHS_CFG_TEMPORALK | UHS_CFG_NT_STORE;
static const uhs_config kBuiltinMultiStepF64[12] = {
    {242, 8, 64, 1, TKNT},      // 2d1r_star
    {236, 6, 64, 1, TK},        // 2d2r_star
    {238, 4, 64, 1, TK},        // 2d3r_star
    {246, 6, 64, 1, TKNT},      // 2d1r_box (window form)
    {244, 4, 64, 1, TKNT},      // 2d2r_box (sum form)
    kMultiStepNone,             // 2d3r_box
    {62, 14, 32, 2, T2},        // 3d1r_star
    {28, 12, 64, 2, T2NT},      // 3d2r_star
    kMultiStepNone,             // 3d3r_star
    {62, 14, 32, 2, T2NT},      // 3d1r_box
    kMultiStepNone, kMultiStepNone,
};
static const uhs_config kBuiltinMultiStepF32[12] = {
    {242, 8, 64, 1, TK},        // 2d1r_star
    {236, 6, 64, 1, TK},        // 2d2r_star
    {488, 5, 64, 1, TK},        // 2d3r_star
    {118, 6, 64, 1, TKNT},      // 2d1r_box (window form)
    {252, 1, 64, 1, T2 | RR},   // 2d2r_box
    kMultiStepNone,             // 2d3r_box
    {62, 14, 32, 2, T2},        // 3d1r_star
    kMultiStepNone, kMultiStepNone,
    {30, 14, 32, 2, T2NT},      // 3d1r_box
    kMultiStepNone, kMultiStepNone,
};
static const uhs_config kBuiltinMultiStepF32K[12] = {
    {242, 8, 64, 1, TKNT},      // 2d1r_star
    {236, 6, 64, 1, TK},        // 2d2r_star
    {488, 5, 64, 1, TKNT},      // 2d3r_star
    {250, 4, 64, 1, TKNT},      // 2d1r_box (window form)
    {124, 1, 64, 1, T2NT | RR}, // 2d2r_box
    kMultiStepNone, kMultiStepNone, kMultiStepNone, kMultiStepNone, kMultiStepNone, kMultiStepNone,
    kMultiStepNone,
};

uhs_config uhs_heuristic_config(const Plan &pl, bool one_step)
{
    uhs_config c{};
    const SweepArgs &a = pl.args;
    if (pl.family == FAM_GATHER) return c;
    c = a.dtype == DT_F64 ? kBuiltinDefaults[pl.prog] : kBuiltinDefaultsF32[pl.prog];
    char why[128];
    const uhs_config *two = a.dtype == DT_F64   ? &kBuiltinMultiStepF64[pl.prog]
                            : a.dtype == DT_F32 ? &kBuiltinMultiStepF32[pl.prog]
                                                : &kBuiltinMultiStepF32K[pl.prog];
    if (!one_step && two->flags && uhs_check_config(pl, *two, why, sizeof why) == UHS_OK) c = *two;
    if (uhs_check_config(pl, c, why, sizeof why) != UHS_OK) {
        // e.g. an FP32 form whose vec variant is not compiled
        c.vec = 1;
        c.flags = 0;
    }
    // Small grids: shorten the streamed slab until the launch fills the
    // chip -- one workgroup per CU for the HBM-bound kernels (the defaults
    // above launch >= 256 at 512^3 / 8192^2 and are fastest there), four per
    // CU for the VALU-bound boxes, which need the occupancy -- but not below
    // the point where re-reading (star) or re-computing (box: 2R planes of
    // partial sums per slab) the slab's 2R boundary planes costs > 25%.
    const int tz_min = std::max(8, 8 * pl.radius);
    const long long want = pl.family == FAM_BOX3D ? 1024 : 256;
    while (c.tile_z / 2 >= tz_min && uhs_config_blocks(a, c, pl.radius) < want) c.tile_z /= 2;
    return c;
}

int uhs_make_plan(int dims, const int64_t *ext, int halo, const uhs_stencil *st, int dtype,
                  const uhs_config *cfg, Plan *pl)
{
    if (!ext || !st) return fail(UHS_E_INVAL, "null extents or stencil");
    if (dims != 2 && dims != 3) return fail(UHS_E_INVAL, "dims must be 2 or 3 (got %d)", dims);
    if (st->dims != dims) return fail(UHS_E_INVAL, "stencil has %d dims, grid %d", st->dims, dims);
    if (st->npoints < 1 || st->npoints > UHS_MAX_POINTS)
        return fail(UHS_E_INVAL, "npoints %d outside [1, %d]", st->npoints, UHS_MAX_POINTS);
    if (!st->offsets || !st->coeffs) return fail(UHS_E_INVAL, "null offsets or coefficients");
    if (dtype < DT_F64 || dtype > DT_F32K) return fail(UHS_E_INVAL, "bad dtype %d", dtype);
    if (halo < 0) return fail(UHS_E_INVAL, "negative halo");
    for (int k = 0; k < dims; k++)
        if (ext[k] < 1 || ext[k] + 2LL * halo > INT_MAX / 2)
            return fail(UHS_E_INVAL, "extent %lld on axis %d out of range", (long long)ext[k], k);
    int radius = 0;
    for (int i = 0; i < st->npoints * dims; i++) radius = std::max(radius, std::abs(st->offsets[i]));
    if (radius > UHS_MAX_RADIUS) return fail(UHS_E_UNSUPPORTED, "radius %d > %d", radius, UHS_MAX_RADIUS);
    if (radius > halo) return fail(UHS_E_INVAL, "halo %d smaller than the stencil radius %d", halo, radius);

    *pl = Plan{};
    SweepArgs &a = pl->args;
    a.dims = dims;
    a.halo = halo;
    if (dims == 3) {
        a.n0 = (int)ext[0]; a.n1 = (int)ext[1]; a.n2 = (int)ext[2];
        a.sy = ext[2] + 2LL * halo;
        a.sz = (ext[1] + 2LL * halo) * a.sy;
        if (a.sz > INT_MAX) return fail(UHS_E_INVAL, "plane of %lld cells exceeds 2^31", a.sz);
    } else {
        a.n0 = (int)ext[0]; a.n1 = 1; a.n2 = (int)ext[1];
        a.sy = ext[1] + 2LL * halo;
        a.sz = (ext[0] + 2LL * halo) * a.sy;
    }
    a.npoints = st->npoints;
    a.off = st->offsets;
    a.coef = st->coeffs;
    a.dtype = dtype;
    pl->radius = radius;
    pl->prog = uhs_match_builtin(dims, st->npoints, st->offsets);
    if (pl->prog >= 0) {
        const BuiltinStencil &b = kBuiltins[pl->prog];
        pl->family = dims == 2 ? FAM_WIN2D : (b.star ? FAM_STAR3D : FAM_BOX3D);
        pl->radius = b.radius;
    } else {
        pl->family = FAM_GATHER;
        if (st->npoints > kGatherMaxPoints)
            return fail(UHS_E_UNSUPPORTED,
                        "stencil of %d terms has no specialised kernel and exceeds the "
                        "%d-term general kernel",
                        st->npoints, kGatherMaxPoints);
        if ((long long)radius * (dims == 3 ? a.sz : a.sy) > INT_MAX)
            return fail(UHS_E_INVAL, "term offsets exceed 2^31 cells");
    }

    uhs_config c;
    char why[256] = "";
    if (cfg) {
        c = *cfg;
    } else if (!uhs_lookup_tuned(uhs_config_key(pl->family, pl->prog, dtype, dims, ext, halo), &c) ||
               uhs_check_config(*pl, c, why, sizeof why) != UHS_OK) {
        // No tuned entry, or one this build's kernels no longer accept (a
        // database written by another version): the heuristic choice.
        c = uhs_heuristic_config(*pl);
    }
    if (uhs_check_config(*pl, c, why, sizeof why) != UHS_OK)
        return fail(UHS_E_CONFIG, "config {%d,%d,%d,%d,%d} rejected: %s", c.tile_x, c.tile_y,
                    c.tile_z, c.vec, c.flags, why);
    return uhs_plan_apply(*pl, c);
}

int uhs_plan_apply(Plan &pl, const uhs_config &c)
{
    SweepArgs &a = pl.args;
    a.tx = c.tile_x; a.ty = c.tile_y; a.tz = c.tile_z; a.vec = c.vec; a.flags = c.flags;
    pl.cfg = c;
    pl.fn = uhs_pick_fn(pl, c);
    pl.fn1 = nullptr;
    if (!pl.fn) return fail(UHS_E_UNSUPPORTED, "no kernel for this stencil and precision");
    if (c.flags & (UHS_CFG_TEMPORAL2 | UHS_CFG_TEMPORALK)) {
        Plan one = pl;
        const uhs_config c1 = uhs_heuristic_config(one, true);
        one.args.tx = c1.tile_x; one.args.ty = c1.tile_y; one.args.tz = c1.tile_z;
        one.args.vec = c1.vec; one.args.flags = c1.flags;
        pl.fn1 = uhs_pick_fn(one, c1);
        pl.args1 = one.args;
        if (!pl.fn1) return fail(UHS_E_UNSUPPORTED, "no single-step kernel for the leading steps");
    }
    pl.fnk[0] = pl.fnk[1] = nullptr;
    pl.kk[0] = pl.kk[1] = 0;
    if (c.flags & UHS_CFG_TEMPORALK) {
        int n = 0;
        for (int k : {5, 3}) {
            if (k >= c.tile_y) continue;
            const uhs_config ck{c.tile_x, k, c.tile_z, c.vec, c.flags};
            char why[160];
            Plan p2 = pl;
            if (uhs_check_config(p2, ck, why, sizeof why) != UHS_OK) continue;
            p2.args.ty = k;
            pl.fnk[n] = uhs_pick_fn(p2, ck);
            if (!pl.fnk[n]) continue;
            pl.argsk[n] = p2.args;
            pl.kk[n] = k;
            n++;
        }
    }
    return UHS_OK;
}

// Steps advanced per launch of a multi-step plan (UHS_CFG_TEMPORAL2, or
// UHS_CFG_TEMPORALK with K = tile_y):
// `ones` single steps, then `launches` K-step launches.  Every launch reads
// one grid and writes the other, so the result lands where the one-step
// ping-pong leaves it (A for even T, B for odd T) when each launch advances
// an odd number of steps, or, for even K, when the K-step launches are even
// in number.
static void multi_step_schedule(int T, int K, int *ones, int *launches)
{
    *launches = T / K;
    *ones = T % K;
    if (K % 2 == 0 && *launches % 2) {
        (*launches)--;
        *ones += K;
    }
}

int uhs_run_plan(Plan &pl, void *A, void *B, int timesteps, hipStream_t stream)
{
    if (!A || !B) return fail(UHS_E_INVAL, "null grid pointer");
    if (A == B) return fail(UHS_E_INVAL, "A and B must be distinct grids");
    if (timesteps < 0) return fail(UHS_E_INVAL, "negative timestep count");
    if (pl.cfg.flags & (UHS_CFG_TEMPORAL2 | UHS_CFG_TEMPORALK)) {
        // T mod K single steps, then K-step launches.  Launches of an odd
        // number of steps leave every state in the grid of its parity, as in
        // the one-step ping-pong; for even K the K-step launches are even in
        // number.
        const int K = (pl.cfg.flags & UHS_CFG_TEMPORALK) ? pl.cfg.tile_y : 2;
        int ones, twos;
        multi_step_schedule(timesteps, K, &ones, &twos);
        // The leading steps: launches of an odd K' < K of the same tile
        // (largest first), then single sweeps.  Odd launches and single
        // sweeps keep every state in the grid of its parity, so they may
        // precede the K-step launches in any mix.
        int l = 0, k0 = 0; // launches so far, steps so far
        auto multi = [&](LaunchFn fn, SweepArgs &a) {
            a.in = (l & 1) ? B : A;
            a.out = (l & 1) ? A : B;
            a.hA = A;
            a.hB = B;
            a.k0 = k0;
            return fn(a, stream);
        };
        for (int j = 0; j < 2; j++)
            while (pl.fnk[j] && ones >= pl.kk[j]) {
                hipError_t e = multi(pl.fnk[j], pl.argsk[j]);
                if (e != hipSuccess) return hip_fail(e, "sweep launch");
                ones -= pl.kk[j];
                k0 += pl.kk[j];
                l++;
            }
        for (; ones > 0; ones--) {
            pl.args1.in = (l & 1) ? B : A;
            pl.args1.out = (l & 1) ? A : B;
            hipError_t e = pl.fn1(pl.args1, stream);
            if (e != hipSuccess) return hip_fail(e, "sweep launch");
            k0++;
            l++;
        }
        // Launch l reads the grid written by launch l-1; halo cells of each
        // state come from the grid the one-step ping-pong would hold it in.
        for (int j = 0; j < twos; j++) {
            hipError_t e = multi(pl.fn, pl.args);
            if (e != hipSuccess) return hip_fail(e, "sweep launch");
            k0 += K;
            l++;
        }
        return UHS_OK;
    }
    for (int k = 1; k <= timesteps; k++) {
        pl.args.in = (k & 1) ? A : B;
        pl.args.out = (k & 1) ? B : A;
        hipError_t e = pl.fn(pl.args, stream);
        if (e != hipSuccess) return hip_fail(e, "sweep launch");
    }
    return UHS_OK;
}

// Stencils outside the built-in forms with at most UHS_MAX_PROGRAM_TERMS
// terms run as a one-kernel program over the interior: the streaming program
// kernel computes the same sum in the same order (first term a product, then
// fused multiply-adds, or their compensated form) as the one-thread-per-point
// general kernel, staged through LDS instead of the caches.
static bool as_program(const Plan &pl, const uhs_stencil *st)
{
    return pl.family == FAM_GATHER && st->npoints <= UHS_MAX_PROGRAM_TERMS;
}

static int sweep_as_program(void *A, void *B, int dims, const int64_t *ext, int halo, const uhs_stencil *st,
                            int timesteps, int dtype, hipStream_t stream)
{
    int src[UHS_MAX_PROGRAM_TERMS];
    for (int t = 0; t < st->npoints; t++) src[t] = -1;
    uhs_program pg{};
    pg.dims = dims;
    pg.output = *st;
    pg.sources = src;
    int64_t shape[3];
    for (int a = 0; a < dims; a++) {
        shape[a] = ext[a] + 2 * (int64_t)halo;
        pg.lo[a] = halo;
        pg.hi[a] = halo + ext[a];
    }
    if (dtype == DT_F64)
        return uhs_program_sweep_ex_f64((double *)A, (double *)B, shape, &pg, timesteps, UHS_PROGRAM_AUTO, nullptr,
                                        stream);
    return uhs_program_sweep_ex_f32((float *)A, (float *)B, shape, &pg, timesteps, dtype == DT_F32K ? 1 : 0,
                                    UHS_PROGRAM_AUTO, nullptr, stream);
}

static int sweep_or_program(Plan &pl, void *A, void *B, int dims, const int64_t *ext, int halo,
                            const uhs_stencil *st, int timesteps, hipStream_t stream)
{
    if (as_program(pl, st)) {
        if (!A || !B) return fail(UHS_E_INVAL, "null grid pointer");
        if (A == B) return fail(UHS_E_INVAL, "A and B must be distinct grids");
        if (timesteps < 0) return fail(UHS_E_INVAL, "negative timestep count");
        return sweep_as_program(A, B, dims, ext, halo, st, timesteps, pl.args.dtype, stream);
    }
    return uhs_run_plan(pl, A, B, timesteps, stream);
}

// Setters of the LDS-poison flag, one per translation unit that has kernels
// (each registers its own when the library loads).
static std::vector<hipError_t (*)(int)> &lds_poison_setters()
{
    static std::vector<hipError_t (*)(int)> v;
    return v;
}

int uhs_register_lds_poison_setter(hipError_t (*set)(int on))
{
    lds_poison_setters().push_back(set);
    return 0;
}

extern "C" {

int uhs_sweep_f64(double *A, double *B, int dims, const int64_t *ext, int halo,
                  const uhs_stencil *st, int timesteps, const uhs_config *cfg, uhs_stream_t stream)
{
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, DT_F64, cfg, &pl);
    if (rc) return rc;
    return sweep_or_program(pl, A, B, dims, ext, halo, st, timesteps, (hipStream_t)stream);
}

int uhs_sweep_f32(float *A, float *B, int dims, const int64_t *ext, int halo,
                  const uhs_stencil *st, int timesteps, int compensated, const uhs_config *cfg,
                  uhs_stream_t stream)
{
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, compensated ? DT_F32K : DT_F32, cfg, &pl);
    if (rc) return rc;
    return sweep_or_program(pl, A, B, dims, ext, halo, st, timesteps, (hipStream_t)stream);
}

int uhs_plan_kind(int dims, const int64_t *ext, int halo, const uhs_stencil *st, int dtype,
                  const uhs_config *cfg, char *buf, size_t buflen)
{
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, cfg, &pl);
    if (rc) return rc;
    if (buf && buflen) snprintf(buf, buflen, "%s", as_program(pl, st) ? "program" : family_name(pl.family));
    return UHS_OK;
}

int uhs_default_config(int dims, const int64_t *ext, int halo, const uhs_stencil *st, int dtype,
                       uhs_config *out)
{
    if (!out) return fail(UHS_E_INVAL, "null output");
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, nullptr, &pl);
    if (rc) return rc;
    *out = pl.cfg;
    return UHS_OK;
}

int uhs_one_step_config(int dims, const int64_t *ext, int halo, const uhs_stencil *st, int dtype,
                        uhs_config *out)
{
    if (!out) return fail(UHS_E_INVAL, "null output");
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, nullptr, &pl);
    if (rc) return rc;
    *out = uhs_heuristic_config(pl, true);
    return UHS_OK;
}

int uhs_builtin_count(void) { return (int)(sizeof(kBuiltins) / sizeof(kBuiltins[0])); }

int uhs_builtin_info(int i, const char **case_name, const char **symbol, int *dims, int *radius,
                     int *npoints, const int **offsets, const double **coeffs)
{
    if (i < 0 || i >= uhs_builtin_count()) return fail(UHS_E_INVAL, "no built-in stencil %d", i);
    const BuiltinStencil &b = kBuiltins[i];
    if (case_name) *case_name = b.case_name;
    if (symbol) *symbol = b.symbol;
    if (dims) *dims = b.dims;
    if (radius) *radius = b.radius;
    if (npoints) *npoints = b.npoints;
    if (offsets) *offsets = b.off;
    if (coeffs) *coeffs = b.coef;
    return UHS_OK;
}

int uhs_builtin_find(const char *case_name)
{
    if (!case_name) return fail(UHS_E_INVAL, "null name");
    for (int i = 0; i < uhs_builtin_count(); i++)
        if (strcmp(kBuiltins[i].case_name, case_name) == 0) return i;
    return fail(UHS_E_INVAL, "no built-in stencil named '%s'", case_name);
}

static int grid_shape(int dims, const int64_t *ext, int halo, GridShape *s)
{
    if (!ext || (dims != 2 && dims != 3) || halo < 0) return fail(UHS_E_INVAL, "bad grid shape");
    for (int k = 0; k < dims; k++)
        if (ext[k] < 1 || ext[k] + 2LL * halo > INT_MAX / 2) return fail(UHS_E_INVAL, "bad extent");
    if (dims == 3) *s = {(int)ext[0] + 2 * halo, (int)ext[1] + 2 * halo, (int)ext[2] + 2 * halo, halo, halo};
    else *s = {1, (int)ext[0] + 2 * halo, (int)ext[1] + 2 * halo, 0, halo};
    return UHS_OK;
}

int uhs_fill_uniform_f64(double *grid, int dims, const int64_t *ext, int halo, uint64_t seed,
                         uhs_stream_t stream)
{
    GridShape s;
    if (int rc = grid_shape(dims, ext, halo, &s)) return rc;
    if (!grid) return fail(UHS_E_INVAL, "null grid");
    hipError_t e = uhs_launch_fill(grid, 0, s, seed, (hipStream_t)stream);
    return e == hipSuccess ? UHS_OK : hip_fail(e, "fill");
}

int uhs_fill_uniform_f32(float *grid, int dims, const int64_t *ext, int halo, uint64_t seed,
                         uhs_stream_t stream)
{
    GridShape s;
    if (int rc = grid_shape(dims, ext, halo, &s)) return rc;
    if (!grid) return fail(UHS_E_INVAL, "null grid");
    hipError_t e = uhs_launch_fill(grid, 1, s, seed, (hipStream_t)stream);
    return e == hipSuccess ? UHS_OK : hip_fail(e, "fill");
}

static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

int uhs_make_coeffs(int dims, int npoints, const int *offsets, uint64_t seed, double *coeffs)
{
    if ((dims != 2 && dims != 3) || npoints < 1 || !offsets || !coeffs)
        return fail(UHS_E_INVAL, "bad coefficient request");
    uint64_t draws = 0;
    for (int t = 0; t < npoints; t++) {
        int partner = -1;
        for (int s = 0; s < t && partner < 0; s++) {
            bool mirror = true;
            for (int k = 0; k < dims; k++)
                if (offsets[s * dims + k] != -offsets[t * dims + k]) mirror = false;
            if (mirror) partner = s;
        }
        if (partner >= 0) {
            coeffs[t] = coeffs[partner];
        } else {
            const uint64_t m = mix64(seed + (draws + 1) * 0x9E3779B97F4A7C15ULL);
            coeffs[t] = (double)(m >> 11) * 0x1.0p-53;
            draws++;
        }
    }
    double sum = 0.0;
    for (int t = 0; t < npoints; t++) sum += coeffs[t];
    for (int t = 0; t < npoints; t++) coeffs[t] /= sum;
    return UHS_OK;
}

static int norms(const void *a, int a_f32, const double *b, int dims, const int64_t *ext, int halo,
                 double out[3], hipStream_t st)
{
    GridShape s;
    if (int rc = grid_shape(dims, ext, halo, &s)) return rc;
    if (!a || !b || !out) return fail(UHS_E_INVAL, "null argument");
    const int nparts = kNormParts;
    double *dev = nullptr;
    hipError_t e = hipMalloc(&dev, sizeof(double) * (3 * nparts + 3));
    if (e != hipSuccess) return hip_fail(e, "norm workspace");
    e = uhs_launch_norms(a, a_f32, b, s, dev, nparts, dev + 3 * nparts, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dev + 3 * nparts, 3 * sizeof(double),
                                            hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dev);
    return e == hipSuccess ? UHS_OK : hip_fail(e, "error norms");
}

int uhs_error_norms_f32_f64(const float *a, const double *b, int dims, const int64_t *ext, int halo,
                            double out[3], uhs_stream_t stream)
{
    return norms(a, 1, b, dims, ext, halo, out, (hipStream_t)stream);
}

int uhs_error_norms_f64(const double *a, const double *b, int dims, const int64_t *ext, int halo,
                        double out[3], uhs_stream_t stream)
{
    return norms(a, 0, b, dims, ext, halo, out, (hipStream_t)stream);
}

int uhs_error_norms_f32_f64_noalloc(const float *a, const double *b, int dims, const int64_t *ext,
                                    int halo, double out[3], uhs_stream_t stream)
{
    GridShape s;
    if (int rc = grid_shape(dims, ext, halo, &s)) return rc;
    if (!a || !b || !out) return fail(UHS_E_INVAL, "null argument");
    // The reduction's scratch need is fixed (kNormParts partial triples and
    // the result), whatever the grid size; it must fit the static array.
    const int nparts = kNormParts;
    if (3 * nparts + 3 > kNormScratchDoubles)
        return fail(UHS_E_NOMEM, "norm scratch of %d doubles exceeds the static %d", 3 * nparts + 3,
                    kNormScratchDoubles);
    double *scr = nullptr;
    hipError_t e = uhs_norms_scratch(&scr);
    if (e != hipSuccess) return hip_fail(e, "norm scratch address");
    hipStream_t st = (hipStream_t)stream;
    // Every k_norms workgroup writes its own triple and k_norms_final writes
    // the result, so nothing left by an earlier call needs clearing.
    e = uhs_launch_norms(a, 1, b, s, scr, nparts, scr + 3 * nparts, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, scr + 3 * nparts, 3 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? UHS_OK : hip_fail(e, "error norms");
}

int uhs_copy_f64(double *dst, const double *src, size_t n, uhs_stream_t stream)
{
    if (!dst || !src) return fail(UHS_E_INVAL, "null pointer");
    if (n == 0) return UHS_OK;
    if (n > (size_t)(1LL << 40)) return fail(UHS_E_INVAL, "copy too large");
    hipError_t e = uhs_launch_copy(src, dst, (long long)n, (hipStream_t)stream);
    return e == hipSuccess ? UHS_OK : hip_fail(e, "copy");
}

int uhs_fma_f64(double *out, size_t n, double x, double y, int iters, uhs_stream_t stream)
{
    if (!out) return fail(UHS_E_INVAL, "null pointer");
    if (iters < 0) return fail(UHS_E_INVAL, "negative iteration count");
    if (n == 0) return UHS_OK;
    if (n > (size_t)(1LL << 32)) return fail(UHS_E_INVAL, "too many threads");
    hipError_t e = uhs_launch_fma(out, (long long)n, x, y, iters, (hipStream_t)stream);
    return e == hipSuccess ? UHS_OK : hip_fail(e, "fma");
}

const char *uhs_last_error(void) { return g_err.c_str(); }
const char *uhs_version(void) { return "uhs-mi355x 0.1 (gfx950)"; }

int uhs_set_lds_poison(int on)
{
    for (auto set : lds_poison_setters()) {
        const hipError_t e = set(on ? 1 : 0);
        if (e != hipSuccess) return fail(UHS_E_HIP, hipGetErrorString(e));
    }
    return (int)lds_poison_setters().size();
}

int uhs_device_count(void)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    return n;
}

int uhs_device_name(char *buf, size_t buflen)
{
    hipDeviceProp_t p;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    snprintf(buf, buflen, "%s %s %d CUs", p.name, p.gcnArchName, p.multiProcessorCount);
    return UHS_OK;
}

#define HIPCALL(expr, what)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        return e_ == hipSuccess ? UHS_OK : hip_fail(e_, what);                                    \
    } while (0)

int uhs_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(UHS_E_INVAL, "null pointer");
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return fail(UHS_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return UHS_OK;
}
int uhs_free(void *ptr) { HIPCALL(hipFree(ptr), "hipFree"); }
int uhs_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    HIPCALL(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
int uhs_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    HIPCALL(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
int uhs_memset(void *dst, int value, size_t bytes) { HIPCALL(hipMemset(dst, value, bytes), "hipMemset"); }
int uhs_synchronize(void) { HIPCALL(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
int uhs_stream_create(uhs_stream_t *s)
{
    if (!s) return fail(UHS_E_INVAL, "null pointer");
    HIPCALL(hipStreamCreate((hipStream_t *)s), "hipStreamCreate");
}
int uhs_stream_destroy(uhs_stream_t s) { HIPCALL(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy"); }
int uhs_stream_synchronize(uhs_stream_t s)
{
    HIPCALL(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize");
}
int uhs_event_create(uhs_event_t *e)
{
    if (!e) return fail(UHS_E_INVAL, "null pointer");
    HIPCALL(hipEventCreate((hipEvent_t *)e), "hipEventCreate");
}
int uhs_event_destroy(uhs_event_t e) { HIPCALL(hipEventDestroy((hipEvent_t)e), "hipEventDestroy"); }
int uhs_event_record(uhs_event_t e, uhs_stream_t s)
{
    HIPCALL(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord");
}
int uhs_event_elapsed_ms(uhs_event_t start, uhs_event_t stop, float *ms)
{
    if (!ms) return fail(UHS_E_INVAL, "null pointer");
    hipError_t e = hipEventSynchronize((hipEvent_t)stop);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    HIPCALL(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

} // extern "C"
