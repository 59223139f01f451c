// Host-side planning shared by the C ABI, the batch runner and the autotuner.
#pragma once
#include "uhs.h"
#include "uhs_internal.h"
#include <string>

enum Family { FAM_STAR3D = 0, FAM_BOX3D = 1, FAM_WIN2D = 2, FAM_GATHER = 3 };

// A resolved sweep: which kernel, with which arguments and configuration.
struct Plan {
    int family;
    int prog;     // index into kBuiltins, -1 for a general stencil
    int radius;
    uhs_config cfg;
    LaunchFn fn;
    SweepArgs args;
    // UHS_CFG_TEMPORAL2 / UHS_CFG_TEMPORALK: the single-step sweep that runs
    // the leading steps.
    LaunchFn fn1;
    SweepArgs args1;
    // UHS_CFG_TEMPORALK: the same tile with an odd K' < K steps per launch
    // (K' = 5 and 3, where they pass the config check), which cover the
    // leading steps in fewer launches.
    LaunchFn fnk[2];
    SweepArgs argsk[2];
    int kk[2];
};

int uhs_set_error(int code, const char *msg);
int uhs_match_builtin(int dims, int npoints, const int *off);
int uhs_make_plan(int dims, const int64_t *ext, int halo, const uhs_stencil *st, int dtype,
                  const uhs_config *cfg, Plan *pl);
int uhs_run_plan(Plan &pl, void *A, void *B, int timesteps, hipStream_t stream);
// Sets a checked configuration on a plan: tiling arguments, kernel, and for
// UHS_CFG_TEMPORAL2 the single-step kernel of the leading steps.
int uhs_plan_apply(Plan &pl, const uhs_config &c);
int uhs_check_config(const Plan &pl, const uhs_config &c, char *why, size_t whylen);
// The built-in default configuration for a plan; one_step excludes the
// two-steps-per-launch defaults (for the leading single steps of such a
// plan, and for one-launch-per-step baselines).
uhs_config uhs_heuristic_config(const Plan &pl, bool one_step = false);
// The launcher for a plan under configuration c (nullptr: not compiled).
LaunchFn uhs_pick_fn(const Plan &pl, const uhs_config &c);
long long uhs_config_blocks(const SweepArgs &a, const uhs_config &c, int radius);

// Tuned configurations, keyed by kernel family, stencil, precision and extents.
std::string uhs_config_key(int family, int prog, int dtype, int dims, const int64_t *ext, int halo);
void uhs_store_tuned(const std::string &key, const uhs_config &c);
bool uhs_lookup_tuned(const std::string &key, uhs_config *c);
void uhs_clear_tuned();
