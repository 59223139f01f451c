/*
 * uhs_tune.h -- autotuning of the sweep's runtime tiling, and the learned
 * cost model that prunes the search.
 *
 * The reference tunes by random search over compile-time tiling parameters,
 * recompiling and profiling every point (scripts/evaluation_scripts_GPU/
 * test_driver.py:36-99,101-164,169-285; collect.py:30-69).  Here the tiling is
 * a runtime argument of one kernel family, so the tuner times configurations
 * directly.  Timing runs sweeps on the caller's grids: their contents are
 * overwritten.
 */
#ifndef UHS_TUNE_H
#define UHS_TUNE_H
#include <stddef.h>
#include <stdint.h>
#include "uhs.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct uhs_tune_result {
    uhs_config best;
    double best_ms;     /* per sweep */
    int evaluated;      /* configurations timed or read from the database */
    int failed;         /* configurations that could not run (skipped) */
    int reused;         /* timings taken from the database instead of re-run */
    double seconds;     /* wall time of the search */
} uhs_tune_result;

/* Exhaustive timed search over the legal configuration space of this
 * problem.  Every timing is appended to the CSV database at db_path (NULL:
 * none); timings already there for this device and problem are reused, so
 * an interrupted search resumes.  The winner becomes the configuration
 * uhs_sweep_* use for this problem when cfg is NULL (this process). */
int uhs_tune_sweep(void *A, void *B, int dims, const int64_t *ext, int halo,
                   const uhs_stencil *st, int dtype, int reps,
                   const char *db_path, uhs_tune_result *out,
                   uhs_stream_t stream);
/* The same search over an explicit candidate list (e.g. a random sample of
 * the space); illegal candidates are recorded as failed and skipped. */
int uhs_tune_sweep_list(void *A, void *B, int dims, const int64_t *ext, int halo,
                        const uhs_stencil *st, int dtype, const uhs_config *cands,
                        int ncand, int reps, const char *db_path,
                        uhs_tune_result *out, uhs_stream_t stream);
/* The legal configurations the search visits: returns their number and
 * copies up to cap of them into out (out may be NULL). */
int uhs_tune_space(int dims, const int64_t *ext, int halo, const uhs_stencil *st,
                   int dtype, uhs_config *out, int cap);
/* Milliseconds per sweep of one configuration (one warm-up sweep, then reps). */
int uhs_tune_time(void *A, void *B, int dims, const int64_t *ext, int halo,
                  const uhs_stencil *st, int dtype, const uhs_config *cfg,
                  int reps, double *ms, uhs_stream_t stream);
/* Mock timing backend.  With fn set, the searches (uhs_tune_sweep,
 * uhs_tune_sweep_list, uhs_tune_sweep_model[s]) call fn for every candidate
 * that passes the configuration check, instead of timing sweeps on the GPU:
 * fn returns a UHS_* status (UHS_E_HIP acts as a launch failure: the search
 * records it and stops) and writes the milliseconds per sweep.  The search
 * logic -- database reuse, skipping failed candidates, stopping on a launch
 * failure, installing the winner -- then runs without a GPU and the grids are
 * never touched, so tests can inject timings and failures.  NULL restores
 * the GPU timer.  Process-wide; set it while no search runs. */
typedef int (*uhs_tune_timer_fn)(const uhs_config *cfg, int reps, double *ms, void *user);
int uhs_tune_set_timer(uhs_tune_timer_fn fn, void *user);
/* Installs, for every built-in problem in the database measured on this
 * device, its fastest configuration as the default.  Returns the count. */
int uhs_tune_load_db(const char *db_path);
int uhs_tune_clear(void);

/* Timed search over batch configurations (temporal depth {1,2,4} x star
 * tiles x box tiles) for a batch of grids: each candidate is planned,
 * captured and run for `timesteps` steps (grid contents are overwritten);
 * the fastest goes to *best.  ms_out (optional, ncand_cap entries) receives
 * each candidate's time, <0 for candidates that could not be built. */
int uhs_batch_tune(const uhs_grid_desc *grids, int n, int timesteps,
                   uhs_batch_config *best, double *best_ms,
                   uhs_batch_config *cands_out, double *ms_out, int ncand_cap,
                   int *ncand, uhs_stream_t stream);

/* ---- learned cost model ---------------------------------------------------
 * An MLP (UHS_MLP_INPUTS -> 128 -> 128 -> 128 -> 1, ReLU, FP32) mapping a
 * problem and configuration to log(ms per sweep).  Forward, backward, Adam and
 * inference run as HIP kernels on the GPU.  Inputs are standardised inside
 * the model with the training set's mean and deviation. */
#define UHS_MLP_INPUTS 20
#define UHS_MLP_HIDDEN 128

typedef struct uhs_mlp *uhs_mlp_t;

/* Feature vector of (problem, configuration), 20 inputs: 3D flag (1),
 * radius one-hot (3), star flag (1), log2 extents (3, axis 0 first, 0 for a
 * missing axis), log2 tile x/y/z (3), log2 threads per workgroup (1), points
 * per thread (1), temporal depth one-hot {1, 2, 3 or more} (3;
 * UHS_CFG_TEMPORAL2 counts as depth 2, UHS_CFG_TEMPORALK as K = tile_y), log2 of the stencil's term count (1; 2 dims r + 1 for a
 * star, (2r+1)^dims for a box), precision (1: 0 FP64, 1 FP32, 2 FP32
 * compensated), non-temporal store flag (1), kernel variant (1: 0 LDS-staged,
 * 1 deeper prefetch (two planes for stars and two rows for 2D, three cells
 * for boxes), 2 four rows of prefetch (2D), 3 register-tiled, 4 two rows per
 * thread, 5 two planes of box loads in flight, 6 both 4 and 5, 7 a register
 * ring of rows (2D)).  Threads per workgroup follow the variant (ghost zone
 * included for UHS_CFG_TEMPORAL2 and UHS_CFG_TEMPORALK). */
int uhs_mlp_features(int dims, int radius, int star, int dtype, const int64_t *ext,
                     const uhs_config *cfg, int depth, float *out20);

int uhs_mlp_create(uint64_t seed, uhs_mlp_t *out);
int uhs_mlp_destroy(uhs_mlp_t m);
int uhs_mlp_num_params(void);
/* Parameters as one FP32 vector: W1[20][128], b1[128], W2[128][128], b2,
 * W3[128][128], b3, W4[128], b4 (row-major, input index first). */
int uhs_mlp_get_params(uhs_mlp_t m, float *out, int n);
int uhs_mlp_set_params(uhs_mlp_t m, const float *in, int n);
/* Standardisation: x_std = (x - mean) / scale per input; target likewise. */
int uhs_mlp_set_norm(uhs_mlp_t m, const float *mean20, const float *scale20,
                     float tmean, float tscale);
/* Predicted log(ms) for n raw feature rows (host arrays). */
int uhs_mlp_predict(uhs_mlp_t m, const float *x, int n, float *y);
/* Loss 0.5*mean((pred_std - t_std)^2) and its gradient w.r.t. the
 * parameters for n rows (host arrays), computed by the backward kernels. */
int uhs_mlp_grad(uhs_mlp_t m, const float *x, const float *t, int n,
                 float *grad, float *loss);
/* Adam (b1 0.9, b2 0.999, eps 1e-8) for `epochs` passes over (x, t) in
 * shuffled minibatches, all on the GPU.  Sets the standardisation from the
 * data first when fit_norm != 0.  Returns the last epoch's mean loss. */
int uhs_mlp_train(uhs_mlp_t m, const float *x, const float *t, int n,
                  int epochs, int batch, float lr, int fit_norm,
                  uint64_t seed, float *final_loss);
/* One Adam step with an explicit gradient (host array), step count t >= 1. */
int uhs_mlp_adam_step(uhs_mlp_t m, const float *grad, float lr, int t);
int uhs_mlp_save(uhs_mlp_t m, const char *path);
int uhs_mlp_load(uhs_mlp_t m, const char *path);

/* Model-pruned search: predict every legal configuration, time the topk
 * predicted fastest, keep the best (becomes the default like
 * uhs_tune_sweep).  res->seconds includes the inference. */
int uhs_tune_sweep_model(void *A, void *B, int dims, const int64_t *ext,
                         int halo, const uhs_stencil *st, int dtype,
                         uhs_mlp_t model, int topk, int reps,
                         const char *db_path, uhs_tune_result *out,
                         uhs_stream_t stream);
/* The same search ranked by an ensemble of nmodels models: the mean of
 * their predicted log times.  uhs_tune_sweep_model is the one-model case. */
int uhs_tune_sweep_models(void *A, void *B, int dims, const int64_t *ext,
                          int halo, const uhs_stencil *st, int dtype,
                          const uhs_mlp_t *models, int nmodels, int topk,
                          int reps, const char *db_path,
                          uhs_tune_result *out, uhs_stream_t stream);

/* ---- batch configurations: candidate space, timed search, model --------- */

/* The candidate batch configurations uhs_batch_tune times: returns their
 * number and copies up to cap of them into out (out may be NULL). */
int uhs_batch_space(uhs_batch_config *out, int cap);
/* Timed search over an explicit list of batch configurations (each planned,
 * captured, run once to warm up and once timed for `timesteps` steps; grid
 * contents are overwritten).  ms_out (optional, ncand entries) receives each
 * candidate's time, < 0 for candidates that could not be built. */
int uhs_batch_tune_list(const uhs_grid_desc *grids, int n, int timesteps,
                        const uhs_batch_config *cands, int ncand, uhs_batch_config *best,
                        double *best_ms, double *ms_out, uhs_stream_t stream);
/* Feature vector of (batch, configuration) for the cost model, 20 inputs:
 * star depth one-hot {1,2,4} (3), log2 star tile x/y/z (3), log2 box tile
 * x/y/z (3), box variant (1: two rows per thread + 1, two planes in flight
 * + 2), split mode (1: 0 shared launches, 1 UHS_BATCH_SPLIT, 2 with
 * UHS_BATCH_CONCURRENT), log2 interior
 * points of the star grids and of the box grids (2, 0 when none), log2
 * number of grids (1), mean log2 extent per axis (3), spread (standard
 * deviation) of the log2 extents (1), log2 (1 + mean terms per point of the
 * star grids + of the box grids) (1), star tile takes the register-queue
 * two-step form (1).  The model maps it to log(ms per step). */
int uhs_batch_features(const uhs_grid_desc *grids, int n, const uhs_batch_config *cfg,
                       float *out20);
/* Model-pruned batch search: predicts every candidate of uhs_batch_space,
 * times the topk predicted fastest (as uhs_batch_tune_list) and returns the
 * fastest; *seconds (optional) is the wall time including the inference. */
int uhs_batch_tune_model(const uhs_grid_desc *grids, int n, int timesteps, uhs_mlp_t model,
                         int topk, uhs_batch_config *best, double *best_ms, double *seconds,
                         uhs_stream_t stream);

/* Timed search over the streaming program kernel's tiles
 * (uhs_program_sweep_ex_f64 with UHS_PROGRAM_TILED): every candidate tile
 * {tile_x, tile_y, tile_z} whose LDS rings fit runs one warm-up sweep and
 * then `reps` timed sweeps on the caller's grids (overwritten).  tile_out[3]
 * receives the fastest tile, *best_ms its time per sweep, *evaluated the
 * number of tiles timed (any output pointer but tile_out may be NULL).  The
 * winner becomes the tile uhs_program_sweep_f64 uses for this program and
 * grid shape when no tile is given (this process).  UHS_E_CONFIG when no
 * candidate fits: the pointwise kernel is then the only one. */
int uhs_tune_program_f64(double *A, double *B, const int64_t *shape, const uhs_program *pg,
                         int reps, int *tile_out, double *best_ms, int *evaluated,
                         uhs_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* UHS_TUNE_H */
