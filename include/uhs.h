/*
 * uhs.h -- C ABI of the MI355X (gfx950) structured-grid stencil sweep.
 *
 * This replaces the entry the reference's GPU backend generates for each
 * benchmark stencil,
 *     extern "C" void _mlir_ciface_stencil_<name>_iteration(MemRef *A, MemRef *B)
 * (declared by the callers, e.g. UH-GPU/3d1r_star/3d7pt_main.cpp:25, called at
 * :150; emitted because ConvertStencilToStandard.cpp:80 sets
 * llvm.emit_c_interface).  In the reference, extents, halo, coefficients and
 * the timestep count are compile-time constants of the generated kernel; here
 * they are arguments, so one library serves every stencil.  The per-case
 * `_mlir_ciface_*` symbols themselves are provided by uhs_compat.h.
 *
 * Conventions shared by every entry point:
 *   - Grids are dense, row-major and halo-inclusive: axis 0 is outermost and
 *     the last axis has unit stride; axis k holds ext[k] + 2*halo cells
 *     (the reference's MemRef layout, 3d7pt_main.cpp:17-23,147-148).
 *   - One sweep writes the interior [halo, ext+halo) of the destination:
 *         out[p] = c[0]*in[p+o[0]] + c[1]*in[p+o[1]] + ...
 *     summed left to right in term order (the DSL's order,
 *     UHStencil/include/Parser/Parser.h:437-468, MLIRGen.cpp:452-501).  Halo
 *     cells are never written (ConvertStencilToStandard.cpp:339-381).
 *     The kernels evaluate each term after the first as one fused
 *     multiply-add; the result differs from an unfused sum only by rounding
 *     (within 1e-12 relative in FP64 for every benchmark stencil).
 *   - Time stepping ping-pongs A and B: sweep k (k = 1..T) reads A and writes
 *     B when k is odd, reads B and writes A when k is even
 *     (ConvertStencilToStandard.cpp:298-337).  T = 1 leaves the result in B,
 *     even T in A (the reference's only legal counts, MLIRGen.cpp:317-321);
 *     odd T > 1 leaves it in B.
 *   - Every call is asynchronous on `stream` (NULL = the default stream);
 *     the caller owns all grid memory.  Calls return UHS_OK or a negative
 *     UHS_E_* code and never throw; uhs_last_error() describes the last
 *     failure of the calling thread.
 */
#ifndef UHS_H
#define UHS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *uhs_stream_t; /* == hipStream_t */

enum {
    UHS_OK = 0,
    UHS_E_INVAL = -1,       /* bad argument (null pointer, extent, halo, ...) */
    UHS_E_UNSUPPORTED = -2, /* stencil/precision combination not provided */
    UHS_E_CONFIG = -3,      /* tiling configuration illegal for this stencil */
    UHS_E_HIP = -4,         /* HIP runtime error */
    UHS_E_NOMEM = -5,       /* allocation failed */
    UHS_E_IO = -6,          /* file error (tuning database, DSL file) */
    UHS_E_PARSE = -7        /* DSL text could not be read */
};

#define UHS_MAX_POINTS 343 /* largest stencil (7x7x7 box) */
#define UHS_MAX_RADIUS 4

/* A linear stencil: npoints terms, offsets[npoints*dims] (axis 0 first) and
 * coeffs[npoints], both in term order.  Mirrors what the DSL `expr` block
 * defines (the UH-GPU DSL files; Parser.h:525-606). */
typedef struct uhs_stencil {
    int dims;            /* 2 or 3 */
    int npoints;
    const int *offsets;  /* [npoints][dims] */
    const double *coeffs;/* [npoints]; rounded to float by the FP32 forms */
} uhs_stencil;

/* Runtime tiling parameters of the sweep kernel family.  They replace the
 * reference's compile-time tile_sizes_l0/l1 and inner_tile pass options
 * (UHStencil/include/Conversion/UHAffineToParallel/Passes.td:6-26,85-99).
 *   tile_x  points along the unit-stride axis per workgroup
 *   tile_y  rows per workgroup (3D only; must be 1 for 2D)
 *   tile_z  length of the slab streamed through registers along axis 0
 *   vec     points per thread along the unit-stride axis (1, 2 or 4; 8 for
 *           3D stars)
 *   flags   UHS_CFG_NT_STORE: output stores bypass the caches;
 *           UHS_CFG_PREFETCH2: 3D stars keep two planes of loads in flight
 *           (2D stencils two rows), 3D boxes fetch staged cells three ahead
 *           instead of one;
 *           UHS_CFG_PREFETCH4: 2D stencils keep four rows in flight;
 *           UHS_CFG_ROWRING: 2D stencils keep 2r+1 rows in flight in a
 *           register ring (not with UHS_CFG_PREFETCH2/4); with
 *           UHS_CFG_TEMPORAL2 it selects the window form of the 2D two-step
 *           kernel (any 2D stencil), without it a 2D star takes the ring
 *           form and a 2D box the sum form (one or two columns per
 *           thread);
 *           UHS_CFG_ROWS2: 3D boxes compute two rows per thread (vec must
 *           be 1, tile_y even; tile_x * tile_y / 2 threads), halving the
 *           LDS reads per output;
 *           UHS_CFG_REGTILE: 3D stars use the register-tiled kernel (no LDS,
 *           no barriers): tile_x must be 64 (one wavefront across, 64 - 2r
 *           columns computed), vec = rows per lane (2, 4 or 8), tile_y a
 *           multiple of vec (tile_y / vec wavefronts per workgroup);
 *           UHS_CFG_PLANES2: 3D boxes keep two planes of tile loads in
 *           flight instead of one (not with UHS_CFG_PREFETCH2);
 *           UHS_CFG_TEMPORAL2: 3D stars, the radius-1 3D box and 2D
 *           stencils advance two timesteps per launch (temporal blocking,
 *           each cell read and written once per two updates): the workgroup
 *           covers the tile_x * tile_y output tile plus an r-wide ghost
 *           zone; 3D stars (tile_x + 2r) / vec * (tile_y + 2r) threads, vec
 *           1 or 2 dividing tile_x + 2r; the radius-1 box the same threads,
 *           vec 1 or 2 (and 4 in FP32); 2D stencils (tile_x + 2r) / vec
 *           threads, tile_y 1,
 *           vec 1, 2 or 4 dividing tile_x + 2r (FP64: vec 1 for radius 2
 *           and 3, no radius-3 box; FP32: vec <= 2 for radius 2, 1 for
 *           radius 3).  Up to three leading
 *           timesteps run as single sweeps (default configuration) so the
 *           result lands in the grid the one-step ping-pong leaves it in;
 *           results are bitwise those of single sweeps.  The other grid's
 *           interior then holds state T-2, not T-1, when a two-step
 *           launch ran;
 *           UHS_CFG_TEMPORALK: 2D stars and the radius-1 and radius-2 2D
 *           boxes advance K = tile_y timesteps per launch, K = 3, 4, 5, 6 or 8
 *           (not with UHS_CFG_TEMPORAL2): (tile_x + 2(K-1)r) / vec threads,
 *           vec 1 or 2 (4 also for the radius-1 star; the FP64 radius-2
 *           box 1);
 *           the leading T mod K steps (K more when K is even and T / K odd,
 *           so that even-K launches are even in number) run as 5- and 3-step
 *           launches of the same tile and single sweeps, then the K-step
 *           launches; the other grid's interior then holds state T-K.
 * Threads per workgroup = tile_x / vec * tile_y (64 * tile_y / vec with
 * UHS_CFG_REGTILE). */
#define UHS_CFG_NT_STORE 1
#define UHS_CFG_PREFETCH2 2
#define UHS_CFG_REGTILE 4
#define UHS_CFG_PREFETCH4 8
#define UHS_CFG_ROWS2 16
#define UHS_CFG_TEMPORAL2 32
#define UHS_CFG_PLANES2 64
#define UHS_CFG_ROWRING 128
#define UHS_CFG_TEMPORALK 256
typedef struct uhs_config {
    int tile_x, tile_y, tile_z, vec, flags;
} uhs_config;

/* ---- the sweep ---------------------------------------------------------- */

/* FP64 sweep, T timesteps.  ext[dims] = interior extents, halo >= the
 * stencil radius.  cfg NULL = the tuned (or default) configuration; the
 * FP64 defaults of the radius-1 and radius-2 3D stars (3d1r_star,
 * 3d2r_star), the radius-1 3D box (3d1r_box) and the 2D stencils but the
 * radius-3 box advance two steps per launch (UHS_CFG_TEMPORAL2), with
 * results bitwise those of single sweeps. */
int uhs_sweep_f64(double *A, double *B, int dims, const int64_t *ext, int halo,
                  const uhs_stencil *st, int timesteps, const uhs_config *cfg,
                  uhs_stream_t stream);

/* FP32 storage and arithmetic.  compensated = 0: plain fused sum (the
 * reference's `float` DSL mode, Parser.h:86-91); compensated = 1:
 * Kahan-compensated sum of the terms.  cfg NULL = the tuned (or default)
 * configuration; the plain FP32 defaults of 3d1r_star and 3d1r_box and the
 * FP32 defaults (plain and compensated) of the 2D stencils but the radius-3 box
 * advance two steps per launch (UHS_CFG_TEMPORAL2), bitwise the same as
 * single sweeps. */
int uhs_sweep_f32(float *A, float *B, int dims, const int64_t *ext, int halo,
                  const uhs_stencil *st, int timesteps, int compensated,
                  const uhs_config *cfg, uhs_stream_t stream);

/* Which kernel a stencil/precision/config maps to, without launching:
 * writes a short name into buf ("star3d", "box3d", "win2d", "program" --
 * a stencil of other forms with at most UHS_MAX_PROGRAM_TERMS terms, run by
 * the streaming program kernel --, "gather" -- the general one-thread-per-
 * point kernel for larger ones).
 * Returns UHS_OK, or the error the sweep would return. */
int uhs_plan_kind(int dims, const int64_t *ext, int halo, const uhs_stencil *st,
                  int dtype /* 0 f64, 1 f32, 2 f32 compensated */,
                  const uhs_config *cfg, char *buf, size_t buflen);

/* The configuration used when cfg is NULL. */
int uhs_default_config(int dims, const int64_t *ext, int halo,
                       const uhs_stencil *st, int dtype, uhs_config *out);

/* The default configuration with one launch per timestep: the one-step
 * kernel and tile that the leading single steps of a two-step default use,
 * and that uhs_sweep_each_f64 runs.  Equals uhs_default_config's result for
 * every stencil whose default is one step per launch. */
int uhs_one_step_config(int dims, const int64_t *ext, int halo,
                        const uhs_stencil *st, int dtype, uhs_config *out);

/* ---- the 12 benchmark stencils ------------------------------------------ */

/* Number of built-in stencils (the 12 UH-GPU cases). */
int uhs_builtin_count(void);
/* Describe built-in stencil i: case name ("3d1r_star"), reference symbol
 * ("stencil_3d7pt"), dims, radius, npoints, and pointers to its offsets and
 * DSL coefficients in DSL term order (static storage). */
int uhs_builtin_info(int i, const char **case_name, const char **symbol,
                     int *dims, int *radius, int *npoints,
                     const int **offsets, const double **coeffs);
/* Index of the built-in stencil with this case name, or UHS_E_INVAL. */
int uhs_builtin_find(const char *case_name);

/* ---- synthetic inputs (seeded, bitwise reproducible on host) ------------- */

/* Interior cell with halo-inclusive linear index L gets
 * u01(mix64(seed + (L+1)*0x9E3779B97F4A7C15)), halo cells 0.  mix64 is the
 * splitmix64 finaliser, u01(z) = (z >> 11) * 2^-53. */
int uhs_fill_uniform_f64(double *grid, int dims, const int64_t *ext, int halo,
                         uint64_t seed, uhs_stream_t stream);
int uhs_fill_uniform_f32(float *grid, int dims, const int64_t *ext, int halo,
                         uint64_t seed, uhs_stream_t stream);
/* Symmetric (c(o) = c(-o)) U[0,1) coefficients normalised to sum 1, drawn
 * in term order from the same generator (host only). */
int uhs_make_coeffs(int dims, int npoints, const int *offsets, uint64_t seed,
                    double *coeffs);

/* ---- error norms (FP64 reductions on the GPU) --------------------------- */

/* Over the interior: max |a - b|, sqrt(sum (a-b)^2), sqrt(sum b^2), where a
 * is an FP32 grid and b an FP64 grid of the same shape.  Synchronous. */
int uhs_error_norms_f32_f64(const float *a, const double *b, int dims,
                            const int64_t *ext, int halo, double out[3],
                            uhs_stream_t stream);
/* Same with both grids FP64. */
int uhs_error_norms_f64(const double *a, const double *b, int dims,
                        const int64_t *ext, int halo, double out[3],
                        uhs_stream_t stream);
/* uhs_error_norms_f32_f64 without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The same kernels and the same
 * result, bit for bit; the partial sums live in a static per-device array of
 * 3 * 1024 + 3 doubles (24.6 KB, held until the process ends), whose size
 * does not depend on the grid.  Every call on a device shares that array:
 * calls that overlap on one device (other threads or streams) race.
 * Synchronous. */
int uhs_error_norms_f32_f64_noalloc(const float *a, const double *b, int dims,
                                    const int64_t *ext, int halo, double out[3],
                                    uhs_stream_t stream);

/* ---- batched grids: one hipGraph, persistent work-queue launches -------- */

/* One grid of a batch: halo-inclusive FP64 grids A and B, interior extents,
 * halo, and its stencil (one of the six built-in 3D stencils). */
typedef struct uhs_grid_desc {
    double *A, *B;
    int dims;            /* 3 */
    int64_t ext[3];
    int halo;
    uhs_stencil stencil;
} uhs_grid_desc;

/* Tiling of the batched launches (256 threads per workgroup).  Star
 * stencils are advanced star_depth steps per launch (temporal blocking with
 * ghost zones, 1, 2 or 4); box stencils one step per launch.
 *   Stars: a 512-point tile (2 points per thread) runs single steps in
 *   registers; with star_depth 2, a tile whose extension by the radius r on
 *   each side, (tile_x + 2r) / 2 * (tile_y + 2r), fits 256 threads (tile_x
 *   even) runs the two-step register-queue kernel; other tiles use the LDS
 *   ring form.
 *   Boxes: box_tile_x * box_tile_y must be 256, or 512 with
 *   UHS_BATCH_BOX_ROWS2 (two rows per thread, box_tile_y even);
 *   UHS_BATCH_BOX_PLANES2 keeps two planes of tile loads in flight. */
#define UHS_BATCH_BOX_ROWS2 1
#define UHS_BATCH_BOX_PLANES2 2
/* UHS_BATCH_SPLIT: star grids and box grids run in launches of their own
 * instead of sharing each launch.  For a batch of 7-point stars and 125-point
 * boxes (the benchmark mix) the star launches may then take star tiles that
 * need 512-thread workgroups: the star tile rules above hold with 512 for
 * 256 (single steps: a 1024-point tile; two steps: (tile_x + 2r) / 2 *
 * (tile_y + 2r) <= 512), and such tiles get 512-thread workgroups.
 * UHS_BATCH_CONCURRENT (with UHS_BATCH_SPLIT): the star launches form a
 * second chain of the graph beside the box launches (forked after the
 * counter reset, joined at the end), so HBM-bound star tiles and
 * FP64-bound box tiles run at the same time. */
#define UHS_BATCH_SPLIT 4
#define UHS_BATCH_CONCURRENT 8
typedef struct uhs_batch_config {
    int star_depth;
    int star_tile_x, star_tile_y, star_tile_z;
    int box_tile_x, box_tile_y, box_tile_z;
    int flags; /* UHS_BATCH_* */
} uhs_batch_config;

typedef struct uhs_batch *uhs_batch_t;

/* Plans T timesteps of every grid, uploads the task lists and captures the
 * whole run as one hipGraph (a counter-reset node, then one persistent
 * launch per time block in which workgroups take tiles of any grid from a
 * device-side queue, largest first; with UHS_BATCH_SPLIT the star and box
 * grids have launches of their own).  Grids keep the ping-pong convention:
 * after the run each grid's result is in A for even T, in B for odd T; the
 * other buffer holds an earlier state.  cfg NULL = the default. */
int uhs_batch_create(const uhs_grid_desc *grids, int n, int timesteps,
                     const uhs_batch_config *cfg, uhs_batch_t *out);
/* Launches the captured graph on `stream` (asynchronous).  Reusable: every
 * run advances each grid by T more steps, continuing from the state the
 * previous run left, with the ping-pong convention carried on (after r runs
 * the state r*T is in A when r*T is even, in B when odd; for odd T the
 * batch holds a second graph with the roles of A and B swapped and runs
 * alternate between the two).  Runs of one batch must not overlap. */
int uhs_batch_run(uhs_batch_t b, uhs_stream_t stream);
int uhs_batch_destroy(uhs_batch_t b);
int uhs_batch_info(uhs_batch_t b, int *launches, long long *tasks,
                   int *max_workgroups, size_t *max_lds_bytes);
/* The default for a batch of 7-point stars and 125-point boxes (the
 * benchmark mix): {2, 62,14,32, 32,16,128, ROWS2 | PLANES2 | SPLIT |
 * CONCURRENT}.  uhs_batch_create with cfg NULL uses it for such a batch and
 * {2, 32,12,64, 32,16,128, ROWS2 | PLANES2} for other stencil sets. */
int uhs_batch_default_config(uhs_batch_config *out);
/* The unbatched baseline: uhs_sweep_f64 on each grid in turn, same stream,
 * one launch per timestep (the one-step kernel and its default tile). */
int uhs_sweep_each_f64(const uhs_grid_desc *grids, int n, int timesteps,
                       uhs_stream_t stream);
/* Splits a batch over `ngpus` GPUs, one batch per GPU and no exchange between
 * grids: longest-processing-time-first bin packing (each grid, in order of
 * decreasing cost, goes to the GPU with the least load so far; ties go to
 * the lower GPU index and equal costs keep grid order), within 4/3 of the
 * best split.  cost[i] is grid i's cost in any unit (measured or predicted
 * time); cost NULL estimates it, for the built-in 3D stencils, as interior
 * points x max(1, npoints / 64) (memory-bound up to a few tens of terms,
 * then proportional to the terms; fitted to the measured FP64 rates).
 * Writes the GPU of every grid to gpu_of[n] and every GPU's
 * summed cost to load[ngpus] (NULL: not written).  Host only; no device
 * call. */
int uhs_batch_partition(const uhs_grid_desc *grids, int n, int ngpus, const double *cost,
                        int *gpu_of, double *load);

/* ---- fused multi-kernel programs ---------------------------------------- */

/* A stencil program of several kernels after kernel fusion (the reference's
 * StencilKernelFusion.cpp:66-180 inlines a producer kernel into its single
 * consumer at every access offset, repeatedly, so chains fuse completely):
 * the output kernel's term j reads either the input grid at p + o_j or
 * producer k evaluated at p + o_j, and producer k's terms read the input
 * grid or an earlier producer (index < k) the same way.  So
 *     out[p] = c_0*x_0 + c_1*x_1 + ...,   x_j = in[p+o_j]  or  P_k(p+o_j)
 *     P_k(q) = d_0*y_0 + d_1*y_1 + ...,   y_i = in[q+e_i]  or  P_s(q+e_i)
 * each sum left to right in term order (first term a product, later terms
 * fused multiply-adds, as in uhs_sweep_f64).  Up to UHS_MAX_PRODUCERS
 * producers and UHS_MAX_PROGRAM_TERMS terms per kernel. */
#define UHS_MAX_PRODUCERS 4
#define UHS_MAX_PROGRAM_TERMS 32
typedef struct uhs_program {
    int dims;                      /* 2 or 3 */
    int nproducers;
    const uhs_stencil *producers;  /* [nproducers], in dependency order */
    uhs_stencil output;            /* the output ("operation") kernel */
    const int *sources;            /* [output.npoints]: -1 = the input grid, k = producers[k] */
    int64_t lo[3], hi[3];          /* cells written: [lo, hi) per axis, in halo-inclusive
                                      grid coordinates (the output kernel's domain) */
    const int *producer_sources;   /* [nproducers][UHS_MAX_PROGRAM_TERMS]: term i of
                                      producer k reads the input (-1) or producer s < k;
                                      NULL = every producer reads only the input grid */
} uhs_program;

/* T sweeps of a program over halo-inclusive FP64 grids of `shape` cells per
 * axis, ping-ponging A and B as uhs_sweep_f64 does; cells outside [lo, hi)
 * are never written.  Every read must fall inside the grid (UHS_E_INVAL
 * otherwise). */
int uhs_program_sweep_f64(double *A, double *B, const int64_t *shape, const uhs_program *pg,
                          int timesteps, uhs_stream_t stream);

/* The same with the kernel and its tile chosen by the caller.  Both kernels
 * give bitwise the same result: a producer's value at a cell does not
 * depend on which access evaluates it. */
#define UHS_PROGRAM_AUTO 0      /* tiled when its LDS frame fits, else pointwise */
#define UHS_PROGRAM_TILED 1     /* a block of the input staged in LDS once, each producer
                                   evaluated once per cell of the block */
#define UHS_PROGRAM_POINTWISE 2 /* one thread per cell, producers re-evaluated per access */
/* tile (tiled kernel): {tile_x, tile_y, tile_z}.  3D: a workgroup owns
 * tile_x x tile_y cells of axes 2 and 1 and walks a slab of tile_z planes of
 * axis 0; 2D: tile_x columns and a slab of tile_y rows (tile_z 1).  NULL =
 * chosen from the program's reach (uhs_tune_program_f64 in uhs_tune.h times
 * the candidates).  The LDS rings (staged planes of the tile plus reach, for
 * the input and each producer) must fit in 64 KiB (UHS_E_CONFIG otherwise). */
int uhs_program_sweep_ex_f64(double *A, double *B, const int64_t *shape, const uhs_program *pg,
                             int timesteps, int kernel, const int *tile, uhs_stream_t stream);

/* FP32 storage and arithmetic (the reference DSL's `float` mode), as
 * uhs_sweep_f32: coefficients rounded to float; compensated = 1 sums every
 * kernel's terms, each producer's included, with Kahan compensation. */
int uhs_program_sweep_f32(float *A, float *B, const int64_t *shape, const uhs_program *pg,
                          int timesteps, int compensated, uhs_stream_t stream);
int uhs_program_sweep_ex_f32(float *A, float *B, const int64_t *shape, const uhs_program *pg,
                             int timesteps, int compensated, int kernel, const int *tile,
                             uhs_stream_t stream);

/* ---- plumbing: device memory, streams, timing ---------------------------- */

/* Streaming device-to-device copy of n doubles, one element per thread: the
 * attainable-HBM-bandwidth reference the benchmark reports, and the
 * calibration kernel for the traffic counters. */
int uhs_copy_f64(double *dst, const double *src, size_t n, uhs_stream_t stream);
/* FP64 FMA throughput probe: n threads each run `iters` rounds of 32
 * independent chains a_i <- fma(a_i, x, y), a_i starting at i, and write
 * sum_i a_i to out[thread] (out holds n doubles).  32 * iters * n FMAs per
 * call, no memory traffic inside the loop: the attainable FP64 vector rate,
 * which the benchmark reports next to the FP64-ALU-bound stencils' roof. */
int uhs_fma_f64(double *out, size_t n, double x, double y, int iters, uhs_stream_t stream);
/* Debug mode, on the current device: with on != 0, every kernel that stages
 * through LDS fills the LDS span of each tile with a NaN pattern (and waits
 * for its workgroup) before staging anything, so a read of a cell a tile
 * never wrote gives NaN instead of a stale value.  Results are otherwise
 * bitwise unchanged.  Read by the kernels when they run (captured graphs
 * included); off by default.  Returns the number of kernel source files
 * whose flag was set. */
int uhs_set_lds_poison(int on);
const char *uhs_last_error(void);
const char *uhs_version(void);
int uhs_device_count(void);
int uhs_device_name(char *buf, size_t buflen);
int uhs_malloc(void **ptr, size_t bytes);
int uhs_free(void *ptr);
int uhs_memcpy_h2d(void *dst, const void *src, size_t bytes);
int uhs_memcpy_d2h(void *dst, const void *src, size_t bytes);
int uhs_memset(void *dst, int value, size_t bytes);
int uhs_synchronize(void);
int uhs_stream_create(uhs_stream_t *s);
int uhs_stream_destroy(uhs_stream_t s);
int uhs_stream_synchronize(uhs_stream_t s);

/* Events for timing on a given stream (hipEvent_t underneath). */
typedef struct ihipEvent_t *uhs_event_t;
int uhs_event_create(uhs_event_t *e);
int uhs_event_destroy(uhs_event_t e);
int uhs_event_record(uhs_event_t e, uhs_stream_t s);
int uhs_event_elapsed_ms(uhs_event_t start, uhs_event_t stop, float *ms);

#ifdef __cplusplus
}
#endif
#endif /* UHS_H */
