/*
 * CPU oracle for the structured-grid stencil sweep -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or as the timed CPU baseline).  The
 * product path (libuhs.so) never links or calls it.
 *
 * It restates the reference's serial sweep `cpu_verify`
 * (UH-GPU/3d1r_star/3d7pt_main.cpp:50-78, the same loop in all 12 UH-GPU/<case>
 * drivers) and the DSL semantics it mirrors:
 *   - out[p] = c0*in[p+o0] + c1*in[p+o1] + ...  left-associative in DSL term
 *     order, '*' binding tighter than '+'  (UHStencil/include/Parser/Parser.h:437-468,
 *     591-606; UHStencil/lib/Parser/MLIRGen.cpp:452-501);
 *   - offsets are positional: the k-th bracket offsets array axis k
 *     (Parser.h:525-575; ConvertStencilToStandard.cpp:228-272);
 *   - the domain is the interior [h, n+h) per axis; halo cells of `out` are
 *     never written (ConvertStencilToStandard.cpp:91-149,339-381);
 *   - iteration(T): T/2 pairs of {sweep(A->B); sweep(B->A)}, result in A
 *     (MLIRGen.cpp:317-349; ConvertStencilToStandard.cpp:298-337).  T=1 is
 *     one sweep A->B.  Odd T>1 (an extension) leaves the result in B.
 *
 * Grids are dense, row-major and halo-inclusive: axis 0 outermost, the last
 * axis unit-stride, each axis n_k + 2*halo long.
 *
 * Arithmetic modes (f64 and f32):
 *   ORC_PLAIN  separate multiply and add per term, no contraction (the
 *              reference's cpu_verify compiled with -ffp-contract=off);
 *   ORC_FUSED  first term a product, every later term one fma(c, v, acc) --
 *              the exact operation sequence of the HIP kernels, so the
 *              comparison with them is bitwise;
 *   ORC_KAHAN  (f32 only) Kahan-compensated sum, the sequence of the FP32
 *              compensated HIP kernels:
 *                s = c0*v0; e = 0;
 *                for t>=1: y = fma(ct, vt, -e); u = s + y; e = (u - s) - y; s = u;
 */
#ifndef UHS_ORACLE_H
#define UHS_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PLAIN = 0, ORC_FUSED = 1, ORC_KAHAN = 2 };

/* One sweep over the interior.  ext[dims] = interior extents (axis 0 first),
 * off[P*dims] = per-term offsets in term order, coef[P].  nthreads<=0: all
 * cores.  Returns 0, or -1 on bad arguments. */
int orc_sweep_f64(const double *in, double *out, int dims, const int64_t *ext,
                  int halo, int P, const int *off, const double *coef,
                  int mode, int nthreads);
int orc_sweep_f32(const float *in, float *out, int dims, const int64_t *ext,
                  int halo, int P, const int *off, const float *coef,
                  int mode, int nthreads);

/* T sweeps with ping-pong between A and B (see the header comment). */
int orc_iterate_f64(double *A, double *B, int dims, const int64_t *ext,
                    int halo, int P, const int *off, const double *coef,
                    int T, int mode, int nthreads);
int orc_iterate_f32(float *A, float *B, int dims, const int64_t *ext,
                    int halo, int P, const int *off, const float *coef,
                    int T, int mode, int nthreads);

/* Fused multi-kernel programs (the reference's multi-kernel stencils after
 * kernel fusion, UHStencil/lib/Dialect/Stencil/StencilKernelFusion.cpp:66-180:
 * a producer with a single consumer is cloned into it at every access
 * offset, repeatedly, so chains fuse completely):
 *   out[p] = sum_j c_j * x_j,  x_j = in[p+o_j]         (src[j] == -1)
 *                              x_j = P_k(p+o_j)         (src[j] == k)
 *   P_k(q) = sum_i d_ki * y_ki, y_ki = in[q+e_ki] or P_s(q+e_ki) (s = prod_src[i] < k)
 * each sum left to right in term order, in `mode` (ORC_PLAIN or ORC_FUSED).
 * Grids are `shape` cells per axis, halo included; cells [lo, hi) are
 * written, the rest never.  Producers: nprod of them, producer k has
 * prod_np[k] terms; their offsets, coefficients and sources are
 * concatenated in producer order (prod_off[(sum of earlier np + i)*dims +
 * a]); prod_src NULL = every producer term reads the input.  T sweeps
 * ping-pong A and B as orc_iterate does.  Returns 0, or -1 on bad
 * arguments (a read outside the grid or a producer reading itself or a
 * later producer included). */
int orc_program_f64(double *A, double *B, int dims, const int64_t *shape,
                    int nprod, const int *prod_np, const int *prod_off, const double *prod_coef,
                    const int *prod_src,
                    int np, const int *off, const double *coef, const int *src,
                    const int64_t *lo, const int64_t *hi, int T, int mode);
/* The same in FP32 storage and arithmetic (coefficients already rounded to
 * float), mode ORC_PLAIN, ORC_FUSED or ORC_KAHAN: every kernel's sum, each
 * producer's included, compensated as in orc_sweep_f32. */
int orc_program_f32(float *A, float *B, int dims, const int64_t *shape,
                    int nprod, const int *prod_np, const int *prod_off, const float *prod_coef,
                    const int *prod_src,
                    int np, const int *off, const float *coef, const int *src,
                    const int64_t *lo, const int64_t *hi, int T, int mode);

/* Seeded synthetic inputs.  Interior cell with halo-inclusive linear index L
 * gets u01(mix64(seed + (L+1)*0x9E3779B97F4A7C15)), halo cells get 0. */
void orc_fill_uniform_f64(double *a, int dims, const int64_t *ext, int halo,
                          uint64_t seed);
void orc_fill_uniform_f32(float *a, int dims, const int64_t *ext, int halo,
                          uint64_t seed);
/* Symmetric (c(o) == c(-o)) i.i.d. U[0,1) coefficients normalised to sum 1. */
void orc_make_coeffs(int dims, int P, const int *off, uint64_t seed,
                     double *coef);
/* Reference-workload fill: every cell, halo included, rand() % 50 in
 * row-major order with libc's default seed (3d7pt_main.cpp:133-141). */
void orc_fill_rand50_f64(double *a, size_t n);

uint64_t orc_mix64(uint64_t z);

#ifdef __cplusplus
}
#endif
#endif
