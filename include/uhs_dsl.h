/*
 * uhs_dsl.h -- reading the reference's stencil DSL (.dsl) into a uhs_stencil.
 *
 * Accepts the form every UH-GPU benchmark .dsl file uses: one
 * `stencil` block declaring a double or float array, iteration(T), and one
 * kernel whose expr is a sum of coefficient*access terms (the reference
 * parser: UHStencil/include/Parser/Parser.h:66-409 and :412-606; lexer
 * UHStencil/include/Parser/Lexer.h:31-59,155-215).  Offsets are positional
 * (the k-th bracket is axis k) and terms keep their DSL order, which is the
 * summation order.  uhs_parse_dsl reads one kernel (programs with several
 * kernels return UHS_E_UNSUPPORTED there; uhs_parse_dsl_program below reads
 * them); general arithmetic (parentheses, division) returns UHS_E_PARSE.
 */
#ifndef UHS_DSL_H
#define UHS_DSL_H
#include <stdint.h>
#include "uhs.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct uhs_dsl_info {
    char name[64];          /* stencil name, e.g. "stencil_3d7pt" */
    char kernel[64];        /* kernel name */
    int dims;               /* 2 or 3 */
    int is_float;           /* 1 for a `float` array, 0 for `double` */
    int iteration;          /* iteration(T); 1 when absent */
    int halo;               /* the domain's lower bound (= its margin on every side) */
    int npoints;
    int64_t grid[3];        /* declared array extents, halo included */
    int64_t domain_lo[3], domain_hi[3];
    int64_t extents[3];     /* interior extents = domain_hi - domain_lo */
    int offsets[UHS_MAX_POINTS * 3]; /* [npoints][dims], term order */
    double coeffs[UHS_MAX_POINTS];
} uhs_dsl_info;

int uhs_parse_dsl(const char *text, uhs_dsl_info *out);
int uhs_parse_dsl_file(const char *path, uhs_dsl_info *out);

/* Programs of several kernels (the fused "nested" examples,
 * UHStencil/examples/{2d5pt_nested,3d7pt9pt_nested}/).  The operation kernel
 * (the last one when `operation` is absent) becomes the program's output;
 * every kernel it reads, directly or through other kernels, becomes a
 * producer, in definition order (a kernel may read the array and kernels
 * defined before it); kernels it does not read are dropped.  The output
 * domain is the operation kernel's domain, or, without one, the
 * cells whose reads stay inside the array (one-sided programs keep the
 * side they do not read).  Up to UHS_MAX_PRODUCERS producers and
 * UHS_MAX_PROGRAM_TERMS terms per kernel (UHS_E_UNSUPPORTED beyond). */
typedef struct uhs_dsl_program {
    char name[64];
    char operation[64];
    int dims, is_float, iteration;
    int64_t grid[3];                 /* declared array extents, halo included */
    int64_t domain_lo[3], domain_hi[3];
    int nproducers;
    char producer_names[UHS_MAX_PRODUCERS][64];
    int producer_npoints[UHS_MAX_PRODUCERS];
    int producer_offsets[UHS_MAX_PRODUCERS][UHS_MAX_PROGRAM_TERMS * 3];
    double producer_coeffs[UHS_MAX_PRODUCERS][UHS_MAX_PROGRAM_TERMS];
    int producer_sources[UHS_MAX_PRODUCERS][UHS_MAX_PROGRAM_TERMS]; /* -1 the array, s producer s */
    int npoints;                     /* output kernel terms */
    int offsets[UHS_MAX_PROGRAM_TERMS * 3];
    double coeffs[UHS_MAX_PROGRAM_TERMS];
    int sources[UHS_MAX_PROGRAM_TERMS]; /* -1 the array, k producer k */
} uhs_dsl_program;

int uhs_parse_dsl_program(const char *text, uhs_dsl_program *out);
int uhs_parse_dsl_program_file(const char *path, uhs_dsl_program *out);
/* Fills *pg (and producers[0..nproducers-1]) with pointers into *dp, for
 * uhs_program_sweep_f64 on grids of dp->grid cells. */
int uhs_dsl_program_view(const uhs_dsl_program *dp, uhs_stencil *producers, uhs_program *pg);

#ifdef __cplusplus
}
#endif
#endif /* UHS_DSL_H */
