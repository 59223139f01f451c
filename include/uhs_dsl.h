/*
 * uhs_dsl.h -- reading the reference's stencil DSL (.dsl) into a uhs_stencil.
 *
 * Accepts the form every UH-GPU benchmark .dsl file uses: one
 * `stencil` block declaring a double or float array, iteration(T), and one
 * kernel whose expr is a sum of coefficient*access terms (the reference
 * parser: UHStencil/include/Parser/Parser.h:66-409 and :412-606; lexer
 * UHStencil/include/Parser/Lexer.h:31-59,155-215).  Offsets are positional
 * (the k-th bracket is axis k) and terms keep their DSL order, which is the
 * summation order.  Programs with several kernels (the fused "nested"
 * examples) return UHS_E_UNSUPPORTED; general arithmetic (parentheses,
 * division) returns UHS_E_PARSE.
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

#ifdef __cplusplus
}
#endif
#endif /* UHS_DSL_H */
