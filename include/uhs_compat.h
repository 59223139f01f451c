/*
 * uhs_compat.h -- the reference's own kernel entry, kept symbol for symbol.
 *
 * The reference's GPU drivers declare and call, per benchmark case,
 *     extern "C" void _mlir_ciface_stencil_<name>_iteration(MemRef *A, MemRef *B);
 * (UH-GPU/3d1r_star/3d7pt_main.cpp:25 and :150; UH-GPU/2d1r_star/2d5pt_main.cpp:23;
 * the symbol exists because ConvertStencilToStandard.cpp:80 sets
 * llvm.emit_c_interface on the generated function).  The generated code also
 * exports the one-sweep function _mlir_ciface_stencil_<name>(A, B)
 * (FuncOpLowering, ConvertStencilToStandard.cpp:51-89).
 *
 * Each libuhs_compat_<case>.so (uxxxxxxxxh_amd/compat/) exports exactly those
 * two symbols for one UH-GPU case, with that case's DSL coefficients and
 * iteration count baked in, like the per-case reference binaries.  Several
 * cases share a reference name (stencil_2d9pt is both 2d1r_box and 2d2r_star;
 * stencil_2d3r and stencil_3d3r each serve a star and a box), which is why
 * the shims are separate shared objects.
 *
 * Semantics (as the reference):
 *   - A and B are halo-inclusive grids described by the MemRef descriptors
 *     (3d7pt_main.cpp:17-23,147-148): data at aligned + offset, shape[k] =
 *     n_k + 2r, row-major strides {.., shape[R-1], 1}.  The halo width is
 *     the stencil radius r (UHAffineToParallel.cpp:1674-1680).
 *   - _iteration runs the DSL's iteration(T) (T = 4 in all 12 cases):
 *     T/2 pairs of sweep(A->B), sweep(B->A); the result is in A
 *     (MLIRGen.cpp:317-349, ConvertStencilToStandard.cpp:298-337).
 *   - the one-sweep symbol writes B = sweep(A).
 *   - the reference returns nothing and reports nothing; these shims also
 *     return void, run on the default stream and wait for completion, and
 *     on a bad descriptor print a message to stderr and leave the grids
 *     unchanged.  uhs_compat_last_status() returns the UHS_* code of the last
 *     call from the calling thread (an addition, not in the reference).
 */
#ifndef UHS_COMPAT_H
#define UHS_COMPAT_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* The reference drivers name both structs TwoDMemrefF64 (3d7pt_main.cpp:17). */
typedef struct uhs_memref2d_f64 {
    double *allocated;
    double *aligned;
    int64_t offset;
    int64_t shape[2];
    int64_t stride[2];
} uhs_memref2d_f64;

typedef struct uhs_memref3d_f64 {
    double *allocated;
    double *aligned;
    int64_t offset;
    int64_t shape[3];
    int64_t stride[3];
} uhs_memref3d_f64;

/* 2D cases: 2d1r_star stencil_2d5pt, 2d2r_star/2d1r_box stencil_2d9pt,
 * 2d2r_box stencil_2d25pt, 2d3r_star/2d3r_box stencil_2d3r.  Each shim
 * library defines one pair:
 *   void _mlir_ciface_stencil_<name>_iteration(uhs_memref2d_f64 *A, uhs_memref2d_f64 *B);
 *   void _mlir_ciface_stencil_<name>(uhs_memref2d_f64 *A, uhs_memref2d_f64 *B);
 * 3D cases: 3d1r_star stencil_3d7pt, 3d2r_star stencil_3d13pt,
 * 3d1r_box stencil_3d27pt, 3d2r_box stencil_3d125pt,
 * 3d3r_star/3d3r_box stencil_3d3r, with uhs_memref3d_f64. */

int uhs_compat_last_status(void);
/* The UH-GPU case ("3d1r_star", ...) this shim library was built for. */
const char *uhs_compat_case(void);

#ifdef __cplusplus
}
#endif
#endif /* UHS_COMPAT_H */
