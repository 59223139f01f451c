// One reference-symbol shim library per UH-GPU case (see include/uhs_compat.h).
// Built once per case with -DUHS_CASE=<case> -DUHS_SYMBOL=<reference name>
// -DUHS_DIMS=<2|3>; forwards to uhs_sweep_f64 in libuhs.so.
#include "uhs.h"
#include "uhs_compat.h"

#include <cstdio>
#include <cstring>

#define UHS_STR2(x) #x
#define UHS_STR(x) UHS_STR2(x)
#define UHS_CAT2(a, b) a##b
#define UHS_CAT(a, b) UHS_CAT2(a, b)

#if UHS_DIMS == 3
typedef uhs_memref3d_f64 MemRef;
#else
typedef uhs_memref2d_f64 MemRef;
#endif

namespace {
thread_local int g_status = UHS_OK;

// iteration(T) of the case's DSL; 4 for all 12 UH-GPU cases.
constexpr int kIteration = UHS_ITERATION;

int run(MemRef *A, MemRef *B, int timesteps)
{
    const int idx = uhs_builtin_find(UHS_STR(UHS_CASE));
    int dims = 0, radius = 0, npoints = 0;
    const int *off = nullptr;
    const double *coef = nullptr;
    if (idx < 0 || uhs_builtin_info(idx, nullptr, nullptr, &dims, &radius, &npoints, &off, &coef))
        return UHS_E_INVAL;
    if (!A || !B || !A->aligned || !B->aligned) {
        fprintf(stderr, "%s: null MemRef\n", UHS_STR(UHS_SYMBOL));
        return UHS_E_INVAL;
    }
    int64_t ext[3];
    for (int k = 0; k < dims; k++) {
        if (A->shape[k] != B->shape[k] || A->shape[k] <= 2 * radius) {
            fprintf(stderr, "%s: grid shapes must match and exceed twice the halo\n",
                    UHS_STR(UHS_SYMBOL));
            return UHS_E_INVAL;
        }
        ext[k] = A->shape[k] - 2 * radius;
    }
    int64_t stride = 1;
    for (int k = dims - 1; k >= 0; k--) {
        if (A->stride[k] != stride || B->stride[k] != stride) {
            fprintf(stderr, "%s: only dense row-major grids are supported\n", UHS_STR(UHS_SYMBOL));
            return UHS_E_INVAL;
        }
        stride *= A->shape[k];
    }
    uhs_stencil st = {dims, npoints, off, coef};
    int rc = uhs_sweep_f64(A->aligned + A->offset, B->aligned + B->offset, dims, ext, radius, &st,
                           timesteps, nullptr, nullptr);
    if (rc == UHS_OK) rc = uhs_synchronize();
    if (rc != UHS_OK) fprintf(stderr, "%s: %s\n", UHS_STR(UHS_SYMBOL), uhs_last_error());
    return rc;
}
} // namespace

extern "C" {

void UHS_CAT(UHS_CAT(_mlir_ciface_, UHS_SYMBOL), _iteration)(MemRef *A, MemRef *B)
{
    g_status = run(A, B, kIteration);
}

void UHS_CAT(_mlir_ciface_, UHS_SYMBOL)(MemRef *A, MemRef *B) { g_status = run(A, B, 1); }

int uhs_compat_last_status(void) { return g_status; }
const char *uhs_compat_case(void) { return UHS_STR(UHS_CASE); }
}
