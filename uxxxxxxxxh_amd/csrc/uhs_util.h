// Host-callable launchers of the utility kernels (k_util.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Halo-inclusive grid shape: planes x rows x cols cells, halo h0 on axis 0
// (0 for 2D grids, which are one plane) and h on the other axes.
struct GridShape {
    int planes, rows, cols;
    int h0, h;
};

hipError_t uhs_launch_fill(void *grid, int is_f32, const GridShape &s, uint64_t seed,
                           hipStream_t st);
// partial must hold 3*nparts doubles, out3 three doubles (device memory).
hipError_t uhs_launch_norms(const void *a, int a_is_f32, const double *b, const GridShape &s,
                            double *partial, int nparts, double *out3, hipStream_t st);
// Partial triples of one norm reduction (one per workgroup of k_norms).
constexpr int kNormParts = 1024;
// Device scratch of the allocation-free norm entry: the partial triples and
// the three results (24.6 KB per device).
constexpr int kNormScratchDoubles = 3 * kNormParts + 3;
// This device's copy of that scratch (looked up on the first call per device).
hipError_t uhs_norms_scratch(double **ptr);
hipError_t uhs_launch_copy(const double *src, double *dst, long long n, hipStream_t st);
hipError_t uhs_launch_fma(double *out, long long n, double x, double y, int iters, hipStream_t st);
