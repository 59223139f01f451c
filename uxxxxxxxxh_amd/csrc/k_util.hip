// Utility kernels: seeded synthetic grids and FP64 error-norm reductions.
#include "uhs_internal.h"
#include "uhs_util.h"

#include <mutex>

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// One thread per cell of the halo-inclusive grid: interior cells get
// u01(mix64(seed + (L+1)*golden)), halo cells 0.  Counter-based, so the
// values do not depend on the launch shape.
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *g, GridShape s, uint64_t seed)
{
    const long long total = (long long)s.planes * s.rows * s.cols;
    for (long long L = (long long)blockIdx.x * blockDim.x + threadIdx.x; L < total;
         L += (long long)gridDim.x * blockDim.x) {
        const long long plane = (long long)s.rows * s.cols;
        const int z = (int)(L / plane);
        const long long r = L - (long long)z * plane;
        const int y = (int)(r / s.cols), x = (int)(r - (long long)y * s.cols);
        const bool in = z >= s.h0 && z < s.planes - s.h0 && y >= s.h && y < s.rows - s.h &&
                        x >= s.h && x < s.cols - s.h;
        T v = T(0);
        if (in) {
            const uint64_t m = mix64(seed + ((uint64_t)L + 1) * 0x9E3779B97F4A7C15ULL);
            v = (T)((double)(m >> 11) * 0x1.0p-53);
        }
        g[L] = v;
    }
}

// Per-block partial norms over the interior: max|a-b|, sum (a-b)^2, sum b^2.
template <typename TA>
__global__ void __launch_bounds__(256) k_norms(const TA *a, const double *b, GridShape s,
                                               double *partial)
{
    __shared__ double smax[256], sd[256], sr[256];
    const int ni0 = s.planes - 2 * s.h0, ni1 = s.rows - 2 * s.h, ni2 = s.cols - 2 * s.h;
    const long long total = (long long)ni0 * ni1 * ni2;
    double m = 0.0, d2 = 0.0, r2 = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long plane = (long long)ni1 * ni2;
        const int z = (int)(i / plane);
        const long long r = i - (long long)z * plane;
        const int y = (int)(r / ni2), x = (int)(r - (long long)y * ni2);
        const long long L = ((long long)(z + s.h0) * s.rows + (y + s.h)) * s.cols + (x + s.h);
        const double bv = b[L];
        const double d = (double)a[L] - bv;
        m = fmax(m, fabs(d));
        d2 = __builtin_fma(d, d, d2);
        r2 = __builtin_fma(bv, bv, r2);
    }
    const int t = threadIdx.x;
    smax[t] = m; sd[t] = d2; sr[t] = r2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            smax[t] = fmax(smax[t], smax[t + w]);
            sd[t] += sd[t + w];
            sr[t] += sr[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        partial[3 * blockIdx.x] = smax[0];
        partial[3 * blockIdx.x + 1] = sd[0];
        partial[3 * blockIdx.x + 2] = sr[0];
    }
}

// Folds nparts partial triples (in a fixed order, so the result is
// deterministic) into out[0..2].
__global__ void __launch_bounds__(256) k_norms_final(const double *partial, int nparts, double *out)
{
    __shared__ double smax[256], sd[256], sr[256];
    const int t = threadIdx.x;
    double m = 0.0, d2 = 0.0, r2 = 0.0;
    for (int i = t; i < nparts; i += 256) {
        m = fmax(m, partial[3 * i]);
        d2 += partial[3 * i + 1];
        r2 += partial[3 * i + 2];
    }
    smax[t] = m; sd[t] = d2; sr[t] = r2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            smax[t] = fmax(smax[t], smax[t + w]);
            sd[t] += sd[t + w];
            sr[t] += sr[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = smax[0];
        out[1] = sqrt(sd[0]);
        out[2] = sqrt(sr[0]);
    }
}

// Streaming copy, one element per thread: the reference point for the HBM
// bandwidth this device reaches (and the calibration target of the traffic
// counters: it reads and writes exactly n * 8 bytes).
__global__ void __launch_bounds__(256) k_copy(const double *__restrict__ a, double *__restrict__ b, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// FP64 FMA throughput: 32 independent chains per thread, no memory traffic
// inside the loop (x, y and iters are runtime values, so nothing folds).
__global__ void __launch_bounds__(256) k_fma(double *__restrict__ out, long long n, double x, double y, int iters)
{
    constexpr int NA = 32;
    double a[NA];
#pragma unroll
    for (int i = 0; i < NA; i++) a[i] = (double)i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < NA; i++) a[i] = __builtin_fma(a[i], x, y);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NA; i++) s += a[i];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = s;
}

} // namespace

hipError_t uhs_launch_fma(double *out, long long n, double x, double y, int iters, hipStream_t st)
{
    hipLaunchKernelGGL(k_fma, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n, x, y, iters);
    return hipGetLastError();
}

hipError_t uhs_launch_copy(const double *src, double *dst, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

hipError_t uhs_launch_fill(void *grid, int is_f32, const GridShape &s, uint64_t seed,
                           hipStream_t st)
{
    const long long total = (long long)s.planes * s.rows * s.cols;
    const long long want = (total + 255) / 256;
    const unsigned blocks = (unsigned)(want < 65536 ? (want > 0 ? want : 1) : 65536);
    if (is_f32)
        hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(256), 0, st, (float *)grid, s, seed);
    else
        hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(256), 0, st, (double *)grid, s, seed);
    return hipGetLastError();
}

hipError_t uhs_launch_norms(const void *a, int a_is_f32, const double *b, const GridShape &s,
                            double *partial, int nparts, double *out3, hipStream_t st)
{
    if (a_is_f32)
        hipLaunchKernelGGL(k_norms<float>, dim3(nparts), dim3(256), 0, st, (const float *)a, b, s,
                           partial);
    else
        hipLaunchKernelGGL(k_norms<double>, dim3(nparts), dim3(256), 0, st, (const double *)a, b,
                           s, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_norms_final, dim3(1), dim3(256), 0, st, partial, nparts, out3);
    return hipGetLastError();
}

// The partial-norm triples and the result of the allocation-free norm entry
// (uhs_error_norms_f32_f64_noalloc): one copy per device, held until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: only the address hipGetSymbolAddress returns for the
// current device is ever passed to a kernel or a copy.
static __device__ __align__(256) double scratch[kNormScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
} // namespace

hipError_t uhs_norms_scratch(double **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        // First call on this device (it is current): look its copy up once.
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = (double *)p;
    }
    *ptr = g_scratch[dev];
    return hipSuccess;
}
