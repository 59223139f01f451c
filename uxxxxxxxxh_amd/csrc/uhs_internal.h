// Internal declarations shared by the HIP kernel files and the host library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "uhs.h"
#include <stddef.h>
#include <utility>

struct BuiltinStencil {
    const char *case_name; // UH-GPU case directory, e.g. "3d1r_star"
    const char *symbol;    // reference stencil name, e.g. "stencil_3d7pt"
    int dims, radius, npoints, star;
    const int *off;        // [npoints][dims], DSL term order
    const double *coef;    // DSL coefficients
};

#include "shape_table.inc"

enum Dtype { DT_F64 = 0, DT_F32 = 1, DT_F32K = 2 };

// Everything a launcher needs; built by the host dispatcher.
struct SweepArgs {
    const void *in;        // halo-inclusive source grid (base pointer)
    void *out;             // halo-inclusive destination grid
    int dims;
    int n0, n1, n2;        // interior extents; 2D grids use n0 = rows, n1 = 1, n2 = cols
    int halo;
    long long sy, sz;      // row and plane strides in elements
    int tx, ty, tz, vec;   // tiling configuration
    int flags;             // UHS_CFG_* bits
    int npoints;
    const int *off;        // host copy of the offsets (gather path)
    const double *coef;    // host copy of the coefficients
    int dtype;
    // Multi-step kernels only: the grids the one-step ping-pong calls A and B
    // (base pointers; their halos hold every state's halo) and the global
    // step of the input state.
    const void *hA, *hB;
    int k0;
};

typedef hipError_t (*LaunchFn)(const SweepArgs &, hipStream_t);

// Per-family kernel tables (nullptr when the combination is not compiled).
// ring_cells: halo-ring cells each thread stages per plane (ceil(ring / threads)).
LaunchFn uhs_find_star3d(int prog, int dtype, int vec, int prefetch, int ring_cells);
LaunchFn uhs_find_star3d_rt(int prog, int dtype, int rows);
LaunchFn uhs_find_star3d_tb2(int prog, int dtype, int vec, int prefetch); // two steps per launch
LaunchFn uhs_find_box3d(int prog, int dtype, int vec, int deep, int planes);
LaunchFn uhs_find_box3d_tb2(int prog, int dtype, int vec); // two steps per launch (radius-1 box)
LaunchFn uhs_find_win2d(int prog, int dtype, int vec, int prefetch);
// Two steps per launch, 2D: window = 1 the window form (any stencil), 0 the
// ring form (stars).
LaunchFn uhs_find_win2d_tb2(int prog, int dtype, int vec, int window);
// K steps per launch (2D stars), one function per compiled K; steps 3, 4, 5,
// 6 or 8 (uhs_find_win2d_tbk dispatches).
LaunchFn uhs_find_win2d_tb3(int prog, int dtype, int vec);
LaunchFn uhs_find_win2d_tb4(int prog, int dtype, int vec);
LaunchFn uhs_find_win2d_tb5(int prog, int dtype, int vec);
LaunchFn uhs_find_win2d_tb6(int prog, int dtype, int vec);
LaunchFn uhs_find_win2d_tb8(int prog, int dtype, int vec);
inline LaunchFn uhs_find_win2d_tbk(int prog, int dtype, int vec, int steps)
{
    switch (steps) {
    case 3: return uhs_find_win2d_tb3(prog, dtype, vec);
    case 4: return uhs_find_win2d_tb4(prog, dtype, vec);
    case 5: return uhs_find_win2d_tb5(prog, dtype, vec);
    case 6: return uhs_find_win2d_tb6(prog, dtype, vec);
    case 8: return uhs_find_win2d_tb8(prog, dtype, vec);
    }
    return nullptr;
}
LaunchFn uhs_find_gather(int dtype);

// Limits baked into the kernels; the host rejects configurations beyond them.
constexpr int kMaxThreads = 512;      // __launch_bounds__ of every sweep kernel
constexpr int kStarHaloPerThread = 4; // star3d: halo cells staged per thread per plane
// box3d: plane-tile cells staged per thread.  Radius 1 keeps 8 (its
// best tiles run faster at the occupancy that register count gives);
// radius 2 and 3 keep 4, which frees 16-32 VGPRs for more waves.
constexpr int box_cells_per_thread(int radius) { return radius == 1 ? 8 : 4; }
constexpr int kTb2RingCells = 2;       // star3d_tb2: ring cells staged per thread per plane
constexpr int kBoxTb2RingCells = 2;    // box3d_tb2: level-0 ring cells staged per thread per plane
constexpr int kGatherMaxPoints = 256; // gather: terms passed in kernel arguments
constexpr size_t kMaxLdsBytes = 64 * 1024;

// Number of XCDs workgroups are dealt across (round-robin by workgroup id).
constexpr int kNumXcd = 8;

#if defined(__HIPCC__)
// Calls f(std::integral_constant<int, I>) for I = 0..N-1, fully unrolled at
// compile time, so term offsets and register indices are constants.
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (f(std::integral_constant<int, I>{}), ...);
    }(std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

// LDS poisoning, a debug mode (uhs_set_lds_poison).  With the flag set, a
// tile first fills the LDS span it uses with a NaN pattern and waits for its
// whole workgroup, before it stages anything: a read of a cell the tile never
// wrote then shows up as NaN in a result instead of as a stale value.  The
// flag is read at run time (one uniform load per tile, graphs included).
// Every translation unit has its own copy (no relocatable device code) and
// registers a setter for it with the host library when the library loads.
static __device__ int g_lds_poison;
template <typename T> __device__ __forceinline__ T lds_poison_value()
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, 0x7FF4DEAD7FF4DEADull);
    else return __builtin_bit_cast(T, 0x7FA0DEADu);
}
template <typename T> __device__ __forceinline__ void lds_poison(T *lds, int cells)
{
    if (__builtin_expect(*(volatile int *)&g_lds_poison != 0, 0)) {
        const T v = lds_poison_value<T>();
        for (int i = threadIdx.x; i < cells; i += blockDim.x) lds[i] = v;
        __syncthreads();
    }
}
int uhs_register_lds_poison_setter(hipError_t (*set)(int on));
namespace {
hipError_t uhs_set_lds_poison_here(int on)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(g_lds_poison), &on, sizeof on);
}
[[maybe_unused]] const int g_lds_poison_registered = uhs_register_lds_poison_setter(uhs_set_lds_poison_here);
} // namespace

// Workgroups are dealt round-robin to the XCDs, each with its own L2.  Give
// every XCD one contiguous range of logical tiles, so tiles that share halo
// rows also share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int per = nblocks / kNumXcd, rem = nblocks % kNumXcd;
    const int x = bid % kNumXcd, i = bid / kNumXcd;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
}

// pin(): an empty asm that takes the running sum as an operand and clobbers
// memory.  Placed after a group of terms, it makes the compiler finish those
// terms before any later LDS read, which bounds how many staged values are
// live at once.
//
// Term accumulation.  The first term is a product; every later term is one
// fused multiply-add (FP64/FP32), or a Kahan-compensated add of the fused
// product (FP32 compensated form):
//     y = fma(c, v, -e); u = s + y; e = (u - s) - y; s = u.
template <typename T, bool KAHAN> struct Acc;

template <> struct Acc<double, false> {
    double s;
    __device__ __forceinline__ void zero() { s = 0.0; }
    __device__ __forceinline__ void first(double c, double v) { s = c * v; }
    __device__ __forceinline__ void add(double c, double v) { s = __builtin_fma(c, v, s); }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ void pin() { asm volatile("" : "+v"(s)::"memory"); }
};
template <> struct Acc<float, false> {
    float s;
    __device__ __forceinline__ void zero() { s = 0.0f; }
    __device__ __forceinline__ void first(float c, float v) { s = c * v; }
    __device__ __forceinline__ void add(float c, float v) { s = __builtin_fmaf(c, v, s); }
    __device__ __forceinline__ float value() const { return s; }
    __device__ __forceinline__ void pin() { asm volatile("" : "+v"(s)::"memory"); }
};
template <> struct Acc<float, true> {
    float s, e;
    __device__ __forceinline__ void zero() { s = 0.0f; e = 0.0f; }
    __device__ __forceinline__ void first(float c, float v) { s = c * v; e = 0.0f; }
    __device__ __forceinline__ void add(float c, float v)
    {
        const float y = __builtin_fmaf(c, v, -e);
        const float u = s + y;
        e = (u - s) - y;
        s = u;
    }
    __device__ __forceinline__ float value() const { return s; }
    __device__ __forceinline__ void pin() { asm volatile("" : "+v"(s), "+v"(e)::"memory"); }
};

// Kernel parameters, passed by value (coefficients travel in the kernel
// argument segment: no device allocation, nothing to copy per call).
template <typename T, int P> struct KParams {
    const T *in;           // interior origin (0,0,0) of the source grid
    T *out;                // interior origin of the destination grid
    int n0, n1, n2;
    int sy;
    long long sz;
    int TX, TY, ZC, BXT;
    int tiles_x, tiles_y, nblocks;
    int flags;
    T c[P];
};

// Coefficient access.  Up to kHoistPoints coefficients stay in SGPRs for the
// whole kernel.  Beyond that they cannot (a 343-term FP64 stencil needs 686
// SGPRs): the compiler would spill them into VGPR lanes and pay a v_readlane
// per term.  Large stencils therefore re-read their coefficients with scalar
// loads from the kernel-argument segment every step, through a pointer the
// compiler cannot treat as loop-invariant.
constexpr int kHoistPoints = 32;

template <typename T, int P> struct Coefs {
    typedef const __attribute__((address_space(4))) T *kptr;
    kptr ptr;
    T hoisted[P <= kHoistPoints ? P : 1];
    __device__ __forceinline__ explicit Coefs(const KParams<T, P> &p)
    {
        if constexpr (P <= kHoistPoints) {
            for (int t = 0; t < P; t++) hoisted[t] = p.c[t];
            ptr = nullptr;
        } else {
            auto base = (const __attribute__((address_space(4))) char *)
                __builtin_amdgcn_kernarg_segment_ptr();
            using KP = KParams<T, P>;
            ptr = (kptr)(base + __builtin_offsetof(KP, c));
        }
    }
    // Coefficients stored in device memory (the batched path's per-grid
    // table), read through the constant address space so the loads are scalar.
    __device__ __forceinline__ explicit Coefs(kptr table)
    {
        if constexpr (P <= kHoistPoints) {
            for (int t = 0; t < P; t++) hoisted[t] = table[t];
            ptr = nullptr;
        } else {
            ptr = table;
        }
    }
    // Call once per step: hides the pointer's value from loop-invariant code
    // motion so the scalar loads are issued inside the step.
    __device__ __forceinline__ void refresh()
    {
        if constexpr (P > kHoistPoints) {
            // Through readfirstlane first: the value is then uniform by
            // construction, so the scalar-register operand never needs a
            // vector-to-scalar copy (which the backend cannot emit).
            const unsigned long long v = (unsigned long long)ptr;
            unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
            unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
            asm volatile("" : "+s"(lo), "+s"(hi));
            ptr = (kptr)(((unsigned long long)hi << 32) | lo);
        }
    }
    template <int t> __device__ __forceinline__ T get() const
    {
        if constexpr (P <= kHoistPoints) return hoisted[t];
        else return ptr[t];
    }
};

template <typename T, int P>
inline KParams<T, P> make_params(const SweepArgs &a)
{
    KParams<T, P> p;
    const long long h = a.halo;
    const long long origin = (a.dims == 3 ? h * a.sz : 0) + h * a.sy + h;
    p.in = static_cast<const T *>(a.in) + origin;
    p.out = static_cast<T *>(a.out) + origin;
    p.n0 = a.n0; p.n1 = a.n1; p.n2 = a.n2;
    p.sy = (int)a.sy; p.sz = a.sz;
    p.TX = a.tx; p.TY = a.ty; p.ZC = a.tz; p.BXT = a.tx / a.vec;
    p.flags = a.flags;
    p.tiles_x = (a.n2 + a.tx - 1) / a.tx;
    p.tiles_y = a.dims == 3 ? (a.n1 + a.ty - 1) / a.ty : 1;
    const int chunks = (a.n0 + a.tz - 1) / a.tz;
    p.nblocks = p.tiles_x * p.tiles_y * chunks;
    for (int t = 0; t < P; t++) p.c[t] = (T)a.coef[t];
    return p;
}
#endif
