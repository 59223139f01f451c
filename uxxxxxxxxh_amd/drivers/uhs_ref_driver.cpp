// The reference-workload driver: what each UH-GPU case's main program does
// (UH-GPU/<case>/*_main.cpp, e.g. UH-GPU/3d1r_star/3d7pt_main.cpp:105-160),
// for any of the 12 cases, on the MI355X library through the reference's own
// entry symbol.
//
//     uhs_ref_driver <case> [--size N] [--reps R] [--no-verify]
//
//  1. Loads the case's shim library (compat/libuhs_compat_<case>.so), which
//     exports _mlir_ciface_stencil_<name>_iteration with the case's DSL
//     coefficients and iteration(4) baked in.
//  2. Fills every cell of the grid, halo included, with glibc rand() % 50 in
//     row-major order, no srand (3d7pt_main.cpp:133-141); the second grid is
//     a copy (the reference leaves it uninitialised, which its own check
//     only tolerates when the halo happens to be zero).
//  3. Copies both grids to the device, calls the entry through MemRef
//     descriptors (3d7pt_main.cpp:17-23,147-151), timed with HIP events on
//     the stream the shim runs on, and with the wall clock as the reference
//     does (3d7pt_main.cpp:149-153); prints the device time.
//  4. Runs the driver's own serial sweep on the host (cpu_verify,
//     3d7pt_main.cpp:50-103: iteration/2 pairs of sweeps, each term a
//     product then an addition, left to right in DSL order), OpenMP over
//     rows, and applies the reference's check |cpu - gpu| <= 1e-6
//     (3d7pt_main.cpp:84-96) -- at the full size for all 12 cases, also
//     those whose reference driver comments the check out.
//  5. Prints two 64-bit checksums of the result grid's bit patterns (their
//     sum and their position-weighted sum), so a test can compare the result
//     with another computation's without writing the grid out.
//
// Sizes default to the reference's: 8192^2 (2D) and 512^3 (3D) interior.
#include "uhs.h"
#include "uhs_compat.h"

#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <omp.h>
#include <sys/time.h>
#include <unistd.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

double wall_ms()
{
    timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec * 1e3 + t.tv_usec * 1e-3;
}

std::string exe_dir()
{
    char buf[4096];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (n <= 0) return ".";
    buf[n] = 0;
    std::string p(buf);
    const size_t s = p.rfind('/');
    return s == std::string::npos ? "." : p.substr(0, s);
}

// Two checksums of a grid's bit patterns (mod 2^64): their plain sum, and
// the sum weighted by position (i + 1), which also catches permutations.
void checksums(const double *g, int64_t n, uint64_t *sum, uint64_t *wsum)
{
    uint64_t s = 0, w = 0;
    for (int64_t i = 0; i < n; i++) {
        uint64_t b;
        memcpy(&b, g + i, sizeof b);
        s += b;
        w += b * (uint64_t)(i + 1);
    }
    *sum = s;
    *wsum = w;
}

#define HIPCHECK(x)                                                                      \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));               \
            exit(2);                                                                     \
        }                                                                                \
    } while (0)

// One sweep in -> out over the interior [r, n + r) of a halo-inclusive grid
// of (planes, rows, cols) cells (2D: one plane, no plane halo): each term a
// rounded product added to the running sum, left to right in term order.
void cpu_sweep(const double *in, double *out, int dims, int64_t planes, int64_t rows, int64_t cols, int r,
               int P, const int *off, const double *coef)
{
    std::vector<int64_t> lin(P);
    for (int t = 0; t < P; t++) {
        const int dz = dims == 3 ? off[3 * t] : 0;
        const int dy = off[dims * t + dims - 2], dx = off[dims * t + dims - 1];
        lin[t] = (int64_t)dz * rows * cols + (int64_t)dy * cols + dx;
    }
    const int64_t z0 = dims == 3 ? r : 0, z1 = dims == 3 ? planes - r : 1;
    const int64_t nrows = (z1 - z0) * (rows - 2 * r);
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < nrows; row++) {
        const int64_t z = z0 + row / (rows - 2 * r), y = r + row % (rows - 2 * r);
        const int64_t base = (z * rows + y) * cols;
        for (int64_t x = r; x < cols - r; x++) {
            const double *p = in + base + x;
            double s = coef[0] * p[lin[0]];
            for (int t = 1; t < P; t++) {
                const double prod = coef[t] * p[lin[t]];
                s = s + prod;
            }
            out[base + x] = s;
        }
    }
}

int usage()
{
    fprintf(stderr, "usage: uhs_ref_driver <case> [--size N] [--reps R] [--no-verify]\n"
                    "  case: one of the 12 UH-GPU cases (2d1r_star ... 3d3r_box)\n");
    return 1;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) return usage();
    const char *case_name = argv[1];
    int64_t size = 0;
    int reps = 1;
    bool verify = true;
    for (int i = 2; i < argc; i++) {
        if (!strcmp(argv[i], "--size") && i + 1 < argc) size = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--reps") && i + 1 < argc) reps = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--no-verify")) verify = false;
        else return usage();
    }
    const int idx = uhs_builtin_find(case_name);
    const char *symbol = nullptr;
    int dims = 0, r = 0, P = 0;
    const int *off = nullptr;
    const double *coef = nullptr;
    if (idx < 0 || uhs_builtin_info(idx, nullptr, &symbol, &dims, &r, &P, &off, &coef) != UHS_OK) {
        fprintf(stderr, "unknown case '%s'\n", case_name);
        return usage();
    }
    if (size <= 0) size = dims == 2 ? 8192 : 512;
    if (reps < 1) reps = 1;

    // The case's shim library and the reference's entry symbol.
    const std::string lib = exe_dir() + "/../compat/libuhs_compat_" + case_name + ".so";
    void *h = dlopen(lib.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "cannot load %s: %s\n", lib.c_str(), dlerror());
        return 2;
    }
    const std::string sym = std::string("_mlir_ciface_") + symbol + "_iteration";
    typedef void (*Entry)(void *, void *);
    Entry entry = (Entry)dlsym(h, sym.c_str());
    typedef int (*StatusFn)(void);
    StatusFn status = (StatusFn)dlsym(h, "uhs_compat_last_status");
    if (!entry || !status) {
        fprintf(stderr, "%s does not export %s\n", lib.c_str(), sym.c_str());
        return 2;
    }
    const int T = 4; // iteration(4) in every UH-GPU DSL

    const int64_t Q = size + 2 * r; // cells per axis, halo included
    const int64_t planes = dims == 3 ? Q : 1, cells = planes * Q * Q;
    const size_t bytes = sizeof(double) * (size_t)cells;
    printf("%s (%s, %d terms): %lld^%d interior, halo %d, iteration(%d)\n", case_name, symbol, P,
           (long long)size, dims, r, T);

    // Initialising data, as the reference: rand() % 50 over every cell.
    printf("Initializing data...\n");
    std::vector<double> input_cpu(cells), tmp_cpu(cells), input_gpu(cells);
    for (int64_t i = 0; i < cells; i++) {
        const double v = rand() % 50;
        input_cpu[i] = v;
        input_gpu[i] = v;
    }
    tmp_cpu = input_cpu; // the second grid's halo: a copy of the first's

    double *dA = nullptr, *dB = nullptr;
    HIPCHECK(hipMalloc(&dA, bytes));
    HIPCHECK(hipMalloc(&dB, bytes));
    HIPCHECK(hipMemcpy(dA, input_gpu.data(), bytes, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dB, input_gpu.data(), bytes, hipMemcpyHostToDevice));
    HIPCHECK(hipDeviceSynchronize());

    uhs_memref3d_f64 A3{dA, dA, 0, {Q, Q, Q}, {Q * Q, Q, 1}}, B3{dB, dB, 0, {Q, Q, Q}, {Q * Q, Q, 1}};
    uhs_memref2d_f64 A2{dA, dA, 0, {Q, Q}, {Q, 1}}, B2{dB, dB, 0, {Q, Q}, {Q, 1}};
    void *mA = dims == 3 ? (void *)&A3 : (void *)&A2;
    void *mB = dims == 3 ? (void *)&B3 : (void *)&B2;

    // The first call is the reference's own (one call on fresh inputs); later
    // repetitions (--reps) restore the inputs first and only add timings.
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    double best_ev = 1e30, first_wall = 0;
    for (int rep = 0; rep < reps; rep++) {
        if (rep) {
            HIPCHECK(hipMemcpy(dA, input_cpu.data(), bytes, hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(dB, input_cpu.data(), bytes, hipMemcpyHostToDevice));
            HIPCHECK(hipDeviceSynchronize());
        }
        const double t0 = wall_ms();
        HIPCHECK(hipEventRecord(e0, nullptr));
        entry(mA, mB); // the shim runs on the default stream and waits for completion
        HIPCHECK(hipEventRecord(e1, nullptr));
        HIPCHECK(hipDeviceSynchronize());
        const double wall = wall_ms() - t0;
        float ev = 0.f;
        HIPCHECK(hipEventElapsedTime(&ev, e0, e1));
        if (status() != UHS_OK) {
            fprintf(stderr, "%s returned status %d\n", sym.c_str(), status());
            return 2;
        }
        if (rep == 0) first_wall = wall;
        if (ev < best_ev) best_ev = ev;
    }
    HIPCHECK(hipMemcpy(input_gpu.data(), dA, bytes, hipMemcpyDeviceToHost));
    printf("device elapsed Time: %lf (ms)\n", first_wall);
    const double updates = (double)size * size * (dims == 3 ? size : 1) * T;
    printf("device event time: %.4f ms (best of %d), %.2f GStencils/s\n", best_ev, reps,
           updates / (best_ev * 1e-3) / 1e9);

    int rc = 0;
    if (verify) {
        printf("cpu computing\n");
        const double t0 = wall_ms();
        for (int it = 0; it < T / 2; it++) {
            cpu_sweep(input_cpu.data(), tmp_cpu.data(), dims, planes, Q, Q, r, P, off, coef);
            cpu_sweep(tmp_cpu.data(), input_cpu.data(), dims, planes, Q, Q, r, P, off, coef);
        }
        printf("cpu elapsed Time: %lf (ms), %d threads\n", wall_ms() - t0, omp_get_max_threads());
        double maxerr = 0.0;
        int64_t bad = -1;
        for (int64_t i = 0; i < cells; i++) {
            const double err = std::fabs(input_cpu[i] - input_gpu[i]);
            if (!(err <= maxerr)) maxerr = err;
            if (!(err <= 1e-6) && bad < 0) bad = i;
        }
        if (bad < 0) {
            printf("Verify Success (max |cpu - gpu| = %.3e)\n", maxerr);
        } else {
            printf("Verify Failed at cell %lld: cpu %lf gpu %lf\n", (long long)bad, input_cpu[bad],
                   input_gpu[bad]);
            rc = 3;
        }
    }
    uint64_t cs = 0, cw = 0;
    checksums(input_gpu.data(), cells, &cs, &cw);
    printf("result checksum: sum %016llx weighted %016llx\n", (unsigned long long)cs, (unsigned long long)cw);
    HIPCHECK(hipFree(dA));
    HIPCHECK(hipFree(dB));
    dlclose(h);
    return rc;
}
