// The learned cost model: a 20 -> 128 -> 128 -> 128 -> 1 ReLU MLP in FP32
// whose forward, backward, Adam and inference passes are HIP kernels.
// Sizes are tiny (35,841 parameters, minibatches of a few hundred rows), so
// the kernels are plain row-blocked FMA loops with the activations staged in
// LDS; every reduction has a fixed order, so training is deterministic.
#include "uhs.h"
#include "uhs_host.h"
#include "uhs_tune.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

constexpr int IN = UHS_MLP_INPUTS, H = UHS_MLP_HIDDEN;
constexpr int ROWS = 8; // rows per workgroup in the dense kernels

// Parameter layout (floats): W1 b1 W2 b2 W3 b3 W4 b4, W row-major [in][out].
constexpr int OW1 = 0, OB1 = OW1 + IN * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H * H, OW4 = OB3 + H, OB4 = OW4 + H, NPARAM = OB4 + 1;
constexpr int LIN[4] = {IN, H, H, H}, LOUT[4] = {H, H, H, 1};
constexpr int OW[4] = {OW1, OW2, OW3, OW4}, OB[4] = {OB1, OB2, OB3, OB4};

// Y[r][o] = act(sum_i X[r][i] W[i][o] + b[o]) for ROWS rows per workgroup.
__global__ void __launch_bounds__(128) k_dense_fwd(const float *X, const float *W, const float *b,
                                                   float *Y, int n, int in, int out, int relu)
{
    __shared__ float xs[ROWS][H];
    const int r0 = blockIdx.x * ROWS;
    for (int k = threadIdx.x; k < ROWS * in; k += blockDim.x) {
        const int r = k / in, i = k - r * in;
        xs[r][i] = r0 + r < n ? X[(size_t)(r0 + r) * in + i] : 0.f;
    }
    __syncthreads();
    const int o = threadIdx.x;
    if (o >= out) return;
    float acc[ROWS];
    for (int r = 0; r < ROWS; r++) acc[r] = b[o];
    for (int i = 0; i < in; i++) {
        const float w = W[i * out + o];
        for (int r = 0; r < ROWS; r++) acc[r] = __builtin_fmaf(xs[r][i], w, acc[r]);
    }
    for (int r = 0; r < ROWS && r0 + r < n; r++)
        Y[(size_t)(r0 + r) * out + o] = relu ? fmaxf(acc[r], 0.f) : acc[r];
}

// dX[r][i] = (sum_o dY[r][o] W[i][o]) * (Xact[r][i] > 0): the gradient into
// the previous layer's (post-ReLU) activations.
__global__ void __launch_bounds__(128) k_dense_bwd_data(const float *dY, const float *W, const float *Xact,
                                                        float *dX, int n, int in, int out)
{
    __shared__ float gs[ROWS][H];
    const int r0 = blockIdx.x * ROWS;
    for (int k = threadIdx.x; k < ROWS * out; k += blockDim.x) {
        const int r = k / out, o = k - r * out;
        gs[r][o] = r0 + r < n ? dY[(size_t)(r0 + r) * out + o] : 0.f;
    }
    __syncthreads();
    const int i = threadIdx.x;
    if (i >= in) return;
    float acc[ROWS];
    for (int r = 0; r < ROWS; r++) acc[r] = 0.f;
    for (int o = 0; o < out; o++) {
        const float w = W[i * out + o];
        for (int r = 0; r < ROWS; r++) acc[r] = __builtin_fmaf(gs[r][o], w, acc[r]);
    }
    for (int r = 0; r < ROWS && r0 + r < n; r++) {
        const size_t k = (size_t)(r0 + r) * in + i;
        dX[k] = Xact[k] > 0.f ? acc[r] : 0.f;
    }
}

// dW[i][o] = sum_r X[r][i] dY[r][o]; db[o] = sum_r dY[r][o] (one thread per
// weight, rows summed in order).
__global__ void __launch_bounds__(256) k_dense_bwd_weight(const float *X, const float *dY, float *dW,
                                                          float *db, int n, int in, int out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < in * out) {
        const int i = k / out, o = k - i * out;
        float acc = 0.f;
        for (int r = 0; r < n; r++) acc = __builtin_fmaf(X[(size_t)r * in + i], dY[(size_t)r * out + o], acc);
        dW[k] = acc;
    }
    if (k < out) {
        float acc = 0.f;
        for (int r = 0; r < n; r++) acc += dY[(size_t)r * out + k];
        db[k] = acc;
    }
}

// dy = (y - t) / n and loss = 0.5 * mean((y - t)^2), one workgroup.
__global__ void __launch_bounds__(256) k_out_grad(const float *y, const float *t, float *dy, float *loss, int n)
{
    __shared__ float s[256];
    float acc = 0.f;
    for (int r = threadIdx.x; r < n; r += 256) {
        const float d = y[r] - t[r];
        dy[r] = d / n;
        acc = __builtin_fmaf(d, d, acc);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = 0.5f * s[0] / n;
}

// Gathers rows idx[0..n) of the raw data and standardises them.
__global__ void __launch_bounds__(256) k_gather_std(const float *X, const float *T, const int *idx, int n,
                                                   const float *mean, const float *scale, float tmean,
                                                   float tscale, float *Xo, float *To)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n * IN) {
        const int r = k / IN, i = k - r * IN;
        const int src = idx ? idx[r] : r;
        Xo[k] = (X[(size_t)src * IN + i] - mean[i]) / scale[i];
    }
    if (T && k < n) To[k] = (T[idx ? idx[k] : k] - tmean) / tscale;
}

__global__ void __launch_bounds__(256) k_adam(float *p, const float *g, float *m, float *v, int n, float lr,
                                              float b1, float b2, float c1, float c2, float eps)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

__global__ void k_destd(float *y, int n, float tmean, float tscale)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = y[k] * tscale + tmean;
}

uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

} // namespace

struct uhs_mlp {
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr; // parameters, gradient, Adam moments
    float *mean = nullptr, *scale = nullptr;                      // device copies of the standardisation
    float hmean[IN], hscale[IN];
    float tmean = 0.f, tscale = 1.f;
    int adam_t = 0;
    // workspace for up to `cap` rows
    int cap = 0;
    float *xs = nullptr, *ts = nullptr, *a[4] = {}, *d[4] = {}, *dy = nullptr, *loss = nullptr;
    int *idx = nullptr;
    hipStream_t st = nullptr;
};

namespace {

int herr(hipError_t e, const char *what)
{
    char b[160];
    snprintf(b, sizeof b, "%s: %s", what, hipGetErrorString(e));
    return uhs_set_error(UHS_E_HIP, b);
}

#define HCK(x, what)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return herr(e_, what);                                               \
    } while (0)

int reserve(uhs_mlp *M, int n)
{
    if (n <= M->cap) return UHS_OK;
    for (float *&q : {std::ref(M->xs), std::ref(M->ts), std::ref(M->dy), std::ref(M->a[0]), std::ref(M->a[1]),
                      std::ref(M->a[2]), std::ref(M->a[3]), std::ref(M->d[0]), std::ref(M->d[1]),
                      std::ref(M->d[2]), std::ref(M->d[3])})
        if (q) { (void)hipFree(q); q = nullptr; }
    if (M->idx) { (void)hipFree(M->idx); M->idx = nullptr; }
    HCK(hipMalloc(&M->xs, sizeof(float) * n * IN), "mlp workspace");
    HCK(hipMalloc(&M->ts, sizeof(float) * n), "mlp workspace");
    HCK(hipMalloc(&M->dy, sizeof(float) * n), "mlp workspace");
    for (int l = 0; l < 4; l++) {
        HCK(hipMalloc(&M->a[l], sizeof(float) * n * LOUT[l]), "mlp workspace");
        HCK(hipMalloc(&M->d[l], sizeof(float) * n * LOUT[l]), "mlp workspace");
    }
    HCK(hipMalloc(&M->idx, sizeof(int) * n), "mlp workspace");
    M->cap = n;
    return UHS_OK;
}

// Forward over the standardised rows in M->xs; a[3] holds the output.
void forward(uhs_mlp *M, int n)
{
    const float *x = M->xs;
    for (int l = 0; l < 4; l++) {
        hipLaunchKernelGGL(k_dense_fwd, dim3((n + ROWS - 1) / ROWS), dim3(128), 0, M->st, x, M->p + OW[l],
                           M->p + OB[l], M->a[l], n, LIN[l], LOUT[l], l < 3 ? 1 : 0);
        x = M->a[l];
    }
}

// Backward from the standardised targets in M->ts into M->g; loss -> M->loss.
void backward(uhs_mlp *M, int n)
{
    hipLaunchKernelGGL(k_out_grad, dim3(1), dim3(256), 0, M->st, M->a[3], M->ts, M->d[3], M->loss, n);
    for (int l = 3; l >= 0; l--) {
        const float *xin = l ? M->a[l - 1] : M->xs;
        const int nw = LIN[l] * LOUT[l];
        hipLaunchKernelGGL(k_dense_bwd_weight, dim3((std::max(nw, LOUT[l]) + 255) / 256), dim3(256), 0, M->st,
                           xin, M->d[l], M->g + OW[l], M->g + OB[l], n, LIN[l], LOUT[l]);
        if (l)
            hipLaunchKernelGGL(k_dense_bwd_data, dim3((n + ROWS - 1) / ROWS), dim3(128), 0, M->st, M->d[l],
                               M->p + OW[l], M->a[l - 1], M->d[l - 1], n, LIN[l], LOUT[l]);
    }
}

void adam(uhs_mlp *M, float lr)
{
    M->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)M->adam_t), c2 = 1.f - powf(b2, (float)M->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, M->st, M->p, M->g, M->m, M->v, NPARAM,
                       lr, b1, b2, c1, c2, 1e-8f);
}

int upload_rows(uhs_mlp *M, const float *x, const float *t, int n, float **dx, float **dt)
{
    HCK(hipMalloc(dx, sizeof(float) * n * IN), "mlp data");
    HCK(hipMemcpy(*dx, x, sizeof(float) * n * IN, hipMemcpyHostToDevice), "mlp data");
    if (t) {
        HCK(hipMalloc(dt, sizeof(float) * n), "mlp data");
        HCK(hipMemcpy(*dt, t, sizeof(float) * n, hipMemcpyHostToDevice), "mlp data");
    }
    return UHS_OK;
}

} // namespace

extern "C" {

int uhs_mlp_num_params(void) { return NPARAM; }

int uhs_mlp_features(int dims, int radius, int star, int dtype, const int64_t *ext, const uhs_config *c,
                     int depth, float *f)
{
    if (!ext || !c || !f || (dims != 2 && dims != 3) || radius < 1 || radius > 3 || dtype < 0 || dtype > 2)
        return uhs_set_error(UHS_E_INVAL, "bad feature request");
    memset(f, 0, sizeof(float) * IN);
    f[0] = dims == 3 ? 1.f : 0.f;
    f[1 + radius - 1] = 1.f;
    f[4] = star ? 1.f : 0.f;
    for (int k = 0; k < dims; k++) f[5 + k] = log2f((float)ext[k]);
    f[8] = log2f((float)c->tile_x);
    f[9] = log2f((float)c->tile_y);
    f[10] = log2f((float)c->tile_z);
    const int v = std::max(1, c->vec);
    int threads = c->tile_x / v * c->tile_y;
    if (c->flags & UHS_CFG_REGTILE) threads = 64 * (c->tile_y / v);
    else if (c->flags & UHS_CFG_ROWS2) threads = c->tile_x * c->tile_y / 2;
    else if (c->flags & UHS_CFG_TEMPORAL2)
        threads = (c->tile_x + 2 * radius) / v * (dims == 3 ? c->tile_y + 2 * radius : 1);
    else if (c->flags & UHS_CFG_TEMPORALK)
        threads = (c->tile_x + 2 * (c->tile_y - 1) * radius) / v;
    f[11] = log2f((float)std::max(1, threads));
    f[12] = (float)c->vec;
    if (c->flags & UHS_CFG_TEMPORAL2) depth = std::max(depth, 2);
    if (c->flags & UHS_CFG_TEMPORALK) depth = std::max(depth, c->tile_y);
    f[13 + (depth >= 3 ? 2 : depth >= 2 ? 1 : 0)] = 1.f; // steps per launch: 1, 2, 3 or more
    // Terms of the stencil: a full star has 2 dims r + 1, a full box (2r+1)^dims.
    const int side = 2 * radius + 1;
    const int P = star ? 2 * dims * radius + 1 : (dims == 3 ? side * side * side : side * side);
    f[16] = log2f((float)P);
    f[17] = (float)dtype;
    f[18] = (c->flags & UHS_CFG_NT_STORE) ? 1.f : 0.f;
    const bool rows2 = c->flags & UHS_CFG_ROWS2, planes2 = c->flags & UHS_CFG_PLANES2;
    f[19] = (c->flags & UHS_CFG_ROWRING) ? 7.f : rows2 && planes2 ? 6.f : planes2 ? 5.f : rows2 ? 4.f
          : (c->flags & UHS_CFG_REGTILE) ? 3.f
          : (c->flags & UHS_CFG_PREFETCH4) ? 2.f : (c->flags & UHS_CFG_PREFETCH2) ? 1.f : 0.f;
    return UHS_OK;
}

int uhs_mlp_create(uint64_t seed, uhs_mlp_t *out)
{
    if (!out) return uhs_set_error(UHS_E_INVAL, "null output");
    auto *M = new uhs_mlp();
    auto cleanup = [&](int rc) {
        uhs_mlp_destroy(M);
        return rc;
    };
    if (hipStreamCreate(&M->st) != hipSuccess) return cleanup(uhs_set_error(UHS_E_HIP, "mlp stream"));
    for (float **q : {&M->p, &M->g, &M->m, &M->v})
        if (hipMalloc(q, sizeof(float) * NPARAM) != hipSuccess || hipMemset(*q, 0, sizeof(float) * NPARAM) != hipSuccess)
            return cleanup(uhs_set_error(UHS_E_NOMEM, "mlp parameters"));
    if (hipMalloc(&M->mean, sizeof(float) * IN) != hipSuccess || hipMalloc(&M->scale, sizeof(float) * IN) != hipSuccess ||
        hipMalloc(&M->loss, sizeof(float)) != hipSuccess)
        return cleanup(uhs_set_error(UHS_E_NOMEM, "mlp parameters"));
    // He-uniform weights (ReLU layers), zero biases, from a seeded splitmix64.
    std::vector<float> h(NPARAM, 0.f);
    uint64_t k = 0;
    for (int l = 0; l < 4; l++) {
        const float lim = sqrtf(6.f / LIN[l]);
        for (int i = 0; i < LIN[l] * LOUT[l]; i++) {
            const double u = (double)(mix(seed + (++k) * 0x9E3779B97F4A7C15ULL) >> 11) * 0x1.0p-53;
            h[OW[l] + i] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    if (hipMemcpy(M->p, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(uhs_set_error(UHS_E_HIP, "mlp upload"));
    for (int i = 0; i < IN; i++) { M->hmean[i] = 0.f; M->hscale[i] = 1.f; }
    if (uhs_mlp_set_norm(M, M->hmean, M->hscale, 0.f, 1.f) != UHS_OK) return cleanup(UHS_E_HIP);
    *out = M;
    return UHS_OK;
}

int uhs_mlp_destroy(uhs_mlp_t M)
{
    if (!M) return UHS_OK;
    for (float *q : {M->p, M->g, M->m, M->v, M->mean, M->scale, M->xs, M->ts, M->dy, M->loss, M->a[0], M->a[1],
                     M->a[2], M->a[3], M->d[0], M->d[1], M->d[2], M->d[3]})
        if (q) (void)hipFree(q);
    if (M->idx) (void)hipFree(M->idx);
    if (M->st) (void)hipStreamDestroy(M->st);
    delete M;
    return UHS_OK;
}

int uhs_mlp_get_params(uhs_mlp_t M, float *out, int n)
{
    if (!M || !out || n < NPARAM) return uhs_set_error(UHS_E_INVAL, "bad parameter buffer");
    HCK(hipMemcpy(out, M->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost), "mlp params");
    return UHS_OK;
}

int uhs_mlp_set_params(uhs_mlp_t M, const float *in, int n)
{
    if (!M || !in || n < NPARAM) return uhs_set_error(UHS_E_INVAL, "bad parameter buffer");
    HCK(hipMemcpy(M->p, in, sizeof(float) * NPARAM, hipMemcpyHostToDevice), "mlp params");
    return UHS_OK;
}

int uhs_mlp_set_norm(uhs_mlp_t M, const float *mean, const float *scale, float tmean, float tscale)
{
    if (!M || !mean || !scale || tscale == 0.f) return uhs_set_error(UHS_E_INVAL, "bad standardisation");
    for (int i = 0; i < IN; i++) {
        if (scale[i] == 0.f) return uhs_set_error(UHS_E_INVAL, "zero scale");
        M->hmean[i] = mean[i];
        M->hscale[i] = scale[i];
    }
    M->tmean = tmean;
    M->tscale = tscale;
    HCK(hipMemcpy(M->mean, M->hmean, sizeof(float) * IN, hipMemcpyHostToDevice), "mlp norm");
    HCK(hipMemcpy(M->scale, M->hscale, sizeof(float) * IN, hipMemcpyHostToDevice), "mlp norm");
    return UHS_OK;
}

int uhs_mlp_predict(uhs_mlp_t M, const float *x, int n, float *y)
{
    if (!M || !x || !y || n < 1) return uhs_set_error(UHS_E_INVAL, "bad prediction request");
    if (int rc = reserve(M, n)) return rc;
    float *dx = nullptr, *dt = nullptr;
    if (int rc = upload_rows(M, x, nullptr, n, &dx, &dt)) return rc;
    hipLaunchKernelGGL(k_gather_std, dim3((n * IN + 255) / 256), dim3(256), 0, M->st, dx, nullptr, nullptr, n,
                       M->mean, M->scale, 0.f, 1.f, M->xs, nullptr);
    forward(M, n);
    hipLaunchKernelGGL(k_destd, dim3((n + 255) / 256), dim3(256), 0, M->st, M->a[3], n, M->tmean, M->tscale);
    hipError_t e = hipMemcpyAsync(y, M->a[3], sizeof(float) * n, hipMemcpyDeviceToHost, M->st);
    if (e == hipSuccess) e = hipStreamSynchronize(M->st);
    (void)hipFree(dx);
    if (e != hipSuccess) return herr(e, "mlp predict");
    return UHS_OK;
}

int uhs_mlp_grad(uhs_mlp_t M, const float *x, const float *t, int n, float *grad, float *loss)
{
    if (!M || !x || !t || n < 1) return uhs_set_error(UHS_E_INVAL, "bad gradient request");
    if (int rc = reserve(M, n)) return rc;
    float *dx = nullptr, *dt = nullptr;
    if (int rc = upload_rows(M, x, t, n, &dx, &dt)) return rc;
    hipLaunchKernelGGL(k_gather_std, dim3((n * IN + 255) / 256), dim3(256), 0, M->st, dx, dt, nullptr, n, M->mean,
                       M->scale, M->tmean, M->tscale, M->xs, M->ts);
    forward(M, n);
    backward(M, n);
    hipError_t e = hipStreamSynchronize(M->st);
    if (e == hipSuccess && grad) e = hipMemcpy(grad, M->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dt);
    if (e != hipSuccess) return herr(e, "mlp grad");
    return UHS_OK;
}

int uhs_mlp_adam_step(uhs_mlp_t M, const float *grad, float lr, int t)
{
    if (!M || !grad || t < 1) return uhs_set_error(UHS_E_INVAL, "bad Adam step");
    HCK(hipMemcpy(M->g, grad, sizeof(float) * NPARAM, hipMemcpyHostToDevice), "mlp grad upload");
    M->adam_t = t - 1;
    adam(M, lr);
    HCK(hipStreamSynchronize(M->st), "mlp adam");
    return UHS_OK;
}

int uhs_mlp_train(uhs_mlp_t M, const float *x, const float *t, int n, int epochs, int batch, float lr,
                  int fit_norm, uint64_t seed, float *final_loss)
{
    if (!M || !x || !t || n < 2 || epochs < 1 || batch < 1) return uhs_set_error(UHS_E_INVAL, "bad training request");
    if (fit_norm) {
        float mean[IN], scale[IN];
        for (int i = 0; i < IN; i++) {
            double s = 0, s2 = 0;
            for (int r = 0; r < n; r++) {
                s += x[(size_t)r * IN + i];
                s2 += (double)x[(size_t)r * IN + i] * x[(size_t)r * IN + i];
            }
            const double mu = s / n, var = std::max(0.0, s2 / n - mu * mu);
            mean[i] = (float)mu;
            scale[i] = var > 1e-12 ? (float)sqrt(var) : 1.f;
        }
        double s = 0, s2 = 0;
        for (int r = 0; r < n; r++) { s += t[r]; s2 += (double)t[r] * t[r]; }
        const double mu = s / n, sd = sqrt(std::max(1e-12, s2 / n - mu * mu));
        if (int rc = uhs_mlp_set_norm(M, mean, scale, (float)mu, (float)sd)) return rc;
    }
    batch = std::min(batch, n);
    if (int rc = reserve(M, batch)) return rc;
    float *dx = nullptr, *dt = nullptr;
    if (int rc = upload_rows(M, x, t, n, &dx, &dt)) return rc;
    int *didx = nullptr;
    if (hipMalloc(&didx, sizeof(int) * n) != hipSuccess) {
        (void)hipFree(dx); (void)hipFree(dt);
        return uhs_set_error(UHS_E_NOMEM, "mlp index");
    }
    float *dloss = nullptr;
    const int nb = n / batch; // whole minibatches per epoch
    if (hipMalloc(&dloss, sizeof(float) * nb) != hipSuccess) {
        (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(didx);
        return uhs_set_error(UHS_E_NOMEM, "mlp loss");
    }
    std::vector<int> perm(n);
    std::vector<float> hl(nb);
    float last = 0.f;
    hipError_t e = hipSuccess;
    for (int ep = 0; ep < epochs && e == hipSuccess; ep++) {
        std::iota(perm.begin(), perm.end(), 0);
        for (int i = n - 1; i > 0; i--) { // seeded Fisher-Yates
            const int j = (int)(mix(seed + (uint64_t)ep * 0x100000001ULL + i) % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        e = hipMemcpyAsync(didx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, M->st);
        for (int b = 0; b < nb && e == hipSuccess; b++) {
            hipLaunchKernelGGL(k_gather_std, dim3((batch * IN + 255) / 256), dim3(256), 0, M->st, dx, dt,
                               didx + (size_t)b * batch, batch, M->mean, M->scale, M->tmean, M->tscale, M->xs, M->ts);
            forward(M, batch);
            backward(M, batch);
            (void)hipMemcpyAsync(dloss + b, M->loss, sizeof(float), hipMemcpyDeviceToDevice, M->st);
            adam(M, lr);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hl.data(), dloss, sizeof(float) * nb, hipMemcpyDeviceToHost, M->st);
        if (e == hipSuccess) e = hipStreamSynchronize(M->st);
        double s = 0;
        for (float v : hl) s += v;
        last = (float)(s / nb);
    }
    (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(didx); (void)hipFree(dloss);
    if (e != hipSuccess) return herr(e, "mlp train");
    if (final_loss) *final_loss = last;
    return UHS_OK;
}

int uhs_mlp_save(uhs_mlp_t M, const char *path)
{
    if (!M || !path) return uhs_set_error(UHS_E_INVAL, "null argument");
    std::vector<float> h(NPARAM);
    if (int rc = uhs_mlp_get_params(M, h.data(), NPARAM)) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return uhs_set_error(UHS_E_IO, "cannot write the model file");
    const char magic[8] = {'U', 'H', 'S', 'M', 'L', 'P', '1', 0};
    int np = NPARAM;
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(&np, sizeof np, 1, f) == 1 &&
              fwrite(M->hmean, sizeof(float), IN, f) == (size_t)IN && fwrite(M->hscale, sizeof(float), IN, f) == (size_t)IN &&
              fwrite(&M->tmean, sizeof(float), 1, f) == 1 && fwrite(&M->tscale, sizeof(float), 1, f) == 1 &&
              fwrite(h.data(), sizeof(float), NPARAM, f) == (size_t)NPARAM;
    ok = (fclose(f) == 0) && ok;
    return ok ? UHS_OK : uhs_set_error(UHS_E_IO, "model write failed");
}

int uhs_mlp_load(uhs_mlp_t M, const char *path)
{
    if (!M || !path) return uhs_set_error(UHS_E_INVAL, "null argument");
    FILE *f = fopen(path, "rb");
    if (!f) return uhs_set_error(UHS_E_IO, "cannot open the model file");
    char magic[8];
    int np = 0;
    float mean[IN], scale[IN], tm, ts;
    std::vector<float> h(NPARAM);
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "UHSMLP1", 8) == 0 && fread(&np, sizeof np, 1, f) == 1 &&
              np == NPARAM && fread(mean, sizeof(float), IN, f) == (size_t)IN &&
              fread(scale, sizeof(float), IN, f) == (size_t)IN && fread(&tm, sizeof(float), 1, f) == 1 &&
              fread(&ts, sizeof(float), 1, f) == 1 && fread(h.data(), sizeof(float), NPARAM, f) == (size_t)NPARAM;
    fclose(f);
    if (!ok) return uhs_set_error(UHS_E_IO, "not a model file of this network");
    if (int rc = uhs_mlp_set_norm(M, mean, scale, tm, ts)) return rc;
    return uhs_mlp_set_params(M, h.data(), NPARAM);
}

} // extern "C"

// Model-pruned search (declared in uhs_tune.h).
std::vector<uhs_config> uhs_config_space(const Plan &pl);
int uhs_search(Plan &pl, void *A, void *B, int dims, const int64_t *ext, int halo, int dtype,
               const std::vector<uhs_config> &cands, int reps, const char *db_path, hipStream_t st,
               uhs_tune_result *res);

extern "C" int uhs_tune_sweep_models(void *A, void *B, int dims, const int64_t *ext, int halo,
                                     const uhs_stencil *st, int dtype, const uhs_mlp_t *models, int nmodels,
                                     int topk, int reps, const char *db_path, uhs_tune_result *out,
                                     uhs_stream_t stream)
{
    if (!models || nmodels < 1 || topk < 1) return uhs_set_error(UHS_E_INVAL, "model search needs models and topk >= 1");
    for (int m = 0; m < nmodels; m++)
        if (!models[m]) return uhs_set_error(UHS_E_INVAL, "null model in the ensemble");
    const auto t0 = std::chrono::steady_clock::now();
    Plan pl;
    int rc = uhs_make_plan(dims, ext, halo, st, dtype, nullptr, &pl);
    if (rc) return rc;
    if (pl.family == FAM_GATHER) return uhs_set_error(UHS_E_UNSUPPORTED, "the general kernel has no tiling to tune");
    std::vector<uhs_config> sp = uhs_config_space(pl);
    std::vector<float> feat(sp.size() * IN), pred(sp.size());
    const BuiltinStencil &b = kBuiltins[pl.prog];
    for (size_t i = 0; i < sp.size(); i++)
        if ((rc = uhs_mlp_features(dims, b.radius, b.star, dtype, ext, &sp[i], 1, &feat[i * IN]))) return rc;
    // The ensemble's prediction is the mean of its members' predicted log
    // times: members trained from different seeds disagree most where the
    // data leave the ranking open, and the mean damps those disagreements.
    std::vector<float> one(sp.size());
    for (int m = 0; m < nmodels; m++) {
        if ((rc = uhs_mlp_predict(models[m], feat.data(), (int)sp.size(), one.data()))) return rc;
        for (size_t i = 0; i < sp.size(); i++) pred[i] = (m ? pred[i] : 0.f) + one[i] / (float)nmodels;
    }
    std::vector<size_t> ord(sp.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t c) { return pred[a] < pred[c]; });
    std::vector<uhs_config> top;
    for (size_t i = 0; i < ord.size() && (int)top.size() < topk; i++) top.push_back(sp[ord[i]]);
    uhs_tune_result r{};
    rc = uhs_search(pl, A, B, dims, ext, halo, dtype, top, reps, db_path, (hipStream_t)stream, &r);
    if (rc) return rc;
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (out) *out = r;
    return UHS_OK;
}

extern "C" int uhs_tune_sweep_model(void *A, void *B, int dims, const int64_t *ext, int halo,
                                    const uhs_stencil *st, int dtype, uhs_mlp_t model, int topk, int reps,
                                    const char *db_path, uhs_tune_result *out, uhs_stream_t stream)
{
    return uhs_tune_sweep_models(A, B, dims, ext, halo, st, dtype, &model, 1, topk, reps, db_path, out, stream);
}
