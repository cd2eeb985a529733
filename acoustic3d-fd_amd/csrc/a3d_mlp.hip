// a3d_mlp.hip -- training of the absorbing-strip network on the GPU: forward,
// backward and weight-gradient kernels, an Adam step, and the a3d_mlp_* C API.
//
// One thread per sample for the forward and backward passes (parameters staged
// in LDS, read as broadcasts); the weight gradients sum over the samples of a
// minibatch in a fixed order (per-thread partial sums over a strided sample
// range, then a wave-shuffle and LDS reduction), so a gradient is the same on
// every run. Loss: mean over samples of the squared error of the damping head's
// prediction summed over the two outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/acoustic3d.h"
#include "a3d_internal.h"
#include "a3d_mlp.h"

using namespace a3d;

namespace {

constexpr int MB = 256;   // threads per block of the per-sample kernels

#define MLP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return report_error(A3D_EHIP, #expr, hipGetErrorString(e_)); \
    } while (0)

__device__ __forceinline__ void stage_params(float* sp, const float* __restrict__ p) {
    for (int i = threadIdx.x; i < LB_NPAR; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
}

// Forward pass of the minibatch idx[0..nb): activations, the gradient of the loss
// with respect to the raw outputs y (through the damping head) and per-block sums
// of the squared error.
__global__ __launch_bounds__(MB) void k_mlp_fwd(const float* __restrict__ p, const float* __restrict__ X,
                                                const float* __restrict__ T, const int* __restrict__ idx, int nb,
                                                float inv_n, float* __restrict__ H1, float* __restrict__ H2,
                                                float* __restrict__ DY, float* __restrict__ loss_part, int head) {
    __shared__ float sp[LB_NPAR];
    __shared__ float red[MB / 64];
    stage_params(sp, p);
    const int s = blockIdx.x * MB + threadIdx.x;
    float err = 0.f;
    if (s < nb) {
        const int r = idx ? idx[s] : s;
        float x[LB_NIN], h1[LB_NH], h2[LB_NH], y[LB_NOUT];
#pragma unroll
        for (int i = 0; i < LB_NIN; ++i) x[i] = X[(size_t)r * LB_NIN + i];
        mlp_forward(sp, x, h1, h2, y);
#pragma unroll
        for (int i = 0; i < LB_NH; ++i) {
            H1[(size_t)s * LB_NH + i] = h1[i];
            H2[(size_t)s * LB_NH + i] = h2[i];
        }
        if (head == 2) {
            float pred[LB_NOUT], dm;
            coef_head(y, pred, dm);
            const float e = pred[0] - T[(size_t)r * LB_NOUT];
            err = e * e;
            DY[(size_t)s * LB_NOUT] = 2.f * e * inv_n * dm;
            DY[(size_t)s * LB_NOUT + 1] = 0.f;
        } else {
            float g[LB_NOUT], pred[LB_NOUT];
            strip_head(y, x, g, pred);
#pragma unroll
            for (int o = 0; o < LB_NOUT; ++o) {
                const float e = pred[o] - T[(size_t)r * LB_NOUT + o];
                err += e * e;
                // d pred / d y = -x_k * g * (1 - g / GMAX)
                DY[(size_t)s * LB_NOUT + o] =
                    2.f * e * inv_n * (-x[LB_HEAD_IN[o]] * g[o] * (1.f - g[o] * (1.f / STRIP_GMAX)));
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) err += __shfl_xor(err, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = err;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < MB / 64; ++w) t += red[w];
        loss_part[blockIdx.x] = t;
    }
}

// Backward pass through the activations: D2 = (W3^T dy) * (1 - h2^2), D1 = (W2^T D2) * (1 - h1^2).
__global__ __launch_bounds__(MB) void k_mlp_bwd(const float* __restrict__ p, const float* __restrict__ H1,
                                                const float* __restrict__ H2, const float* __restrict__ DY, int nb,
                                                float* __restrict__ D1, float* __restrict__ D2) {
    __shared__ float sp[LB_NPAR];
    stage_params(sp, p);
    const int s = blockIdx.x * MB + threadIdx.x;
    if (s >= nb) return;
    float dy[LB_NOUT], d2[LB_NH], d1[LB_NH];
#pragma unroll
    for (int o = 0; o < LB_NOUT; ++o) dy[o] = DY[(size_t)s * LB_NOUT + o];
#pragma unroll
    for (int j = 0; j < LB_NH; ++j) {
        float g = 0.f;
#pragma unroll
        for (int o = 0; o < LB_NOUT; ++o) g += sp[LB_W3 + o * LB_NH + j] * dy[o];
        const float h = H2[(size_t)s * LB_NH + j];
        d2[j] = g * (1.f - h * h);
        d1[j] = 0.f;
    }
#pragma unroll
    for (int j = 0; j < LB_NH; ++j) {
#pragma unroll
        for (int i = 0; i < LB_NH; ++i) d1[i] += sp[LB_W2 + j * LB_NH + i] * d2[j];
    }
#pragma unroll
    for (int i = 0; i < LB_NH; ++i) {
        const float h = H1[(size_t)s * LB_NH + i];
        D1[(size_t)s * LB_NH + i] = d1[i] * (1.f - h * h);
        D2[(size_t)s * LB_NH + i] = d2[i];
    }
}

// Weight and bias gradients of one output row per block:
// rows [0,64) layer 1 (delta D1, input X), [64,128) layer 2 (D2, H1), [128,130) layer 3 (DY, H2).
template <int NC>
__device__ __forceinline__ void row_grad(const float* __restrict__ A, int astride, int acol,
                                         const float* __restrict__ In, const int* __restrict__ idx, int nb,
                                         float* __restrict__ gw, float* __restrict__ gb) {
    __shared__ float red[MB / 64][NC + 1];
    float acc[NC + 1];
#pragma unroll
    for (int c = 0; c <= NC; ++c) acc[c] = 0.f;
    for (int s = threadIdx.x; s < nb; s += MB) {
        const float a = A[(size_t)s * astride + acol];
        const int r = idx ? idx[s] : s;
        const float* in = In + (size_t)r * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a * in[c];
        acc[NC] += a;
    }
#pragma unroll
    for (int c = 0; c <= NC; ++c) {
        float v = acc[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = v;
    }
    __syncthreads();
    for (int c = threadIdx.x; c <= NC; c += MB) {
        float t = 0.f;
        for (int w = 0; w < MB / 64; ++w) t += red[w][c];
        if (c < NC) gw[c] = t;
        else *gb = t;
    }
}

__global__ __launch_bounds__(MB) void k_mlp_wgrad(const float* __restrict__ X, const int* __restrict__ idx,
                                                  const float* __restrict__ H1, const float* __restrict__ H2,
                                                  const float* __restrict__ D1, const float* __restrict__ D2,
                                                  const float* __restrict__ DY, int nb, float* __restrict__ G) {
    const int r = blockIdx.x;
    if (r < LB_NH)
        row_grad<LB_NIN>(D1, LB_NH, r, X, idx, nb, G + LB_W1 + r * LB_NIN, G + LB_B1 + r);
    else if (r < 2 * LB_NH)
        row_grad<LB_NH>(D2, LB_NH, r - LB_NH, H1, nullptr, nb, G + LB_W2 + (r - LB_NH) * LB_NH, G + LB_B2 + r - LB_NH);
    else
        row_grad<LB_NH>(DY, LB_NOUT, r - 2 * LB_NH, H2, nullptr, nb, G + LB_W3 + (r - 2 * LB_NH) * LB_NH,
                        G + LB_B3 + r - 2 * LB_NH);
}

// Adam (Kingma and Ba 2015) with bias corrections bc1 = 1 - b1^t, bc2 = 1 - b2^t.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1,
                           float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ __launch_bounds__(MB) void k_mlp_predict(const float* __restrict__ p, const float* __restrict__ X, int n,
                                                    float* __restrict__ Y, int head) {
    __shared__ float sp[LB_NPAR];
    stage_params(sp, p);
    const int s = blockIdx.x * MB + threadIdx.x;
    if (s >= n) return;
    float x[LB_NIN], h1[LB_NH], h2[LB_NH], y[LB_NOUT];
#pragma unroll
    for (int i = 0; i < LB_NIN; ++i) x[i] = X[(size_t)s * LB_NIN + i];
    mlp_forward(sp, x, h1, h2, y);
    float g[LB_NOUT], pred[LB_NOUT], dm;
    if (head == 2) coef_head(y, pred, dm);
    else strip_head(y, x, g, pred);
#pragma unroll
    for (int o = 0; o < LB_NOUT; ++o) Y[(size_t)s * LB_NOUT + o] = pred[o];
}

}  // namespace

struct a3d_mlp {
    int device = 0;
    hipStream_t st = nullptr;
    float* p = nullptr;      // parameters
    float* g = nullptr;      // gradient
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    long long adam_t = 0;
    int head = 0;            // training target: 0 the damping head's change, 2 the coefficient head's multiplier
    // per-minibatch work buffers
    int cap = 0;
    float *H1 = nullptr, *H2 = nullptr, *D1 = nullptr, *D2 = nullptr, *DY = nullptr, *LP = nullptr;
    int* idx = nullptr;
};

namespace {

void free_work(a3d_mlp* m) {
    for (float** q : {&m->H1, &m->H2, &m->D1, &m->D2, &m->DY, &m->LP}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    if (m->idx) (void)hipFree(m->idx);
    m->idx = nullptr;
    m->cap = 0;
}

int ensure_work(a3d_mlp* m, int nb) {
    if (nb <= m->cap) return A3D_OK;
    free_work(m);
    const size_t h = (size_t)nb * LB_NH * sizeof(float);
    const int nblk = (nb + MB - 1) / MB;
    if (hipMalloc(&m->H1, h) != hipSuccess || hipMalloc(&m->H2, h) != hipSuccess || hipMalloc(&m->D1, h) != hipSuccess ||
        hipMalloc(&m->D2, h) != hipSuccess || hipMalloc(&m->DY, (size_t)nb * LB_NOUT * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->LP, (size_t)nblk * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->idx, (size_t)nb * sizeof(int)) != hipSuccess) {
        free_work(m);
        return report_error(A3D_ENOMEM, "network work buffers", "allocation failed");
    }
    m->cap = nb;
    return A3D_OK;
}

// Loss and gradient of the samples idx[0..nb) (idx null: 0..nb) into m->g; the
// squared-error sum is left in m->LP (one value per block).
int loss_grad(a3d_mlp* m, const float* X, const float* T, const int* idx, int nb, float inv_n) {
    const int nblk = (nb + MB - 1) / MB;
    k_mlp_fwd<<<nblk, MB, 0, m->st>>>(m->p, X, T, idx, nb, inv_n, m->H1, m->H2, m->DY, m->LP, m->head);
    MLP_TRY(hipGetLastError());
    k_mlp_bwd<<<nblk, MB, 0, m->st>>>(m->p, m->H1, m->H2, m->DY, nb, m->D1, m->D2);
    MLP_TRY(hipGetLastError());
    k_mlp_wgrad<<<2 * LB_NH + LB_NOUT, MB, 0, m->st>>>(X, idx, m->H1, m->H2, m->D1, m->D2, m->DY, nb, m->g);
    MLP_TRY(hipGetLastError());
    return A3D_OK;
}

}  // namespace

extern "C" {

int a3d_mlp_nparams(void) { return LB_NPAR; }

int a3d_mlp_create(int device, a3d_mlp** out) {
    if (!out) return report_error(A3D_EINVAL, "a3d_mlp_create", "null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return report_error(A3D_EINVAL, "a3d_mlp_create", "no such device");
    a3d_mlp* m = new a3d_mlp();
    m->device = device;
    const size_t b = LB_NPAR * sizeof(float);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&m->p, b) != hipSuccess || hipMalloc(&m->g, b) != hipSuccess || hipMalloc(&m->m, b) != hipSuccess ||
        hipMalloc(&m->v, b) != hipSuccess) {
        a3d_mlp_destroy(m);
        return report_error(A3D_ENOMEM, "a3d_mlp_create", "allocation failed");
    }
    (void)hipMemset(m->p, 0, b);
    (void)hipMemset(m->m, 0, b);
    (void)hipMemset(m->v, 0, b);
    *out = m;
    return A3D_OK;
}

int a3d_mlp_destroy(a3d_mlp* m) {
    if (!m) return A3D_OK;
    (void)hipSetDevice(m->device);
    if (m->st) (void)hipStreamSynchronize(m->st);
    free_work(m);
    for (float* q : {m->p, m->g, m->m, m->v})
        if (q) (void)hipFree(q);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return A3D_OK;
}

int a3d_mlp_set_params(a3d_mlp* m, const float* params, int n) {
    if (!m || !params || n != LB_NPAR) return report_error(A3D_EINVAL, "a3d_mlp_set_params", "bad argument");
    MLP_TRY(hipSetDevice(m->device));
    MLP_TRY(hipMemcpy(m->p, params, LB_NPAR * sizeof(float), hipMemcpyHostToDevice));
    MLP_TRY(hipMemset(m->m, 0, LB_NPAR * sizeof(float)));
    MLP_TRY(hipMemset(m->v, 0, LB_NPAR * sizeof(float)));
    m->adam_t = 0;
    return A3D_OK;
}

int a3d_mlp_set_head(a3d_mlp* m, int head) {
    if (!m || (head != 0 && head != 2)) return report_error(A3D_EINVAL, "a3d_mlp_set_head", "head must be 0 or 2");
    m->head = head;
    return A3D_OK;
}

int a3d_mlp_get_params(a3d_mlp* m, float* params, int n) {
    if (!m || !params || n != LB_NPAR) return report_error(A3D_EINVAL, "a3d_mlp_get_params", "bad argument");
    MLP_TRY(hipSetDevice(m->device));
    MLP_TRY(hipStreamSynchronize(m->st));
    MLP_TRY(hipMemcpy(params, m->p, LB_NPAR * sizeof(float), hipMemcpyDeviceToHost));
    return A3D_OK;
}

int a3d_mlp_predict(a3d_mlp* m, const float* X, int n, float* Y) {
    if (!m || !X || !Y || n <= 0) return report_error(A3D_EINVAL, "a3d_mlp_predict", "bad argument");
    MLP_TRY(hipSetDevice(m->device));
    float *dx = nullptr, *dy = nullptr;
    if (hipMalloc(&dx, (size_t)n * LB_NIN * sizeof(float)) != hipSuccess ||
        hipMalloc(&dy, (size_t)n * LB_NOUT * sizeof(float)) != hipSuccess) {
        if (dx) (void)hipFree(dx);
        return report_error(A3D_ENOMEM, "a3d_mlp_predict", "allocation failed");
    }
    int rc = A3D_OK;
    if (hipMemcpy(dx, X, (size_t)n * LB_NIN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = A3D_EHIP;
    if (!rc) {
        k_mlp_predict<<<(n + MB - 1) / MB, MB, 0, m->st>>>(m->p, dx, n, dy, m->head);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(m->st) != hipSuccess ||
            hipMemcpy(Y, dy, (size_t)n * LB_NOUT * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = A3D_EHIP;
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc ? report_error(rc, "a3d_mlp_predict", "HIP call failed") : A3D_OK;
}

int a3d_mlp_loss_grad(a3d_mlp* m, const float* X, const float* T, int n, double* loss, float* grad) {
    if (!m || !X || !T || n <= 0) return report_error(A3D_EINVAL, "a3d_mlp_loss_grad", "bad argument");
    MLP_TRY(hipSetDevice(m->device));
    int rc = ensure_work(m, n);
    if (rc) return rc;
    float *dx = nullptr, *dt = nullptr;
    if (hipMalloc(&dx, (size_t)n * LB_NIN * sizeof(float)) != hipSuccess ||
        hipMalloc(&dt, (size_t)n * LB_NOUT * sizeof(float)) != hipSuccess) {
        if (dx) (void)hipFree(dx);
        return report_error(A3D_ENOMEM, "a3d_mlp_loss_grad", "allocation failed");
    }
    const int nblk = (n + MB - 1) / MB;
    std::vector<float> lp(nblk);
    if (hipMemcpy(dx, X, (size_t)n * LB_NIN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, T, (size_t)n * LB_NOUT * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        rc = report_error(A3D_EHIP, "a3d_mlp_loss_grad", "upload failed");
    if (!rc) rc = loss_grad(m, dx, dt, nullptr, n, 1.f / n);
    if (!rc && (hipStreamSynchronize(m->st) != hipSuccess ||
                hipMemcpy(lp.data(), m->LP, nblk * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
                (grad && hipMemcpy(grad, m->g, LB_NPAR * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)))
        rc = report_error(A3D_EHIP, "a3d_mlp_loss_grad", "download failed");
    (void)hipFree(dx);
    (void)hipFree(dt);
    if (!rc && loss) {
        double t = 0.0;
        for (float v : lp) t += v;
        *loss = t / n;
    }
    return rc;
}

int a3d_mlp_train(a3d_mlp* m, const float* X, const float* T, int n, int epochs, int batch, float lr, float lr_end,
                  unsigned seed, double* loss_per_epoch) {
    if (!m || !X || !T || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0))
        return report_error(A3D_EINVAL, "a3d_mlp_train", "bad argument");
    MLP_TRY(hipSetDevice(m->device));
    batch = std::min(batch, n);
    int rc = ensure_work(m, batch);
    if (rc) return rc;
    float *dx = nullptr, *dt = nullptr;
    int* perm_d = nullptr;
    if (hipMalloc(&dx, (size_t)n * LB_NIN * sizeof(float)) != hipSuccess ||
        hipMalloc(&dt, (size_t)n * LB_NOUT * sizeof(float)) != hipSuccess ||
        hipMalloc(&perm_d, (size_t)n * sizeof(int)) != hipSuccess) {
        for (void* q : {(void*)dx, (void*)dt, (void*)perm_d})
            if (q) (void)hipFree(q);
        return report_error(A3D_ENOMEM, "a3d_mlp_train", "allocation failed");
    }
    auto cleanup = [&](int r) {
        (void)hipStreamSynchronize(m->st);
        (void)hipFree(dx);
        (void)hipFree(dt);
        (void)hipFree(perm_d);
        return r;
    };
    if (hipMemcpy(dx, X, (size_t)n * LB_NIN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, T, (size_t)n * LB_NOUT * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(report_error(A3D_EHIP, "a3d_mlp_train", "upload failed"));
    const int nbatch = n / batch;   // full minibatches; the remainder of each shuffle is left out
    const int nblk = (batch + MB - 1) / MB;
    std::vector<float> lp((size_t)nblk);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937 rng(seed);
    const long long total_steps = (long long)epochs * nbatch;
    long long step = 0;
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpyAsync(perm_d, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, m->st) != hipSuccess)
            return cleanup(report_error(A3D_EHIP, "a3d_mlp_train", "permutation upload failed"));
        double esum = 0.0;
        for (int b = 0; b < nbatch; ++b, ++step) {
            if ((rc = loss_grad(m, dx, dt, perm_d + (size_t)b * batch, batch, 1.f / batch))) return cleanup(rc);
            // cosine decay from lr to lr_end over all steps
            const double f = 0.5 * (1.0 + std::cos(M_PI * (double)step / std::max(1LL, total_steps)));
            const float lr_t = (float)(lr_end + (lr - lr_end) * f);
            ++m->adam_t;
            const float b1 = 0.9f, b2 = 0.999f;
            const float bc1 = 1.f - (float)std::pow(b1, (double)m->adam_t);
            const float bc2 = 1.f - (float)std::pow(b2, (double)m->adam_t);
            k_mlp_adam<<<(LB_NPAR + 255) / 256, 256, 0, m->st>>>(m->p, m->g, m->m, m->v, LB_NPAR, lr_t, b1, b2, 1e-8f,
                                                                 bc1, bc2);
            if (hipGetLastError() != hipSuccess) return cleanup(report_error(A3D_EHIP, "a3d_mlp_train", "launch failed"));
            if (loss_per_epoch) {
                if (hipMemcpyAsync(lp.data(), m->LP, nblk * sizeof(float), hipMemcpyDeviceToHost, m->st) != hipSuccess ||
                    hipStreamSynchronize(m->st) != hipSuccess)
                    return cleanup(report_error(A3D_EHIP, "a3d_mlp_train", "loss download failed"));
                for (float v : lp) esum += v;
            }
        }
        if (loss_per_epoch) loss_per_epoch[e] = esum / ((double)nbatch * batch);
    }
    return cleanup(A3D_OK);
}

}  // extern "C"
