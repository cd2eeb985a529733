// a3d_mlp.h -- the absorbing-strip network: layout of its parameter vector, the
// per-sample forward pass and the damping head, shared by the training kernels
// (a3d_mlp.hip) and the strip-correction kernel (a3d_strip.hip). Not part of
// the public ABI.
//
// y = W3 tanh(W2 tanh(W1 x + b1) + b2) + b3, x in R^16, hidden 64 + 64, y in R^2.
// All parameters FP32, one flat vector:
//   W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]   (row-major, out x in)
// Damping head: g_o = STRIP_GMAX * sigmoid(y_o) in (0, STRIP_GMAX), the fraction of
// the cell's P (o = 0) and Q (o = 1) removed this step. The trained quantity is the
// change it makes in units of the local amplitude s: pred = (-x[0] g_0, -x[7] g_1),
// x[0] = P/s and x[7] = Q/s being the cell's own values among the features.
// The coefficient head (training mode 2) is trained on its multiplier instead (coef_head).
#pragma once

#include <hip/hip_runtime.h>

namespace a3d {

constexpr int LB_NIN = 16, LB_NH = 64, LB_NOUT = 2;
constexpr int LB_W1 = 0;
constexpr int LB_B1 = LB_W1 + LB_NH * LB_NIN;
constexpr int LB_W2 = LB_B1 + LB_NH;
constexpr int LB_B2 = LB_W2 + LB_NH * LB_NH;
constexpr int LB_W3 = LB_B2 + LB_NH;
constexpr int LB_B3 = LB_W3 + LB_NOUT * LB_NH;
constexpr int LB_NPAR = LB_B3 + LB_NOUT;   // 5378
constexpr float STRIP_GMAX = 0.25f;
constexpr int LB_HEAD_IN[LB_NOUT] = {0, 7};   // features holding the cell's own P/s and Q/s

// tanh(a) = 1 - 2 / (exp(2a) + 1): one exponential and one division, exact limits
// +-1 for large |a| (an absolute error of a few 1e-7 near 0).
__device__ __forceinline__ float act_tanh(float a) { return 1.0f - __fdividef(2.0f, __expf(2.0f * a) + 1.0f); }

// Forward pass of one sample with the parameters p (LDS or global). h1 and h2
// receive the hidden activations (the training kernels store them).
__device__ __forceinline__ void mlp_forward(const float* p, const float (&x)[LB_NIN], float (&h1)[LB_NH],
                                            float (&h2)[LB_NH], float (&y)[LB_NOUT]) {
#pragma unroll
    for (int j = 0; j < LB_NH; ++j) {
        float acc = p[LB_B1 + j];
#pragma unroll
        for (int i = 0; i < LB_NIN; ++i) acc = fmaf(p[LB_W1 + j * LB_NIN + i], x[i], acc);
        h1[j] = act_tanh(acc);
    }
#pragma unroll
    for (int j = 0; j < LB_NH; ++j) {
        float acc = p[LB_B2 + j];
#pragma unroll
        for (int i = 0; i < LB_NH; ++i) acc = fmaf(p[LB_W2 + j * LB_NH + i], h1[i], acc);
        h2[j] = act_tanh(acc);
    }
#pragma unroll
    for (int o = 0; o < LB_NOUT; ++o) {
        float acc = p[LB_B3 + o];
#pragma unroll
        for (int i = 0; i < LB_NH; ++i) acc = fmaf(p[LB_W3 + o * LB_NH + i], h2[i], acc);
        y[o] = acc;
    }
}

__device__ __forceinline__ void strip_head(const float (&y)[LB_NOUT], const float (&x)[LB_NIN], float (&g)[LB_NOUT],
                                           float (&pred)[LB_NOUT]) {
#pragma unroll
    for (int o = 0; o < LB_NOUT; ++o) {
        g[o] = STRIP_GMAX / (1.0f + expf(-y[o]));
        pred[o] = -x[LB_HEAD_IN[o]] * g[o];
    }
}

// Training target of the coefficient head (training mode 2): the multiplier itself,
// pred = (m, 0) with m = STRIP_MMAX sigmoid(y_0) for every sample (whatever the local
// amplitude); output 1 is unused and carries no loss. dm/dy_0 = m (1 - m / STRIP_MMAX).
__device__ __forceinline__ void coef_head(const float (&y)[LB_NOUT], float (&pred)[LB_NOUT], float& dm) {
    const float m = STRIP_MMAX / (1.0f + expf(-y[0]));
    pred[0] = m;
    pred[1] = 0.f;
    dm = m * (1.f - m * (1.f / STRIP_MMAX));
}

}  // namespace a3d
