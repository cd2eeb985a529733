// a3d_mlp.h -- the absorbing-strip network: layout of its parameter vector and
// the per-sample forward pass shared by the training kernels (a3d_mlp.hip) and
// the strip-correction kernel (a3d_kernels.hip). Not part of the public ABI.
//
// y = W3 tanh(W2 tanh(W1 x + b1) + b2) + b3, x in R^16, hidden 64 + 64, y in R^2.
// All parameters FP32, one flat vector:
//   W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]   (row-major, out x in)
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

// Forward pass of one sample with the parameters p (LDS or global). h1 and h2
// receive the hidden activations (the training kernels store them).
__device__ __forceinline__ void mlp_forward(const float* p, const float (&x)[LB_NIN], float (&h1)[LB_NH],
                                            float (&h2)[LB_NH], float (&y)[LB_NOUT]) {
#pragma unroll
    for (int j = 0; j < LB_NH; ++j) {
        float acc = p[LB_B1 + j];
#pragma unroll
        for (int i = 0; i < LB_NIN; ++i) acc = fmaf(p[LB_W1 + j * LB_NIN + i], x[i], acc);
        h1[j] = tanhf(acc);
    }
#pragma unroll
    for (int j = 0; j < LB_NH; ++j) {
        float acc = p[LB_B2 + j];
#pragma unroll
        for (int i = 0; i < LB_NH; ++i) acc = fmaf(p[LB_W2 + j * LB_NH + i], h1[i], acc);
        h2[j] = tanhf(acc);
    }
#pragma unroll
    for (int o = 0; o < LB_NOUT; ++o) {
        float acc = p[LB_B3 + o];
#pragma unroll
        for (int i = 0; i < LB_NH; ++i) acc = fmaf(p[LB_W3 + o * LB_NH + i], h2[i], acc);
        y[o] = acc;
    }
}

}  // namespace a3d
