// a3d_strip.hip -- the learned absorbing strip of the VTI operator (FP32).
//
// The strip is the outer STRIP_CELLS updated cells of every face of the
// padded grid (padding nb = STRIP_CELLS + 4, the 4-cell ring staying zero as
// in the reference's PML), run with the reference split-PML update at that
// width. After each stress sweep a network damps each strip cell's P and Q:
//   features  f = local state / s, s = max |local state| (amplitude-free)
//   factors   (g_P, g_Q) = STRIP_GMAX * sigmoid(net(f)) in (0, STRIP_GMAX)
//   update    P, px, py, pz, u, v *= 1 - g_P;  Q, qx, qy, qz, w *= 1 - g_Q
// (the velocities with their stresses, as a sponge damps every field: damping
// the stresses alone would leave a net inflow into the strip and a drifting
// static field). A factor in (0, 1) only removes energy, so the strip stays
// stable for any parameters; the network decides where and how strongly to damp.
// Local state of a cell with outward normal axis a (the axis of its nearest
// face): P at the 7-point star (centre, inward, outward, tangential pairs),
// Q at centre, inward, outward, and the outward normal velocity times the
// impedance rho*vp; model features vp*dt/d_a, epsilon, delta and the depth in
// the strip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "a3d_internal.h"
#include "a3d_mlp.h"

namespace a3d {
namespace {

constexpr int SB = 256;   // threads per block

__device__ __forceinline__ float strip_features(const StripArgs& a, int id, int meta, float (&x)[LB_NIN]) {
    const int axis = meta & 3, side = (meta >> 2) & 1, dist = meta >> 3;
    const int sa = axis == 0 ? 1 : (axis == 1 ? a.g.ld : a.g.pl);
    const int sn = side ? sa : -sa;   // one cell outward
    const int s1 = axis == 0 ? a.g.ld : 1;
    const int s2 = axis == 2 ? a.g.ld : a.g.pl;
    float f[11];
    f[0] = a.P[id];
    f[1] = a.P[id - sn];
    f[2] = a.P[id + sn];
    f[3] = a.P[id - s1];
    f[4] = a.P[id + s1];
    f[5] = a.P[id - s2];
    f[6] = a.P[id + s2];
    f[7] = a.Q[id];
    f[8] = a.Q[id - sn];
    f[9] = a.Q[id + sn];
    const float* vel = axis == 0 ? a.u : (axis == 1 ? a.v : a.w);
    const float un = 0.5f * (vel[id] + vel[id - sa]);   // velocity sample k sits at k + 1/2
    f[10] = (side ? un : -un) * (a.rho[id] * a.vp[id]);
    float s = 1e-30f;
#pragma unroll
    for (int k = 0; k < 11; ++k) s = fmaxf(s, fabsf(f[k]));
    const float inv = 1.0f / s;
#pragma unroll
    for (int k = 0; k < 11; ++k) x[k] = f[k] * inv;
    x[11] = a.vp[id] * a.cdt[axis];
    x[12] = a.eps[id];
    x[13] = a.del[id];
    x[14] = (float)dist * (1.0f / STRIP_CELLS);
    x[15] = 0.f;
    return s;
}

__global__ __launch_bounds__(SB) void k_strip_infer(const StripArgs a, const float* __restrict__ net,
                                                    float2* __restrict__ delta) {
    __shared__ float sp[LB_NPAR];
    for (int i = threadIdx.x; i < LB_NPAR; i += SB) sp[i] = net[i];
    __syncthreads();
    const int k = blockIdx.x * SB + threadIdx.x;
    if (k >= a.n) return;
    float x[LB_NIN], h1[LB_NH], h2[LB_NH], y[LB_NOUT], g[LB_NOUT], pred[LB_NOUT];
    (void)strip_features(a, a.ids[k], a.meta[k], x);
    mlp_forward(sp, x, h1, h2, y);
    strip_head(y, x, g, pred);
    delta[k] = make_float2(g[0], g[1]);
}

// The same network on the matrix cores: one wavefront evaluates 64 strip cells as
// a 64 x 16 by 16 x 64 and a 64 x 64 by 64 x 64 product with v_mfma_f32_32x32x8_f16
// (operands rounded to binary16, products exact, sums in FP32), the 2-output last
// layer per lane in FP32. Weights, features and hidden activations pass through LDS
// in binary16. MFMA 32x32x8 operand layout (lane l, r = 0..15): A[i][k] with
// i = l % 32, k = 4 (l / 32) + 0..3; B[k][j] with j = l % 32, same k; D[i][j] with
// j = l % 32, i = 8 (r / 4) + 4 (l / 32) + r % 4.
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef float float16v __attribute__((ext_vector_type(16)));
constexpr int XP = LB_NIN + 4;   // padded LDS rows (halves), 8-byte aligned
constexpr int HP = LB_NH + 8;

__global__ __launch_bounds__(SB) void k_strip_infer_mfma(const StripArgs a, const float* __restrict__ net,
                                                         float2* __restrict__ delta) {
    __shared__ _Float16 sW1[LB_NH][XP];
    __shared__ _Float16 sW2[LB_NH][HP];
    __shared__ float sB1[LB_NH], sB2[LB_NH], sW3[LB_NOUT][LB_NH], sB3[LB_NOUT];
    __shared__ _Float16 sX[SB / 64][64][XP];
    __shared__ _Float16 sH[SB / 64][64][HP];
    for (int i = threadIdx.x; i < LB_NH * LB_NIN; i += SB) sW1[i / LB_NIN][i % LB_NIN] = (_Float16)net[LB_W1 + i];
    for (int i = threadIdx.x; i < LB_NH * LB_NH; i += SB) sW2[i / LB_NH][i % LB_NH] = (_Float16)net[LB_W2 + i];
    for (int i = threadIdx.x; i < LB_NH; i += SB) {
        sB1[i] = net[LB_B1 + i];
        sB2[i] = net[LB_B2 + i];
        sW3[0][i] = net[LB_W3 + i];
        sW3[1][i] = net[LB_W3 + LB_NH + i];
    }
    if (threadIdx.x < LB_NOUT) sB3[threadIdx.x] = net[LB_B3 + threadIdx.x];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int lr = l & 31, lk = (l >> 5) * 4;
    // the block stages the weights once and walks the strip cells 256 at a time
    for (int base = blockIdx.x * SB; base < a.n; base += gridDim.x * SB) {
        const int k = base + threadIdx.x;
        float x[LB_NIN];
#pragma unroll
        for (int i = 0; i < LB_NIN; ++i) x[i] = 0.f;
        if (k < a.n) (void)strip_features(a, a.ids[k], a.meta[k], x);
        __syncthreads();   // the previous group's last reads of sX / sH are done
#pragma unroll
        for (int i = 0; i < LB_NIN; ++i) sX[wv][l][i] = (_Float16)x[i];
        __syncthreads();
        float16v acc[2][2];
        // layer 1: H1 = tanh(X W1^T + b1)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                float16v c = {};
#pragma unroll
                for (int ks = 0; ks < LB_NIN / 8; ++ks) {
                    const half4v av = *(const half4v*)&sX[wv][32 * rb + lr][8 * ks + lk];
                    const half4v bv = *(const half4v*)&sW1[32 * cb + lr][8 * ks + lk];
                    c = __builtin_amdgcn_mfma_f32_32x32x8f16(av, bv, c, 0, 0, 0);
                }
                acc[rb][cb] = c;
            }
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = 32 * rb + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3), j = 32 * cb + lr;
                    sH[wv][i][j] = (_Float16)act_tanh(acc[rb][cb][r] + sB1[j]);
                }
        __syncthreads();
        // layer 2: H2 = tanh(H1 W2^T + b2)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                float16v c = {};
#pragma unroll
                for (int ks = 0; ks < LB_NH / 8; ++ks) {
                    const half4v av = *(const half4v*)&sH[wv][32 * rb + lr][8 * ks + lk];
                    const half4v bv = *(const half4v*)&sW2[32 * cb + lr][8 * ks + lk];
                    c = __builtin_amdgcn_mfma_f32_32x32x8f16(av, bv, c, 0, 0, 0);
                }
                acc[rb][cb] = c;
            }
        __syncthreads();
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = 32 * rb + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3), j = 32 * cb + lr;
                    sH[wv][i][j] = (_Float16)act_tanh(acc[rb][cb][r] + sB2[j]);
                }
        __syncthreads();
        if (k < a.n) {
            // layer 3 (2 outputs) and the damping head, one cell per lane
            float y[LB_NOUT], g[LB_NOUT], pred[LB_NOUT];
#pragma unroll
            for (int o = 0; o < LB_NOUT; ++o) {
                float t = 0.f;
#pragma unroll
                for (int j = 0; j < LB_NH; ++j) t = fmaf(sW3[o][j], (float)sH[wv][l][j], t);
                y[o] = t + sB3[o];
            }
            strip_head(y, x, g, pred);
            delta[k] = make_float2(g[0], g[1]);
        }
    }
}

__global__ __launch_bounds__(SB) void k_strip_apply(const StripArgs a, const float2* __restrict__ delta,
                                                    float* __restrict__ P, float* __restrict__ Q,
                                                    float* __restrict__ px, float* __restrict__ py,
                                                    float* __restrict__ pz, float* __restrict__ qx,
                                                    float* __restrict__ qy, float* __restrict__ qz,
                                                    float* __restrict__ u, float* __restrict__ v,
                                                    float* __restrict__ w) {
    const int k = blockIdx.x * SB + threadIdx.x;
    if (k >= a.n) return;
    const int id = a.ids[k];
    const float2 d = delta[k];
    const float fp = 1.0f - d.x, fq = 1.0f - d.y;
    P[id] *= fp;
    px[id] *= fp;
    py[id] *= fp;
    pz[id] *= fp;
    Q[id] *= fq;
    qx[id] *= fq;
    qy[id] *= fq;
    qz[id] *= fq;
    u[id] *= fp;
    v[id] *= fp;
    w[id] *= fq;
}

// Teacher-forced samples: the strip state a (the student's stress sweep applied to
// the teacher's state) against the teacher's P, Q at the same cells (teacher padded
// index = student index shifted by off on every axis). Cell k of sweep `tag` is kept
// when a hash of (seed, tag, k) falls below keep * 2^32.
__global__ __launch_bounds__(SB) void k_strip_sample(const StripArgs a, const float* __restrict__ Pt,
                                                     const float* __restrict__ Qt, Geo gt, int off,
                                                     unsigned keep, unsigned seed, unsigned tag, int cap,
                                                     float* __restrict__ X, float* __restrict__ T,
                                                     float* __restrict__ S, int* __restrict__ count) {
    const int k = blockIdx.x * SB + threadIdx.x;
    if (k >= a.n) return;
    unsigned h = (unsigned)k * 0x9E3779B1u ^ (seed * 0x85EBCA77u) ^ (tag * 0xC2B2AE3Du);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    if (h >= keep) return;
    const int id = a.ids[k];
    float x[LB_NIN];
    const float s = strip_features(a, id, a.meta[k], x);
    const int z = id / a.g.pl, r = id - z * a.g.pl, y = r / a.g.ld, xx = r - y * a.g.ld;
    const int it = (xx + off) + gt.ld * (y + off) + gt.pl * (z + off);
    const int slot = atomicAdd(count, 1);
    if (slot >= cap) return;
#pragma unroll
    for (int i = 0; i < LB_NIN; ++i) X[(size_t)slot * LB_NIN + i] = x[i];
    const float inv = 1.0f / s;
    T[(size_t)slot * 2 + 0] = (Pt[it] - a.P[id]) * inv;
    T[(size_t)slot * 2 + 1] = (Qt[it] - a.Q[id]) * inv;
    S[slot] = s;
}

// dst (student grid gd) <- src (teacher grid gs) over the student's updated region.
__global__ void k_strip_copy(const float* __restrict__ src, Geo gs, float* __restrict__ dst, Geo gd, int off) {
    const int x = MM + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = MM + blockIdx.y;
    const int z = MM + blockIdx.z;
    if (x >= gd.nnx - MM) return;
    dst[x + gd.ld * y + gd.pl * z] = src[(x + off) + gs.ld * (y + off) + gs.pl * (z + off)];
}

}  // namespace

hipError_t launch_strip_infer(const StripArgs& a, const float* net, float2* delta, int mfma, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    if (mfma)   // two resident blocks per CU (LDS), each walking its share of the cells
        k_strip_infer_mfma<<<std::min((a.n + SB - 1) / SB, 512), SB, 0, st>>>(a, net, delta);
    else
        k_strip_infer<<<(a.n + SB - 1) / SB, SB, 0, st>>>(a, net, delta);
    return hipGetLastError();
}

hipError_t launch_strip_apply(const StripArgs& a, const float2* delta, float* P, float* Q, float* px, float* py,
                              float* pz, float* qx, float* qy, float* qz, float* u, float* v, float* w,
                              hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    k_strip_apply<<<(a.n + SB - 1) / SB, SB, 0, st>>>(a, delta, P, Q, px, py, pz, qx, qy, qz, u, v, w);
    return hipGetLastError();
}

hipError_t launch_strip_sample(const StripArgs& a, const float* Pt, const float* Qt, const Geo& gt, int off,
                               unsigned keep, unsigned seed, unsigned tag, int cap, float* X, float* T, float* S,
                               int* count, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    k_strip_sample<<<(a.n + SB - 1) / SB, SB, 0, st>>>(a, Pt, Qt, gt, off, keep, seed, tag, cap, X, T, S, count);
    return hipGetLastError();
}

hipError_t launch_strip_copy(const float* src, const Geo& gs, float* dst, const Geo& gd, int off, hipStream_t st) {
    const int nx = gd.nnx - 2 * MM, ny = gd.nny - 2 * MM, nz = gd.nnz - 2 * MM;
    if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
    dim3 grid((nx + 255) / 256, ny, nz);
    k_strip_copy<<<grid, 256, 0, st>>>(src, gs, dst, gd, off);
    return hipGetLastError();
}

}  // namespace a3d
