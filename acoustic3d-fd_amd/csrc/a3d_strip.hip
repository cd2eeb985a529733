// a3d_strip.hip -- the learned absorbing strip of the VTI operator (FP32).
//
// The strip is the outer STRIP_CELLS updated cells of every face of the
// padded grid (padding nb = STRIP_CELLS + 4, the 4-cell ring staying zero as
// in the reference's PML), run with the reference split-PML update at that
// width. After each stress sweep a network damps each strip cell's P and Q:
//   features  f = local state / s, s = max |local state| (amplitude-free)
//   factors   (g_P, g_Q) = STRIP_GMAX * sigmoid(net(f)) in (0, STRIP_GMAX)
//   update    P, px, py, pz, u, v *= 1 - g_P;  Q, qx, qy, qz, w *= 1 - g_Q
//             (head mode 1: only the split components along the cell's normal axis a
//             and the normal velocity, d = g p_a, p_a -= d, P -= d, as a local
//             increase of the PML's damping along a)
// (the velocities with their stresses, as a sponge damps every field: damping
// the stresses alone would leave a net inflow into the strip and a drifting
// static field). A factor in (0, 1) only removes energy, so the strip stays
// stable for any parameters; the network decides where and how strongly to damp.
// Local state of a cell with outward normal axis a (the axis of its nearest
// face): P at the 7-point star (centre, inward, outward, tangential pairs),
// Q at centre, inward, outward, and the outward normal velocity times the
// impedance rho*vp; model features vp*dt/d_a, epsilon, delta and the depth in
// the strip; for the coefficient head also vp over the velocity sample the ramp's
// strength d0 is computed from (the reference's get_d0 sample).
// Head mode 2 ("coefficient"): the network's first output gives each strip cell a
// multiplier m = STRIP_MMAX * sigmoid(y_0) of its PML damping d; the velocity and
// stress sweeps recompute the cell's split-PML coefficients from m d inside their
// update, so nothing is applied after the sweeps (m = 1: the plain ramp).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "a3d_internal.h"
#include "a3d_mlp.h"

namespace a3d {
namespace {

constexpr int SB = 256;   // threads per block

__device__ __forceinline__ float strip_features(const StripArgs& a, int id, int meta, float (&x)[LB_NIN]) {
    const int axis = meta & 3, side = (meta >> 2) & 1, dist = (meta >> 3) & 7;
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
    x[15] = a.mfac ? a.vp[id] / a.vref : 0.f;   // coefficient head: velocity over the ramp's d0 sample
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
    if (a.mfac) {   // coefficient head: the cell's damping multiplier
        a.mfac[a.ids[k]] = STRIP_MMAX / (1.0f + expf(-y[0]));
        return;
    }
    strip_head(y, x, g, pred);
    delta[k] = make_float2(g[0], g[1]);
}

// The same network on the matrix cores (binary16 operands, exact products, FP32
// sums), one wavefront per 64 strip cells, in transposed form: every layer computes
// H^T = W X^T with v_mfma_f32_32x32x16_f16, the weights as the A operand (held in
// registers for the whole kernel) and the cells as the B operand. MFMA 32x32 operand
// layout (lane l, group g = l / 32, r = 0..15): A[i][k] with i = l % 32, B[k][j] with
// j = l % 32, a lane supplying 8 values of k for group g; D[i][j] with j = l % 32,
// i = 8 (r / 4) + 4 g + r % 4. The k order within a product is free as long as A and
// B use the same one, so layer 2 takes layer 1's accumulator registers as its B
// operand directly: lane slot (ks, g, e) carries hidden unit
//   pi(ks, g, e) = 16 ks + 8 (e / 4) + 4 g + e % 4,
// which is what the layer-1 tile cb = ks / 2 holds in r = 8 (ks % 2) + e. Layer 3
// (2 outputs) is a per-lane sum over the 32 hidden units a lane holds plus its
// partner lane's (l ^ 32). Activations tanh(a) = 1 - 2 / (2^(2a log2 e) + 1) with the
// hardware exponential and reciprocal, rounded to binary16 like the operands.
typedef _Float16 half8v __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));
union Half8 {
    half8v h;
    int i[4];
};

__device__ __forceinline__ float act_tanh_hw(float a) {
    const float t = __builtin_amdgcn_exp2f(a * 2.8853900817779268f);   // 2 log2(e)
    return fmaf(-2.0f, __builtin_amdgcn_rcpf(t + 1.0f), 1.0f);
}

__global__ __launch_bounds__(SB) void k_strip_infer_mfma(const StripArgs a, const float* __restrict__ net,
                                                         float2* __restrict__ delta) {
    __shared__ float sB1[LB_NH], sB2[LB_NH], sW3[LB_NOUT][LB_NH];
    for (int i = threadIdx.x; i < LB_NH; i += SB) {
        sB1[i] = net[LB_B1 + i];
        sB2[i] = net[LB_B2 + i];
        sW3[0][i] = net[LB_W3 + i];
        sW3[1][i] = net[LB_W3 + LB_NH + i];
    }
    const float b3[LB_NOUT] = {net[LB_B3], net[LB_B3 + 1]};
    const int l = threadIdx.x & 63, g = l >> 5, lr = l & 31;
    // A operands: W1 row 32 cb + lr, inputs 8 g + e; W2 row 32 cb + lr, inputs pi(ks, g, e)
    half8v w1[2], w2[2][4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const float* r1 = net + LB_W1 + (32 * cb + lr) * LB_NIN + 8 * g;
        const float* r2 = net + LB_W2 + (32 * cb + lr) * LB_NH + 4 * g;
#pragma unroll
        for (int e = 0; e < 8; ++e) w1[cb][e] = (_Float16)r1[e];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) w2[cb][ks][e] = (_Float16)r2[16 * ks + 8 * (e >> 2) + (e & 3)];
    }
    __syncthreads();
    const int nwave = gridDim.x * (SB / 64);
    for (int base = (blockIdx.x * SB + threadIdx.x - l); base < a.n; base += nwave * 64) {
        const int k = base + l;   // this lane's cell
        float x[LB_NIN];
#pragma unroll
        for (int i = 0; i < LB_NIN; ++i) x[i] = 0.f;
        if (k < a.n) (void)strip_features(a, a.ids[k], a.meta[k], x);
        // layer-1 B operands: cells 32 rb + lr, inputs 8 g + e. Cell 32 g + lr is this
        // lane's own; the other half comes from the partner lane, which sends the
        // inputs of group g of its own cell.
        Half8 own, send, recv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            own.h[e] = (_Float16)(g ? x[8 + e] : x[e]);
            send.h[e] = (_Float16)(g ? x[e] : x[8 + e]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) recv.i[q] = __shfl_xor(send.i[q], 32);
        float y[LB_NOUT] = {0.f, 0.f};
#pragma nounroll
        for (int rb = 0; rb < 2; ++rb) {
            Half8 xb;
#pragma unroll
            for (int q = 0; q < 4; ++q) xb.i[q] = (g == rb) ? own.i[q] : recv.i[q];
            // biases and last-layer weights are re-read from LDS where they are used
            // (kept live across the loop they would take 128 registers)
            __asm__ volatile("" ::: "memory");
            // layer 1: H1^T tiles cb (hidden 32 cb ..) x cells 32 rb ..
            half8v hb[4];
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const float16v zero = {};
                const float16v d = __builtin_amdgcn_mfma_f32_32x32x16_f16(w1[cb], xb.h, zero, 0, 0, 0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 bq = *(const float4*)&sB1[32 * cb + 8 * q + 4 * g];
                    const float bb[4] = {bq.x, bq.y, bq.z, bq.w};
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int r = 4 * q + t;
                        hb[2 * cb + (r >> 3)][r & 7] = (_Float16)act_tanh_hw(d[r] + bb[t]);
                    }
                }
            }
            // layer 2: H2^T tiles cb (hidden 32 cb ..) x cells 32 rb .., then the partial
            // last layer over the hidden units this lane holds
            float part[LB_NOUT] = {0.f, 0.f};
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                float16v d = {};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) d = __builtin_amdgcn_mfma_f32_32x32x16_f16(w2[cb][ks], hb[ks], d, 0, 0, 0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int j = 32 * cb + 8 * q + 4 * g;
                    const float4 bq = *(const float4*)&sB2[j];
                    const float4 u0 = *(const float4*)&sW3[0][j];
                    const float4 u1 = *(const float4*)&sW3[1][j];
                    const float bb[4] = {bq.x, bq.y, bq.z, bq.w};
                    const float c0[4] = {u0.x, u0.y, u0.z, u0.w};
                    const float c1[4] = {u1.x, u1.y, u1.z, u1.w};
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const float h = (float)(_Float16)act_tanh_hw(d[4 * q + t] + bb[t]);
                        part[0] = fmaf(c0[t], h, part[0]);
                        part[1] = fmaf(c1[t], h, part[1]);
                    }
                }
            }
#pragma unroll
            for (int o = 0; o < LB_NOUT; ++o) {
                const float full = part[o] + __shfl_xor(part[o], 32);
                if (g == rb) y[o] = full + b3[o];
            }
        }
        if (k < a.n) {
            if (a.mfac) {   // coefficient head: the cell's damping multiplier
                a.mfac[a.ids[k]] = STRIP_MMAX / (1.0f + expf(-y[0]));
            } else {
                float gg[LB_NOUT], pred[LB_NOUT];
                strip_head(y, x, gg, pred);
                delta[k] = make_float2(gg[0], gg[1]);
            }
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
    if (a.normal) {
        // the split components along the outward normal axis and the normal velocity:
        // a local increase of the PML's damping along that axis
        const int meta = a.meta[k], axis = meta & 3;
        if (!((meta >> (8 + axis)) & 1)) return;   // stress factor 1 along the normal
        float* pa = axis == 0 ? px : (axis == 1 ? py : pz);
        float* qa = axis == 0 ? qx : (axis == 1 ? qy : qz);
        float* va = axis == 0 ? u : (axis == 1 ? v : w);
        const float dp = d.x * pa[id], dq = d.y * qa[id];
        pa[id] = pa[id] - dp;
        P[id] = P[id] - dp;
        qa[id] = qa[id] - dq;
        Q[id] = Q[id] - dq;
        va[id] *= 1.0f - (axis == 2 ? d.y : d.x);
        return;
    }
    // split fields of the directions whose stress factor is 1 stay zero when the stress
    // sweep splits only the damped directions: scaling them is skipped
    const int dirs = a.dirsplit ? (a.meta[k] >> 8) & 7 : 7;
    const float fp = 1.0f - d.x, fq = 1.0f - d.y;
    P[id] *= fp;
    Q[id] *= fq;
    if (dirs & 1) {
        px[id] *= fp;
        qx[id] *= fq;
    }
    if (dirs & 2) {
        py[id] *= fp;
        qy[id] *= fq;
    }
    if (dirs & 4) {
        pz[id] *= fp;
        qz[id] *= fq;
    }
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
    if (mfma)   // a few resident blocks per CU, each wave walking its share of 64-cell groups
        k_strip_infer_mfma<<<std::min((a.n + SB - 1) / SB, 1024), SB, 0, st>>>(a, net, delta);
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
