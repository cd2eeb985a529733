// a3d_strip.hip -- the learned absorbing strip of the VTI operator (FP32).
//
// The strip is the outer STRIP_CELLS updated cells of every face of the
// padded grid (padding nb = STRIP_CELLS + 4, the 4-cell ring staying zero as
// in the reference's PML), run with the reference split-PML update at that
// width. After each stress sweep a network corrects the normal-axis split
// pressure components (and so P and Q) of every strip cell:
//   features  f = local state / s, s = max |local state| (amplitude-free)
//   delta     (dP, dQ) = s * net(f)
// so the correction is positively homogeneous of degree 1 in the wavefield.
// Local state of a cell with outward normal axis a (the axis of its nearest
// face): P at the 7-point star (centre, inward, outward, tangential pairs),
// Q at centre, inward, outward, and the outward normal velocity times the
// impedance rho*vp; model features vp*dt/d_a, epsilon, delta and the depth in
// the strip.
#include <hip/hip_runtime.h>

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
    float x[LB_NIN], h1[LB_NH], h2[LB_NH], y[LB_NOUT];
    const float s = strip_features(a, a.ids[k], a.meta[k], x);
    mlp_forward(sp, x, h1, h2, y);
    delta[k] = make_float2(s * y[0], s * y[1]);
}

__global__ __launch_bounds__(SB) void k_strip_apply(const StripArgs a, const float2* __restrict__ delta,
                                                    float* __restrict__ P, float* __restrict__ Q,
                                                    float* __restrict__ px, float* __restrict__ py,
                                                    float* __restrict__ pz, float* __restrict__ qx,
                                                    float* __restrict__ qy, float* __restrict__ qz) {
    const int k = blockIdx.x * SB + threadIdx.x;
    if (k >= a.n) return;
    const int id = a.ids[k], axis = a.meta[k] & 3;
    const float2 d = delta[k];
    P[id] += d.x;
    Q[id] += d.y;
    float* sp = axis == 0 ? px : (axis == 1 ? py : pz);
    float* sq = axis == 0 ? qx : (axis == 1 ? qy : qz);
    sp[id] += d.x;
    sq[id] += d.y;
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

hipError_t launch_strip_infer(const StripArgs& a, const float* net, float2* delta, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    k_strip_infer<<<(a.n + SB - 1) / SB, SB, 0, st>>>(a, net, delta);
    return hipGetLastError();
}

hipError_t launch_strip_apply(const StripArgs& a, const float2* delta, float* P, float* Q, float* px, float* py,
                              float* pz, float* qx, float* qy, float* qz, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    k_strip_apply<<<(a.n + SB - 1) / SB, SB, 0, st>>>(a, delta, P, Q, px, py, pz, qx, qy, qz);
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
