// a3d_fused.hip -- one-pass VTI time step (FP32): the velocity and stress
// updates of the staggered scheme in a single z-streaming kernel.
//
// The two-sweep step reads P, Q and writes u, v, w in one pass, then reads
// u, v, w back in the next; here a workgroup (64 x TY tile, z segment) makes
// the new velocities it needs itself, so they never leave the chip before
// the stress update that consumes them:
//   plane zc:   P(zc) staged in LDS with a 7-cell halo (x for u, y for v):
//               u_new(zc) on the tile's rows over x0-4 .. x0+66,
//               v_new(zc) on the tile's columns over y0-4 .. y0+TY+2;
//   column:     w_new(zc+3) from the register window Q(zc .. zc+7);
//   stress:     P, Q at zc from u_new, v_new (LDS) and w_new(zc-4 .. zc+3)
//               (register window), as the scheme-1 stress sweep.
// The halo velocities are recomputed by every tile that needs them (the same
// arithmetic, so the same bits) and stored only by their owner. Old and new
// P, Q, u, v, w live in separate buffers (ping-pong across steps): a tile
// must read its neighbours' old values after they may have written new ones.
// Every operation is the scheme-1 two-sweep step's, in the same order: the
// results are bitwise equal to it.
#include <hip/hip_runtime.h>

#include "a3d_internal.h"

namespace a3d {
namespace {

constexpr float kC0 = 1.196289f, kC1 = -0.0797526f, kC2 = 0.009570313f, kC3 = -0.0006975447f;

// Staggered difference over w(0..7), centre pair (w(4), w(3)); accumulation as the
// reference loop m = 0..3 (same as the two-sweep kernels).
template <class F>
__device__ __forceinline__ float d8(F w) {
    float r = 0;
    r += kC0 * (w(4) - w(3));
    r += kC1 * (w(5) - w(2));
    r += kC2 * (w(6) - w(1));
    r += kC3 * (w(7) - w(0));
    return r;
}

__device__ __forceinline__ int remap(const Launch& l) {
    const int per = l.grid >> 3;
    return (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
}

template <int TY>
__global__ __launch_bounds__(TX* TY) void k_fused_vti(const FusedArgs a) {
    constexpr int HP = 2 * MM - 1;          // 7: P halo (u at x-4 .. x+3 needs P at x-7 .. x+7)
    __shared__ float sP[TY + 2 * HP][TX + 2 * HP];
    __shared__ float sU[TY][TX + 2 * MM - 1];   // u_new at x0-4 .. x0+66
    __shared__ float sV[TY + 2 * MM - 1][TX];   // v_new at y0-4 .. y0+TY+2

    const Launch& L = a.l;
    const int t = remap(L);
    if (t >= L.ntiles) return;
    const int tix = t % L.ntx, r0t = t / L.ntx, tiy = r0t % L.nty, seg = r0t / L.nty;
    const Geo g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = tix * TX, y0 = MM + tiy * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool yin = y < g.nny;
    const bool xupd = x >= MM && x < g.nnx - MM;
    const bool valid = y < g.nny - MM && xupd;
    const int yl = yin ? y : g.nny - 1;
    const int col = x + g.ld * yl;
    const int zb = L.zbeg + seg * L.zlen;
    const int ze = min(zb + L.zlen, L.zend);
    const float* T = (const float*)a.tab;
    const float dtx = (float)a.dtx, dty = (float)a.dty, dtz = (float)a.dtz;
    const Src S0 = a.s;

    // the step's source is added to P and Q on every load (never stored)
    auto ldP = [&](int i) {
        float v = a.P0[i];
        if (i == S0.s0) v += (float)S0.p0;
        if (i == S0.s1) v += (float)S0.p1;
        return v;
    };
    auto ldQ = [&](int i) {
        float v = a.Q0[i];
        if (i == S0.s0) v += (float)S0.q0;
        if (i == S0.s1) v += (float)S0.q1;
        return v;
    };
    auto ldQz = [&](int z) { return (z >= 0 && z < g.nnz) ? ldQ(z * g.pl + col) : 0.f; };

    // w_new at plane p of this column from wQ = Q(p-3 .. p+4); 0 where the
    // velocity sweep never writes (outside the updated region)
    float wQ[8], wW[8];
    auto w_at = [&](int p) -> float {
        if (!(valid && p >= MM && p < g.nnz - MM)) return 0.f;
        const int id = p * g.pl + col;
        float zz = d8([&](int k) { return wQ[k]; });
        zz /= a.rho[id];
        return T[a.c.cz2 + p] * a.w0[id] - T[a.c.cz1 + p] * dtz * zz;
    };

#pragma unroll
    for (int k = 0; k < 7; ++k) wQ[k] = ldQz(zb - 7 + k);
    // warm-up: w_new at zb-4 .. zb+2 (stored by this segment only on its own planes)
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const int p = zb - 4 + j;
        wQ[7] = ldQz(p + 4);
        const float wn = w_at(p);
        wW[j] = wn;
        if (p >= zb && p < ze && valid) a.w1[p * g.pl + col] = wn;
#pragma unroll
        for (int k = 0; k < 7; ++k) wQ[k] = wQ[k + 1];
    }

    const int xs = x < g.nnx ? x : 0;
    const float ax1 = T[a.c.ax1 + xs], ax2 = T[a.c.ax2 + xs];
    const float ay1 = T[a.c.ay1 + yl], ay2 = T[a.c.ay2 + yl];
    const bool recx = a.rec && valid && x >= a.nb && x < a.nb + a.nx && y >= a.nb && y < a.nb + a.ny;

    constexpr int NYH = (2 * HP * TX + TX * TY - 1) / (TX * TY);   // y-halo P loads per thread
    constexpr int NV = (TY + 2 * MM - 1 + TY - 1) / TY;           // v rows per thread
    const bool xh = tx < 2 * HP;                                  // loads an x-halo P column
    const int xxh = tx < HP ? x0 - HP + tx : x0 + TX + (tx - HP);
    const int sch = tx < HP ? tx : TX + tx;
    const bool uh = tx < 2 * MM - 1;                              // computes an x-halo u column
    const int xcu = tx < MM ? x0 - MM + tx : x0 + TX + (tx - MM);
    const int scu = tx < MM ? tx : TX + tx;
    const bool uh_upd = uh && y < g.nny - MM && xcu >= MM && xcu < g.nnx - MM;
    const bool u_upd = y < g.nny - MM && xupd;

    for (int zc = zb; zc < ze; ++zc) {
        const int zo = zc * g.pl;
        const int id = zo + col;
        // ---- every global load of this plane first (one exposed latency per plane)
        const int pw = zc + 3;
        const bool w_upd = valid && pw < g.nnz - MM;
        const float q7 = ldQz(pw + 4);
        const float w0p = w_upd ? a.w0[pw * g.pl + col] : 0.f;
        const float rhop = w_upd ? a.rho[pw * g.pl + col] : 1.f;
        const float pc = (yin && x < g.nnx) ? ldP(zo + yl * g.ld + x) : 0.f;
        const float phx = (xh && yin && xxh >= 0 && xxh < g.nnx) ? ldP(zo + yl * g.ld + xxh) : 0.f;
        float phy[NYH];
#pragma unroll
        for (int j = 0; j < NYH; ++j) {
            const int k = ty * TX + tx + j * TX * TY;
            const int r = k >> 6, c = k & 63;
            const int yy = r < HP ? y0 - HP + r : y0 + TY + (r - HP);
            phy[j] = (k < 2 * HP * TX && yy >= 0 && yy < g.nny && x0 + c < g.nnx) ? ldP(zo + yy * g.ld + x0 + c) : 0.f;
        }
        const float u0c = u_upd ? a.u0[id] : 0.f;
        const float rhoc = u_upd ? a.rho[id] : 1.f;
        const int iuh = zo + yl * g.ld + xcu;
        const float u0h = uh_upd ? a.u0[iuh] : 0.f;
        const float rhoh = uh_upd ? a.rho[iuh] : 1.f;
        float v0r[NV], rhov[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int rr = ty + j * TY;
            const int yv = y0 - MM + rr;
            const bool ok = rr < TY + 2 * MM - 1 && xupd && yv >= MM && yv < g.nny - MM;
            v0r[j] = ok ? a.v0[zo + yv * g.ld + x] : 0.f;
            rhov[j] = ok ? a.rho[zo + yv * g.ld + x] : 1.f;
        }
        const float az1 = T[a.c.az1 + zc], az2 = T[a.c.az2 + zc];
        const bool split = !(ax1 == 1.0f && ay1 == 1.0f && az1 == 1.0f);
        float c0 = 0.f, e0 = 0.f, d0 = 0.f, P0c = 0.f, Q0c = 0.f;
        float px0 = 0.f, py0 = 0.f, pz0 = 0.f, qx0 = 0.f, qy0 = 0.f, qz0 = 0.f;
        if (valid) {
            c0 = a.vp[id];
            e0 = a.eps[id];
            d0 = a.del[id];
            if (split) {
                px0 = a.px[id];
                py0 = a.py[id];
                pz0 = a.pz[id];
                qx0 = a.qx[id];
                qy0 = a.qy[id];
                qz0 = a.qz[id];
            } else {
                P0c = a.P0[id];
                Q0c = a.Q0[id];
            }
        }
        // ---- column: w_new(zc+3) from the Q window
        {
            wQ[7] = q7;
            float wn = 0.f;
            if (w_upd) {
                float zz = d8([&](int k) { return wQ[k]; });
                zz /= rhop;
                wn = T[a.c.cz2 + pw] * w0p - T[a.c.cz1 + pw] * dtz * zz;
                if (pw < ze) a.w1[pw * g.pl + col] = wn;
            }
            wW[7] = wn;
#pragma unroll
            for (int k = 0; k < 7; ++k) wQ[k] = wQ[k + 1];
        }
        // ---- P(zc) with halos into LDS
        __syncthreads();
        sP[ty + HP][tx + HP] = pc;
        if (xh) sP[ty + HP][sch] = phx;
#pragma unroll
        for (int j = 0; j < NYH; ++j) {
            const int k = ty * TX + tx + j * TX * TY;
            if (k < 2 * HP * TX) {
                const int r = k >> 6, c = k & 63;
                sP[r < HP ? r : TY + r][c + HP] = phy[j];
            }
        }
        __syncthreads();
        // ---- u_new(zc) on the tile's rows (x0-4 .. x0+66), v_new(zc) on its columns (y0-4 .. y0+TY+2)
        {
            float un = 0.f;
            if (u_upd) {
                float xx = d8([&](int k) { return sP[ty + HP][tx + HP - 3 + k]; });
                xx /= rhoc;
                un = T[a.c.cx2 + x] * u0c - T[a.c.cx1 + x] * dtx * xx;
                if (valid) a.u1[id] = un;
            }
            sU[ty][tx + MM] = un;
            if (uh) {
                float uhn = 0.f;
                if (uh_upd) {
                    const int spc = xcu - (x0 - HP);
                    float xx = d8([&](int k) { return sP[ty + HP][spc - 3 + k]; });
                    xx /= rhoh;
                    uhn = T[a.c.cx2 + xcu] * u0h - T[a.c.cx1 + xcu] * dtx * xx;
                }
                sU[ty][scu] = uhn;
            }
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int rr = ty + j * TY;
            if (rr < TY + 2 * MM - 1) {
                const int yv = y0 - MM + rr;
                float vn = 0.f;
                if (xupd && yv >= MM && yv < g.nny - MM) {
                    float yy = d8([&](int k) { return sP[rr + k][tx + HP]; });
                    yy /= rhov[j];
                    vn = T[a.c.cy2 + yv] * v0r[j] - T[a.c.cy1 + yv] * dty * yy;
                    if (rr >= MM && rr < TY + MM) a.v1[zo + yv * g.ld + x] = vn;
                }
                sV[rr][tx] = vn;
            }
        }
        __syncthreads();
        // ---- stress at (x, y, zc): the scheme-1 update (split fields in PML cells only)
        if (valid) {
            const float r0 = rhoc;
            float ee = e0, dd = d0;
            if (a.sv) {
                const int rx = x - a.sxp, ry = y - a.syp, rz = zc - a.szp;
                const int r2 = rx * rx + ry * ry + rz * rz;
                if (r2 <= 225) {
                    if (r2 <= 16) {
                        ee = 0;
                        dd = 0;
                    } else {
                        const double f = a.taper[r2];
                        ee = (float)(f * (double)e0);
                        dd = (float)(f * (double)d0);
                    }
                }
            }
            const float yy = d8([&](int k) { return sV[ty + k][tx]; });
            const float xx = d8([&](int k) { return sU[ty][tx + k]; });
            const float zz = d8([&](int k) { return wW[k]; });
            const float e2 = 1 + 2 * ee;
            const float s2 = sqrtf(1 + 2 * dd);
            float Pn, Qn;
            if (split) {
                const float npx = ax2 * px0 - ax1 * r0 * c0 * c0 * e2 * dtx * xx;
                const float npy = ay2 * py0 - ay1 * r0 * c0 * c0 * e2 * dty * yy;
                const float npz = az2 * pz0 - az1 * r0 * c0 * c0 * s2 * dtz * zz;
                const float nqx = ax2 * qx0 - ax1 * r0 * c0 * c0 * s2 * dtx * xx;
                const float nqy = ay2 * qy0 - ay1 * r0 * c0 * c0 * s2 * dty * yy;
                const float nqz = az2 * qz0 - az1 * r0 * c0 * c0 * dtz * zz;
                a.px[id] = npx;
                a.py[id] = npy;
                a.pz[id] = npz;
                a.qx[id] = nqx;
                a.qy[id] = nqy;
                a.qz[id] = nqz;
                Pn = npx + npy + npz;
                Qn = nqx + nqy + nqz;
            } else {
                const float K = r0 * c0 * c0;
                const float hxy = dtx * xx + dty * yy, hz = dtz * zz;
                Pn = P0c - K * (e2 * hxy + s2 * hz);
                Qn = Q0c - K * (s2 * hxy + hz);
            }
            a.P1[id] = Pn;
            a.Q1[id] = Qn;
            if (recx && zc == a.zrec) a.rec[(y - a.nb) * a.nx + (x - a.nb)] = Pn;
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) wW[k] = wW[k + 1];
    }
}

}  // namespace

hipError_t launch_fused_vti(int ty, const FusedArgs& a, hipStream_t st) {
    switch (ty) {
        case 2: hipLaunchKernelGGL(k_fused_vti<2>, dim3(a.l.grid), dim3(TX, 2), 0, st, a); break;
        case 4: hipLaunchKernelGGL(k_fused_vti<4>, dim3(a.l.grid), dim3(TX, 4), 0, st, a); break;
        case 8: hipLaunchKernelGGL(k_fused_vti<8>, dim3(a.l.grid), dim3(TX, 8), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int fused_blocks_per_cu(int ty) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (ty) {
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_vti<2>, TX * 2, 0); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_vti<4>, TX * 4, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused_vti<8>, TX * 8, 0); break;
    }
    return e == hipSuccess ? n : 1;
}

}  // namespace a3d
