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
//   column:     w_new(zc+3) from the old Q at zc .. zc+7 (register window);
//   stress:     P, Q at zc from u_new, v_new (LDS) and w_new(zc-4 .. zc+3)
//               (register window), as the scheme-1 stress sweep.
// The halo velocities are recomputed by every tile that needs them (the same
// arithmetic, so the same bits) and stored only by their owner: the x-halo
// columns of u by the first 7*TY threads (one value per lane), the y-halo rows
// of v by whole waves. Old and new P, Q, u, v, w live in separate buffers
// (ping-pong across steps): a tile reads its neighbours' old values after they
// may have written new ones. Every operation is the two-sweep scheme-1 step's,
// in the same order: the results are bitwise equal to it.
//
// Load phase as in the sweeps: every global load of a plane is issued before
// the workgroup's first barrier, none in a branch (clamped indices), through
// buffer resources with the plane base in scalar registers; single-use streams
// carry the non-temporal hint; the model is read as the interleaved
// (vp, rho, eps, delta) float4 of plane zc+3 and kept 4 planes deep.
#include <hip/hip_runtime.h>

#include "a3d_device.h"
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
    constexpr int HP = 2 * MM - 1;                                   // 7: P halo (u at x-4 .. x+3 needs P at x-7 .. x+7)
    constexpr int NH = 2 * MM - 1;                                   // 7 halo columns of u, rows of v
    constexpr int NYH = (2 * HP * TX + TX * TY - 1) / (TX * TY);    // y-halo P loads per thread
    constexpr int NPV = (NH + TY - 1) / TY;                          // halo v rows per thread
    static_assert(2 * HP * TY <= TX * TY, "one x-halo P value per thread");
    __shared__ float sP[TY + 2 * HP][TX + 2 * HP];   // row ri: y0-7+ri, column ci: x0-7+ci
    __shared__ float sU[TY][TX + NH];                // u_new, column cu: x0-4+cu
    __shared__ float sV[TY + NH][TX];                // v_new, row rv: y0-4+rv

    const Launch& L = a.l;
    const int tile = remap(L);
    if (tile >= L.ntiles) return;
    const int tix = tile % L.ntx, r0t = tile / L.ntx, tiy = r0t % L.nty, seg = r0t / L.nty;
    const Geo g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * TX + tx;
    const int x0 = tix * TX, y0 = MM + tiy * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool xupd = x >= MM && x < g.nnx - MM;
    const bool valid = y < g.nny - MM && xupd;
    const int yl = min(y, g.nny - 1);
    const int col = x + g.ld * yl;
    const int colc = min(max(x, MM), g.nnx - MM - 1) + g.ld * min(y, g.nny - MM - 1);
    const int zb = L.zbeg + seg * L.zlen;
    const int ze = min(zb + L.zlen, L.zend);
    const float* T = (const float*)a.tab;
    const float dtx = (float)a.dtx, dty = (float)a.dty, dtz = (float)a.dtz;
    const Src S0 = a.s;
    const f32x4* m4 = (const f32x4*)a.m4;

    // the step's source is added to P and Q where the velocity updates read them (never stored)
    auto injP = [&](float v, int i) {
        if (i == S0.s0) v += (float)S0.p0;
        if (i == S0.s1) v += (float)S0.p1;
        return v;
    };
    auto injQ = [&](float v, int i) {
        if (i == S0.s0) v += (float)S0.q0;
        if (i == S0.s1) v += (float)S0.q1;
        return v;
    };
    const bool srccol = (S0.s0 >= 0 && S0.s0 % g.pl == col) || (S0.s1 >= 0 && S0.s1 % g.pl == col);

    // fF = old Q (without the source) at planes zc .. zc+7, wW = new w at zc-4 .. zc+3,
    // mq = model at zc .. zc+3
    float fF[8], wW[8];
    f32x4 mq[4];
    // new w at plane p (velocity sweep arithmetic) from Fj(j) = old Q at plane p-3+j
    auto wcalc = [&](int p, auto Fj, float wo, float r, float cz1, float cz2) -> float {
        float zz = d8([&](int j) {
            float v = Fj(j);
            if (srccol) v = injQ(v, (p - 3 + j) * g.pl + col);
            return v;
        });
        zz /= r;
        return cz2 * wo - cz1 * dtz * zz;
    };
    {
        // warm-up: new w at zb-4 .. zb+2 (stored by this segment on its own planes only;
        // planes outside the updated range keep their value)
        float fw[14];   // old Q at zb-7 .. zb+6
#pragma unroll
        for (int i = 0; i < 14; ++i) fw[i] = bld(a.Q0, min(max(zb - 7 + i, 0), g.nnz - 1) * g.pl, col);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int p = zb - 4 + k;
            const float wo = bld(a.w0, p * g.pl, col);
            if (p >= MM && p < g.nnz - MM) {
                const float r = bld(a.rho, p * g.pl, colc);
                const float wn = wcalc(p, [&](int j) { return fw[k + j]; }, wo, r, T[a.c.cz1 + p], T[a.c.cz2 + p]);
                if (valid && p >= zb && p < ze) bst<NT>(a.w1, p * g.pl, col, wn);
                wW[k] = wn;
            } else {
                wW[k] = wo;
            }
        }
#pragma unroll
        for (int j = 0; j < 7; ++j) fF[j] = fw[7 + j];
#pragma unroll
        for (int j = 0; j < 3; ++j) mq[j] = bld4(m4, (zb + j) * g.pl, colc);
    }

    // per-thread constants: PML factors of the own point, the halo u column, the halo v rows
    const int xs = x < g.nnx ? x : 0;
    const float ax1 = T[a.c.ax1 + xs], ax2 = T[a.c.ax2 + xs];
    const float ay1 = T[a.c.ay1 + yl], ay2 = T[a.c.ay2 + yl];
    const float cx1 = T[a.c.cx1 + xs], cx2 = T[a.c.cx2 + xs];
    const float cy1 = T[a.c.cy1 + yl], cy2 = T[a.c.cy2 + yl];
    const bool recx = a.rec && valid && x >= a.nb && x < a.nb + a.nx && y >= a.nb && y < a.nb + a.ny;
    // P x-halo value of this thread: row pxr, column pxj of the 14 halo columns
    const bool pxh = t < 2 * HP * TY;
    const int pxr = pxh ? t / (2 * HP) : ty, pxj = t % (2 * HP);
    const int pxc = pxj < HP ? pxj : TX + pxj;   // sP column
    const int pxo = min(max(x0 - HP + pxc, 0), g.nnx - 1) + g.ld * min(y0 + pxr, g.nny - 1);
    // halo u of this thread: row hr, column hj of x0-4 .. x0-1, x0+64 .. x0+66
    const bool uh = t < NH * TY;
    const int hr = uh ? t / NH : 0, hj = t % NH;
    const int xcu = hj < MM ? x0 - MM + hj : x0 + TX + (hj - MM);
    const int cu = hj < MM ? hj : TX + hj;       // sU column
    const bool uh_upd = uh && y0 + hr < g.nny - MM && xcu >= MM && xcu < g.nnx - MM;
    const int xcs = min(max(xcu, 0), g.nnx - 1);
    const int uho = min(max(xcu, MM), g.nnx - MM - 1) + g.ld * min(y0 + hr, g.nny - MM - 1);
    const float hcx1 = T[a.c.cx1 + xcs], hcx2 = T[a.c.cx2 + xcs];
    // halo v rows of this thread: hv = ty + j*TY < 7
    int vho[NPV], vrv[NPV];
    bool vh_upd[NPV];
    float hcy1[NPV], hcy2[NPV];
#pragma unroll
    for (int j = 0; j < NPV; ++j) {
        const int hv = ty + j * TY;
        const int yv = hv < MM ? y0 - MM + hv : y0 + TY + (hv - MM);
        vrv[j] = hv < MM ? hv : TY + hv;
        vh_upd[j] = hv < NH && xupd && yv >= MM && yv < g.nny - MM;
        vho[j] = min(max(x, MM), g.nnx - MM - 1) + g.ld * min(max(yv, MM), g.nny - MM - 1);
        const int yvs = min(max(yv, 0), g.nny - 1);
        hcy1[j] = T[a.c.cy1 + yvs];
        hcy2[j] = T[a.c.cy2 + yvs];
    }

    for (int zc = zb; zc < ze; ++zc) {
        const int zo = zc * g.pl;
        // ---- every global load of this plane first
        fF[7] = bld<NT>(a.Q0, min(zc + 7, g.nnz - 1) * g.pl, col);
        const float wo = bld<NT>(a.w0, zo + 3 * g.pl, col);
        mq[3] = bld4<NT>(m4, zo + 3 * g.pl, colc);
        const float czw1 = T[a.c.cz1 + zc + 3], czw2 = T[a.c.cz2 + zc + 3];
        const float az1 = T[a.c.az1 + zc], az2 = T[a.c.az2 + zc];
        const bool split = !(ax1 == 1.0f && ay1 == 1.0f && zc >= a.zi0 && zc < a.zi1);
        float f0, f1, f2, f3, f4, f5;   // split fields, loaded in PML cells only
        if (valid && split) {
            f0 = bld<NT>(a.px, zo, col);
            f1 = bld<NT>(a.py, zo, col);
            f2 = bld<NT>(a.pz, zo, col);
            f3 = bld<NT>(a.qx, zo, col);
            f4 = bld<NT>(a.qy, zo, col);
            f5 = bld<NT>(a.qz, zo, col);
        }
        const float u0c = bld<NT>(a.u0, zo, colc);
        const float v0c = bld<NT>(a.v0, zo, colc);
        const float u0h = bld(a.u0, zo, uho);
        const float rhoh = bld(a.rho, zo, uho);
        float v0h[NPV], rhov[NPV];
#pragma unroll
        for (int j = 0; j < NPV; ++j) {
            v0h[j] = bld(a.v0, zo, vho[j]);
            rhov[j] = bld(a.rho, zo, vho[j]);
        }
        const float pc = bld(a.P0, zo, col);   // old P of the own point (without the source)
        const float phx = bld(a.P0, zo, pxo);
        float phy[NYH];
        int phyo[NYH];
#pragma unroll
        for (int j = 0; j < NYH; ++j) {
            const int k = min(t + j * TX * TY, 2 * HP * TX - 1);
            const int rr = k >> 6, cc = k & 63;
            const int yy = rr < HP ? y0 - HP + rr : y0 + TY + (rr - HP);
            phyo[j] = (x0 + cc) + g.ld * min(max(yy, 0), g.nny - 1);
            phy[j] = bld(a.P0, zo, phyo[j]);
        }
        // keep the halo loads ahead of the barrier: a load whose only use sits in a
        // branch below would otherwise be moved into it and waited for there
#pragma unroll
        for (int j = 0; j < NPV; ++j) asm volatile("" ::"v"(v0h[j]), "v"(rhov[j]));
        asm volatile("" ::"v"(u0h), "v"(rhoh));
        // ---- P(zc) with the step's source, and its halos, into LDS. Threads past the
        // halo counts write a halo cell another thread also writes, with the same value
        // (their indices were clamped onto it), so the stores need no branch.
        sP[ty + HP][tx + HP] = injP(pc, zo + col);
        sP[pxr + HP][pxc] = injP(phx, zo + pxo);
#pragma unroll
        for (int j = 0; j < NYH; ++j) {
            const int k = min(t + j * TX * TY, 2 * HP * TX - 1);
            const int rr = k >> 6, cc = k & 63;
            sP[rr < HP ? rr : TY + rr][cc + HP] = injP(phy[j], zo + phyo[j]);
        }
        __syncthreads();
        // ---- u_new(zc) on the tile's rows and the 7 halo columns, v_new(zc) on its
        // columns and the 7 halo rows (0 where the velocity sweep never writes)
        // (the plane's global stores all come after the stress update: a store here
        // would make later register reuse wait for it)
        const float r0 = mq[0].y;
        float un = 0.f, vn = 0.f;
        {
            if (valid) {
                float xx = d8([&](int k) { return sP[ty + HP][tx + HP - 3 + k]; });
                xx /= r0;
                un = cx2 * u0c - cx1 * dtx * xx;
            }
            sU[ty][tx + MM] = un;
            if (uh) {
                float uhn = 0.f;
                if (uh_upd) {
                    const int spc = xcu - (x0 - HP);
                    float xx = d8([&](int k) { return sP[hr + HP][spc - 3 + k]; });
                    xx /= rhoh;
                    uhn = hcx2 * u0h - hcx1 * dtx * xx;
                }
                sU[hr][cu] = uhn;
            }
            if (valid) {
                float yy = d8([&](int k) { return sP[ty + HP - 3 + k][tx + HP]; });
                yy /= r0;
                vn = cy2 * v0c - cy1 * dty * yy;
            }
            sV[ty + MM][tx] = vn;
#pragma unroll
            for (int j = 0; j < NPV; ++j) {
                if (ty + j * TY < NH) {
                    float vhn = 0.f;
                    if (vh_upd[j]) {
                        const int spr = vrv[j] + (HP - MM);   // sP row of the halo v row
                        float yy = d8([&](int k) { return sP[spr - 3 + k][tx + HP]; });
                        yy /= rhov[j];
                        vhn = hcy2[j] * v0h[j] - hcy1[j] * dty * yy;
                    }
                    sV[vrv[j]][tx] = vhn;
                }
            }
        }
        // ---- column: w_new(zc+3) completes the backward window
        const bool wupd = valid && zc + 3 < g.nnz - MM;
        const float wz3 = wupd ? wcalc(zc + 3, [&](int j) { return fF[j]; }, wo, mq[3].y, czw1, czw2) : wo;
        wW[7] = wz3;
        const float Wz = d8([&](int k) { return wW[k]; });
        const float q0 = fF[0];
        const f32x4 m = mq[0];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            wW[k] = wW[k + 1];
            fF[k] = fF[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) mq[k] = mq[k + 1];
        __syncthreads();
        // ---- stress at (x, y, zc): the scheme-1 update (split fields in PML cells only)
        if (valid) {
            const float c0 = m.x, e0 = m.z, d0 = m.w;
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
            const float zz = Wz;
            const float e2 = 1 + 2 * ee;
            const float s2 = sqrtf(1 + 2 * dd);
            float Pn, Qn;
            if (split) {
                const float npx = ax2 * f0 - ax1 * r0 * c0 * c0 * e2 * dtx * xx;
                const float npy = ay2 * f1 - ay1 * r0 * c0 * c0 * e2 * dty * yy;
                const float npz = az2 * f2 - az1 * r0 * c0 * c0 * s2 * dtz * zz;
                const float nqx = ax2 * f3 - ax1 * r0 * c0 * c0 * s2 * dtx * xx;
                const float nqy = ay2 * f4 - ay1 * r0 * c0 * c0 * s2 * dty * yy;
                const float nqz = az2 * f5 - az1 * r0 * c0 * c0 * dtz * zz;
                bst<NT>(a.px, zo, col, npx);
                bst<NT>(a.py, zo, col, npy);
                bst<NT>(a.pz, zo, col, npz);
                bst<NT>(a.qx, zo, col, nqx);
                bst<NT>(a.qy, zo, col, nqy);
                bst<NT>(a.qz, zo, col, nqz);
                Pn = npx + npy + npz;
                Qn = nqx + nqy + nqz;
            } else {
                const float K = r0 * c0 * c0;
                const float hxy = dtx * xx + dty * yy, hz = dtz * zz;
                Pn = pc - K * (e2 * hxy + s2 * hz);
                Qn = q0 - K * (s2 * hxy + hz);
            }
            bst<NT>(a.P1, zo, col, Pn);
            bst<NT>(a.Q1, zo, col, Qn);
            bst<NT>(a.u1, zo, col, un);
            bst<NT>(a.v1, zo, col, vn);
            if (recx && zc == a.zrec) a.rec[(y - a.nb) * a.nx + (x - a.nb)] = Pn;
        }
        if (wupd && zc + 3 < ze) bst<NT>(a.w1, zo + 3 * g.pl, col, wz3);
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
