// a3d_kernels.hip -- CDNA4 (gfx950) time-stepping kernels of the acoustic
// forward modeller.
//
// One kernel per sweep and operator family:
//   k_step2   2nd-order isotropic leapfrog with the sponge folded in
//             (reference cuda_step_fd3d + cuda_absorb_bndr + cuda_add_dipole +
//             cuda_record, src/Toa_3dfd_cuda.cu:123-298);
//   k_vel     1st-order velocity sweep, ISO / VTI / TTI
//             (reference update_vel + add_source: iso:52-142, vti:54-128, tti:31-112);
//   k_stress  1st-order stress sweep with split PML, SV taper and the
//             receiver-plane write (reference update_stress + shot_record:
//             iso:145-190,294-304; vti:131-203; tti:115-201).
//
// Every kernel is HBM-bandwidth bound (no dense contraction; ~1-3 flop/B
// against a 20 flop/B ridge). Design: workgroup = 64 x TY threads owns an
// x-y tile and streams a z segment. Fields differentiated along z live in a
// per-thread register window; fields differentiated along x or y are staged
// per plane in LDS with a 4-cell halo on the needed sides only; pointwise
// fields are read once. Workgroup ids are remapped so each XCD (ids b with
// equal b % 8) walks a contiguous band of tiles and its L2 serves the halo
// rows its neighbours share. Per plane, every global load is issued before
// the workgroup barrier (none inside a branch), through buffer resources with
// the plane base in scalar registers; single-use streams carry the
// non-temporal hint. For ISO and VTI the stress sweep can also update w from
// its own column (the WZ instances), which the velocity sweep then skips.
//
// Precision (template PR): Fp32 (wavefields, model and arithmetic in float,
// written operation for operation as the reference -- float/double mix,
// evaluation order -- and compiled with -ffp-contract=off and IEEE division
// and sqrt, so results equal the CPU restatement bit for bit); Fp16 (wavefields
// stored as binary16 times a power-of-two scale per field, model and arithmetic
// in float); Fp64 (everything in double: the GPU error reference).
#include "a3d_internal.h"
#include "a3d_device.h"

#include <algorithm>
#include <type_traits>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace a3d {
namespace {

struct Fp32 {
    using S = float;    // wavefield storage (P, Q, u, v, w; 2nd-order pressure)
    using SP = float;   // split PML field storage
    using M = float;    // model storage
    using C = float;    // arithmetic
};
// binary16 wavefields; split PML fields stay FP32: their components with PML factor 1
// drift secularly under rounding-excited static modes, which binary16 cannot absorb.
struct Fp16 {
    using S = __half;
    using SP = float;
    using M = float;
    using C = float;
};
struct Fp64 {
    using S = double;
    using SP = double;
    using M = double;
    using C = double;
};

// Staggered 8th-order coefficients (reference iso:37, vti:40, tti:17) and the
// 2nd-derivative coefficients (Toa_3dfd_cuda.cu:32, double literals stored to
// float in the reference; exact doubles in the Fp64 instance).
template <class C> struct Coef;
template <> struct Coef<float> {
    static constexpr float c0 = 1.196289f, c1 = -0.0797526f, c2 = 0.009570313f, c3 = -0.0006975447f;
    static constexpr float s0 = (float)(-205.0 / 72.0), s1 = (float)(8.0 / 5.0), s2 = (float)(-1.0 / 5.0),
                           s3 = (float)(8.0 / 315.0), s4 = (float)(-1.0 / 560.0);
};
template <> struct Coef<double> {
    static constexpr double c0 = 1.196289, c1 = -0.0797526, c2 = 0.009570313, c3 = -0.0006975447;
    static constexpr double s0 = -205.0 / 72.0, s1 = 8.0 / 5.0, s2 = -1.0 / 5.0, s3 = 8.0 / 315.0,
                            s4 = -1.0 / 560.0;
};

__device__ __forceinline__ float csqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double csqrt(double v) { return sqrt(v); }

// Wavefield load with the per-field scale of binary16 storage (element i).
__device__ __forceinline__ float ldw(const float* p, int i, float) { return p[i]; }
__device__ __forceinline__ double ldw(const double* p, int i, float) { return p[i]; }
__device__ __forceinline__ float ldw(const __half* p, int i, float sc) { return __half2float(p[i]) * sc; }

// The same per plane (element o of the plane at offset zo), and the matching stores
// (non-temporal by default: no sweep re-reads what it writes).
template <int A = 0>
__device__ __forceinline__ float ldw(const float* p, int zo, int o, float) { return bld<A>(p, zo, o); }
template <int A = 0>
__device__ __forceinline__ double ldw(const double* p, int zo, int o, float) { return bld<A>(p, zo, o); }
template <int A = 0>
__device__ __forceinline__ float ldw(const __half* p, int zo, int o, float sc) {
    return (float)bld<A>((const _Float16*)p, zo, o) * sc;
}
template <int A = NT>
__device__ __forceinline__ void stw(float* p, int zo, int o, float v, float) { bst<A>(p, zo, o, v); }
template <int A = NT>
__device__ __forceinline__ void stw(double* p, int zo, int o, double v, float) { bst<A>(p, zo, o, v); }
template <int A = NT>
__device__ __forceinline__ void stw(__half* p, int zo, int o, float v, float inv) {
    bst<A>((_Float16*)p, zo, o, __builtin_bit_cast(_Float16, __float2half_rn(v * inv)));
}
// The value a store of v followed by a load returns (binary16 storage rounds).
__device__ __forceinline__ float stored(const float*, float v, float, float) { return v; }
__device__ __forceinline__ double stored(const double*, double v, float, float) { return v; }
__device__ __forceinline__ float stored(const __half*, float v, float inv, float sc) {
    return __half2float(__float2half_rn(v * inv)) * sc;
}

// Staggered difference over an 8-point window w(0..7) whose centre pair is
// (w(4), w(3)): forward windows hold F[-3..+4], backward windows F[-4..+3].
// Accumulation order m = 0..3 as the reference loop.
template <class C, class F>
__device__ __forceinline__ C dst8(F w) {
    using K = Coef<C>;
    C r = 0;
    r += K::c0 * (w(4) - w(3));
    r += K::c1 * (w(5) - w(2));
    r += K::c2 * (w(6) - w(1));
    r += K::c3 * (w(7) - w(0));
    return r;
}

// 2nd derivative over a 9-point window w(0..8) centred at w(4)
// (reference Toa_3dfd_cuda.cu:282-289).
template <class C, class F>
__device__ __forceinline__ C lap9(F w) {
    using K = Coef<C>;
    C c = K::s0 * w(4);
    c += K::s1 * (w(3) + w(5));
    c += K::s2 * (w(2) + w(6));
    c += K::s3 * (w(1) + w(7));
    c += K::s4 * (w(0) + w(8));
    return c;
}

// Workgroup -> tile. The hardware deals workgroup b to XCD b % 8; XCD k owns the
// contiguous run of tiles [l.band[k], l.band[k+1]) (equal runs, or runs of equal
// estimated bytes), its workgroups in order; a workgroup past its run has no tile.
// Tile order: x tile fastest, then y tile, then z segment.
__device__ __forceinline__ int remap_tile(const Launch& l) {
    const int k = (int)(blockIdx.x & 7u);
    const int t = l.band[k] + (int)(blockIdx.x >> 3);
    return t < l.band[k + 1] ? t : l.ntiles;
}

__device__ __forceinline__ void split_tile(const Launch& l, int t, int& tix, int& tiy, int& seg) {
    tix = t % l.ntx;
    const int r = t / l.ntx;
    tiy = r % l.nty;
    seg = r / l.nty;
}

__device__ __forceinline__ bool decode_tile(const Launch& l, int& tix, int& tiy, int& seg) {
    const int t = remap_tile(l);
    if (t >= l.ntiles) return false;
    split_tile(l, t, tix, tiy, seg);
    return true;
}

// Batched launches: global tile t belongs to window w with base[w] <= t < base[w+1].
__device__ __forceinline__ int find_window(const int* __restrict__ base, int nwin, int t) {
    int lo = 0, hi = nwin - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class C>
__device__ __forceinline__ C inj(C v, int i, int s0, double a0, int s1, double a1) {
    if (i == s0) v += (C)a0;
    if (i == s1) v += (C)a1;
    return v;
}

// Sponge of the 2nd-order operator (reference cuda_absorb_bndr, y then x then z,
// each product in double and stored to the working type). Factor 1.0 leaves v unchanged.
template <class C>
__device__ __forceinline__ C damp3(C v, double fy, double fx, double fz) {
    v = (C)(fy * (double)v);
    v = (C)(fx * (double)v);
    v = (C)(fz * (double)v);
    return v;
}

// Halo of one plane for an LDS tile s[TY+8][72]: HY the 4 rows above and below
// the tile, HX the 4 columns left and right of each tile row. The values are
// loaded into registers with every other load of the plane, before the
// workgroup's barrier, and written to LDS after it, so a plane costs one
// exposed memory latency (a load that fed LDS directly would wait inside its
// own branch). Every lane loads: indices outside the grid are clamped to its
// edge (entries no updated point reads), and lanes without an x-halo column
// re-read their own point.
// Table entry at a wave-uniform index (a per-plane factor): the tables are constant while
// a sweep runs, so the read goes through the scalar data cache into scalar registers.
template <class C>
__device__ __forceinline__ C ztab(const C* T, int i) {
    return ((const __attribute__((address_space(4))) C*)T)[i];
}

// Split-PML coefficients of a learned-strip cell from half its damping step h = dt d / 2
// and the network's multiplier m: c1 = 1 / (1 + m h), c2 = c1 (1 - m h) (h = 0: 1, 1).
template <class C>
__device__ __forceinline__ void strip_coef(C m, C h, C& c1, C& c2) {
    const C t = m * h;
    c1 = (C)1 / ((C)1 + t);
    c2 = c1 * ((C)1 - t);
}

template <int TY, bool HX, bool HY, class C>
struct Halo {
    static_assert((2 * MM * TX) % (TX * TY) == 0, "the halo rows split evenly over the workgroup");
    static constexpr int NY = HY ? 2 * MM * TX / (TX * TY) : 0;   // halo-row values per thread
    C y[NY > 0 ? NY : 1];
    C x;
};

template <int TY, bool HX, bool HY, class C, class LD>
__device__ __forceinline__ void load_halo(Halo<TY, HX, HY, C>& h, const Geo& g, int zo, int x0, int y0, int tx,
                                          int ty, int yl, LD ld) {
    if constexpr (HY) {
#pragma unroll
        for (int j = 0; j < Halo<TY, HX, HY, C>::NY; ++j) {
            const int k = ty * TX + tx + j * TX * TY;
            const int r = k >> 6, c = k & 63;
            const int yy = r < MM ? y0 - MM + r : y0 + TY + (r - MM);
            h.y[j] = ld(zo, min(yy, g.nny - 1) * g.ld + x0 + c);
        }
    }
    if constexpr (HX) {
        const int xx = tx < MM ? x0 - MM + tx : (tx < 2 * MM ? x0 + TX + (tx - MM) : x0 + tx);
        h.x = ld(zo, yl * g.ld + min(max(xx, 0), g.nnx - 1));
    }
}

template <int TY, bool HX, bool HY, class C>
__device__ __forceinline__ void store_halo(C (*s)[TX + 2 * MM], const Halo<TY, HX, HY, C>& h, int tx, int ty) {
    if constexpr (HY) {
#pragma unroll
        for (int j = 0; j < Halo<TY, HX, HY, C>::NY; ++j) {
            const int k = ty * TX + tx + j * TX * TY;
            const int r = k >> 6, c = k & 63;
            s[r < MM ? r : TY + r][c + MM] = h.y[j];
        }
    }
    if constexpr (HX) {
        if (tx < 2 * MM) s[ty + MM][tx < MM ? tx : TX + tx] = h.x;
    }
}

// --------------------------------------------------------------------------
// 1st-order velocity sweep
// --------------------------------------------------------------------------
// WZ (ISO, VTI): the stress sweep makes w (from its own column's P or Q), so this
// sweep updates u and v only and reads neither the z window nor w.
// SM (VTI, FP32, w in this sweep): learned strip with the coefficient head, whose
// cells take their damping coefficients from their multiplier in a.mfac.
template <class PR, int OP, int TY, bool WZ = false, bool SM = false>
__device__ __forceinline__ void vel_body(const VelArgs& a, const Src S0, int tix, int tiy, int seg) {
    using S = typename PR::S;
    using M = typename PR::M;
    using C = typename PR::C;
    // OP 4: TTI with the rotation terms evaluated here from the angles (float model)
    constexpr bool ISO = OP == 1, VTI = OP == 2, TTI = OP == 3 || OP == 4, ANG = OP == 4;
    static_assert(!(WZ && TTI), "TTI's w needs x-y differences of Q");
    constexpr bool P_Z = !VTI && !WZ;   // P differentiated along z (ISO: all axes; TTI: all axes)
    constexpr bool HAS_Q = !ISO && !WZ;
    constexpr bool Q_XY = TTI;
    constexpr int H = TY + 2 * MM, W = TX + 2 * MM;
    __shared__ C sP[H][W];
    __shared__ C sQ[Q_XY ? H : 1][Q_XY ? W : 1];

    const Geo g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = tix * TX, y0 = MM + tiy * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool yin = y < g.nny;
    const bool valid = y < g.nny - MM && x >= MM && x < g.nnx - MM;
    const int yl = yin ? y : g.nny - 1;
    const int col = x + g.ld * yl;
    // column of the nearest updated point: points outside the updated region load
    // there, so no load sits in a branch (a wave then waits once per plane)
    const int colc = min(max(x, MM), g.nnx - MM - 1) + g.ld * min(y, g.nny - MM - 1);
    const int zb = a.l.zbeg + seg * a.l.zlen;
    const int ze = min(zb + a.l.zlen, a.l.zend);
    const Scales sc = a.sc;
    const S* P = (const S*)a.P;
    const S* Q = (const S*)a.Q;
    S* U = (S*)a.u;
    S* V = (S*)a.v;
    S* Wv = (S*)a.w;
    const M* rho = (const M*)a.rho;
    const M* Ct = (const M*)a.ct;
    const M* St = (const M*)a.st;
    const M* Cp = (const M*)a.cp;
    const M* Sp = (const M*)a.sp;
    // TTI with a float model: the four rotation terms as one 16-byte load per point, or
    // (ANG) the tilt and azimuth as one 8-byte load, their sines and cosines computed here
    static_assert(!ANG || sizeof(M) == 4, "angles in the sweep: float model");
    constexpr bool PTRIG = TTI && sizeof(M) == 4 && !ANG;
    const f32x4* T4 = (const f32x4*)a.t4;
    const f32x2* A2 = (const f32x2*)a.t4;
    const C* T = (const C*)a.tab;

    auto ldP = [&](int zo, int o) {
        return inj<C>(ldw(P, zo, o, sc.s[F_P]), zo + o, S0.s0, S0.p0, S0.s1, S0.p1);
    };
    auto ldQ = [&](int zo, int o) {
        return inj<C>(ldw(Q, zo, o, sc.s[F_Q]), zo + o, S0.s0, S0.q0, S0.s1, S0.q1);
    };

    C wP[8], wQ[8];   // forward windows: F(z-3+k), k = 0..7
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int zi = (zb - 3 + k) * g.pl;
        if constexpr (P_Z) wP[k] = ldP(zi, col);
        if constexpr (HAS_Q) wQ[k] = ldQ(zi, col);
    }
    const int xs = x < g.nnx ? x : 0;
    const C cx1 = T[a.c.cx1 + xs], cx2 = T[a.c.cx2 + xs];
    // a wave is one tile row (TX lanes): its y factors are wave-uniform
    const int yw = __builtin_amdgcn_readfirstlane(yl);
    const C cy1 = ztab(T, a.c.cy1 + yw), cy2 = ztab(T, a.c.cy2 + yw);
    const C dtx = (C)a.dtx, dty = (C)a.dty, dtz = (C)a.dtz;
    static_assert(!SM || VTI, "the coefficient head runs the VTI operator");
    C hx = 0, hy = 0;
    if constexpr (SM) {
        hx = T[a.c.hvx + xs];
        hy = ztab(T, a.c.hvy + yw);
    }

    // Without z windows (w made by the stress sweep: no z difference here) the planes are
    // walked downward (a.zrev): the sweep then starts where the previous stress sweep ended,
    // whose last planes may still be in the last-level cache, and ends where the next
    // stress sweep starts (VTI 400x400x300: +1% per step, bit-identical).
    const bool down = !P_Z && !HAS_Q && a.zrev;
    for (int i = 0; i < ze - zb; ++i) {
        const int z = down ? ze - 1 - i : zb + i;
        const int zo = z * g.pl;
        C hz = 0, mc;
        bool scell = false;
        if constexpr (SM) {
            // WZ: u and v only, so only the x and y ramps matter here
            if constexpr (!WZ) hz = ztab(T, a.c.hvz + z);
            scell = hx != (C)0 || hy != (C)0 || hz != (C)0;
            if (scell) mc = (C)bld<NT>(a.mfac, zo, colc);
        }
        if constexpr (P_Z) wP[7] = ldP(zo + 4 * g.pl, col);
        if constexpr (HAS_Q) wQ[7] = ldQ(zo + 4 * g.pl, col);
        C r0 = 1, ct = 0, st = 0, cp = 0, sp = 0;
        f32x2 an;
        const C u0 = ldw<NT>(U, zo, colc, sc.s[F_U]);
        const C v0 = ldw<NT>(V, zo, colc, sc.s[F_V]);
        C w0 = 0;
        if constexpr (!WZ) w0 = ldw<NT>(Wv, zo, colc, sc.s[F_W]);
        if constexpr (!TTI) r0 = (C)bld<NT>(rho, zo, colc);
        if constexpr (PTRIG) {
            const f32x4 t4 = bld4<NT>(T4, zo, colc);
            ct = t4.x;
            st = t4.y;
            cp = t4.z;
            sp = t4.w;
        } else if constexpr (ANG) {
            an = bld<NT>(A2, zo, colc);
        } else if constexpr (TTI) {
            ct = (C)bld<NT>(Ct, zo, colc);
            st = (C)bld<NT>(St, zo, colc);
            cp = (C)bld<NT>(Cp, zo, colc);
            sp = (C)bld<NT>(Sp, zo, colc);
        }
        C pc;
        if constexpr (P_Z) pc = wP[3];
        else pc = ldP(zo, col);
        Halo<TY, true, true, C> hP;
        load_halo(hP, g, zo, x0, y0, tx, ty, yl, ldP);
        Halo<TY, Q_XY, Q_XY, C> hQ;
        if constexpr (Q_XY) load_halo(hQ, g, zo, x0, y0, tx, ty, yl, ldQ);
        const C cz1 = ztab(T, a.c.cz1 + z), cz2 = ztab(T, a.c.cz2 + z);
        __syncthreads();
        sP[ty + MM][tx + MM] = pc;
        store_halo(sP, hP, tx, ty);
        if constexpr (Q_XY) {
            sQ[ty + MM][tx + MM] = wQ[3];
            store_halo(sQ, hQ, tx, ty);
        }
        __syncthreads();
        // z differences, then the window shift, before any store: the loop comes
        // round without waiting for this plane's stores
        C Pz = 0, Qz = 0;
        if constexpr (P_Z) Pz = dst8<C>([&](int k) { return wP[k]; });
        if constexpr (HAS_Q) Qz = dst8<C>([&](int k) { return wQ[k]; });
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            if constexpr (P_Z) wP[k] = wP[k + 1];
            if constexpr (HAS_Q) wQ[k] = wQ[k + 1];
        }
        if (valid) {
            const int r = ty + MM, c = tx + MM;
            C xx, yy, zz;
            if constexpr (!TTI) {
                yy = dst8<C>([&](int k) { return sP[r - 3 + k][c]; });
                xx = dst8<C>([&](int k) { return sP[r][c - 3 + k]; });
                zz = ISO ? Pz : Qz;
                xx /= r0;
                yy /= r0;
                if constexpr (!WZ) zz /= r0;
            } else {
                const C Py = dst8<C>([&](int k) { return sP[r - 3 + k][c]; });
                const C Px = dst8<C>([&](int k) { return sP[r][c - 3 + k]; });
                const C Qy = dst8<C>([&](int k) { return sQ[r - 3 + k][c]; });
                const C Qx = dst8<C>([&](int k) { return sQ[r][c - 3 + k]; });
                if constexpr (ANG) {
                    rot_sincos(an.x, st, ct);
                    rot_sincos(an.y, sp, cp);
                }
                xx = ct * cp * Px - ct * sp * Py - st * Pz;
                yy = sp * Px + cp * Py;
                zz = st * cp * Qx - st * sp * Qy + ct * Qz;
            }
            C kx1 = cx1, kx2 = cx2, ky1 = cy1, ky2 = cy2, kz1 = cz1, kz2 = cz2;
            if constexpr (SM) {
                if (scell) {
                    strip_coef<C>(mc, hx, kx1, kx2);
                    strip_coef<C>(mc, hy, ky1, ky2);
                    strip_coef<C>(mc, hz, kz1, kz2);
                }
            }
            stw(U, zo, col, kx2 * u0 - kx1 * dtx * xx, sc.inv[F_U]);
            stw(V, zo, col, ky2 * v0 - ky1 * dty * yy, sc.inv[F_V]);
            if constexpr (!WZ) stw(Wv, zo, col, kz2 * w0 - kz1 * dtz * zz, sc.inv[F_W]);
        }
    }
}

template <class PR, int OP, int TY, bool WZ = false, bool SM = false>
__global__ __launch_bounds__(TX* TY) void k_vel(const VelArgs a) {
    int tix, tiy, seg;
    if (!decode_tile(a.l, tix, tiy, seg)) return;
    vel_body<PR, OP, TY, WZ, SM>(a, a.s, tix, tiy, seg);
}

// Batched velocity sweep: the tiles of all windows of a batch in one launch; per-window
// arguments in device memory, per-step sources srcs[it * nwin + w]. WZ (ISO, VTI): the
// stress sweep updates w.
template <class PR, int OP, int TY, bool WZ = false, bool SM = false>
__global__ __launch_bounds__(TX* TY) void k_vel_batch(const VelArgs* __restrict__ wins, const int* __restrict__ base,
                                                      int nwin, Launch lb, const Src* __restrict__ srcs, int it) {
    const int t = remap_tile(lb);
    if (t >= lb.ntiles) return;
    const int w = find_window(base, nwin, t);
    const VelArgs a = wins[w];
    int tix, tiy, seg;
    split_tile(a.l, t - base[w], tix, tiy, seg);
    vel_body<PR, OP, TY, WZ, SM>(a, srcs[(size_t)it * nwin + w], tix, tiy, seg);
}

// --------------------------------------------------------------------------
// 1st-order stress sweep
// --------------------------------------------------------------------------
// WZ (ISO, VTI; one z segment): this sweep also makes w, from its own column's old
// P (ISO) or Q (VTI) -- the velocity sweep's arithmetic, bit for bit -- one plane
// ahead of where the backward window needs it. The column reads F = P or Q at
// planes z..z+7 before writing plane z, so no other tile's values are involved.
// SM (VTI, FP32, w in the velocity sweep): learned strip with the coefficient head, whose
// cells take their damping coefficients from their multiplier in a.mfac.
template <class PR, int OP, int TY, bool WZ = false, int DIRS = 0, bool SM = false>
__device__ __forceinline__ void stress_body(const StressArgs& a, float* rec, int tix, int tiy, int seg) {
    using S = typename PR::S;
    using SP = typename PR::SP;
    using M = typename PR::M;
    using C = typename PR::C;
    // OP 4: TTI with the rotation terms evaluated here from the angles (float model)
    constexpr bool ISO = OP == 1, VTI = OP == 2, TTI = OP == 3 || OP == 4, ANG = OP == 4;
    static_assert(!(WZ && TTI), "TTI's w needs x-y differences of Q");
    static_assert(!ANG || sizeof(M) == 4, "angles in the sweep: float model");
    constexpr int H = TY + 2 * MM, W = TX + 2 * MM;
    __shared__ C sU[H][W];
    __shared__ C sV[H][W];
    __shared__ C sW[TTI ? H : 1][TTI ? W : 1];

    const Geo g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = tix * TX, y0 = MM + tiy * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool yin = y < g.nny;
    const bool valid = y < g.nny - MM && x >= MM && x < g.nnx - MM;
    const int yl = yin ? y : g.nny - 1;
    const int col = x + g.ld * yl;
    // column of the nearest updated point: points outside the updated region load
    // there, so no load sits in a branch (a wave then waits once per plane)
    const int colc = min(max(x, MM), g.nnx - MM - 1) + g.ld * min(y, g.nny - MM - 1);
    const int zb = a.l.zbeg + seg * a.l.zlen;
    const int ze = min(zb + a.l.zlen, a.l.zend);
    const Scales sc = a.sc;
    const S* U = (const S*)a.u;
    const S* V = (const S*)a.v;
    const S* Wv = (const S*)a.w;
    SP* px = (SP*)a.px;
    SP* py = (SP*)a.py;
    SP* pz = (SP*)a.pz;
    SP* qx = (SP*)a.qx;
    SP* qy = (SP*)a.qy;
    SP* qz = (SP*)a.qz;
    S* Po = (S*)a.P;
    S* Qo = (S*)a.Q;
    const M* vp = (const M*)a.vp;
    const M* rho = (const M*)a.rho;
    const M* eps = (const M*)a.eps;
    const M* del = (const M*)a.del;
    const M* Ct = (const M*)a.ct;
    const M* St = (const M*)a.st;
    const M* Cp = (const M*)a.cp;
    const M* Sp = (const M*)a.sp;
    const C* T = (const C*)a.tab;

    auto ldU = [&](int zo, int o) { return ldw(U, zo, o, sc.s[F_U]); };
    auto ldV = [&](int zo, int o) { return ldw(V, zo, o, sc.s[F_V]); };
    auto ldW = [&](int zo, int o) { return ldw(Wv, zo, o, sc.s[F_W]); };

    const C dtx = (C)a.dtx, dty = (C)a.dty, dtz = (C)a.dtz;
    C wU[8], wV[8], wW[8];   // backward windows: F(z-4+k), k = 0..7
    // WZ: fF = old F (P for ISO, Q for VTI) at planes z..z+7, fr = density at z..z+3
    C fF[8], fr[4];
    // VTI with a float model: (vp, rho, eps, delta) as one 16-byte load per point
    // (one stream instead of four; the host always provides it); WZ keeps the
    // planes z..z+3 in mq (fr is the density FIFO otherwise)
    constexpr bool PACKED = VTI && sizeof(M) == 4;
    // TTI with a float model: the four rotation terms as one 16-byte load per point, or
    // (ANG) the tilt and azimuth as one 8-byte load, their sines and cosines computed here;
    // ISO with a float model and w in the velocity sweep: (vp, rho) as one 8-byte load
    constexpr bool PTRIG = TTI && sizeof(M) == 4 && !ANG;
    constexpr bool PACK2 = ISO && sizeof(M) == 4 && !WZ;
    const f32x4* m4 = (const f32x4*)a.m4;
    const f32x2* m2 = (const f32x2*)a.m4;
    f32x4 mq[4];
    const S* Fo = ISO ? (const S*)Po : (const S*)Qo;
    const float sF = sc.s[ISO ? F_P : F_Q];
    const double fa0 = ISO ? a.s.p0 : a.s.q0, fa1 = ISO ? a.s.p1 : a.s.q1;
    // lanes of the source columns add this step's source to F as the velocity sweep does
    const bool srccol = WZ && ((a.s.s0 >= 0 && a.s.s0 % g.pl == col) || (a.s.s1 >= 0 && a.s.s1 % g.pl == col));
    // waves holding a source column (wave-uniform: the others skip the per-lane test)
    const bool srcw = WZ && __builtin_amdgcn_ballot_w64(srccol) != 0;
    // new w at plane p (velocity sweep arithmetic) from Fj(j) = F at plane p-3+j
    auto wcalc = [&](int p, auto Fj, C wo, C r, C cz1, C cz2) -> C {
        C zz;
        if (srcw) {
            zz = dst8<C>([&](int j) {
                C v = Fj(j);
                if (srccol) v = inj<C>(v, (p - 3 + j) * g.pl + col, a.s.s0, fa0, a.s.s1, fa1);
                return v;
            });
        } else {
            zz = dst8<C>([&](int j) { return Fj(j); });
        }
        zz /= r;
        return cz2 * wo - cz1 * dtz * zz;
    };
    if constexpr (WZ) {
        // new w at zb-4 .. zb+2 (planes outside the updated range keep their value)
        C fw[14];   // F at zb-7 .. zb+6
#pragma unroll
        for (int i = 0; i < 14; ++i) fw[i] = ldw(Fo, min(max(zb - 7 + i, 0), g.nnz - 1) * g.pl, col, sF);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int p = zb - 4 + k;
            const C wo = ldW(p * g.pl, col);
            if (p >= MM && p < g.nnz - MM) {
                const C r = (C)bld(rho, p * g.pl, colc);
                C c1 = ztab(T, a.c.cz1 + p), c2 = ztab(T, a.c.cz2 + p);
                if constexpr (SM) {
                    // the velocity sweep's coefficients of a strip cell (its z ramp)
                    const C hzp = ztab(T, a.c.hvz + p);
                    if (hzp != (C)0) strip_coef<C>((C)bld(a.mfac, p * g.pl, colc), hzp, c1, c2);
                }
                const C wn = wcalc(p, [&](int j) { return fw[k + j]; }, wo, r, c1, c2);
                if (valid && p >= zb) stw((S*)Wv, p * g.pl, col, wn, sc.inv[F_W]);
                wW[k] = stored(Wv, wn, sc.inv[F_W], sc.s[F_W]);   // as the stress sweep would read it
            } else {
                wW[k] = wo;
            }
        }
#pragma unroll
        for (int j = 0; j < 7; ++j) fF[j] = fw[7 + j];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if constexpr (PACKED) mq[j] = bld4(m4, (zb + j) * g.pl, colc);
            else fr[j] = (C)bld(rho, (zb + j) * g.pl, colc);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int zi = (zb - 4 + k) * g.pl;
            wW[k] = ldW(zi, col);
            if constexpr (TTI) {
                wU[k] = ldU(zi, col);
                wV[k] = ldV(zi, col);
            }
        }
    }
    const int xs = x < g.nnx ? x : 0;
    const C ax1 = T[a.c.ax1 + xs], ax2 = T[a.c.ax2 + xs];
    // A wave is one tile row (TX lanes): the row's factors and conditions are wave-uniform.
    const int yw = __builtin_amdgcn_readfirstlane(yl);
    const C ay1 = ztab(T, a.c.ay1 + yw), ay2 = ztab(T, a.c.ay2 + yw);
    const bool recx = rec && valid && x >= a.nb && x < a.nb + a.nx && y >= a.nb && y < a.nb + a.ny;
    // Directional split (DIRS 1, unsplit == 2): a PML cell keeps the split fields of its
    // damped directions only (factor != 1); P minus those fields (Q likewise) is the
    // undamped remainder, which takes the undamped directions' updates without damping.
    // DIRS 2 (unsplit == 3; VTI, TTI): one field per damped direction. The reference's
    // p_d and q_d obey the same recurrence (same PML factors, both start at 0) and differ
    // only by a per-point constant factor of their source terms, so both are multiples of
    // f_d, the split field without anisotropy factor (f_d <- b2 f_d - b1 K dt_d D_d):
    // p_x,y = e2 f_x,y, p_z = s2 f_z, q_x,y = s2 f_x,y, q_z = f_z (e2 = 1+2eps, s2 = sqrt(1+2delta)).
    constexpr bool dirs = DIRS != 0;
    constexpr bool ONE = DIRS == 2 && !ISO;
    const bool dX = ax1 != (C)1, dY = ay1 != (C)1;
    static_assert(!SM || VTI, "the coefficient head runs the VTI operator");
    C hx = 0, hy = 0;
    if constexpr (SM) {
        hx = T[a.c.hsx + xs];
        hy = ztab(T, a.c.hsy + yw);
    }

    // Unsplit scheme: outside the PML (all stress factors 1) P and Q are updated
    // directly, P -= K (dP_x + dP_y + dP_z), which equals the sum of the split
    // updates in exact arithmetic; the split fields are read and written only in
    // PML cells. split: the split update; rem: the remainder needs the old P (Q);
    // sx, sy, sz: the split fields kept.
    struct Mode {
        bool split, dZ, rem, sx, sy, sz;
    };
    auto mode = [&](int z) {
        Mode m{false, false, false, true, true, true};
        if constexpr (DIRS) {
            m.dZ = !(z >= a.zi0 && z < a.zi1);
            m.split = dX || dY || m.dZ;
            m.rem = !(dX && dY && m.dZ);
            m.sx = dX;
            m.sy = dY;
            m.sz = m.dZ;
        } else {
            m.split = !a.unsplit || !(ax1 == (C)1 && ay1 == (C)1 && z >= a.zi0 && z < a.zi1);
        }
        return m;
    };
    // Everything plane z reads from memory (the z-window leads at planes z+3 and z+7).
    // Members a variant does not use are never touched.
    struct Plane {
        C Fl, wl, ul, vl, r3, czw1, czw2, az1, az2;
        f32x4 m3, m, t;
        f32x2 mm, an;
        C c0, r0, e0, d0, ct, st, cp, sp, uc, vc;
        C f0, f1, f2, f3, f4, f5, p0, q0, mc, mw;
        Halo<TY, true, TTI, C> hU;
        Halo<TY, TTI, true, C> hV;
        Halo<TY, TTI, TTI, C> hW;
    };
    auto load = [&](int z, Plane& L) {
        const int zo = z * g.pl;
        if constexpr (WZ) {
            L.Fl = ldw<NT>(Fo, min(z + 7, g.nnz - 1) * g.pl, col, sF);
            L.wl = ldw<NT>(Wv, zo + 3 * g.pl, col, sc.s[F_W]);
            if constexpr (PACKED) L.m3 = bld4<NT>(m4, zo + 3 * g.pl, colc);
            else L.r3 = (C)bld<NT>(rho, zo + 3 * g.pl, colc);
            L.czw1 = ztab(T, a.c.cz1 + z + 3);
            L.czw2 = ztab(T, a.c.cz2 + z + 3);
            // strip cell of w's plane z+3 (its z ramp): its multiplier (wave-uniform test)
            if constexpr (SM)
                if (ztab(T, a.c.hvz + min(z + 3, g.nnz - 1)) != (C)0)
                    L.mw = (C)bld<NT>(a.mfac, min(z + 3, g.nnz - 1) * g.pl, colc);
        } else if constexpr (!TTI) {
            L.wl = ldw<NT>(Wv, zo + 3 * g.pl, col, sc.s[F_W]);   // TTI reads w's x-y halo too
        } else {
            L.wl = ldW(zo + 3 * g.pl, col);
            L.ul = ldU(zo + 3 * g.pl, col);
            L.vl = ldV(zo + 3 * g.pl, col);
        }
        if constexpr (PACKED) {
            if constexpr (!WZ) L.m = bld4<NT>(m4, zo, colc);
        } else if constexpr (PACK2) {
            L.mm = bld<NT>(m2, zo, colc);
        } else {
            L.c0 = (C)bld<NT>(vp, zo, colc);
            if constexpr (!WZ && !TTI) L.r0 = (C)bld<NT>(rho, zo, colc);
            if constexpr (!ISO) {
                L.e0 = (C)bld<NT>(eps, zo, colc);
                L.d0 = (C)bld<NT>(del, zo, colc);
            }
        }
        if constexpr (PTRIG) {
            L.t = bld4<NT>(m4, zo, colc);
        } else if constexpr (ANG) {
            L.an = bld<NT>(m2, zo, colc);
        } else if constexpr (TTI) {
            L.ct = (C)bld<NT>(Ct, zo, colc);
            L.st = (C)bld<NT>(St, zo, colc);
            L.cp = (C)bld<NT>(Cp, zo, colc);
            L.sp = (C)bld<NT>(Sp, zo, colc);
        }
        if constexpr (!TTI) {
            L.uc = ldU(zo, col);
            L.vc = ldV(zo, col);
        }
        load_halo(L.hU, g, zo, x0, y0, tx, ty, yl, ldU);
        load_halo(L.hV, g, zo, x0, y0, tx, ty, yl, ldV);
        if constexpr (TTI) load_halo(L.hW, g, zo, x0, y0, tx, ty, yl, ldW);
        L.az1 = ztab(T, a.c.az1 + z);
        L.az2 = ztab(T, a.c.az2 + z);
        // One-sided branches with distinct destinations, left unset where not loaded:
        // an if/else writing the same registers, or zeroing them first, makes the wave
        // wait for earlier loads into those registers.
        const Mode md = mode(z);
        if (valid && md.split) {
            if constexpr (SM) L.mc = (C)bld<NT>(a.mfac, zo, col);
            if (md.sx) L.f0 = ldw<NT>(px, zo, col, sc.s[F_PX]);
            if (md.sy) L.f1 = ldw<NT>(py, zo, col, sc.s[F_PY]);
            if (md.sz) L.f2 = ldw<NT>(pz, zo, col, sc.s[F_PZ]);
            if constexpr (!ISO && !ONE) {
                if (md.sx) L.f3 = ldw<NT>(qx, zo, col, sc.s[F_QX]);
                if (md.sy) L.f4 = ldw<NT>(qy, zo, col, sc.s[F_QY]);
                if (md.sz) L.f5 = ldw<NT>(qz, zo, col, sc.s[F_QZ]);
            }
        }
        if (valid && (!md.split || md.rem)) {
            // WZ: the old value of F at plane z is fF[0]
            if constexpr (!(WZ && ISO)) L.p0 = ldw<NT>((const S*)Po, zo, col, sc.s[F_P]);
            if constexpr (!ISO && !WZ) L.q0 = ldw<NT>((const S*)Qo, zo, col, sc.s[F_Q]);
        }
    };

    auto body = [&](int z, const Plane& L) {
        const int zo = z * g.pl;
        C wo = 0, czw1 = 0, czw2 = 0;   // WZ: old w, z velocity factors at plane z+3
        if constexpr (WZ) {
            fF[7] = L.Fl;
            wo = L.wl;
            if constexpr (PACKED) mq[3] = L.m3;
            else fr[3] = L.r3;
            czw1 = L.czw1;
            czw2 = L.czw2;
            if constexpr (SM) {
                const C hzw = ztab(T, a.c.hvz + min(z + 3, g.nnz - 1));
                if (hzw != (C)0) strip_coef<C>(L.mw, hzw, czw1, czw2);
            }
        } else {
            wW[7] = L.wl;
        }
        if constexpr (TTI) {
            wU[7] = L.ul;
            wV[7] = L.vl;
        }
        C c0 = 0, r0 = 0, e0 = 0, d0 = 0, ct = 0, st = 0, cp = 0, sp = 0;
        f32x2 an;
        if constexpr (PACKED) {
            const f32x4 m = WZ ? mq[0] : L.m;
            c0 = (C)m.x;
            r0 = (C)m.y;
            e0 = (C)m.z;
            d0 = (C)m.w;
        } else if constexpr (PACK2) {
            c0 = (C)L.mm.x;
            r0 = (C)L.mm.y;
        } else {
            c0 = L.c0;
            if constexpr (WZ) r0 = fr[0];
            else if constexpr (!TTI) r0 = L.r0;
            if constexpr (!ISO) {
                e0 = L.e0;
                d0 = L.d0;
            }
        }
        if constexpr (PTRIG) {
            ct = L.t.x;
            st = L.t.y;
            cp = L.t.z;
            sp = L.t.w;
        } else if constexpr (ANG) {
            an = L.an;
        } else if constexpr (TTI) {
            ct = L.ct;
            st = L.st;
            cp = L.cp;
            sp = L.sp;
        }
        C uc, vc;
        if constexpr (TTI) {
            uc = wU[4];
            vc = wV[4];
        } else {
            uc = L.uc;
            vc = L.vc;
        }
        const C az1 = L.az1, az2 = L.az2;
        const Mode md = mode(z);
        const bool split = md.split, dZ = md.dZ, rem = md.rem, sx = md.sx, sy = md.sy, sz = md.sz;
        C f0 = L.f0, f1 = L.f1, f2 = L.f2, f3 = L.f3, f4 = L.f4, f5 = L.f5, p0 = L.p0, q0 = L.q0, mc = L.mc;
        if constexpr (WZ) {
            if constexpr (ISO) p0 = fF[0];
            else q0 = fF[0];
        }
        __syncthreads();
        sU[ty + MM][tx + MM] = uc;
        sV[ty + MM][tx + MM] = vc;
        store_halo(sU, L.hU, tx, ty);
        store_halo(sV, L.hV, tx, ty);
        if constexpr (TTI) {
            sW[ty + MM][tx + MM] = wW[4];
            store_halo(sW, L.hW, tx, ty);
        }
        __syncthreads();
        // WZ: new w at plane z+3 completes the backward window
        const bool wupd = WZ && z + 3 < g.nnz - MM;
        C wz3 = wo;
        if constexpr (WZ) {
            const C rw = PACKED ? (C)mq[3].y : fr[3];   // density at plane z+3
            if (wupd) wz3 = wcalc(z + 3, [&](int j) { return fF[j]; }, wo, rw, czw1, czw2);
            wW[7] = stored(Wv, wz3, sc.inv[F_W], sc.s[F_W]);   // as the stress sweep would read it
        }
        // z differences, then the window shift, before any store: the loop comes
        // round without waiting for this plane's stores
        const C Wz = dst8<C>([&](int k) { return wW[k]; });
        C Uz = 0;
        if constexpr (TTI) Uz = dst8<C>([&](int k) { return wU[k]; });
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            wW[k] = wW[k + 1];
            if constexpr (TTI) {
                wU[k] = wU[k + 1];
                wV[k] = wV[k + 1];
            }
            if constexpr (WZ) fF[k] = fF[k + 1];
        }
        if constexpr (WZ) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if constexpr (PACKED) mq[k] = mq[k + 1];
                else fr[k] = fr[k + 1];
            }
        }
        if (valid && wupd) stw((S*)Wv, zo + 3 * g.pl, col, wz3, sc.inv[F_W]);
        if (valid) {
            const int r = ty + MM, c = tx + MM;
            C Pn, Qn = 0;
            if constexpr (ISO) {
                const C yy = dst8<C>([&](int k) { return sV[r - 4 + k][c]; });
                const C xx = dst8<C>([&](int k) { return sU[r][c - 4 + k]; });
                const C zz = Wz;
                if (split) {
                    C px0 = f0, py0 = f1, pz0 = f2, rP = 0;
                    if (dirs) {
                        // undamped directions: no split field (factors 1, old value 0)
                        px0 = dX ? f0 : (C)0;
                        py0 = dY ? f1 : (C)0;
                        pz0 = dZ ? f2 : (C)0;
                        if (rem) rP = p0 - px0 - py0 - pz0;
                    }
                    const C npx = ax2 * px0 - ax1 * r0 * c0 * c0 * dtx * xx;
                    const C npy = ay2 * py0 - ay1 * r0 * c0 * c0 * dty * yy;
                    const C npz = az2 * pz0 - az1 * r0 * c0 * c0 * dtz * zz;
                    if (sx) stw(px, zo, col, npx, sc.inv[F_PX]);
                    if (sy) stw(py, zo, col, npy, sc.inv[F_PY]);
                    if (sz) stw(pz, zo, col, npz, sc.inv[F_PZ]);
                    Pn = npx + npy + npz;
                    if (dirs) Pn += rP;
                } else {
                    const C K = r0 * c0 * c0;
                    Pn = p0 - K * (dtx * xx + dty * yy + dtz * zz);
                }
            } else {
                // SV taper (reference vti:154-175, tti:139-160)
                C ee = e0, dd = d0;
                if (a.sv && abs(z - a.szp) <= 15) {   // r2 <= 225 needs |z - szp| <= 15
                    const int rx = x - a.sxp, ry = y - a.syp, rz = z - a.szp;
                    const int r2 = rx * rx + ry * ry + rz * rz;
                    if (r2 <= 225) {
                        if (r2 <= 16) {
                            ee = 0;
                            dd = 0;
                        } else {
                            const double f = a.taper[r2];
                            ee = (C)(f * (double)e0);
                            dd = (C)(f * (double)d0);
                        }
                    }
                }
                C xx, yy, zz;
                if constexpr (VTI) {
                    yy = dst8<C>([&](int k) { return sV[r - 4 + k][c]; });
                    xx = dst8<C>([&](int k) { return sU[r][c - 4 + k]; });
                    zz = Wz;
                } else {
                    const C uy = dst8<C>([&](int k) { return sU[r - 4 + k][c]; });
                    const C ux = dst8<C>([&](int k) { return sU[r][c - 4 + k]; });
                    const C uz = Uz;
                    const C vy = dst8<C>([&](int k) { return sV[r - 4 + k][c]; });
                    const C vx = dst8<C>([&](int k) { return sV[r][c - 4 + k]; });
                    const C wy = dst8<C>([&](int k) { return sW[r - 4 + k][c]; });
                    const C wx = dst8<C>([&](int k) { return sW[r][c - 4 + k]; });
                    const C wz = Wz;
                    if constexpr (ANG) {
                        rot_sincos(an.x, st, ct);
                        rot_sincos(an.y, sp, cp);
                    }
                    xx = ct * cp * ux - ct * sp * uy - st * uz;
                    yy = sp * vx + cp * vy;
                    zz = st * cp * wx - st * sp * wy + ct * wz;
                }
                const C e2 = 1 + 2 * ee;
                const C s2 = csqrt(1 + 2 * dd);
                if (ONE && split) {
                    const C fx0 = dX ? f0 : (C)0, fy0 = dY ? f1 : (C)0, fz0 = dZ ? f2 : (C)0;
                    C rP = 0, rQ = 0;
                    if (rem) {
                        rP = p0 - e2 * (fx0 + fy0) - s2 * fz0;
                        rQ = q0 - s2 * (fx0 + fy0) - fz0;
                    }
                    C bx1 = ax1, bx2 = ax2, by1 = ay1, by2 = ay2, bz1 = az1, bz2 = az2;
                    if constexpr (SM) {
                        strip_coef<C>(mc, hx, bx1, bx2);
                        strip_coef<C>(mc, hy, by1, by2);
                        strip_coef<C>(mc, ztab(T, a.c.hsz + z), bz1, bz2);
                    }
                    const C K = VTI ? r0 * c0 * c0 : c0 * c0;
                    const C nfx = bx2 * fx0 - bx1 * K * dtx * xx;
                    const C nfy = by2 * fy0 - by1 * K * dty * yy;
                    const C nfz = bz2 * fz0 - bz1 * K * dtz * zz;
                    if (sx) stw(px, zo, col, nfx, sc.inv[F_PX]);
                    if (sy) stw(py, zo, col, nfy, sc.inv[F_PY]);
                    if (sz) stw(pz, zo, col, nfz, sc.inv[F_PZ]);
                    const C hxy = nfx + nfy;
                    Pn = e2 * hxy + s2 * nfz + rP;
                    Qn = s2 * hxy + nfz + rQ;
                } else if (split) {
                    C px0 = f0, py0 = f1, pz0 = f2, qx0 = f3, qy0 = f4, qz0 = f5, rP = 0, rQ = 0;
                    if (dirs) {
                        // undamped directions: no split field (factors 1, old value 0)
                        px0 = dX ? f0 : (C)0;
                        py0 = dY ? f1 : (C)0;
                        pz0 = dZ ? f2 : (C)0;
                        qx0 = dX ? f3 : (C)0;
                        qy0 = dY ? f4 : (C)0;
                        qz0 = dZ ? f5 : (C)0;
                        if (rem) {
                            rP = p0 - px0 - py0 - pz0;
                            rQ = q0 - qx0 - qy0 - qz0;
                        }
                    }
                    C npx, npy, npz, nqx, nqy, nqz;
                    if constexpr (VTI) {
                        C bx1 = ax1, bx2 = ax2, by1 = ay1, by2 = ay2, bz1 = az1, bz2 = az2;
                        if constexpr (SM) {
                            strip_coef<C>(mc, hx, bx1, bx2);
                            strip_coef<C>(mc, hy, by1, by2);
                            strip_coef<C>(mc, ztab(T, a.c.hsz + z), bz1, bz2);
                        }
                        npx = bx2 * px0 - bx1 * r0 * c0 * c0 * e2 * dtx * xx;
                        npy = by2 * py0 - by1 * r0 * c0 * c0 * e2 * dty * yy;
                        npz = bz2 * pz0 - bz1 * r0 * c0 * c0 * s2 * dtz * zz;
                        nqx = bx2 * qx0 - bx1 * r0 * c0 * c0 * s2 * dtx * xx;
                        nqy = by2 * qy0 - by1 * r0 * c0 * c0 * s2 * dty * yy;
                        nqz = bz2 * qz0 - bz1 * r0 * c0 * c0 * dtz * zz;
                    } else {
                        npx = ax2 * px0 - ax1 * c0 * c0 * e2 * dtx * xx;
                        npy = ay2 * py0 - ay1 * c0 * c0 * e2 * dty * yy;
                        npz = az2 * pz0 - az1 * c0 * c0 * s2 * dtz * zz;
                        nqx = ax2 * qx0 - ax1 * c0 * c0 * s2 * dtx * xx;
                        nqy = ay2 * qy0 - ay1 * c0 * c0 * s2 * dty * yy;
                        nqz = az2 * qz0 - az1 * c0 * c0 * dtz * zz;
                    }
                    if (sx) stw(px, zo, col, npx, sc.inv[F_PX]);
                    if (sy) stw(py, zo, col, npy, sc.inv[F_PY]);
                    if (sz) stw(pz, zo, col, npz, sc.inv[F_PZ]);
                    if (sx) stw(qx, zo, col, nqx, sc.inv[F_QX]);
                    if (sy) stw(qy, zo, col, nqy, sc.inv[F_QY]);
                    if (sz) stw(qz, zo, col, nqz, sc.inv[F_QZ]);
                    Pn = npx + npy + npz;
                    Qn = nqx + nqy + nqz;
                    if (dirs) {
                        Pn += rP;
                        Qn += rQ;
                    }
                } else {
                    const C K = VTI ? r0 * c0 * c0 : c0 * c0;
                    const C hxy = dtx * xx + dty * yy, hz = dtz * zz;
                    Pn = p0 - K * (e2 * hxy + s2 * hz);
                    Qn = q0 - K * (s2 * hxy + hz);
                }
                stw(Qo, zo, col, Qn, sc.inv[F_Q]);
            }
            stw(Po, zo, col, Pn, sc.inv[F_P]);
            if (z == a.zrec && recx) rec[(y - a.nb) * a.nx + (x - a.nb)] = (float)Pn;
        }
    };

    for (int z = zb; z < ze; ++z) {
        Plane L;
        load(z, L);
        body(z, L);
    }
}

template <class PR, int OP, int TY, bool WZ = false, int DIRS = 0, bool SM = false>
__global__ __launch_bounds__(TX* TY) void k_stress(const StressArgs a) {
    int tix, tiy, seg;
    if (!decode_tile(a.l, tix, tiy, seg)) return;
    stress_body<PR, OP, TY, WZ, DIRS, SM>(a, a.rec, tix, tiy, seg);
}

// Batched stress sweep; a window's record plane of step it is rec + it * nx * ny. WZ
// (ISO, VTI; one z segment per tile column): this sweep also updates w, with the
// step's source of the window (srcs[it * nwin + w], as the velocity sweep adds it).
template <class PR, int OP, int TY, int DIRS = 0, bool WZ = false, bool SM = false>
__global__ __launch_bounds__(TX* TY) void k_stress_batch(const StressArgs* __restrict__ wins,
                                                         const int* __restrict__ base, int nwin, Launch lb, int it,
                                                         const Src* __restrict__ srcs) {
    const int t = remap_tile(lb);
    if (t >= lb.ntiles) return;
    const int w = find_window(base, nwin, t);
    StressArgs a = wins[w];
    if constexpr (WZ) a.s = srcs[(size_t)it * nwin + w];
    int tix, tiy, seg;
    split_tile(a.l, t - base[w], tix, tiy, seg);
    float* rec = a.rec ? a.rec + (size_t)it * a.nx * a.ny : nullptr;
    stress_body<PR, OP, TY, WZ, DIRS, SM>(a, rec, tix, tiy, seg);
}

// --------------------------------------------------------------------------
// 2nd-order step with sponge, dipole and receivers folded in
// --------------------------------------------------------------------------
// --------------------------------------------------------------------------
// Persistent batched time loop
// --------------------------------------------------------------------------
// One launch runs steps it0 .. it1-1 of every window of a batch. The grid is
// resident (no more workgroups than the GPU holds at once); each sweep's tiles
// are handed out by ticket counters, one per XCD band (an eighth of every
// window, in ticket order: windows with the longest tile columns first, so the
// last tiles taken are short ones), and the grid meets at a barrier after each
// sweep (a stress tile reads its neighbours' velocities, a velocity tile their
// stresses). Every barrier wait is bounded: past wait_ticks the workgroup
// raises *err, and a workgroup that sees *err set leaves at its next
// barrier, so the launch always ends.

// Next ticket of counter c for the whole workgroup. The value is wave-uniform
// (readfirstlane), so the window search and the window's arguments that follow
// are scalar loads held in scalar registers, as in the launched sweeps.
__device__ __forceinline__ int next_ticket(unsigned* c) {
    __shared__ int t;
    __syncthreads();   // every thread is past the previous tile (and its ticket)
    if (threadIdx.x == 0 && threadIdx.y == 0) t = (int)atomicAdd(c, 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(t);
}

// Grid barrier (arrival counter sync[0], never reset within a launch: the k-th
// barrier waits for k * nblocks arrivals). Release: each thread's stores are made
// visible at device scope before its workgroup arrives; acquire: each thread
// fences after the wait, so later loads see the other workgroups' stores.
// Returns false when a wait ran out (here or in another workgroup).
__device__ __forceinline__ bool grid_barrier(unsigned* sync, int* err, unsigned nblocks, long long wait_ticks,
                                             unsigned& target) {
    __shared__ int ok;
    __threadfence();
    __syncthreads();
    target += nblocks;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        int good = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        if (good) {
            atomicAdd(sync, 1u);
            const long long t0 = (long long)wall_clock64();
            int spins = 0;   // second bound, independent of the clock: about 2^22 sleeps
            while (__hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    good = 0;
                    break;
                }
                if ((long long)wall_clock64() - t0 > wait_ticks || ++spins > (1 << 22)) {
                    atomicExch(err, 1);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
        }
        ok = good;
    }
    __syncthreads();
    __threadfence();
    return ok != 0;
}

// The window arguments and tables come in as separate __restrict__ kernel parameters, not
// as fields of one argument struct: the compiler may then take them as unwritten by the
// kernel and load them with scalar loads after the sweeps' stores (through a struct field
// they become vector loads held in vector registers: 125-134 instead of 97-105 VGPRs for
// TTI).
template <class PR, int OP, int TY, int DIRS, bool WZ>
__global__ __launch_bounds__(TX* TY) void k_persist(const VelArgs* __restrict__ va, const StressArgs* __restrict__ sa,
                                                   const int* __restrict__ tabs, const Src* __restrict__ srcs,
                                                   unsigned* sync, int* err, int nwin, int group, unsigned nblocks,
                                                   long long wait_ticks, int it0, int it1) {
    unsigned target = 0;
    const int x = (int)(blockIdx.x & 7u);
    const int n1 = nwin + 1;
    // tabs: vbase[8][n1], sbase[8][n1], vlo[16][n1], slo[16][n1], vwin[nwin], swin[nwin]
    const int* vb = tabs + x * n1;
    const int* sb = tabs + 8 * n1 + x * n1;
    const int* vlo = tabs + 16 * n1 + x * n1;
    const int* slo = tabs + 32 * n1 + x * n1;
    const int* vwin = tabs + 48 * n1;
    const int* swin = vwin + nwin;
    for (int it = it0; it < it1; ++it) {
        const int k = it - it0;
        for (;;) {
            const int t = next_ticket(sync + PERSIST_LINE * (1 + (2 * k) * 8 + x));
            if (t >= vb[nwin]) break;
            const int j = find_window(vb, nwin, t);
            const int w = vwin[j];
            const VelArgs a = va[w];
            const Src s0 = srcs[(size_t)it * nwin + w];
            const int t0 = vlo[j] + (t - vb[j]) * group, t1 = min(t0 + group, vlo[8 * n1 + j]);
            for (int tt = t0; tt < t1; ++tt) {
                int tix, tiy, seg;
                split_tile(a.l, tt, tix, tiy, seg);
                vel_body<PR, OP, TY, WZ, false>(a, s0, tix, tiy, seg);
            }
        }
        if (!grid_barrier(sync, err, nblocks, wait_ticks, target)) return;
        for (;;) {
            const int t = next_ticket(sync + PERSIST_LINE * (1 + (2 * k + 1) * 8 + x));
            if (t >= sb[nwin]) break;
            const int j = find_window(sb, nwin, t);
            const int w = swin[j];
            StressArgs a = sa[w];
            if constexpr (WZ) a.s = srcs[(size_t)it * nwin + w];
            float* rec = a.rec ? a.rec + (size_t)it * a.nx * a.ny : nullptr;
            const int t0 = slo[j] + (t - sb[j]) * group, t1 = min(t0 + group, slo[8 * n1 + j]);
            for (int tt = t0; tt < t1; ++tt) {
                int tix, tiy, seg;
                split_tile(a.l, tt, tix, tiy, seg);
                stress_body<PR, OP, TY, WZ, DIRS, false>(a, rec, tix, tiy, seg);
            }
        }
        if (!grid_barrier(sync, err, nblocks, wait_ticks, target)) return;
    }
}

__global__ void k_zero_u32(unsigned* __restrict__ a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}

// One tile (64 x TY columns, one z segment) of a 2nd-order step.
template <class PR, int TY>
__device__ __forceinline__ void step2_tile(const Step2Args& a, int tix, int tiy, int seg) {
    using S = typename PR::S;
    using M = typename PR::M;
    using C = typename PR::C;
    constexpr int H = TY + 2 * MM, W = TX + 2 * MM;
    __shared__ C sY[H][W];

    const Geo g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = tix * TX, y0 = MM + tiy * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool yin = y < g.nny;
    const bool valid = y < g.nny - MM && x >= MM && x < g.nnx - MM;
    const int yl = yin ? y : g.nny - 1;
    const int col = x + g.ld * yl;
    // column of the nearest updated point: points outside the updated region load
    // there, so no load sits in a branch (a wave then waits once per plane)
    const int colc = min(max(x, MM), g.nnx - MM - 1) + g.ld * min(y, g.nny - MM - 1);
    const int zb = a.l.zbeg + seg * a.l.zlen;
    const int ze = min(zb + a.l.zlen, a.l.zend);
    const float sp = a.sc.s[F_P], ip = a.sc.inv[F_P];
    S* X = (S*)a.X;
    const S* Y = (const S*)a.Y;
    const M* VV = (const M*)a.vv;
    const C dx2 = (C)a.dx2, dy2 = (C)a.dy2, dz2 = (C)a.dz2;

    auto ldY = [&](int zo, int o) {
        return inj<C>(ldw(Y, zo, o, sp), zo + o, a.s0, a.ya0, a.s1, a.ya1);
    };

    C wY[9];   // Y(z-4+k), k = 0..8
#pragma unroll
    for (int k = 0; k < 8; ++k) wY[k] = ldY((zb - 4 + k) * g.pl, col);
    const int xs = x < g.nnx ? x : 0;
    const double fx = a.fx[xs], fy = a.fy[yl];
    const bool recx = a.rec && valid && x >= a.off && x < a.off + a.nx && y >= a.off && y < a.off + a.ny;

    for (int z = zb; z < ze; ++z) {
        const int zo = z * g.pl;
        wY[8] = ldY(zo + 4 * g.pl, col);
        // default cache policy: the 2nd-order grids are small enough to stay in the
        // last-level cache between steps
        const C xo = inj<C>(ldw((const S*)X, zo, colc, sp), zo + colc, a.s0, a.xa0, a.s1, a.xa1);
        const C vv = (C)bld(VV, zo, colc);
        Halo<TY, true, true, C> hY;
        load_halo(hY, g, zo, x0, y0, tx, ty, yl, ldY);
        const double fz = a.fz[z];
        __syncthreads();
        sY[ty + MM][tx + MM] = wY[4];
        store_halo(sY, hY, tx, ty);
        __syncthreads();
        // z difference, then the window shift, before any store: the loop comes round
        // without waiting for this plane's stores
        C c1 = lap9<C>([&](int k) { return wY[k]; });   // z
        const C cur = wY[4];
#pragma unroll
        for (int k = 0; k < 8; ++k) wY[k] = wY[k + 1];
        if (valid) {
            const int r = ty + MM, c = tx + MM;
            C c2 = lap9<C>([&](int k) { return sY[r][c - 4 + k]; });   // x
            C c3 = lap9<C>([&](int k) { return sY[r - 4 + k][c]; });   // y
            c1 *= dz2;
            c2 *= dx2;
            c3 *= dy2;
            const C prev = damp3<C>(xo, fy, fx, fz);
            const C lap = vv * (c1 + c2 + c3);
            const C nv = (C)(2.0 * (double)cur - (double)prev + (double)lap);
            stw<0>(X, zo, col, damp3<C>(nv, fy, fx, fz), ip);
            if (recx && z == a.zrec) a.rec[(y - a.off) * a.nx + (x - a.off)] = (float)damp3<C>(cur, fy, fx, fz);
        }
    }
}

template <class PR, int TY>
__global__ __launch_bounds__(TX* TY) void k_step2(const Step2Args a) {
    int tix, tiy, seg;
    if (!decode_tile(a.l, tix, tiy, seg)) return;
    step2_tile<PR, TY>(a, tix, tiy, seg);
}

// 2nd-order step over z segments of at most ZB planes with every load of the segment
// issued at once (its z column, and each plane's previous level, (v dt)^2 and x-y halo):
// one memory round trip per workgroup instead of one per plane, for grids too small to
// keep the GPU busy with plane-streaming tiles. Each point's arithmetic is the streaming
// kernel's, operation for operation: the same bits.
template <class PR, int TY, int ZB>
__device__ __forceinline__ void step2_blk_tile(const Step2Args& a, int tix, int tiy, int seg) {
    using S = typename PR::S;
    using M = typename PR::M;
    using C = typename PR::C;
    constexpr int H = TY + 2 * MM, W = TX + 2 * MM;
    __shared__ C sY[ZB][H][W];

    const Geo g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = tix * TX, y0 = MM + tiy * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool yin = y < g.nny;
    const bool valid = y < g.nny - MM && x >= MM && x < g.nnx - MM;
    const int yl = yin ? y : g.nny - 1;
    const int col = x + g.ld * yl;
    const int colc = min(max(x, MM), g.nnx - MM - 1) + g.ld * min(y, g.nny - MM - 1);
    const int zb = a.l.zbeg + seg * a.l.zlen;
    const int ze = min(zb + a.l.zlen, a.l.zend);
    const float sp = a.sc.s[F_P], ip = a.sc.inv[F_P];
    S* X = (S*)a.X;
    const S* Y = (const S*)a.Y;
    const M* VV = (const M*)a.vv;
    const C dx2 = (C)a.dx2, dy2 = (C)a.dy2, dz2 = (C)a.dz2;
    auto ldY = [&](int zo, int o) {
        return inj<C>(ldw(Y, zo, o, sp), zo + o, a.s0, a.ya0, a.s1, a.ya1);
    };

    C wY[ZB + 8];   // Y(zb-4+k); planes past the grid (a last, shorter segment) clamped, unused
#pragma unroll
    for (int k = 0; k < ZB + 8; ++k) wY[k] = ldY(min(zb - 4 + k, g.nnz - 1) * g.pl, col);
    C xo[ZB], vq[ZB];
    Halo<TY, true, true, C> hY[ZB];
#pragma unroll
    for (int j = 0; j < ZB; ++j) {
        const int zo = min(zb + j, g.nnz - 1) * g.pl;
        xo[j] = inj<C>(ldw((const S*)X, zo, colc, sp), zo + colc, a.s0, a.xa0, a.s1, a.xa1);
        vq[j] = (C)bld(VV, zo, colc);
        load_halo(hY[j], g, zo, x0, y0, tx, ty, yl, ldY);
    }
    const int xs = x < g.nnx ? x : 0;
    const double fx = a.fx[xs], fy = a.fy[yl];
    const bool recx = a.rec && valid && x >= a.off && x < a.off + a.nx && y >= a.off && y < a.off + a.ny;
#pragma unroll
    for (int j = 0; j < ZB; ++j) {
        sY[j][ty + MM][tx + MM] = wY[4 + j];
        store_halo(sY[j], hY[j], tx, ty);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ZB; ++j) {
        const int z = zb + j;
        if (z >= ze) break;   // uniform: the segment's end
        const int zo = z * g.pl;
        const double fz = a.fz[z];
        C c1 = lap9<C>([&](int k) { return wY[j + k]; });   // z
        const C cur = wY[j + 4];
        if (valid) {
            const int r = ty + MM, c = tx + MM;
            C c2 = lap9<C>([&](int k) { return sY[j][r][c - 4 + k]; });   // x
            C c3 = lap9<C>([&](int k) { return sY[j][r - 4 + k][c]; });   // y
            c1 *= dz2;
            c2 *= dx2;
            c3 *= dy2;
            const C prev = damp3<C>(xo[j], fy, fx, fz);
            const C lap = vq[j] * (c1 + c2 + c3);
            const C nv = (C)(2.0 * (double)cur - (double)prev + (double)lap);
            stw<0>(X, zo, col, damp3<C>(nv, fy, fx, fz), ip);
            if (recx && z == a.zrec) a.rec[(y - a.off) * a.nx + (x - a.off)] = (float)damp3<C>(cur, fy, fx, fz);
        }
    }
}

template <class PR, int TY, int ZB>
__global__ __launch_bounds__(TX* TY) void k_step2_blk(const Step2Args a) {
    int tix, tiy, seg;
    if (!decode_tile(a.l, tix, tiy, seg)) return;
    step2_blk_tile<PR, TY, ZB>(a, tix, tiy, seg);
}

// Several shots on one model: tile t is spatial tile t / n of shot t % n, so the shots'
// copies of one tile run side by side on one XCD and share its L2 rows of (v dt)^2 and
// the sponge tables. Each tile runs the single-shot step's code on its shot's levels:
// every shot's results equal its own run bit for bit.
template <class PR, int TY, bool BLK>
__global__ __launch_bounds__(TX* TY) void k_step2_shots(const Step2Args a, const Step2Shots s) {
    const int t = remap_tile(a.l);
    if (t >= a.l.ntiles) return;
    const int k = t % s.n;
    int tix, tiy, seg;
    split_tile(a.l, t / s.n, tix, tiy, seg);
    Step2Args b = a;
    b.X = (char*)a.X + k * s.stride;
    b.Y = (const char*)a.Y + k * s.stride;
    b.s0 = s.s0[k];
    b.s1 = s.s1[k];
    b.rec = a.rec ? a.rec + k * s.rstride : nullptr;
    if constexpr (BLK) step2_blk_tile<PR, TY, 4>(b, tix, tiy, seg);
    else step2_tile<PR, TY>(b, tix, tiy, seg);
}

// --------------------------------------------------------------------------
// Setup and output kernels
// --------------------------------------------------------------------------
// Pads an interior float model into the device layout (float or double), edge replication.
template <class M>
__global__ void k_pad(const float* __restrict__ src, int nx, int ny, int nz, int off, Geo g, M* __restrict__ dst,
                      int mode, float dt, Window win) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int z = blockIdx.z;
    if (x >= g.ld) return;
    M v = 0;
    if (x < g.nnx) {
        const int cx = min(max(x - off, 0), nx - 1) + win.ox;
        const int cy = min(max(y - off, 0), ny - 1) + win.oy;
        const int cz = min(max(z - off, 0), nz - 1) + win.oz;
        v = (M)src[cz + (size_t)win.nz * cx + (size_t)win.nz * win.nx * cy];
        if (mode == 1) {
            const M t = v * (M)dt;
            v = t * t;
        }
    }
    dst[x + (size_t)g.ld * y + (size_t)g.pl * z] = v;
}

// One thread per (ix, iy) column of the interior; loops z.
template <class S>
__global__ void k_snap(const S* __restrict__ src, Geo g, int off, int nx, int ny, int nz, int s0, double a0,
                       int s1, double a1, const double* fx, const double* fy, const double* fz, float scale,
                       float* __restrict__ dst) {
    using C = decltype(ldw(src, 0, 1.0f));
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    if (ix >= nx) return;
    const int x = ix + off, y = iy + off;
    for (int iz = 0; iz < nz; ++iz) {
        const int z = iz + off;
        const int i = x + g.ld * y + g.pl * z;
        C v = inj<C>(ldw(src, i, scale), i, s0, a0, s1, a1);
        if (fx) v = damp3<C>(v, fy[y], fx[x], fz[z]);
        dst[iz + (size_t)nz * ix + (size_t)nz * nx * iy] = (float)v;
    }
}

// [nt][ny][nx] -> [ny][nx][nt] through a 64x64 LDS tile, with the mute
// (reference mute_directwave, iso:322-369).
__global__ void k_shot_transpose(const float* __restrict__ rec, int nt, int nx, int ny, MuteArgs m,
                                 float* __restrict__ shot) {
    __shared__ float t[64][65];
    const int it0 = blockIdx.x * 64, ix0 = blockIdx.y * 64, iy = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;   // 64 x 4
    for (int j = ty; j < 64; j += 4) {
        const int it = it0 + j, ix = ix0 + tx;
        t[j][tx] = (it < nt && ix < nx) ? rec[(size_t)it * nx * ny + (size_t)iy * nx + ix] : 0.0f;
    }
    __syncthreads();
    for (int j = ty; j < 64; j += 4) {
        const int ix = ix0 + j, it = it0 + tx;
        if (ix < nx && it < nt) {
            float v = t[tx][j];
            if (m.mute) {
                const float mx = m.dx * (float)abs(ix - m.sx);
                const float my = m.dy * (float)abs(iy - m.sy);
                const float mz = m.dz * (float)m.sz;
                const float t0 = sqrtf(mx * mx + my * my + mz * mz) / m.vcut;
                const int mu_t = (int)(2.0 / (double)(m.dt * m.fm));
                const int mu_nt = (int)(t0 / m.dt) + mu_t + m.tt;
                if (it < mu_nt) v = 0.0f;
            }
            shot[(size_t)iy * nx * nt + (size_t)ix * nt + it] = v;
        }
    }
}

// TTI rotation terms, once per model: the same (M)cos((double)angle) values the
// reference's kernels compute per point and per sweep (tti:101-103, :184-186).
template <class M>
__global__ void k_trig(M* __restrict__ th, M* __restrict__ st, M* __restrict__ ph, M* __restrict__ sp, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double t = (double)th[i], f = (double)ph[i];
        th[i] = (M)cos(t);
        st[i] = (M)sin(t);
        ph[i] = (M)cos(f);
        sp[i] = (M)sin(f);
    }
}

__global__ void k_pack_angles(const float* __restrict__ ct, const float* __restrict__ st,
                              const float* __restrict__ cp, const float* __restrict__ sp, f32x2* __restrict__ out,
                              size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = f32x2{(float)atan2((double)st[i], (double)ct[i]), (float)atan2((double)sp[i], (double)cp[i])};
}

__global__ void k_pack4(const float* __restrict__ vp, const float* __restrict__ rho, const float* __restrict__ eps,
                        const float* __restrict__ del, f32x4* __restrict__ out, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = f32x4{vp[i], rho[i], eps[i], del[i]};
}

__global__ void k_pack2(const float* __restrict__ vp, const float* __restrict__ rho, f32x2* __restrict__ out,
                        size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = f32x2{vp[i], rho[i]};
}

__global__ void k_fill(float* __restrict__ a, size_t n, float v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_check_finite(const float* __restrict__ a, size_t n, int* flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    bool bad = false;
    for (; i < n; i += stride) bad |= !isfinite(a[i]);
    if (bad) *flag = 1;
}

// Calibration streams for the HBM byte counters and the achievable-bandwidth
// ceiling: a float4 copy and a float4 read-only reduction, grid-stride.
__global__ void k_probe_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) b[i] = a[i];
}

// Ceiling streams: one float4 per thread in address order (copy) or four per thread,
// block-contiguous (read), with non-temporal loads and stores -- the forms that moved
// the most bytes per second among the copy and read variants measured on MI355X.
__global__ void k_probe_copy_flat(const f32x4* __restrict__ a, f32x4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

__global__ void k_probe_read_flat(const f32x4* __restrict__ a, size_t n, float* out) {
    const size_t base = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const size_t i = base + (size_t)k * blockDim.x;
        if (i < n) {
            const f32x4 v = __builtin_nontemporal_load(a + i);
            acc += v.x + v.y + v.z + v.w;
        }
    }
    if (acc == -1.0f) out[0] = acc;   // keeps the loads live; never true for the zeroed input
}

__global__ void k_probe_read(const float4* __restrict__ a, size_t n, float* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t s = (size_t)gridDim.x * blockDim.x;
    float acc = 0.0f;
    for (; i < n; i += s) {
        const float4 v = a[i];
        acc += v.x + v.y + v.z + v.w;
    }
    if (acc == -1.0f) out[0] = acc;   // keeps the loads live; never true for the zeroed input
}

// ---- launch dispatch -------------------------------------------------------
template <class PR, int OP, bool WZ, bool SM = false>
hipError_t vel_ty(int ty, const VelArgs& a, hipStream_t st) {
    const dim3 grid(a.l.grid);
    switch (ty) {
        case 2: hipLaunchKernelGGL((k_vel<PR, OP, 2, WZ, SM>), grid, dim3(TX, 2), 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_vel<PR, OP, 4, WZ, SM>), grid, dim3(TX, 4), 0, st, a); break;
        case 8: hipLaunchKernelGGL((k_vel<PR, OP, 8, WZ, SM>), grid, dim3(TX, 8), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class PR, int OP, bool WZ, int DIRS, bool SM = false>
hipError_t stress_ty(int ty, const StressArgs& a, hipStream_t st) {
    const dim3 grid(a.l.grid);
    switch (ty) {
        case 2: hipLaunchKernelGGL((k_stress<PR, OP, 2, WZ, DIRS, SM>), grid, dim3(TX, 2), 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_stress<PR, OP, 4, WZ, DIRS, SM>), grid, dim3(TX, 4), 0, st, a); break;
        case 8: hipLaunchKernelGGL((k_stress<PR, OP, 8, WZ, DIRS, SM>), grid, dim3(TX, 8), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class PR>
hipError_t step2_ty(int ty, const Step2Args& a, hipStream_t st) {
    const dim3 grid(a.l.grid);
    if (a.l.zlen <= 4) {   // short segments: every load of a segment at once
        switch (ty) {
            case 2: hipLaunchKernelGGL((k_step2_blk<PR, 2, 4>), grid, dim3(TX, 2), 0, st, a); break;
            case 4: hipLaunchKernelGGL((k_step2_blk<PR, 4, 4>), grid, dim3(TX, 4), 0, st, a); break;
            case 8: hipLaunchKernelGGL((k_step2_blk<PR, 8, 4>), grid, dim3(TX, 8), 0, st, a); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (ty) {
        case 2: hipLaunchKernelGGL((k_step2<PR, 2>), grid, dim3(TX, 2), 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_step2<PR, 4>), grid, dim3(TX, 4), 0, st, a); break;
        case 8: hipLaunchKernelGGL((k_step2<PR, 8>), grid, dim3(TX, 8), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class PR, int OP, bool WZ, bool SM = false>
hipError_t vel_batch_ty(int ty, const VelArgs* w, const int* base, int nwin, const Launch& lb, const Src* srcs, int it,
                        hipStream_t st) {
    const dim3 grid(lb.grid);
    switch (ty) {
        case 2: hipLaunchKernelGGL((k_vel_batch<PR, OP, 2, WZ, SM>), grid, dim3(TX, 2), 0, st, w, base, nwin, lb, srcs, it); break;
        case 4: hipLaunchKernelGGL((k_vel_batch<PR, OP, 4, WZ, SM>), grid, dim3(TX, 4), 0, st, w, base, nwin, lb, srcs, it); break;
        case 8: hipLaunchKernelGGL((k_vel_batch<PR, OP, 8, WZ, SM>), grid, dim3(TX, 8), 0, st, w, base, nwin, lb, srcs, it); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class PR, int OP, int DIRS, bool WZ, bool SM = false>
hipError_t stress_batch_ty(int ty, const StressArgs* w, const int* base, int nwin, const Launch& lb, int it,
                           const Src* srcs, hipStream_t st) {
    const dim3 grid(lb.grid);
    switch (ty) {
        case 2: hipLaunchKernelGGL((k_stress_batch<PR, OP, 2, DIRS, WZ, SM>), grid, dim3(TX, 2), 0, st, w, base, nwin, lb, it, srcs); break;
        case 4: hipLaunchKernelGGL((k_stress_batch<PR, OP, 4, DIRS, WZ, SM>), grid, dim3(TX, 4), 0, st, w, base, nwin, lb, it, srcs); break;
        case 8: hipLaunchKernelGGL((k_stress_batch<PR, OP, 8, DIRS, WZ, SM>), grid, dim3(TX, 8), 0, st, w, base, nwin, lb, it, srcs); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class PR>
hipError_t vel_batch_op(int op, int ty, const VelArgs* w, const int* base, int nwin, const Launch& lb, const Src* srcs,
                        int it, bool wz, hipStream_t st) {
    switch (op) {
        case 1: return wz ? vel_batch_ty<PR, 1, true>(ty, w, base, nwin, lb, srcs, it, st)
                          : vel_batch_ty<PR, 1, false>(ty, w, base, nwin, lb, srcs, it, st);
        case 2: return wz ? vel_batch_ty<PR, 2, true>(ty, w, base, nwin, lb, srcs, it, st)
                          : vel_batch_ty<PR, 2, false>(ty, w, base, nwin, lb, srcs, it, st);
        case 3: return wz ? hipErrorInvalidValue : vel_batch_ty<PR, 3, false>(ty, w, base, nwin, lb, srcs, it, st);
        case 4:
            if constexpr (sizeof(typename PR::M) == 4)
                return wz ? hipErrorInvalidValue : vel_batch_ty<PR, 4, false>(ty, w, base, nwin, lb, srcs, it, st);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template <class PR>
hipError_t stress_batch_op(int op, int ty, const StressArgs* w, const int* base, int nwin, const Launch& lb, int it,
                           int dirs, bool wz, const Src* srcs, hipStream_t st) {
    // dirs: 0 split everywhere / unsplit interior, 1 directional split, 2 one field per
    // damped direction (ISO has no q fields: 2 is 1 there)
    switch (op) {
        case 1:
            if (wz)
                return dirs ? stress_batch_ty<PR, 1, 1, true>(ty, w, base, nwin, lb, it, srcs, st)
                            : stress_batch_ty<PR, 1, 0, true>(ty, w, base, nwin, lb, it, srcs, st);
            return dirs ? stress_batch_ty<PR, 1, 1, false>(ty, w, base, nwin, lb, it, srcs, st)
                        : stress_batch_ty<PR, 1, 0, false>(ty, w, base, nwin, lb, it, srcs, st);
        case 2:
            if (wz)
                return dirs == 2 ? stress_batch_ty<PR, 2, 2, true>(ty, w, base, nwin, lb, it, srcs, st)
                     : dirs      ? stress_batch_ty<PR, 2, 1, true>(ty, w, base, nwin, lb, it, srcs, st)
                                 : stress_batch_ty<PR, 2, 0, true>(ty, w, base, nwin, lb, it, srcs, st);
            return dirs == 2 ? stress_batch_ty<PR, 2, 2, false>(ty, w, base, nwin, lb, it, srcs, st)
                 : dirs      ? stress_batch_ty<PR, 2, 1, false>(ty, w, base, nwin, lb, it, srcs, st)
                             : stress_batch_ty<PR, 2, 0, false>(ty, w, base, nwin, lb, it, srcs, st);
        case 3:
            if (wz) return hipErrorInvalidValue;
            return dirs == 2 ? stress_batch_ty<PR, 3, 2, false>(ty, w, base, nwin, lb, it, srcs, st)
                 : dirs      ? stress_batch_ty<PR, 3, 1, false>(ty, w, base, nwin, lb, it, srcs, st)
                             : stress_batch_ty<PR, 3, 0, false>(ty, w, base, nwin, lb, it, srcs, st);
        case 4:
            if constexpr (sizeof(typename PR::M) == 4) {
                if (wz) return hipErrorInvalidValue;
                return dirs == 2 ? stress_batch_ty<PR, 4, 2, false>(ty, w, base, nwin, lb, it, srcs, st)
                     : dirs      ? stress_batch_ty<PR, 4, 1, false>(ty, w, base, nwin, lb, it, srcs, st)
                                 : stress_batch_ty<PR, 4, 0, false>(ty, w, base, nwin, lb, it, srcs, st);
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template <class PR>
hipError_t vel_op(int op, int ty, const VelArgs& a, hipStream_t st) {
    switch (op) {
        case 1: return a.wz ? vel_ty<PR, 1, true>(ty, a, st) : vel_ty<PR, 1, false>(ty, a, st);
        case 2: return a.wz ? vel_ty<PR, 2, true>(ty, a, st) : vel_ty<PR, 2, false>(ty, a, st);
        case 3: return a.wz ? hipErrorInvalidValue : vel_ty<PR, 3, false>(ty, a, st);
        case 4:
            if constexpr (sizeof(typename PR::M) == 4) return a.wz ? hipErrorInvalidValue : vel_ty<PR, 4, false>(ty, a, st);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template <class PR>
hipError_t stress_op(int op, int ty, const StressArgs& a, hipStream_t st) {
    // unsplit: 0, 1 -> DIRS 0; 2 -> DIRS 1 (directional split); 3 -> DIRS 2 (one field per
    // damped direction; ISO has no q fields, so 3 is 2 there)
    switch (op) {
        case 1:
            if (a.unsplit >= 2) return a.wz ? stress_ty<PR, 1, true, 1>(ty, a, st) : stress_ty<PR, 1, false, 1>(ty, a, st);
            return a.wz ? stress_ty<PR, 1, true, 0>(ty, a, st) : stress_ty<PR, 1, false, 0>(ty, a, st);
        case 2:
            if (a.unsplit == 3) return a.wz ? stress_ty<PR, 2, true, 2>(ty, a, st) : stress_ty<PR, 2, false, 2>(ty, a, st);
            if (a.unsplit == 2) return a.wz ? stress_ty<PR, 2, true, 1>(ty, a, st) : stress_ty<PR, 2, false, 1>(ty, a, st);
            return a.wz ? stress_ty<PR, 2, true, 0>(ty, a, st) : stress_ty<PR, 2, false, 0>(ty, a, st);
        case 3:
            if (a.wz) return hipErrorInvalidValue;
            return a.unsplit == 3   ? stress_ty<PR, 3, false, 2>(ty, a, st)
                   : a.unsplit == 2 ? stress_ty<PR, 3, false, 1>(ty, a, st)
                                    : stress_ty<PR, 3, false, 0>(ty, a, st);
        case 4:
            if constexpr (sizeof(typename PR::M) == 4) {
                if (a.wz) return hipErrorInvalidValue;
                return a.unsplit == 3   ? stress_ty<PR, 4, false, 2>(ty, a, st)
                       : a.unsplit == 2 ? stress_ty<PR, 4, false, 1>(ty, a, st)
                                        : stress_ty<PR, 4, false, 0>(ty, a, st);
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template <class K>
int occ(K kern, int ty) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, TX * ty, 0) != hipSuccess) return 1;
    return nb;
}

template <class PR, int OP>
int occ_op(int kind, int ty) {
    if (kind == 0) {
        switch (ty) {
            case 2: return occ(k_vel<PR, OP, 2>, 2);
            case 4: return occ(k_vel<PR, OP, 4>, 4);
            default: return occ(k_vel<PR, OP, 8>, 8);
        }
    }
    switch (ty) {
        case 2: return occ(k_stress<PR, OP, 2>, 2);
        case 4: return occ(k_stress<PR, OP, 4>, 4);
        default: return occ(k_stress<PR, OP, 8>, 8);
    }
}

template <class PR>
int occ_prec(int kind, int op, int ty) {
    switch (op) {
        case 0:
            switch (ty) {
                case 2: return occ(k_step2<PR, 2>, 2);
                case 4: return occ(k_step2<PR, 4>, 4);
                default: return occ(k_step2<PR, 8>, 8);
            }
        case 1: return occ_op<PR, 1>(kind, ty);
        case 2: return occ_op<PR, 2>(kind, ty);
        case 4:
            if constexpr (sizeof(typename PR::M) == 4) return occ_op<PR, 4>(kind, ty);
            return occ_op<PR, 3>(kind, ty);
        default: return occ_op<PR, 3>(kind, ty);
    }
}

}  // namespace

hipError_t launch_vel(int prec, int op, int ty, const VelArgs& a, hipStream_t st) {
    if (a.mfac) {   // learned strip, coefficient head: VTI, FP32
        if (prec != PREC_FP32 || op != 2) return hipErrorInvalidValue;
        return a.wz ? vel_ty<Fp32, 2, true, true>(ty, a, st) : vel_ty<Fp32, 2, false, true>(ty, a, st);
    }
    switch (prec) {
        case PREC_FP32: return vel_op<Fp32>(op, ty, a, st);
        case PREC_FP16: return vel_op<Fp16>(op, ty, a, st);
        case PREC_FP64: return vel_op<Fp64>(op, ty, a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stress(int prec, int op, int ty, const StressArgs& a, hipStream_t st) {
    if (a.mfac) {   // learned strip, coefficient head: VTI, FP32
        if (prec != PREC_FP32 || op != 2) return hipErrorInvalidValue;
        if (a.wz)
            return a.unsplit == 3   ? stress_ty<Fp32, 2, true, 2, true>(ty, a, st)
                   : a.unsplit == 2 ? stress_ty<Fp32, 2, true, 1, true>(ty, a, st)
                                    : stress_ty<Fp32, 2, true, 0, true>(ty, a, st);
        return a.unsplit == 3   ? stress_ty<Fp32, 2, false, 2, true>(ty, a, st)
               : a.unsplit == 2 ? stress_ty<Fp32, 2, false, 1, true>(ty, a, st)
                                : stress_ty<Fp32, 2, false, 0, true>(ty, a, st);
    }
    switch (prec) {
        case PREC_FP32: return stress_op<Fp32>(op, ty, a, st);
        case PREC_FP16: return stress_op<Fp16>(op, ty, a, st);
        case PREC_FP64: return stress_op<Fp64>(op, ty, a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_vel_batch(int prec, int op, int ty, const VelArgs* wins, const int* base, int nwin, const Launch& lb,
                            const Src* srcs, int it, bool wz, hipStream_t st, bool sm) {
    if (sm) {
        if (prec != PREC_FP32 || op != 2 || wz) return hipErrorInvalidValue;
        return vel_batch_ty<Fp32, 2, false, true>(ty, wins, base, nwin, lb, srcs, it, st);
    }
    switch (prec) {
        case PREC_FP32: return vel_batch_op<Fp32>(op, ty, wins, base, nwin, lb, srcs, it, wz, st);
        case PREC_FP16: return vel_batch_op<Fp16>(op, ty, wins, base, nwin, lb, srcs, it, wz, st);
        case PREC_FP64: return vel_batch_op<Fp64>(op, ty, wins, base, nwin, lb, srcs, it, wz, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stress_batch(int prec, int op, int ty, const StressArgs* wins, const int* base, int nwin,
                               const Launch& lb, int it, int dirs, bool wz, const Src* srcs, hipStream_t st, bool sm) {
    if (sm) {
        if (prec != PREC_FP32 || op != 2 || wz) return hipErrorInvalidValue;
        return dirs == 2 ? stress_batch_ty<Fp32, 2, 2, false, true>(ty, wins, base, nwin, lb, it, srcs, st)
             : dirs      ? stress_batch_ty<Fp32, 2, 1, false, true>(ty, wins, base, nwin, lb, it, srcs, st)
                         : stress_batch_ty<Fp32, 2, 0, false, true>(ty, wins, base, nwin, lb, it, srcs, st);
    }
    switch (prec) {
        case PREC_FP32: return stress_batch_op<Fp32>(op, ty, wins, base, nwin, lb, it, dirs, wz, srcs, st);
        case PREC_FP16: return stress_batch_op<Fp16>(op, ty, wins, base, nwin, lb, it, dirs, wz, srcs, st);
        case PREC_FP64: return stress_batch_op<Fp64>(op, ty, wins, base, nwin, lb, it, dirs, wz, srcs, st);
        default: return hipErrorInvalidValue;
    }
}

template <class PR>
hipError_t step2_shots_ty(int ty, const Step2Args& a, const Step2Shots& s, hipStream_t st) {
    const dim3 grid(a.l.grid);
    if (a.l.zlen <= 4) {   // short segments: every load of a segment at once
        switch (ty) {
            case 2: hipLaunchKernelGGL((k_step2_shots<PR, 2, true>), grid, dim3(TX, 2), 0, st, a, s); break;
            case 4: hipLaunchKernelGGL((k_step2_shots<PR, 4, true>), grid, dim3(TX, 4), 0, st, a, s); break;
            case 8: hipLaunchKernelGGL((k_step2_shots<PR, 8, true>), grid, dim3(TX, 8), 0, st, a, s); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (ty) {
        case 2: hipLaunchKernelGGL((k_step2_shots<PR, 2, false>), grid, dim3(TX, 2), 0, st, a, s); break;
        case 4: hipLaunchKernelGGL((k_step2_shots<PR, 4, false>), grid, dim3(TX, 4), 0, st, a, s); break;
        case 8: hipLaunchKernelGGL((k_step2_shots<PR, 8, false>), grid, dim3(TX, 8), 0, st, a, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step2_shots(int prec, int ty, const Step2Args& a, const Step2Shots& s, hipStream_t st) {
    if (s.n < 1 || s.n > MAX_SHOTS) return hipErrorInvalidValue;
    switch (prec) {
        case PREC_FP32: return step2_shots_ty<Fp32>(ty, a, s, st);
        case PREC_FP16: return step2_shots_ty<Fp16>(ty, a, s, st);
        case PREC_FP64: return step2_shots_ty<Fp64>(ty, a, s, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step2(int prec, int ty, const Step2Args& a, hipStream_t st) {
    switch (prec) {
        case PREC_FP32: return step2_ty<Fp32>(ty, a, st);
        case PREC_FP16: return step2_ty<Fp16>(ty, a, st);
        case PREC_FP64: return step2_ty<Fp64>(ty, a, st);
        default: return hipErrorInvalidValue;
    }
}

namespace {
// The persistent kernel instance of (op, ty, dirs, wz), FP32 storage; nullptr if none.
typedef void (*PersistKern)(const VelArgs*, const StressArgs*, const int*, const Src*, unsigned*, int*, int, int,
                            unsigned, long long, int, int);
template <int TY>
PersistKern persist_kern_ty(int op, int dirs, bool wz) {
    switch (op) {
        case 1:
            if (wz) return dirs ? k_persist<Fp32, 1, TY, 1, true> : k_persist<Fp32, 1, TY, 0, true>;
            return dirs ? k_persist<Fp32, 1, TY, 1, false> : k_persist<Fp32, 1, TY, 0, false>;
        case 2:
            if (wz)
                return dirs == 2 ? k_persist<Fp32, 2, TY, 2, true>
                       : dirs    ? k_persist<Fp32, 2, TY, 1, true>
                                 : k_persist<Fp32, 2, TY, 0, true>;
            return dirs == 2 ? k_persist<Fp32, 2, TY, 2, false>
                   : dirs    ? k_persist<Fp32, 2, TY, 1, false>
                             : k_persist<Fp32, 2, TY, 0, false>;
        case 3:
            if (wz) return nullptr;
            return dirs == 2 ? k_persist<Fp32, 3, TY, 2, false>
                   : dirs    ? k_persist<Fp32, 3, TY, 1, false>
                             : k_persist<Fp32, 3, TY, 0, false>;
        case 4:
            if (wz) return nullptr;
            return dirs == 2 ? k_persist<Fp32, 4, TY, 2, false>
                   : dirs    ? k_persist<Fp32, 4, TY, 1, false>
                             : k_persist<Fp32, 4, TY, 0, false>;
        default: return nullptr;
    }
}
PersistKern persist_kern(int op, int ty, int dirs, bool wz) {
    switch (ty) {
        case 4: return persist_kern_ty<4>(op, dirs, wz);
        case 8: return persist_kern_ty<8>(op, dirs, wz);
        default: return nullptr;
    }
}
}  // namespace

int persist_blocks_per_cu(int op, int ty, int dirs, bool wz) {
    const PersistKern k = persist_kern(op, ty, dirs, wz);
    if (!k) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, TX * ty, 0) != hipSuccess) return 0;
    return nb;
}

hipError_t launch_persist(int op, int ty, int dirs, bool wz, const PersistArgs& p, int it0, int it1, hipStream_t st) {
    const PersistKern k = persist_kern(op, ty, dirs, wz);
    if (!k || p.nblocks == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(p.nblocks), dim3(TX, ty), 0, st, p.va, p.sa, p.tabs, p.srcs, p.sync, p.err, p.nwin,
                       p.group, p.nblocks, p.wait_ticks, it0, it1);
    return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned* a, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

int max_blocks_per_cu(int prec, int kind, int op, int ty) {
    switch (prec) {
        case PREC_FP16: return occ_prec<Fp16>(kind, op, ty);
        case PREC_FP64: return occ_prec<Fp64>(kind, op, ty);
        default: return occ_prec<Fp32>(kind, op, ty);
    }
}

hipError_t launch_pad(const float* src, int nx, int ny, int nz, int off, const Geo& g, void* dst, int model_double,
                      int mode, float dt, hipStream_t st, const Window* win) {
    const dim3 grid((g.ld + 255) / 256, g.nny, g.nnz);
    const Window w = win ? *win : Window{0, 0, 0, nx, ny, nz};
    if (model_double)
        hipLaunchKernelGGL(k_pad<double>, grid, dim3(256), 0, st, src, nx, ny, nz, off, g, (double*)dst, mode, dt, w);
    else
        hipLaunchKernelGGL(k_pad<float>, grid, dim3(256), 0, st, src, nx, ny, nz, off, g, (float*)dst, mode, dt, w);
    return hipGetLastError();
}

hipError_t launch_snap(int prec, const void* src, const Geo& g, int off, int nx, int ny, int nz, int s0, double a0,
                       int s1, double a1, const double* fx, const double* fy, const double* fz, float scale,
                       float* dst, hipStream_t st) {
    const dim3 grid((nx + 63) / 64, ny);
    switch (prec) {
        case PREC_FP16:
            hipLaunchKernelGGL(k_snap<__half>, grid, dim3(64), 0, st, (const __half*)src, g, off, nx, ny, nz, s0, a0,
                               s1, a1, fx, fy, fz, scale, dst);
            break;
        case PREC_FP64:
            hipLaunchKernelGGL(k_snap<double>, grid, dim3(64), 0, st, (const double*)src, g, off, nx, ny, nz, s0, a0,
                               s1, a1, fx, fy, fz, scale, dst);
            break;
        default:
            hipLaunchKernelGGL(k_snap<float>, grid, dim3(64), 0, st, (const float*)src, g, off, nx, ny, nz, s0, a0,
                               s1, a1, fx, fy, fz, scale, dst);
    }
    return hipGetLastError();
}

hipError_t launch_shot_transpose(const float* rec, int nt, int nx, int ny, const MuteArgs& m, float* shot,
                                 hipStream_t st) {
    const dim3 grid((nt + 63) / 64, (nx + 63) / 64, ny);
    hipLaunchKernelGGL(k_shot_transpose, grid, dim3(64, 4), 0, st, rec, nt, nx, ny, m, shot);
    return hipGetLastError();
}

hipError_t launch_probe(int read_only, const float* a, float* b, size_t bytes, int grid, hipStream_t st) {
    const size_t n4 = bytes / 16;
    if (grid == 0) {   // the ceiling forms
        if (read_only)
            hipLaunchKernelGGL(k_probe_read_flat, dim3((unsigned)((n4 + 4095) / 4096)), dim3(1024), 0, st,
                               (const f32x4*)a, n4, b);
        else
            hipLaunchKernelGGL(k_probe_copy_flat, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st,
                               (const f32x4*)a, (f32x4*)b, n4);
        return hipGetLastError();
    }
    if (read_only)
        hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(256), 0, st, (const float4*)a, n4, b);
    else
        hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(256), 0, st, (const float4*)a, (float4*)b, n4);
    return hipGetLastError();
}

hipError_t launch_trig(void* th, void* st, void* ph, void* sp, size_t n, int model_double, hipStream_t s) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (model_double)
        k_trig<double><<<grid, 256, 0, s>>>((double*)th, (double*)st, (double*)ph, (double*)sp, n);
    else
        k_trig<float><<<grid, 256, 0, s>>>((float*)th, (float*)st, (float*)ph, (float*)sp, n);
    return hipGetLastError();
}

hipError_t launch_pack_angles(const float* ct, const float* st, const float* cp, const float* sp, void* out, size_t n,
                              hipStream_t s) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pack_angles, dim3(grid), dim3(256), 0, s, ct, st, cp, sp, (f32x2*)out, n);
    return hipGetLastError();
}

hipError_t launch_pack4(const float* vp, const float* rho, const float* eps, const float* del, void* out, size_t n,
                        hipStream_t s) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pack4, dim3(grid), dim3(256), 0, s, vp, rho, eps, del, (f32x4*)out, n);
    return hipGetLastError();
}

hipError_t launch_pack2(const float* vp, const float* rho, void* out, size_t n, hipStream_t s) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pack2, dim3(grid), dim3(256), 0, s, vp, rho, (f32x2*)out, n);
    return hipGetLastError();
}

hipError_t launch_fill(float* a, size_t n, float v, hipStream_t st) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, st, a, n, v);
    return hipGetLastError();
}

hipError_t launch_check_finite(const float* a, size_t n, int* flag, hipStream_t st) {
    hipLaunchKernelGGL(k_check_finite, dim3(1024), dim3(256), 0, st, a, n, flag);
    return hipGetLastError();
}

}  // namespace a3d
