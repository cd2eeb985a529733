/*
 * oracle.cpp -- TEST INFRASTRUCTURE, not product code.
 *
 * A CPU restatement of the reference's forward-modelling arithmetic
 * (sh39o/Acoustic3d-fd), one function per reference kernel, each citing the
 * reference file:line it follows. The layout is the reference's own padded,
 * z-fastest volume (id = iz + ix*nnz + iy*nnz*nnx), so an indexing or
 * layout mistake in the GPU path (which uses x-fastest, z-slowest storage)
 * cannot cancel against the same mistake here.
 *
 * Instances: T = float is the parity oracle; every expression keeps the
 * reference's float/double mix (C++ promotion of float and double operands
 * is the same on the host as in CUDA device code). T = double is the FP64
 * error reference: every value and table in double.
 *
 * Where the reference calls a CUDA device math function on the per-shot or
 * per-step path (exp, expf, cos, log, pow, erf), this file calls glibc; the
 * CUDA Toolkit's results for those calls are not reproduced (parity with
 * CUDA's math library at the last ulp is unpinned). pow(x, 2) is restated
 * as x*x (the correctly rounded square). The per-point float cos/sin of the
 * TTI rotation is restated as the correctly rounded float value,
 * (float)cos((double)x).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * use this file, and only as the checker.
 */
#include "oracle.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int mm = 4;                     /* half stencil length, all files */
constexpr double kPi = 3.141592653;       /* Toa_3dfd_cuda.cu:14, iso:28, vti:31, tti:8 */

/* Staggered 8th-order coefficients: iso:37, vti:40, tti:17 (float constants). */
template <class T> struct Coef;
template <> struct Coef<float> {
    static constexpr float c[4] = {1.196289f, -0.0797526f, 0.009570313f, -0.0006975447f};
    /* Toa_3dfd_cuda.cu:32: double literals rounded to float */
    static constexpr float s[5] = {(float)(-205.0 / 72.0), (float)(8.0 / 5.0), (float)(-1.0 / 5.0),
                                   (float)(8.0 / 315.0), (float)(-1.0 / 560.0)};
};
template <> struct Coef<double> {
    static constexpr double c[4] = {1.196289, -0.0797526, 0.009570313, -0.0006975447};
    static constexpr double s[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
};
constexpr float Coef<float>::c[4];
constexpr float Coef<float>::s[5];
constexpr double Coef<double>::c[4];
constexpr double Coef<double>::s[5];

/* Forward staggered difference (update_vel: iso:125-127, vti:113-115, tti:94-99):
 * r += c[m] * (a[id + (m+1)s] - a[id - m s]), m = 0..3 in order. */
template <class T>
inline T dfwd(const T* a, long id, long s) {
    T r = 0;
    for (int m = 0; m < mm; ++m) r += Coef<T>::c[m] * (a[id + (m + 1) * s] - a[id - m * s]);
    return r;
}
/* Backward staggered difference (update_stress: iso:178-180, vti:183-185, tti:174-182). */
template <class T>
inline T dbwd(const T* a, long id, long s) {
    T r = 0;
    for (int m = 0; m < mm; ++m) r += Coef<T>::c[m] * (a[id + m * s] - a[id - (m + 1) * s]);
    return r;
}

/* 8th-order second derivative along one axis, accumulated as in
 * cuda_step_fd3d (Toa_3dfd_cuda.cu:281-289): c = s0*cur, then c += s[i]*(a[-i]+a[+i]). */
template <class T>
inline T lap1(const T* a, long id, long s) {
    T c = Coef<T>::s[0] * a[id];
    for (int i = 1; i <= mm; ++i) c += Coef<T>::s[i] * (a[id - i * s] + a[id + i * s]);
    return c;
}

template <class T> inline T fsqrt(T x) { return std::sqrt(x); }   /* sqrtf for float */
template <class T> inline T fcos(T x) { return (T)std::cos((double)x); }
template <class T> inline T fsin(T x) { return (T)std::sin((double)x); }

/* pad_vv (iso:198-238, vti:211-251, tti:209-249): edge replication in three
 * passes, x then y then z, over a padded z-fastest volume whose interior was
 * filled by read_file (iso:241-260). */
template <class T>
void pad_vv(int nnx, int nny, int nnz, int npml, T* ee) {
    const long pl = (long)nnz * nnx;
    for (int iy = 0; iy < nny; iy++)
        for (int ix = 0; ix < nnx; ix++)
            for (int iz = 0; iz < nnz; iz++) {
                long id = iz + (long)ix * nnz + iy * pl;
                if (ix < npml) ee[id] = ee[iz + (long)npml * nnz + iy * pl];
                else if (ix >= nnx - npml) ee[id] = ee[iz + (long)(nnx - npml - 1) * nnz + iy * pl];
            }
    for (int iy = 0; iy < nny; iy++)
        for (int ix = 0; ix < nnx; ix++)
            for (int iz = 0; iz < nnz; iz++) {
                long id = iz + (long)ix * nnz + iy * pl;
                if (iy < npml) ee[id] = ee[iz + (long)ix * nnz + npml * pl];
                else if (iy >= nny - npml) ee[id] = ee[iz + (long)ix * nnz + (nny - npml - 1) * pl];
            }
    for (int iy = 0; iy < nny; iy++)
        for (int ix = 0; ix < nnx; ix++)
            for (int iz = 0; iz < nnz; iz++) {
                long id = iz + (long)ix * nnz + iy * pl;
                if (iz < npml) ee[id] = ee[npml + (long)ix * nnz + iy * pl];
                else if (iz >= nnz - npml) ee[id] = ee[nnz - npml - 1 + (long)ix * nnz + iy * pl];
            }
}

template <class T>
void read_interior(const float* src, int nx, int ny, int nz, int nnx, int nnz, int off, T* dst) {
    /* read_file (iso:249-257): interior at offset npml in every axis, file order z fastest */
    const long pl = (long)nnz * nnx;
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++)
            for (int iz = 0; iz < nz; iz++)
                dst[(iz + off) + (long)(ix + off) * nnz + (iy + off) * pl] =
                    (T)src[iz + (long)nz * ix + (long)nz * nx * iy];
}

/* initial_coffe (iso:263-291; identical vti:282-310, tti:282-310): per-axis
 * split-PML profiles. dt*d0 is a float product in the reference, then the
 * expression is evaluated in double and stored to float. */
template <class T>
void initial_coffe(T dt, T d0, int nn, int npml, T* coff1, T* coff2, T* acoff1, T* acoff2, int width) {
    const T dtd0 = dt * d0;
    for (int id = 0; id < nn + 2 * npml; ++id) {
        if (id < npml) {
            double r = (npml - 0.5 - id) / width;
            double ra = ((npml - id) * 1.0) / width;
            coff1[id] = (T)(1.0 / (1.0 + (dtd0 * (r * r)) / 2.0));
            coff2[id] = (T)(coff1[id] * (1.0 - (dtd0 * (r * r)) / 2.0));
            acoff1[id] = (T)(1.0 / (1.0 + (dtd0 * (ra * ra)) / 2.0));
            acoff2[id] = (T)(acoff1[id] * (1.0 - (dtd0 * (ra * ra)) / 2.0));
        } else if (id < npml + nn) {
            coff1[id] = coff2[id] = acoff1[id] = acoff2[id] = (T)1.0;
        } else {
            double r = (0.5 + id - nn - npml) / width;
            double ra = ((id - nn - npml) * 1.0) / width;
            coff1[id] = (T)(1.0 / (1.0 + (dtd0 * (r * r)) / 2.0));
            coff2[id] = (T)(coff1[id] * (1.0 - (dtd0 * (r * r)) / 2.0));
            acoff1[id] = (T)(1.0 / (1.0 + (dtd0 * (ra * ra)) / 2.0));
            acoff2[id] = (T)(acoff1[id] * (1.0 - (dtd0 * (ra * ra)) / 2.0));
        }
    }
}

/* get_d0 with amplitude amp in place of 10 (the learned strip's profile). */
template <class T>
T get_d0_amp(const T* vp, int nnx, int nny, int nnz, int width, float dx, float dy, float dz, double amp) {
    const T dsum = (T)dx + (T)dy + (T)dz;
    return (T)(amp * vp[(long)nny * nnx * nnz / 2] * std::log(100000.0) / (2.0 * width * (dsum / 3.0)));
}

/* get_d0 (iso:193-195): uses the padded z-fastest velocity at linear index N/2. */
template <class T>
T get_d0(const T* vp, int nnx, int nny, int nnz, int npml, float dx, float dy, float dz) {
    const T dsum = (T)dx + (T)dy + (T)dz;   /* float adds in the reference */
    return (T)(10.0 * vp[(long)nny * nnx * nnz / 2] * std::log(100000.0) /
               (2.0 * npml * (dsum / 3.0)));
}

/* add_source wavelet (iso:52-96 types 1-6; vti:54-88 and tti:31-65 types 1-3).
 * t is the float-accumulated time. Returns the value added (before pfac). */
template <class T>
T wavelet(int wtype, T t, T favg) {
    T x_, xx_, tdelay, ts, source = 0;
    tdelay = (T)(1.0 / favg);
    ts = t - tdelay;
    if (wtype == 1) {
        x_ = favg * ts;
        xx_ = x_ * x_;
        source = (T)((1 - 2 * kPi * kPi * (xx_)) * std::exp(-(kPi * kPi * xx_)));
    } else if (wtype == 2) {
        x_ = (T)((-4) * favg * favg * kPi * kPi / std::log(0.1));
        source = (T)((-2) * kPi * kPi * ts * std::exp(-x_ * ts * ts));
    } else if (wtype == 3) {
        x_ = (T)((-1) * favg * favg * kPi * kPi / std::log(0.1));
        source = (T)std::exp(-x_ * ts * ts);
    } else if (wtype == 4) {
        x_ = (T)(std::exp(-favg * favg * kPi * kPi * ts * ts) * ts);
        source = -x_;
    } else if (wtype == 5) {
        x_ = (T)(-1.0 / (2.0 * favg * favg * kPi * kPi));
        x_ = (T)(x_ * std::exp(-favg * favg * kPi * kPi * ts * ts));
        source = x_;
    } else if (wtype == 6) {
        x_ = (T)(-std::erf(favg * kPi * ts) /
                 (4 * favg * favg * favg * kPi * kPi * fsqrt<T>((T)kPi)));
        source = x_;
    }
    return source;
}

/* cuda_ricker_wavelet (Toa_3dfd_cuda.cu:34-43). */
template <class T>
void ricker_table(T fm, T dt, int nt, T* wlt) {
    for (int it = 0; it < nt; ++it) {
        T tmp;
        if (sizeof(T) == 4)
            tmp = (T)(kPi * fm * std::fabs((float)(it * dt - 1.0 / fm)));
        else
            tmp = (T)(kPi * fm * std::fabs(it * dt - 1.0 / fm));
        tmp *= tmp;
        wlt[it] = (T)((1.0 - 2.0 * tmp) * std::exp(-tmp));
    }
}

/* Sponge factor of cuda_absorb_bndr (Toa_3dfd_cuda.cu:138-151) for index i of
 * an axis of interior length n: qp*r^2 + 1 in double, or 1 (no damping). */
inline double sponge_factor(int i, int n, int npd, bool* damped) {
    const double qp = (double)(-0.25f);
    *damped = true;
    if (i < npd) {
        double r = (npd - i) / (1.0 * npd);
        return qp * (r * r) + 1;
    } else if (i >= 2 * mm + npd + n) {
        double r = (i - 2 * mm - npd - n) / (1.0 * npd);
        return qp * (r * r) + 1;
    }
    *damped = false;
    return 1.0;
}

/* SV-suppression taper factor (vti:157-175, tti:142-160) for r^2 in (16, 225]. */
inline double taper_factor(int r2) {
    const int R = 15, r = 4;
    return 0.5 * (1 - std::cos(kPi * ((std::sqrt((float)r2) - r) * 4.0 / (R * 3.0 - 1))));
}

template <class T>
void copy_out(T* dst, const T* src, size_t n) { std::memcpy(dst, src, n * sizeof(T)); }

/* ------------------------------------------------------------------------ */
/* 2nd-order operator: cuda_3dfd (Toa_3dfd_cuda.cu:366-508)                 */
/* ------------------------------------------------------------------------ */
template <class T>
int run_2nd(const OracleShot& p, const float* vel, T* shot, T* snaps, int nsnap_max,
            double* secs, int nsteps_override) {
    const int npd = p.nb, nx = p.nx, ny = p.ny, nz = p.nz;
    const int nnz = nz + 2 * mm + 2 * npd, nnx = nx + 2 * mm + 2 * npd, nny = ny + 2 * mm + 2 * npd;
    const long pl = (long)nnz * nnx;
    const size_t N = (size_t)pl * nny;
    const int off = mm + npd;
    const int nt = p.nt;
    const int nsteps = nsteps_override > 0 ? nsteps_override : nt;

    /* velocity_transform (Toa_3dfd_cuda.cu:300-347): vv = (v*dt)^2, edge replication
     * z (top/down), then x (left/right), then y (front/back). */
    std::vector<T> vv(N, (T)0);
    for (int i3 = 0; i3 < ny; i3++)
        for (int i2 = 0; i2 < nx; i2++)
            for (int i1 = 0; i1 < nz; i1++) {
                T tmp = (T)vel[i1 + (long)nz * i2 + (long)nz * nx * i3] * (T)p.dt;
                vv[(i1 + off) + (long)nnz * (i2 + off) + pl * (i3 + off)] = tmp * tmp;
            }
    for (int i3 = 0; i3 < nny; i3++)
        for (int i2 = 0; i2 < nnx; i2++)
            for (int i1 = 0; i1 < off; i1++) {
                vv[i1 + (long)nnz * i2 + pl * i3] = vv[off + (long)nnz * i2 + pl * i3];
                vv[(nnz - i1 - 1) + (long)nnz * i2 + pl * i3] = vv[(nnz - off - 1) + (long)nnz * i2 + pl * i3];
            }
    for (int i3 = 0; i3 < nny; i3++)
        for (int i2 = 0; i2 < off; i2++)
            for (int i1 = 0; i1 < nnz; i1++) {
                vv[i1 + (long)nnz * i2 + pl * i3] = vv[i1 + (long)nnz * off + pl * i3];
                vv[i1 + (long)nnz * (nnx - i2 - 1) + pl * i3] = vv[i1 + (long)nnz * (nnx - off - 1) + pl * i3];
            }
    for (int i3 = 0; i3 < off; i3++)
        for (int i2 = 0; i2 < nnx; i2++)
            for (int i1 = 0; i1 < nnz; i1++) {
                vv[i1 + (long)nnz * i2 + pl * i3] = vv[i1 + (long)nnz * i2 + pl * off];
                vv[i1 + (long)nnz * i2 + pl * (nny - 1 - i3)] = vv[i1 + (long)nnz * i2 + pl * (nny - off - 1)];
            }

    /* Toa_3dfd_cuda.cu:387-389: 1.0/(d*d) with a float product, stored to float */
    const T dz2 = (T)(1.0 / ((T)p.dz * (T)p.dz));
    const T dx2 = (T)(1.0 / ((T)p.dx * (T)p.dx));
    const T dy2 = (T)(1.0 / ((T)p.dy * (T)p.dy));

    std::vector<T> wlt(nt > nsteps ? nt : nsteps);
    ricker_table<T>((T)p.fm, (T)p.dt, (int)wlt.size(), wlt.data());

    /* cuda_set_s (Toa_3dfd_cuda.cu:45-52): 0-based interior source, dipole partner at z+2 */
    const long s = (p.sz + off) + (long)nnz * (p.sx + off) + pl * (p.sy + off);

    /* per-axis sponge factors */
    std::vector<double> fz(nnz), fx(nnx), fy(nny);
    std::vector<char> dz_(nnz), dx_(nnx), dy_(nny);
    for (int i = 0; i < nnz; ++i) { bool d; fz[i] = sponge_factor(i, nz, npd, &d); dz_[i] = d; }
    for (int i = 0; i < nnx; ++i) { bool d; fx[i] = sponge_factor(i, nx, npd, &d); dx_[i] = d; }
    for (int i = 0; i < nny; ++i) { bool d; fy[i] = sponge_factor(i, ny, npd, &d); dy_[i] = d; }

    std::vector<T> pa(N, (T)0), pb(N, (T)0);
    T* p0 = pa.data();
    T* p1 = pb.data();
    int nsnap = 0;
    const int zr = off + p.rec_z;

    auto absorb = [&](T* d) {
        /* cuda_absorb_bndr (Toa_3dfd_cuda.cu:123-154), applied per axis in the order y, x, z;
         * each multiplication is rounded to the stored type. Bounds-checked (no aliasing). */
#pragma omp parallel for schedule(static)
        for (int iy = 0; iy < nny; iy++)
            for (int ix = 0; ix < nnx; ix++)
                for (int iz = 0; iz < nnz; iz++) {
                    long id = iz + (long)ix * nnz + iy * pl;
                    if (dy_[iy]) d[id] = (T)(fy[iy] * d[id]);
                    if (dx_[ix]) d[id] = (T)(fx[ix] * d[id]);
                    if (dz_[iz]) d[id] = (T)(fz[iz] * d[id]);
                }
    };

    auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < nsteps; ++it) {
        /* cuda_add_dipole (Toa_3dfd_cuda.cu:177-187): persistent add into p1 */
        p1[s] += wlt[it];
        p1[s + 2] -= wlt[it];
        /* cuda_step_fd3d (Toa_3dfd_cuda.cu:189-298), in place on p0 */
#pragma omp parallel for schedule(static)
        for (int iy = mm; iy < nny - mm; iy++)
            for (int ix = mm; ix < nnx - mm; ix++)
                for (int iz = mm; iz < nnz - mm; iz++) {
                    long id = iz + (long)ix * nnz + iy * pl;
                    T c1 = lap1<T>(p1, id, 1);
                    T c2 = lap1<T>(p1, id, nnz);
                    T c3 = lap1<T>(p1, id, pl);
                    c1 *= dz2;
                    c2 *= dx2;
                    c3 *= dy2;
                    /* Toa_3dfd_cuda.cu:295: 2.0*p1 - p0 in double, plus the float product */
                    T lap = vv[id] * (c1 + c2 + c3);
                    p0[id] = (T)(2.0 * (double)p1[id] - (double)p0[id] + (double)lap);
                }
        T* tmp = p0; p0 = p1; p1 = tmp;               /* :445-447 */
        absorb(p0);                                   /* :449 */
        absorb(p1);                                   /* :450 */
        if (shot) {                                   /* cuda_record + trans_xy2txy (:452-453) */
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix)
                    shot[it + (long)nt * ix + (long)nt * nx * iy] =
                        p0[zr + (long)nnz * (ix + off) + pl * (iy + off)];
        }
        if (p.snap_interval > 0 && it % p.snap_interval == 0 && !(p.snap_skip0 && it == 0)) {
            if (snaps && nsnap < nsnap_max) {         /* window3d (:350-363) */
                T* o = snaps + (size_t)nsnap * nz * nx * ny;
                for (int iy = 0; iy < ny; ++iy)
                    for (int ix = 0; ix < nx; ++ix)
                        for (int iz = 0; iz < nz; ++iz)
                            o[iz + (long)nz * ix + (long)nz * nx * iy] =
                                p0[(iz + off) + (long)nnz * (ix + off) + pl * (iy + off)];
            }
            ++nsnap;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();

    if (shot && p.mute) {
        /* mute_directwave (Toa_3dfd_cuda.cu:73-121) is never called by the reference's
         * 2nd-order entry; applied here only when asked, with v = v0[1] (:403). */
        const T v = p.v_mute > 0 ? (T)p.v_mute : (T)vel[1];
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                T mx = (T)p.dx * (T)std::abs(ix - p.sx);
                T my = (T)p.dy * (T)std::abs(iy - p.sy);
                T mz = (T)p.dz * (T)p.sz;
                T t0m = fsqrt<T>(mx * mx + my * my + mz * mz) / v;
                int mu_t = (int)(2.0 / ((T)p.dt * (T)p.fm));
                int mu_nt = (int)(t0m / (T)p.dt) + mu_t + p.mute_tt;
                for (int it = 0; it < nt && it < mu_nt; ++it) shot[it + (long)nt * ix + (long)nt * nx * iy] = 0;
            }
    }
    return nsnap;
}

/* ------------------------------------------------------------------------ */
/* 1st-order family: cuda_3dfd_1order (iso:372-671), cuda_3dfd_vti          */
/* (vti:389-688), cuda_3dfd_tti (tti:390-676)                               */
/* ------------------------------------------------------------------------ */
/* Learned absorbing strip (no counterpart in the reference; restated from the
 * strip's definition, FP32 only). Grid padding nb = 10: the 6 outer updated
 * cells of every face are strip cells, each owned by its nearest face (ties:
 * z, then y, then x). After every stress sweep each strip cell is damped:
 *   g = 0.25 * sigmoid(net(f)),  f = local state / s,  s = max(1e-30, |local state|),
 *   P, px, py, pz, u, v *= 1 - g_0;  Q, qx, qy, qz, w *= 1 - g_1.
 * Local state: P at the centre, one cell inward, one outward and the two
 * tangential pairs of the owning face's axis; Q at the centre, inward, outward;
 * the outward normal velocity (mean of the two staggered samples around the
 * cell) times rho*vp. Then vp*dt/d_axis, epsilon, delta, depth/6 (1 =
 * outermost), 0. Network: 16 -> 64 -> 64 -> 2, tanh (as 1 - 2/(exp(2a) + 1)), parameters W1[64][16] b1
 * W2[64][64] b2 W3[2][64] b3, sums in input order with fused multiply-adds. */
constexpr int kStripCells = 6, kNin = 16, kNh = 64, kNout = 2;
constexpr float kGmax = 0.25f;
constexpr int kW1 = 0, kB1 = kW1 + kNh * kNin, kW2 = kB1 + kNh, kB2 = kW2 + kNh * kNh, kW3 = kB2 + kNh,
              kB3 = kW3 + kNout * kNh, kNpar = kB3 + kNout;
std::vector<float> g_strip;   /* empty: plain PML */
double g_strip_amp = 10.0;    /* PML amplitude of the strip's 6-cell ramp */
bool g_strip_half = true;     /* network operands rounded to binary16 (matrix-core evaluation) */
int g_strip_head = 0;         /* 0: damping of every field; 1: normal-axis split components and velocity;
                                 2: multiplier of the PML damping inside the update (coefficient head) */
int g_strip_period = 1;       /* head 2: the network runs after the stress update of every period-th step
                                 (0: after the first step's only) */
constexpr float kMmax = 3.0f; /* head 2: m = kMmax * sigmoid(y_0) */

/* Planted mistake in the TTI rotation, for tests that must show they catch one (0: none).
 * 1..8 flip the sign of velocity-sweep term k-1, 9..16 of stress-sweep term k-9, in the
 * order ct*cp*d_x, ct*sp*d_y, st*d_z (xx), sp*d_x, cp*d_y (yy), st*cp*d_x, st*sp*d_y,
 * ct*d_z (zz); 17: the velocity sweep's zz takes P's x derivative in place of Q's;
 * 18: the stress sweep's xx takes v's x derivative in place of u's. A sign of +1 is an
 * exact multiplication, so fault 0 leaves every result bit for bit unchanged. */
int g_tti_fault = 0;

/* Nearest binary16 value (ties to even), as a float; finite inputs below 65504 in magnitude. */
float round_half(float f) {
    const float a = std::fabs(f);
    if (a < 6.103515625e-05f) return std::nearbyint(f * 16777216.0f) / 16777216.0f;   /* subnormal step 2^-24 */
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x0FFFu + ((u >> 13) & 1u);
    u &= 0xFFFFE000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

/* tanh(a) = 1 - 2 / (exp(2a) + 1), the network's activation */
inline float act_tanh(float a) { return 1.0f - 2.0f / (std::exp(2.0f * a) + 1.0f); }

void strip_net(const float* p, const float* x, float* y) {
    float h1[kNh], h2[kNh];
    if (!g_strip_half) {
        for (int j = 0; j < kNh; ++j) {
            float a = p[kB1 + j];
            for (int i = 0; i < kNin; ++i) a = std::fma(p[kW1 + j * kNin + i], x[i], a);
            h1[j] = act_tanh(a);
        }
        for (int j = 0; j < kNh; ++j) {
            float a = p[kB2 + j];
            for (int i = 0; i < kNh; ++i) a = std::fma(p[kW2 + j * kNh + i], h1[i], a);
            h2[j] = act_tanh(a);
        }
        for (int o = 0; o < kNout; ++o) {
            float a = p[kB3 + o];
            for (int i = 0; i < kNh; ++i) a = std::fma(p[kW3 + o * kNh + i], h2[i], a);
            y[o] = a;
        }
        return;
    }
    /* binary16 operands of the two hidden layers (products exact in float), sums from 0,
     * bias added after; the last layer in float over the binary16 activations */
    float xh[kNin];
    for (int i = 0; i < kNin; ++i) xh[i] = round_half(x[i]);
    for (int j = 0; j < kNh; ++j) {
        float a = 0.f;
        for (int i = 0; i < kNin; ++i) a += round_half(p[kW1 + j * kNin + i]) * xh[i];
        h1[j] = round_half(act_tanh(a + p[kB1 + j]));
    }
    for (int j = 0; j < kNh; ++j) {
        float a = 0.f;
        for (int i = 0; i < kNh; ++i) a += round_half(p[kW2 + j * kNh + i]) * h1[i];
        h2[j] = round_half(act_tanh(a + p[kB2 + j]));
    }
    for (int o = 0; o < kNout; ++o) {
        float a = 0.f;
        for (int i = 0; i < kNh; ++i) a = std::fma(p[kW3 + o * kNh + i], h2[i], a);
        y[o] = a + p[kB3 + o];
    }
}

/* Oracle layout: z fastest, id = iz + nnz*ix + pl*iy. */
/* Head 2 (mfac != NULL): each strip cell's damping multiplier m = kMmax * sigmoid(y_0) is
 * written to mfac instead; the update loops rebuild the cell's split-PML coefficients from
 * it (pml_coef). */
void strip_correct(int nnx, int nny, int nnz, float dt, float dx, float dy, float dz, float* P, float* Q,
                   float* u, float* v, float* w, float* px, float* py, float* pz, float* qx,
                   float* qy, float* qz, const float* vp, const float* rho, const float* eps, const float* del,
                   float* mfac = nullptr, float vref = 0.0f) {
    const long pl = (long)nnz * nnx;
    const long stride[3] = {nnz, pl, 1};   /* x, y, z */
    const int n3[3] = {nnx, nny, nnz};
    const float cdt[3] = {dt / dx, dt / dy, dt / dz};
    struct Cell { long id; float gp, gq; int axis; bool normal_damped; };
    const int nb = kStripCells + mm;   /* the strip grid's padding */
    std::vector<std::vector<Cell>> rows(nny);
#pragma omp parallel for schedule(static)
    for (int iy = mm; iy < nny - mm; ++iy)
        for (int ix = mm; ix < nnx - mm; ++ix)
            for (int iz = mm; iz < nnz - mm; ++iz) {
                const int pp[3] = {ix, iy, iz};
                int best = 1 << 30, axis = 0, side = 0;
                for (int a = 2; a >= 0; --a) {
                    const int lo = pp[a] - mm + 1, hi = n3[a] - mm - pp[a];
                    if (lo < best) { best = lo; axis = a; side = 0; }
                    if (hi < best) { best = hi; axis = a; side = 1; }
                }
                if (best > kStripCells) continue;
                const long id = iz + (long)nnz * ix + pl * iy;
                const long sa = stride[axis], sn = side ? sa : -sa;
                const long s1 = axis == 0 ? stride[1] : stride[0];
                const long s2 = axis == 2 ? stride[1] : stride[2];
                const float* vel = axis == 0 ? u : (axis == 1 ? v : w);
                float f[11] = {P[id], P[id - sn], P[id + sn], P[id - s1], P[id + s1], P[id - s2], P[id + s2],
                               Q[id], Q[id - sn], Q[id + sn], 0.f};
                const float un = 0.5f * (vel[id] + vel[id - sa]);
                f[10] = (side ? un : -un) * (rho[id] * vp[id]);
                float s = 1e-30f;
                for (int k = 0; k < 11; ++k) s = std::max(s, std::fabs(f[k]));
                const float inv = 1.0f / s;
                float x[kNin], y[kNout];
                for (int k = 0; k < 11; ++k) x[k] = f[k] * inv;
                x[11] = vp[id] * cdt[axis];
                x[12] = eps[id];
                x[13] = del[id];
                x[14] = (float)best * (1.0f / kStripCells);
                x[15] = mfac ? vp[id] / vref : 0.f;   /* head 2: velocity over the d0 sample */
                strip_net(g_strip.data(), x, y);
                if (mfac) {
                    mfac[id] = kMmax / (1.0f + std::exp(-y[0]));
                    continue;
                }
                /* stress PML factor along the normal differs from 1 (the ramp's tables) */
                const bool nd = pp[axis] < nb || pp[axis] > n3[axis] - nb;
                rows[iy].push_back(Cell{id, kGmax / (1.0f + std::exp(-y[0])), kGmax / (1.0f + std::exp(-y[1])), axis,
                                        nd});
            }
    for (const auto& r : rows)
        for (const Cell& c : r) {
            if (g_strip_head == 1) {
                /* normal head: the split components along the normal axis and the normal
                 * velocity (a local increase of the PML damping along that axis) */
                if (!c.normal_damped) continue;
                float* pa = c.axis == 0 ? px : (c.axis == 1 ? py : pz);
                float* qa = c.axis == 0 ? qx : (c.axis == 1 ? qy : qz);
                float* va = c.axis == 0 ? u : (c.axis == 1 ? v : w);
                const float dp = c.gp * pa[c.id], dq = c.gq * qa[c.id];
                pa[c.id] = pa[c.id] - dp;
                P[c.id] = P[c.id] - dp;
                qa[c.id] = qa[c.id] - dq;
                Q[c.id] = Q[c.id] - dq;
                va[c.id] *= 1.0f - (c.axis == 2 ? c.gq : c.gp);
                continue;
            }
            const float fp = 1.0f - c.gp, fq = 1.0f - c.gq;
            P[c.id] *= fp;
            px[c.id] *= fp;
            py[c.id] *= fp;
            pz[c.id] *= fp;
            Q[c.id] *= fq;
            qx[c.id] *= fq;
            qy[c.id] *= fq;
            qz[c.id] *= fq;
            u[c.id] *= fp;
            v[c.id] *= fp;
            w[c.id] *= fq;
        }
}

/* Half the damping step of each index on initial_coffe's velocity (half-cell) and stress
 * nodes, h = (dt*d0) r^2 / 2 with r as there (0 inside the interior), in float: the
 * coefficient head's per-cell coefficients are c1 = 1/(1 + m h), c2 = c1 (1 - m h). */
void pml_half(float dt, float d0, int n, int npml, int width, float* hv, float* hs) {
    const float dtd0 = dt * d0;
    for (int id = 0; id < n + 2 * npml; ++id) {
        double r, ra;
        if (id < npml) {
            r = (npml - 0.5 - id) / width;
            ra = ((npml - id) * 1.0) / width;
        } else if (id < npml + n) {
            hv[id] = hs[id] = 0.0f;
            continue;
        } else {
            r = (0.5 + id - n - npml) / width;
            ra = ((id - n - npml) * 1.0) / width;
        }
        hv[id] = (float)((dtd0 * (r * r)) / 2.0);
        hs[id] = (float)((dtd0 * (ra * ra)) / 2.0);
    }
}

inline void pml_coef(float m, float h, float& c1, float& c2) {
    const float t = m * h;
    c1 = 1.0f / (1.0f + t);
    c2 = c1 * (1.0f - t);
}

template <class T>
int run_1st(const OracleShot& p, const float* vel, const float* rho_in, const float* eps_in,
            const float* del_in, const float* th_in, const float* ph_in, T* shot, T* snaps,
            int nsnap_max, double* secs, int nsteps_override) {
    const int op = p.op;
    const int npml = p.nb, nx = p.nx, ny = p.ny, nz = p.nz;
    const int nnx = nx + 2 * npml, nny = ny + 2 * npml, nnz = nz + 2 * npml;
    const long pl = (long)nnz * nnx;
    const size_t N = (size_t)pl * nny;
    const int nt = p.nt;
    const int nsteps = nsteps_override > 0 ? nsteps_override : nt;
    const bool iso = op == ORC_OP_ISO, vti = op == ORC_OP_VTI, tti = op == ORC_OP_TTI;

    std::vector<T> vp(N, 0), rho, eps, del, th, ph;
    read_interior<T>(vel, nx, ny, nz, nnx, nnz, npml, vp.data());
    pad_vv<T>(nnx, nny, nnz, npml, vp.data());
    if (iso || vti) {
        rho.assign(N, 0);
        read_interior<T>(rho_in, nx, ny, nz, nnx, nnz, npml, rho.data());
        pad_vv<T>(nnx, nny, nnz, npml, rho.data());
    }
    if (vti || tti) {
        eps.assign(N, 0);
        del.assign(N, 0);
        read_interior<T>(eps_in, nx, ny, nz, nnx, nnz, npml, eps.data());
        read_interior<T>(del_in, nx, ny, nz, nnx, nnz, npml, del.data());
        pad_vv<T>(nnx, nny, nnz, npml, eps.data());
        pad_vv<T>(nnx, nny, nnz, npml, del.data());
    }
    if (tti) {
        th.assign(N, 0);
        ph.assign(N, 0);
        read_interior<T>(th_in, nx, ny, nz, nnx, nnz, npml, th.data());
        read_interior<T>(ph_in, nx, ny, nz, nnx, nnz, npml, ph.data());
        pad_vv<T>(nnx, nny, nnz, npml, th.data());
        pad_vv<T>(nnx, nny, nnz, npml, ph.data());
    }

    const T dt = (T)p.dt;
    /* the learned strip (FP32 VTI, network set) ramps its PML over its 6 cells */
    const bool strip = std::is_same<T, float>::value && vti && !g_strip.empty();
    const int width = strip ? kStripCells : npml;
    const T d0 = strip ? get_d0_amp<T>(vp.data(), nnx, nny, nnz, width, p.dx, p.dy, p.dz, g_strip_amp)
                       : get_d0<T>(vp.data(), nnx, nny, nnz, npml, p.dx, p.dy, p.dz);
    std::vector<T> cx1(nnx), cx2(nnx), ax1(nnx), ax2(nnx);
    std::vector<T> cy1(nny), cy2(nny), ay1(nny), ay2(nny);
    std::vector<T> cz1(nnz), cz2(nnz), az1(nnz), az2(nnz);
    initial_coffe<T>(dt, d0, nx, npml, cx1.data(), cx2.data(), ax1.data(), ax2.data(), width);
    initial_coffe<T>(dt, d0, ny, npml, cy1.data(), cy2.data(), ay1.data(), ay2.data(), width);
    initial_coffe<T>(dt, d0, nz, npml, cz1.data(), cz2.data(), az1.data(), az2.data(), width);

    const T dtx = dt / (T)p.dx, dty = dt / (T)p.dy, dtz = dt / (T)p.dz;   /* iso:109-111 */
    /* learned strip, coefficient head: half damping steps per index and the multipliers */
    const bool strip2 = strip && g_strip_head == 2;
    std::vector<float> hvx(nnx), hsx(nnx), hvy(nny), hsy(nny), hvz(nnz), hsz(nnz), mfac;
    if (strip2) {
        pml_half((float)dt, (float)d0, nx, npml, width, hvx.data(), hsx.data());
        pml_half((float)dt, (float)d0, ny, npml, width, hvy.data(), hsy.data());
        pml_half((float)dt, (float)d0, nz, npml, width, hvz.data(), hsz.data());
        mfac.assign(N, 1.0f);
    }
    /* signs of the TTI rotation terms: all +1 unless a test plants a mistake */
    T vsgn[8], ssgn[8];
    for (int k = 0; k < 8; ++k) {
        vsgn[k] = g_tti_fault == k + 1 ? (T)-1 : (T)1;
        ssgn[k] = g_tti_fault == k + 9 ? (T)-1 : (T)1;
    }

    std::vector<T> u(N, 0), v(N, 0), w(N, 0), P(N, 0), Q, px(N, 0), py(N, 0), pz(N, 0), qx, qy, qz;
    if (!iso) { Q.assign(N, 0); qx.assign(N, 0); qy.assign(N, 0); qz.assign(N, 0); }
    T* Pp = P.data();
    T* Qp = iso ? P.data() : Q.data();

    /* add_source: 1-based source, iso:91-94 */
    const long s = (p.sz + npml - 1) + (long)(p.sx + npml - 1) * nnz + (long)(p.sy + npml - 1) * pl;
    const T favg = (T)p.fm;
    const T pfac = (T)p.pfac;
    int nsnap = 0;
    const int sxp = p.sx + npml - 1, syp = p.sy + npml - 1, szp = p.sz + npml - 1;

    auto t0 = std::chrono::steady_clock::now();
    /* The time grid is the reference's float accumulation t = dt, t += dt (iso:521) and
     * the gate t <= 2*tdelay is taken in float, in both instances: the FP64 instance
     * differs from FP32 only by the precision of the wavelet value and the propagation. */
    const float tdelay_f = (float)(1.0 / p.fm);
    float t = p.dt;
    for (int it = 0; it < nsteps; it++, t += p.dt) {
        /* add_source (iso:523-530: -pfac at zs, +pfac at zs+2; vti:557, tti:571: +pfac to P and Q) */
        T src = wavelet<T>(p.wtype, (T)t, favg);
        if (t <= 2 * tdelay_f) {
            if (iso) {
                T mpf = (T)(-1.0 * pfac);
                Pp[s] += mpf * src;
                Pp[s + 2] += pfac * src;
            } else {
                Pp[s] += pfac * src;
                Qp[s] += pfac * src;
            }
        }
        /* update_vel (iso:99-142, vti:91-128, tti:68-112); in place, which equals the
         * reference's ping-pong because every update reads only its own point of u0/v0/w0. */
#pragma omp parallel for schedule(static)
        for (int iy = mm; iy < nny - mm; iy++)
            for (int ix = mm; ix < nnx - mm; ix++)
                for (int iz = mm; iz < nnz - mm; iz++) {
                    long id = iz + (long)ix * nnz + iy * pl;
                    T xx, yy, zz;
                    if (!tti) {
                        yy = dfwd<T>(Pp, id, pl);
                        xx = dfwd<T>(Pp, id, nnz);
                        zz = dfwd<T>(Qp, id, 1);
                        xx /= rho[id];
                        yy /= rho[id];
                        zz /= rho[id];
                    } else {
                        T Py = dfwd<T>(Pp, id, pl), Px = dfwd<T>(Pp, id, nnz), Pz = dfwd<T>(Pp, id, 1);
                        T Qy = dfwd<T>(Qp, id, pl), Qx = dfwd<T>(Qp, id, nnz), Qz = dfwd<T>(Qp, id, 1);
                        T ct = fcos<T>(th[id]), st = fsin<T>(th[id]);
                        T cp = fcos<T>(ph[id]), sp = fsin<T>(ph[id]);
                        const T* f = vsgn;
                        if (g_tti_fault == 17) Qx = Px;
                        xx = f[0] * ct * cp * Px - f[1] * ct * sp * Py - f[2] * st * Pz;     /* tti:101 */
                        yy = f[3] * sp * Px + f[4] * cp * Py;                                /* tti:102 */
                        zz = f[5] * st * cp * Qx - f[6] * st * sp * Qy + f[7] * ct * Qz;     /* tti:103 */
                    }
                    T kx1 = cx1[ix], kx2 = cx2[ix], ky1 = cy1[iy], ky2 = cy2[iy], kz1 = cz1[iz], kz2 = cz2[iz];
                    if constexpr (std::is_same<T, float>::value) {
                        if (strip2 && (hvx[ix] != 0.0f || hvy[iy] != 0.0f || hvz[iz] != 0.0f)) {
                            const float m = mfac[id];
                            pml_coef(m, hvx[ix], kx1, kx2);
                            pml_coef(m, hvy[iy], ky1, ky2);
                            pml_coef(m, hvz[iz], kz1, kz2);
                        }
                    }
                    u[id] = kx2 * u[id] - kx1 * dtx * xx;
                    v[id] = ky2 * v[id] - ky1 * dty * yy;
                    w[id] = kz2 * w[id] - kz1 * dtz * zz;
                }
        /* update_stress (iso:145-190, vti:131-203, tti:115-201) */
#pragma omp parallel for schedule(static)
        for (int iy = mm; iy < nny - mm; iy++)
            for (int ix = mm; ix < nnx - mm; ix++)
                for (int iz = mm; iz < nnz - mm; iz++) {
                    long id = iz + (long)ix * nnz + iy * pl;
                    if (iso) {
                        T yy = dbwd<T>(v.data(), id, pl);
                        T xx = dbwd<T>(u.data(), id, nnz);
                        T zz = dbwd<T>(w.data(), id, 1);
                        px[id] = ax2[ix] * px[id] - ax1[ix] * rho[id] * vp[id] * vp[id] * dtx * xx;
                        py[id] = ay2[iy] * py[id] - ay1[iy] * rho[id] * vp[id] * vp[id] * dty * yy;
                        pz[id] = az2[iz] * pz[id] - az1[iz] * rho[id] * vp[id] * vp[id] * dtz * zz;
                        Pp[id] = px[id] + py[id] + pz[id];
                        continue;
                    }
                    /* SV taper (vti:154-175) */
                    T ee, dd;
                    int rx = ix - sxp, ry = iy - syp, rz = iz - szp;
                    int r2 = rx * rx + ry * ry + rz * rz;
                    if (p.sv && r2 <= 15 * 15) {
                        if (r2 <= 4 * 4) {
                            ee = 0;
                            dd = 0;
                        } else {
                            double f = (sizeof(T) == 4) ? taper_factor(r2)
                                : 0.5 * (1 - std::cos(kPi * ((std::sqrt((double)r2) - 4) * 4.0 / 44.0)));
                            ee = (T)(f * eps[id]);
                            dd = (T)(f * del[id]);
                        }
                    } else {
                        ee = eps[id];
                        dd = del[id];
                    }
                    T xx, yy, zz;
                    if (vti) {
                        yy = dbwd<T>(v.data(), id, pl);
                        xx = dbwd<T>(u.data(), id, nnz);
                        zz = dbwd<T>(w.data(), id, 1);
                        const T r = rho[id], c = vp[id];
                        T bx1 = ax1[ix], bx2 = ax2[ix], by1 = ay1[iy], by2 = ay2[iy], bz1 = az1[iz], bz2 = az2[iz];
                        if constexpr (std::is_same<T, float>::value) {
                            if (strip2 && (hsx[ix] != 0.0f || hsy[iy] != 0.0f || hsz[iz] != 0.0f)) {
                                const float m = mfac[id];
                                pml_coef(m, hsx[ix], bx1, bx2);
                                pml_coef(m, hsy[iy], by1, by2);
                                pml_coef(m, hsz[iz], bz1, bz2);
                            }
                        }
                        px[id] = bx2 * px[id] - bx1 * r * c * c * (1 + 2 * ee) * dtx * xx;
                        py[id] = by2 * py[id] - by1 * r * c * c * (1 + 2 * ee) * dty * yy;
                        pz[id] = bz2 * pz[id] - bz1 * r * c * c * fsqrt<T>(1 + 2 * dd) * dtz * zz;
                        qx[id] = bx2 * qx[id] - bx1 * r * c * c * fsqrt<T>(1 + 2 * dd) * dtx * xx;
                        qy[id] = by2 * qy[id] - by1 * r * c * c * fsqrt<T>(1 + 2 * dd) * dty * yy;
                        qz[id] = bz2 * qz[id] - bz1 * r * c * c * dtz * zz;
                    } else {
                        T uy = dbwd<T>(u.data(), id, pl), ux = dbwd<T>(u.data(), id, nnz), uz = dbwd<T>(u.data(), id, 1);
                        T vy = dbwd<T>(v.data(), id, pl), vx = dbwd<T>(v.data(), id, nnz), vz = dbwd<T>(v.data(), id, 1);
                        T wy = dbwd<T>(w.data(), id, pl), wx = dbwd<T>(w.data(), id, nnz), wz = dbwd<T>(w.data(), id, 1);
                        (void)vz;   /* computed but unused by the reference (tti:179, :185) */
                        T ct = fcos<T>(th[id]), st = fsin<T>(th[id]);
                        T cp = fcos<T>(ph[id]), sp = fsin<T>(ph[id]);
                        const T* f = ssgn;
                        if (g_tti_fault == 18) ux = vx;
                        xx = f[0] * ct * cp * ux - f[1] * ct * sp * uy - f[2] * st * uz;     /* tti:184 */
                        yy = f[3] * sp * vx + f[4] * cp * vy;                                /* tti:185 */
                        zz = f[5] * st * cp * wx - f[6] * st * sp * wy + f[7] * ct * wz;     /* tti:186 */
                        const T c = vp[id];
                        px[id] = ax2[ix] * px[id] - ax1[ix] * c * c * (1 + 2 * ee) * dtx * xx;
                        py[id] = ay2[iy] * py[id] - ay1[iy] * c * c * (1 + 2 * ee) * dty * yy;
                        pz[id] = az2[iz] * pz[id] - az1[iz] * c * c * fsqrt<T>(1 + 2 * dd) * dtz * zz;
                        qx[id] = ax2[ix] * qx[id] - ax1[ix] * c * c * fsqrt<T>(1 + 2 * dd) * dtx * xx;
                        qy[id] = ay2[iy] * qy[id] - ay1[iy] * c * c * fsqrt<T>(1 + 2 * dd) * dty * yy;
                        qz[id] = az2[iz] * qz[id] - az1[iz] * c * c * dtz * zz;
                    }
                    Pp[id] = px[id] + py[id] + pz[id];
                    Qp[id] = qx[id] + qy[id] + qz[id];
                }
        if constexpr (std::is_same<T, float>::value) {
            if (vti && !g_strip.empty() &&
                (!strip2 || (g_strip_period > 0 ? it % g_strip_period == 0 : it == 0)))
                strip_correct(nnx, nny, nnz, p.dt, p.dx, p.dy, p.dz, Pp, Qp, u.data(), v.data(), w.data(),
                              px.data(), py.data(), pz.data(), qx.data(), qy.data(), qz.data(), vp.data(),
                              rho.data(), eps.data(), del.data(), strip2 ? mfac.data() : nullptr,
                              (float)vp[(long)nny * nnx * nnz / 2]);
        }
        /* shot_record (iso:294-304): interior z = rec_z (reference: 0) */
        if (shot) {
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix)
                    shot[it + (long)nt * ix + (long)nt * nx * iy] =
                        Pp[(npml + p.rec_z) + (long)nnz * (ix + npml) + pl * (iy + npml)];
        }
        /* snapshots (iso:568-584, vti:601-617, tti:593-608) */
        if (p.snap_interval > 0 && it % p.snap_interval == 0 && !(p.snap_skip0 && it == 0)) {
            if (snaps && nsnap < nsnap_max) {
                T* o = snaps + (size_t)nsnap * nz * nx * ny;
                for (int iy = 0; iy < ny; ++iy)
                    for (int ix = 0; ix < nx; ++ix)
                        for (int iz = 0; iz < nz; ++iz)
                            o[iz + (long)nz * ix + (long)nz * nx * iy] =
                                Pp[(iz + npml) + (long)nnz * (ix + npml) + pl * (iy + npml)];
            }
            ++nsnap;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();

    if (shot && p.mute) {
        /* mute_directwave (iso:322-369, vti:339-386): v = vp[1] (padded index 1, i.e. the
         * model's first sample), times sqrtf(1+2*epsilon[1]) for VTI/TTI. */
        T vcut;
        if (p.v_mute > 0) vcut = (T)p.v_mute;
        else if (iso) vcut = vp[1];
        else vcut = vp[1] * fsqrt<T>(1 + 2 * eps[1]);
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                T mx = (T)p.dx * (T)std::abs(ix - p.sx);
                T my = (T)p.dy * (T)std::abs(iy - p.sy);
                T mz = (T)p.dz * (T)p.sz;
                T t0m = fsqrt<T>(mx * mx + my * my + mz * mz) / vcut;
                int mu_t = (int)(2.0 / (dt * favg));
                int mu_nt = (int)(t0m / dt) + mu_t + p.mute_tt;
                for (int it = 0; it < nt && it < mu_nt; ++it) shot[it + (long)nt * ix + (long)nt * nx * iy] = 0;
            }
    }
    return nsnap;
}

template <class T>
int dispatch(const OracleShot* p, const float* vel, const float* rho, const float* eps,
             const float* del, const float* th, const float* ph, T* shot, T* snaps, int nsnap_max,
             double* secs, int nsteps_override) {
    if (p->op == ORC_OP_2ND) return run_2nd<T>(*p, vel, shot, snaps, nsnap_max, secs, nsteps_override);
    return run_1st<T>(*p, vel, rho, eps, del, th, ph, shot, snaps, nsnap_max, secs, nsteps_override);
}

int check(const OracleShot* p, const float* vel, const float* rho, const float* eps,
          const float* del, const float* th, const float* ph) {
    if (!p || !vel) return -1;
    if (p->op < 0 || p->op > 3) return -2;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1 || p->nt < 1) return -3;
    if ((p->op == ORC_OP_ISO || p->op == ORC_OP_VTI) && !rho) return -4;
    if ((p->op == ORC_OP_VTI || p->op == ORC_OP_TTI) && (!eps || !del)) return -4;
    if (p->op == ORC_OP_TTI && (!th || !ph)) return -4;
    if (p->op != ORC_OP_2ND && p->nb < 2) return -5;
    if (p->op != ORC_OP_2ND) {
        if (p->sx < 1 || p->sx > p->nx || p->sy < 1 || p->sy > p->ny || p->sz < 1 || p->sz > p->nz) return -6;
        if (p->op == ORC_OP_ISO && (p->sz + p->nb - 1 + 2) >= p->nz + 2 * p->nb - mm) return -6;
    } else {
        if (p->sx < 0 || p->sx >= p->nx || p->sy < 0 || p->sy >= p->ny || p->sz < 0 || p->sz + 2 >= p->nz + p->nb) return -6;
    }
    if (p->rec_z < 0 || p->rec_z >= p->nz) return -7;
    return 0;
}

}  // namespace

extern "C" {

int oracle_run(int prec, const OracleShot* p, const float* vel, const float* rho, const float* eps,
               const float* del, const float* theta, const float* phi, void* shot, void* snaps,
               int nsnap_max, double* loop_seconds) {
    int rc = check(p, vel, rho, eps, del, theta, phi);
    if (rc) return rc;
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
    if (prec == 32)
        return dispatch<float>(p, vel, rho, eps, del, theta, phi, (float*)shot, (float*)snaps, nsnap_max,
                               loop_seconds, 0);
    if (prec == 64)
        return dispatch<double>(p, vel, rho, eps, del, theta, phi, (double*)shot, (double*)snaps, nsnap_max,
                                loop_seconds, 0);
    return -8;
}

int oracle_set_strip_net(const float* params, int n) {
    if (!params) {
        g_strip.clear();
        return 0;
    }
    if (n != kNpar) return -1;
    g_strip.assign(params, params + n);
    return 0;
}

int oracle_set_strip_half(int on) {
    g_strip_half = on != 0;
    return 0;
}

float oracle_round_half(float f) { return round_half(f); }

int oracle_set_strip_head(int mode) {
    if (mode < 0 || mode > 2) return -1;
    g_strip_head = mode;
    return 0;
}

int oracle_set_strip_period(int steps) {
    if (steps < 0) return -1;
    g_strip_period = steps;
    return 0;
}

int oracle_set_tti_fault(int code) {
    if (code < 0 || code > 18) return -1;
    g_tti_fault = code;
    return 0;
}

int oracle_set_strip_profile(float amp) {
    if (!(amp >= 0.0f)) return -1;
    g_strip_amp = amp;
    return 0;
}

int oracle_strip_apply(int nnx, int nny, int nnz, float dt, float dx, float dy, float dz, float* P, float* Q,
                       float* u, float* v, float* w, float* px, float* py, float* pz, float* qx,
                       float* qy, float* qz, const float* vp, const float* rho, const float* eps, const float* del) {
    if (g_strip.empty()) return -1;
    strip_correct(nnx, nny, nnz, dt, dx, dy, dz, P, Q, u, v, w, px, py, pz, qx, qy, qz, vp, rho, eps, del);
    return 0;
}

double oracle_time_steps(const OracleShot* p, const float* vel, const float* rho, const float* eps,
                         const float* del, const float* theta, const float* phi, int nsteps) {
    if (check(p, vel, rho, eps, del, theta, phi)) return -1.0;
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
    double secs = -1.0;
    dispatch<float>(p, vel, rho, eps, del, theta, phi, (float*)nullptr, (float*)nullptr, 0, &secs, nsteps);
    return secs;
}

int oracle_ricker_table(int prec, float fm, float dt, int nt, void* out) {
    if (prec == 32) ricker_table<float>(fm, dt, nt, (float*)out);
    else ricker_table<double>(fm, dt, nt, (double*)out);
    return 0;
}

int oracle_wavelet(int prec, int wtype, float fm, float dt, int nt, void* out) {
    /* the per-step add_source value at float-accumulated t (iso:521) */
    if (prec == 32) {
        float t = dt;
        for (int it = 0; it < nt; ++it, t += dt) ((float*)out)[it] = wavelet<float>(wtype, t, fm);
    } else {
        double t = dt;
        for (int it = 0; it < nt; ++it, t += (double)dt) ((double*)out)[it] = wavelet<double>(wtype, t, fm);
    }
    return 0;
}

int oracle_pml_tables(int prec, int n, int npml, float dt, float d0, void* c1, void* c2, void* a1,
                      void* a2) {
    if (prec == 32)
        initial_coffe<float>(dt, d0, n, npml, (float*)c1, (float*)c2, (float*)a1, (float*)a2, npml);
    else
        initial_coffe<double>(dt, d0, n, npml, (double*)c1, (double*)c2, (double*)a1, (double*)a2, npml);
    return 0;
}

double oracle_d0(const float* vp_padded, int nnx, int nny, int nnz, int npml, float dx, float dy,
                 float dz) {
    return (double)get_d0<float>(vp_padded, nnx, nny, nnz, npml, dx, dy, dz);
}

int oracle_pad(const float* in, int nx, int ny, int nz, int nb, float* out) {
    const int nnx = nx + 2 * nb, nny = ny + 2 * nb, nnz = nz + 2 * nb;
    std::memset(out, 0, sizeof(float) * (size_t)nnx * nny * nnz);
    read_interior<float>(in, nx, ny, nz, nnx, nnz, nb, out);
    pad_vv<float>(nnx, nny, nnz, nb, out);
    return 0;
}

int oracle_sponge_factor(int n, int npd, double* out) {
    for (int i = 0; i < n + 2 * mm + 2 * npd; ++i) {
        bool d;
        out[i] = sponge_factor(i, n, npd, &d);
    }
    return 0;
}

int oracle_taper(int r2, double* out) {
    *out = taper_factor(r2);
    return 0;
}

int oracle_staggered_dx(const float* in, int n, int forward, float* out) {
    for (int i = 0; i < n; ++i) out[i] = 0.0f;
    for (int i = mm; i < n - mm; ++i) out[i] = forward ? dfwd<float>(in, i, 1) : dbwd<float>(in, i, 1);
    return 0;
}

int oracle_laplacian_z(const float* in, int n, float* out) {
    for (int i = 0; i < n; ++i) out[i] = 0.0f;
    for (int i = mm; i < n - mm; ++i) out[i] = lap1<float>(in, i, 1);
    return 0;
}

}  // extern "C"
