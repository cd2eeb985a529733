// a3d_host.hip -- host runtime of libacoustic3d.so: contexts, setup tables,
// the per-shot time loop and the four legacy entry points.
//
// The setup tables (wavelet values, PML profiles, d0, sponge factors, SV
// taper) are time-invariant or per-step scalars; they are computed here on
// the host with the reference's float/double arithmetic and handed to the
// kernels, so no transcendental function runs on the GPU except the per-point
// cos/sin of the TTI rotation.
#include "a3d_internal.h"
#include "a3d_mlp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/acoustic3d.h"

using namespace a3d;

namespace {

constexpr double kPi = 3.141592653;   // reference #define pi 3.141592653

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(A3D_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

}  // namespace

int a3d::report_error(int code, const char* what, const char* detail) { return fail(code, "%s: %s", what, detail); }

namespace {

enum Vol {
    V_VP, V_RHO, V_EPS, V_DEL, V_CT, V_ST, V_CP, V_SP,   // TTI: cos/sin of tilt and azimuth
    V_U, V_V, V_W, V_P, V_Q, V_PX, V_PY, V_PZ, V_QX, V_QY, V_QZ,
    V_VV, V_PA, V_PB,
    V_P2, V_Q2, V_U2, V_V2, V_W2,   // second copies for the one-pass VTI step (scheme 2), allocated on first use
    V_M4,                           // float model, built on first use: VTI (vp, rho, eps, delta) interleaved,
                                    // TTI the rotation terms (cos, sin tilt, cos, sin azimuth), ISO (vp, rho)
    NVOL
};

std::vector<int> volumes_of(int op) {
    switch (op) {
        case A3D_OP_2ND: return {V_VV, V_PA, V_PB};
        case A3D_OP_ISO: return {V_VP, V_RHO, V_U, V_V, V_W, V_P, V_PX, V_PY, V_PZ};
        case A3D_OP_VTI:
            return {V_VP, V_RHO, V_EPS, V_DEL, V_U, V_V, V_W, V_P, V_Q, V_PX, V_PY, V_PZ, V_QX, V_QY, V_QZ};
        case A3D_OP_TTI:
            return {V_VP, V_EPS, V_DEL, V_CT, V_ST, V_CP, V_SP, V_U, V_V, V_W, V_P, V_Q, V_PX, V_PY, V_PZ, V_QX, V_QY, V_QZ};
    }
    return {};
}

bool is_wavefield(int v) { return (v >= V_U && v <= V_QZ) || v == V_PA || v == V_PB || (v >= V_P2 && v <= V_W2); }
bool is_split(int v) { return v >= V_PX && v <= V_QZ; }

// ---------------------------------------------------------------------------
// Setup tables (host), reference arithmetic
// ---------------------------------------------------------------------------

// add_source wavelet value at float-accumulated time t (reference iso:52-89;
// vti:54-79 and tti:31-56 use types 1..3). T = float follows the reference's
// float/double mix; T = double evaluates the same formulas in double.
template <class T>
T wavelet_value(int wtype, float tf, float fm) {
    const T t = (T)tf, favg = (T)fm;
    T x_, xx_, tdelay, ts, source = 0;
    tdelay = (T)(1.0 / favg);
    ts = t - tdelay;
    switch (wtype) {
        case 1:
            x_ = favg * ts;
            xx_ = x_ * x_;
            source = (T)((1 - 2 * kPi * kPi * (xx_)) * std::exp(-(kPi * kPi * xx_)));
            break;
        case 2:
            x_ = (T)((-4) * favg * favg * kPi * kPi / std::log(0.1));
            source = (T)((-2) * kPi * kPi * ts * std::exp(-x_ * ts * ts));
            break;
        case 3:
            x_ = (T)((-1) * favg * favg * kPi * kPi / std::log(0.1));
            source = (T)std::exp(-x_ * ts * ts);
            break;
        case 4:
            x_ = (T)(std::exp(-favg * favg * kPi * kPi * ts * ts) * ts);
            source = -x_;
            break;
        case 5:
            x_ = (T)(-1.0 / (2.0 * favg * favg * kPi * kPi));
            x_ = (T)(x_ * std::exp(-favg * favg * kPi * kPi * ts * ts));
            source = x_;
            break;
        case 6:
            x_ = (T)(-std::erf(favg * kPi * ts) / (4 * favg * favg * favg * kPi * kPi * std::sqrt((T)kPi)));
            source = x_;
            break;
        default: break;
    }
    return source;
}

// cuda_ricker_wavelet (reference Toa_3dfd_cuda.cu:34-43).
template <class T>
T ricker_value(int it, float fm_, float dt_) {
    const T fm = (T)fm_, dt = (T)dt_;
    T tmp;
    if (sizeof(T) == 4) tmp = (T)(kPi * fm * std::fabs((float)(it * dt - 1.0 / fm)));
    else tmp = (T)(kPi * fm * std::fabs(it * dt - 1.0 / fm));
    tmp *= tmp;
    return (T)((1.0 - 2.0 * tmp) * std::exp(-tmp));
}

// initial_coffe (reference iso:263-291) for one axis of interior length n; the
// profile's ramp runs over `width` cells (the reference: width = npml).
template <class T>
void pml_axis(T dt, T d0, int n, int npml, T* c1, T* c2, T* a1, T* a2, int width) {
    const T dtd0 = dt * d0;
    for (int id = 0; id < n + 2 * npml; ++id) {
        double r, ra;
        if (id < npml) {
            r = (npml - 0.5 - id) / width;
            ra = ((npml - id) * 1.0) / width;
        } else if (id < npml + n) {
            c1[id] = c2[id] = a1[id] = a2[id] = (T)1.0;
            continue;
        } else {
            r = (0.5 + id - n - npml) / width;
            ra = ((id - n - npml) * 1.0) / width;
        }
        c1[id] = (T)(1.0 / (1.0 + (dtd0 * (r * r)) / 2.0));
        c2[id] = (T)(c1[id] * (1.0 - (dtd0 * (r * r)) / 2.0));
        a1[id] = (T)(1.0 / (1.0 + (dtd0 * (ra * ra)) / 2.0));
        a2[id] = (T)(a1[id] * (1.0 - (dtd0 * (ra * ra)) / 2.0));
    }
}

// Half the damping step of each index, hv = dt d (r) / 2 on the velocity (half-cell) and
// hs = dt d (ra) / 2 on the stress nodes of initial_coffe's profile, 0 outside the PML: the
// learned strip's coefficient head rebuilds c1, c2 per cell from a multiple of them.
template <class T>
void pml_half_steps(T dt, T d0, int n, int npml, T* hv, T* hs, int width) {
    const T dtd0 = dt * d0;
    for (int id = 0; id < n + 2 * npml; ++id) {
        double r, ra;
        if (id < npml) {
            r = (npml - 0.5 - id) / width;
            ra = ((npml - id) * 1.0) / width;
        } else if (id < npml + n) {
            hv[id] = hs[id] = (T)0;
            continue;
        } else {
            r = (0.5 + id - n - npml) / width;
            ra = ((id - n - npml) * 1.0) / width;
        }
        hv[id] = (T)((dtd0 * (r * r)) / 2.0);
        hs[id] = (T)((dtd0 * (ra * ra)) / 2.0);
    }
}

// cuda_absorb_bndr factor (reference Toa_3dfd_cuda.cu:138-151): qp*r^2 + 1
// inside the npd-wide bands, 1.0 elsewhere.
double sponge_value(int i, int n, int npd) {
    const double qp = (double)(-0.25f);
    if (i < npd) {
        const double r = (npd - i) / (1.0 * npd);
        return qp * (r * r) + 1;
    }
    if (i >= 2 * MM + npd + n) {
        const double r = (i - 2 * MM - npd - n) / (1.0 * npd);
        return qp * (r * r) + 1;
    }
    return 1.0;
}

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Records e on s; a context without loop events (the allocation-free entry point) skips it.
hipError_t mark(hipEvent_t e, hipStream_t s) { return e ? hipEventRecord(e, s) : hipSuccess; }

struct Timing {
    bool on = false;
    int stride = 1;       // time steps sampled: every stride-th one
    bool sample = true;   // the current time step is sampled
    int skip = 0;         // launches of each kernel left out at the start of every shot
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<int, size_t>> marks;   // (kernel index, first event)
    double ms[2] = {0, 0};
    long long n[2] = {0, 0};
};

}  // namespace

struct a3d_ctx {
    a3d_shot_desc d{};
    int device = 0;
    hipStream_t st = nullptr;
    Geo g{};
    size_t nvox = 0;
    int off = 0;
    int prec = PREC_FP32;       // a3d_shot_desc.storage: 0 FP32, 1 FP16 wavefields, 2 FP64
    size_t wsize = 4, msize = 4, csize = 4;   // bytes per wavefield / model / table element
    void* vol[NVOL] = {};
    float* stage = nullptr;     // 2nd order: interior velocity for (v*dt)^2
    float vv_dt = 0.0f;
    void* tab = nullptr;        // 12 PML arrays (float, or double for FP64)
    double* dtab = nullptr;     // sponge fx, fy, fz; SV taper
    float* rec = nullptr;
    size_t rec_cap = 0;
    float* shot_dev = nullptr;
    size_t shot_cap = 0;
    float* snap_dev = nullptr;
    int* flag = nullptr;
    bool model_set = false;
    bool m4_ok = false;         // V_M4 holds the current model
    int tti_angles = 0;         // TTI (float model): 1 the sweeps evaluate the rotation terms from
                                // the angles (8 bytes per point instead of 16), 0 read them
    bool m4_ang = false;        // V_M4 holds the angles (TTI, tti_angles) rather than the terms
    float vel_first = 0.f, eps_first = 0.f, vel_d0 = 0.f, vel_raw1 = 0.f;
    float zmin = 1.0f;          // minimum acoustic impedance rho*v (binary16 velocity-field scale)
    // launch shape per sweep: [0] 2nd-order step / velocity sweep, [1] stress sweep
    int ty[2] = {8, 4}, zseg[2] = {0, 0};
    bool tuned = false;
    int ncu = 256;
    Timing tm;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // learned absorbing strip (VTI, FP32, nb = STRIP_NB): network parameters and strip cells
    float* strip_net = nullptr;
    int* strip_ids = nullptr;
    int* strip_meta = nullptr;
    int strip_n = 0;
    float2* strip_delta = nullptr;
    float strip_amp = 10.0f;    // PML amplitude of the strip's 6-cell ramp
    int strip_mfma = 1;         // network on the matrix cores (binary16 operands) or FP32 per lane
    int strip_head = 0;         // 0: damping of every field (sponge); 1: normal-axis PML components;
                                // 2: multiplier of the PML damping inside the sweeps (coefficient head)
    float* strip_mfac = nullptr;   // head 2: damping multiplier per padded point (1 outside the strip)
    int strip_period = 1;          // head 2: the network runs after the stress sweep of every period-th step
                                   // (0: after the first step's only)
    int zchunk = 0;             // 1st-order two-sweep step in z chunks of this many planes (0: whole grid)
    // which sweep updates w (ISO, VTI): requested mode (-1 tuner, 0 velocity sweep,
    // 1 stress sweep) and the tuner's choice for mode -1
    int wz_mode = -1;
    int wz_tuned = 0;
    bool out_device = false;    // the shot record is transposed into shot_dev, the caller's device buffer
    int graph_steps = 0;        // > 0: the time loop runs as hipGraphs of at most this many steps
    void* multi = nullptr;      // a3d_run_shots: the shots' two pressure levels
    size_t multi_cap = 0;       //   (bytes)
    float* mrec = nullptr;      // a3d_run_shots: the shots' receiver records
    size_t mrec_cap = 0;
    int mty = 4, mzseg = 1;     // a3d_run_shots with several shots: tile rows, z segments
    bool mshape = false;        //   (set by a3d_set_launch or the first multi-shot run's tuning)
};

namespace {

// XCD runs of equal tile counts (the last ones shorter when 8 does not divide the count).
void uniform_bands(Launch& l) {
    const int per = (l.ntiles + 7) / 8;
    for (int k = 0; k <= 8; ++k) l.band[k] = std::min(k * per, l.ntiles);
    l.grid = 8 * per;
}

// XCD runs of about equal summed cost(t) (contiguous, in tile order); grid = 8 x the
// longest run.
template <class F>
void balance_bands(Launch& l, F cost) {
    std::vector<double> pre(l.ntiles + 1, 0.0);
    for (int t = 0; t < l.ntiles; ++t) pre[t + 1] = pre[t] + cost(t);
    l.band[0] = 0;
    l.band[8] = l.ntiles;
    for (int k = 1; k < 8; ++k) {
        const double target = pre[l.ntiles] * k / 8.0;
        int t = (int)(std::lower_bound(pre.begin(), pre.end(), target) - pre.begin());
        if (t > 0 && target - pre[t - 1] < pre[t] - target) --t;
        l.band[k] = std::max(l.band[k - 1], std::min(t, l.ntiles));
    }
    int longest = 1;
    for (int k = 0; k < 8; ++k) longest = std::max(longest, l.band[k + 1] - l.band[k]);
    l.grid = 8 * longest;
}

// Tiles of 64 x ty points over the updated x-y range, the planes [z0, z1) (default: all
// updated planes) cut into zseg segments.
Launch make_launch(const Geo& g, int ty, int zseg, int z0 = MM, int z1 = -1) {
    Launch l{};
    if (z1 < 0) z1 = g.nnz - MM;
    l.ntx = g.ld / TX;
    l.nty = (g.nny - 2 * MM + ty - 1) / ty;
    const int nzi = z1 - z0;
    int nseg = std::max(1, std::min(zseg > 0 ? zseg : 1, nzi));
    l.zlen = (nzi + nseg - 1) / nseg;
    l.nseg = (nzi + l.zlen - 1) / l.zlen;
    l.ntiles = l.ntx * l.nty * l.nseg;
    l.zbeg = z0;
    l.zend = z1;
    uniform_bands(l);
    return l;
}

// Default launch shape: a z segmentation that makes one sweep a single wave of
// resident workgroups (occupancy from the compiled kernel's registers and LDS).
int single_wave_zseg(const a3d_ctx* c, int kind, int ty) {
    const Launch l1 = make_launch(c->g, ty, 1);
    const int per_plane = l1.ntx * l1.nty;
    const int kop = c->d.op == A3D_OP_TTI && c->tti_angles && c->prec != PREC_FP64 ? 4 : c->d.op;
    const int cap = std::max(1, max_blocks_per_cu(c->prec, kind, kop, ty)) * c->ncu;
    const int nzi = c->g.nnz - 2 * MM;
    int nseg = std::max(1, cap / std::max(per_plane, 1));
    nseg = std::min(nseg, std::max(1, nzi / 16));
    return nseg;
}

void default_shapes(a3d_ctx* c) {
    c->ty[0] = 8;
    c->ty[1] = 4;
    c->zseg[0] = single_wave_zseg(c, 0, c->ty[0]);
    c->zseg[1] = single_wave_zseg(c, 1, c->ty[1]);
}

bool same_geometry(const a3d_shot_desc& a, const a3d_shot_desc& b) {
    return a.op == b.op && a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx &&
           a.dy == b.dy && a.dz == b.dz && a.nb == b.nb && a.storage == b.storage;
}

int check_desc_geometry(const a3d_shot_desc* d) {
    if (!d) return fail(A3D_EINVAL, "null shot descriptor");
    if (d->op < A3D_OP_2ND || d->op > A3D_OP_TTI) return fail(A3D_EINVAL, "unknown operator %d", d->op);
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(A3D_EINVAL, "model size must be positive");
    if (!(d->dx > 0) || !(d->dy > 0) || !(d->dz > 0)) return fail(A3D_EINVAL, "grid spacing must be positive");
    if (d->op == A3D_OP_2ND ? d->nb < 1 : d->nb < MM) return fail(A3D_EINVAL, "boundary width %d too small", d->nb);
    if (d->storage < 0 || d->storage > 2) return fail(A3D_EINVAL, "storage mode %d not in 0..2", d->storage);
    const int ring = d->op == A3D_OP_2ND ? 2 * MM + 2 * d->nb : 2 * d->nb;
    // kernels index the padded volume (x pitch rounded up to 64) with 32-bit ints
    const double ld = (double)((d->nx + ring + TX - 1) / TX * TX);
    const double n = ld * (d->ny + ring) * (d->nz + ring);
    if (n > 2.0e9) return fail(A3D_EINVAL, "padded volume too large for 32-bit indexing");
    // buffer loads address a plane with 31-bit byte offsets (16-byte elements at most: the
    // interleaved model and rotation terms)
    if (ld * (d->ny + ring) * 16.0 >= 2147483648.0) return fail(A3D_EINVAL, "padded x-y plane too large");
    return A3D_OK;
}

// the batched stress sweeps' split mode of a stress scheme (3: directional split, 4: one
// field per damped direction)
int dirs_mode(int scheme) { return scheme == 4 ? 2 : scheme == 3 ? 1 : 0; }

// Kernel op code of a context's sweeps: the operator, or 4 for TTI whose sweeps evaluate
// the rotation terms from the angles (float model).
int kernel_op(const a3d_ctx* c) {
    return c->d.op == A3D_OP_TTI && c->tti_angles && c->prec != PREC_FP64 ? 4 : c->d.op;
}

int check_desc_shot(const a3d_ctx* c, const a3d_shot_desc* d) {
    if (!same_geometry(c->d, *d)) return fail(A3D_EINVAL, "shot geometry differs from the context's");
    if (d->nt < 1) return fail(A3D_EINVAL, "nt must be positive");
    if (!(d->dt > 0) || !(d->fm > 0)) return fail(A3D_EINVAL, "dt and fm must be positive");
    if (d->rec_z < 0 || d->rec_z >= d->nz) return fail(A3D_EINVAL, "rec_z %d outside [0,%d)", d->rec_z, d->nz);
    if (d->snap_interval < 0) return fail(A3D_EINVAL, "snap_interval must be >= 0");
    if (d->scheme < 0 || d->scheme > 4) return fail(A3D_EINVAL, "scheme %d not in 0..4", d->scheme);
    if (d->scheme == 4 && c->strip_net && c->strip_head != 2)
        return fail(A3D_EINVAL, "scheme 4 keeps one field per damped direction; the strip's damping heads "
                                "scale the p and q split fields apart (use scheme 3 or the coefficient head)");
    if (d->scheme == 2 && (d->op != A3D_OP_VTI || c->prec != PREC_FP32 || c->strip_net))
        return fail(A3D_EINVAL, "scheme 2 (one-pass step) is built for the VTI operator, FP32 storage, PML boundary");
    if (d->op == A3D_OP_2ND) {
        if (d->sx < 0 || d->sx >= d->nx || d->sy < 0 || d->sy >= d->ny || d->sz < 0 || d->sz + 2 >= d->nz + d->nb)
            return fail(A3D_EINVAL, "source (%d,%d,%d) outside the model", d->sx, d->sy, d->sz);
    } else {
        if (d->sx < 1 || d->sx > d->nx || d->sy < 1 || d->sy > d->ny || d->sz < 1 || d->sz > d->nz)
            return fail(A3D_EINVAL, "source (%d,%d,%d) outside 1..n", d->sx, d->sy, d->sz);
        if (d->op == A3D_OP_ISO && d->sz + d->nb + 1 >= d->nz + 2 * d->nb - MM)
            return fail(A3D_EINVAL, "dipole partner of source z=%d outside the grid", d->sz);
        if (d->op == A3D_OP_ISO && (d->wtype < 1 || d->wtype > 6))
            return fail(A3D_EINVAL, "wavelet type %d not in 1..6", d->wtype);
        if (d->op != A3D_OP_ISO && (d->wtype < 1 || d->wtype > 3))
            return fail(A3D_EINVAL, "wavelet type %d not in 1..3", d->wtype);
    }
    return A3D_OK;
}

int ensure(float** p, size_t* cap, size_t n) {
    if (*cap >= n) return A3D_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, n * sizeof(float)) != hipSuccess) {
        *p = nullptr;
        return fail(A3D_ENOMEM, "device allocation of %zu bytes failed", n * sizeof(float));
    }
    *cap = n;
    return A3D_OK;
}

hipEvent_t timing_event(Timing& t) {
    if (t.used == t.pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        t.pool.push_back(e);
    }
    return t.pool[t.used++];
}

void timing_step(a3d_ctx* c, int it) { c->tm.sample = c->tm.stride <= 1 || it % c->tm.stride == 0; }

void timing_begin(a3d_ctx* c, int k) {
    if (!c->tm.on || !c->tm.sample) return;
    hipEvent_t e = timing_event(c->tm);
    if (!e) return;
    c->tm.marks.push_back({k, c->tm.used - 1});
    (void)hipEventRecord(e, c->st);
}

void timing_end(a3d_ctx* c) {
    if (!c->tm.on || !c->tm.sample || c->tm.marks.empty()) return;
    hipEvent_t e = timing_event(c->tm);
    if (e) (void)hipEventRecord(e, c->st);
}

void timing_collect(a3d_ctx* c) {
    Timing& t = c->tm;
    int seen[2] = {0, 0};
    for (auto& m : t.marks) {
        if (m.second + 1 >= t.used || seen[m.first]++ < t.skip) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.pool[m.second], t.pool[m.second + 1]) == hipSuccess) {
            t.ms[m.first] += ms;
            t.n[m.first] += 1;
        }
    }
    t.marks.clear();
    t.used = 0;
}

// The time loop of one shot captured into hipGraphs of at most c->graph_steps steps, each
// launched as soon as it is instantiated (the host captures the next chunk while the GPU
// runs this one); a chunk ends before a snapshot, whose host copy cannot be captured.
// Off while per-kernel timing is on (its events stay ordinary stream events).
struct StepGraphs {
    a3d_ctx* c;
    bool on;
    bool open = false;
    int n = 0;
    std::vector<hipGraphExec_t> execs;
    explicit StepGraphs(a3d_ctx* ctx) : c(ctx), on(ctx->graph_steps > 0 && !ctx->tm.on) {}
    int begin() {
        if (!on || open) return A3D_OK;
        HIP_TRY(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        open = true;
        n = 0;
        return A3D_OK;
    }
    int end() {
        if (!open) return A3D_OK;
        open = false;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->st, &g);
        if (e != hipSuccess || !g) return fail(A3D_EHIP, "time-loop capture failed: %s", hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(A3D_EHIP, "time-loop graph instantiation failed: %s", hipGetErrorString(ei));
        execs.push_back(ex);
        HIP_TRY(hipGraphLaunch(ex, c->st));
        return A3D_OK;
    }
    // after the launches of one step; cut = a snapshot or the last step follows
    int step_done(bool cut) {
        if (!open) return A3D_OK;
        return (++n >= c->graph_steps || cut) ? end() : A3D_OK;
    }
    ~StepGraphs() {
        if (open) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c->st, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        if (!execs.empty()) (void)hipStreamSynchronize(c->st);
        for (auto e : execs) (void)hipGraphExecDestroy(e);
    }
};

using SnapFn = std::function<void(int k, int it, const float* host)>;

// Power-of-two helper for the binary16 field scales.
float pow2_ceil(double v) { return v > 0 ? (float)std::ldexp(1.0, (int)std::ceil(std::log2(v))) : 1.0f; }

// Scales for binary16 wavefield storage (largest finite half: 65504). Pressures and
// split stresses: the near-source peak is ~1.2 pfac (measured on the oracle), mapped to
// <= 2^11. Particle velocities: |u| <= |P| / (rho v) for a plane wave, so their scale is
// the pressure scale over the model's smallest impedance, leaving >= 16x headroom for
// the near field. Subnormal halves extend the range down by 2^-24 relative to the scale.
Scales field_scales(const a3d_ctx* c, const a3d_shot_desc* d) {
    Scales sc{};
    for (int f = 0; f < NF; ++f) sc.s[f] = sc.inv[f] = 1.0f;
    if (c->prec != PREC_FP16) return sc;
    const double amp = d->op == A3D_OP_2ND ? 1.0 : std::max((double)std::fabs(d->pfac), 1e-30);
    const float sp = pow2_ceil(amp / 2048.0);
    const float su = pow2_ceil((double)sp / std::max((double)c->zmin, 1e-30));
    for (int f : {F_P, F_Q, F_PX, F_PY, F_PZ, F_QX, F_QY, F_QZ}) sc.s[f] = sp;
    for (int f : {F_U, F_V, F_W}) sc.s[f] = su;
    for (int f = 0; f < NF; ++f) sc.inv[f] = 1.0f / sc.s[f];
    return sc;
}

// The stress sweep can update w (from its own column's old P or Q, one z segment)
// for ISO and VTI in the two-sweep step (with the learned strip's coefficient head too:
// a strip cell's w takes the velocity sweep's multiplied z coefficients).
bool wz_eligible(const a3d_ctx* c, const a3d_shot_desc* d) {
    return (d->op == A3D_OP_ISO || d->op == A3D_OP_VTI) && d->scheme != 2 && c->zchunk == 0;
}
bool wz_active(const a3d_ctx* c, const a3d_shot_desc* d) {
    return wz_eligible(c, d) && (c->wz_mode == 1 || (c->wz_mode < 0 && c->wz_tuned));
}

// Tables, zeroed wavefields, kernel arguments and the per-step sources of one
// 1st-order shot (reference iso:372-671, vti:389-688, tti:390-676).
template <class T>
int prepare_first(a3d_ctx* c, const a3d_shot_desc* d, const Scales& sc, VelArgs& va, StressArgs& sa,
                  std::vector<Src>& srcs) {
    const Geo g = c->g;
    const int prec = c->prec;
    const int nx = d->nx, ny = d->ny, nz = d->nz, nb = d->nb, nt = d->nt;
    const size_t wbytes = c->nvox * c->wsize;
    const int op = d->op;
    const T dt = (T)d->dt;
    // get_d0 (reference iso:193-195) from the padded velocity at z-fastest index N/2
    const T dsum = (T)d->dx + (T)d->dy + (T)d->dz;
    // The learned strip ramps its PML over its STRIP_CELLS updated cells with amplitude
    // strip_amp in place of the reference's 10 over npml cells.
    const bool strip = c->strip_net != nullptr;
    const int width = strip ? STRIP_CELLS : nb;
    const double amp = strip ? (double)c->strip_amp : 10.0;
    const T d0 = (T)(amp * (T)c->vel_d0 * std::log(100000.0) / (2.0 * width * (dsum / 3.0)));
    std::vector<T> tabh(6 * (size_t)(g.nnx + g.nny + g.nnz));
    Pml pm{};
    pm.cx1 = 0;
    pm.cx2 = g.nnx;
    pm.ax1 = 2 * g.nnx;
    pm.ax2 = 3 * g.nnx;
    pm.cy1 = 4 * g.nnx;
    pm.cy2 = pm.cy1 + g.nny;
    pm.ay1 = pm.cy1 + 2 * g.nny;
    pm.ay2 = pm.cy1 + 3 * g.nny;
    pm.cz1 = 4 * g.nnx + 4 * g.nny;
    pm.cz2 = pm.cz1 + g.nnz;
    pm.az1 = pm.cz1 + 2 * g.nnz;
    pm.az2 = pm.cz1 + 3 * g.nnz;
    T* t = tabh.data();
    pml_axis<T>(dt, d0, nx, nb, t + pm.cx1, t + pm.cx2, t + pm.ax1, t + pm.ax2, width);
    pml_axis<T>(dt, d0, ny, nb, t + pm.cy1, t + pm.cy2, t + pm.ay1, t + pm.ay2, width);
    pml_axis<T>(dt, d0, nz, nb, t + pm.cz1, t + pm.cz2, t + pm.az1, t + pm.az2, width);
    pm.hvx = 4 * (g.nnx + g.nny + g.nnz);
    pm.hsx = pm.hvx + g.nnx;
    pm.hvy = pm.hsx + g.nnx;
    pm.hsy = pm.hvy + g.nny;
    pm.hvz = pm.hsy + g.nny;
    pm.hsz = pm.hvz + g.nnz;
    pml_half_steps<T>(dt, d0, nx, nb, t + pm.hvx, t + pm.hsx, width);
    pml_half_steps<T>(dt, d0, ny, nb, t + pm.hvy, t + pm.hsy, width);
    pml_half_steps<T>(dt, d0, nz, nb, t + pm.hvz, t + pm.hsz, width);
    // The planes whose z stress factor is exactly 1 form one run (the factor's argument
    // grows away from the interior on both sides); the stress sweep takes the run from
    // its bounds instead of reading the factor per plane.
    int zi0 = -1, zi1 = -1;
    for (int z = 0; z < g.nnz; ++z)
        if (t[pm.az1 + z] == (T)1) {
            if (zi0 < 0) zi0 = z;
            else if (zi1 >= 0) return fail(A3D_EINVAL, "z stress factor: planes of factor 1 are not one run");
        } else if (zi0 >= 0 && zi1 < 0) {
            zi1 = z;
        }
    if (zi0 < 0) zi0 = zi1 = 0;
    else if (zi1 < 0) zi1 = g.nnz;
    HIP_TRY(hipMemcpy(c->tab, tabh.data(), tabh.size() * sizeof(T), hipMemcpyHostToDevice));
    for (int v : volumes_of(op))
        if (is_wavefield(v))
            HIP_TRY(hipMemsetAsync(c->vol[v], 0, is_split(v) ? c->nvox * std::max(c->wsize, (size_t)4) : wbytes,
                                   c->st));

    const T dtx = dt / (T)d->dx, dty = dt / (T)d->dy, dtz = dt / (T)d->dz;
    const int sxp = d->sx + nb - 1, syp = d->sy + nb - 1, szp = d->sz + nb - 1;
    const int s0 = sxp + g.ld * syp + g.pl * szp;
    va = VelArgs{};
    va.g = g;
    va.l = make_launch(g, c->ty[0], c->zseg[0]);
    va.c = pm;
    va.tab = c->tab;
    va.P = c->vol[V_P];
    va.Q = op == A3D_OP_ISO ? c->vol[V_P] : c->vol[V_Q];
    va.rho = c->vol[V_RHO];
    va.ct = c->vol[V_CT];
    va.st = c->vol[V_ST];
    va.cp = c->vol[V_CP];
    va.sp = c->vol[V_SP];
    va.u = c->vol[V_U];
    va.v = c->vol[V_V];
    va.w = c->vol[V_W];
    va.dtx = dtx;
    va.dty = dty;
    va.dtz = dtz;
    va.sc = sc;
    sa = StressArgs{};
    sa.g = g;
    sa.l = make_launch(g, c->ty[1], c->zseg[1]);
    sa.c = pm;
    sa.tab = c->tab;
    sa.u = c->vol[V_U];
    sa.v = c->vol[V_V];
    sa.w = c->vol[V_W];
    sa.px = c->vol[V_PX];
    sa.py = c->vol[V_PY];
    sa.pz = c->vol[V_PZ];
    sa.qx = c->vol[V_QX];
    sa.qy = c->vol[V_QY];
    sa.qz = c->vol[V_QZ];
    sa.vp = c->vol[V_VP];
    sa.rho = c->vol[V_RHO];
    sa.eps = c->vol[V_EPS];
    sa.del = c->vol[V_DEL];
    sa.ct = c->vol[V_CT];
    sa.st = c->vol[V_ST];
    sa.cp = c->vol[V_CP];
    sa.sp = c->vol[V_SP];
    sa.P = c->vol[V_P];
    sa.Q = c->vol[V_Q];
    sa.dtx = dtx;
    sa.dty = dty;
    sa.dtz = dtz;
    sa.sv = d->sv && op != A3D_OP_ISO;
    sa.sxp = sxp;
    sa.syp = syp;
    sa.szp = szp;
    sa.taper = c->dtab + g.nnx + g.nny + g.nnz;
    sa.zrec = nb + d->rec_z;
    sa.nb = nb;
    sa.nx = nx;
    sa.ny = ny;
    sa.sc = sc;
    sa.unsplit = d->scheme == 4 ? 3 : d->scheme == 3 ? 2 : (d->scheme != 0 || prec == PREC_FP16) ? 1 : 0;
    sa.zi0 = zi0;
    sa.zi1 = zi1;
    va.wz = sa.wz = wz_active(c, d) ? 1 : 0;
    va.zrev = va.wz;   // no z window in the velocity sweep: its planes walked downward
    if (sa.wz) sa.l = make_launch(g, c->ty[1], 1);
    // The stress sweep's XCD runs balanced by estimated bytes per tile: a PML cell reads
    // and writes split fields (scheme 1: all six; schemes 3 and 4: those of its damped
    // directions), so the tile rows along the y faces weigh more than interior rows and
    // equal runs would give the XCDs that hold them more bytes. Schemes 0 and 2 move the
    // same bytes in every cell.
    if (sa.unsplit != 0 && d->scheme != 2) {
        const T* axf = t + pm.ax1;
        const T* ayf = t + pm.ay1;
        const T* azf = t + pm.az1;
        auto prefix = [](int n, auto damped) {
            std::vector<int> p(n + 1, 0);
            for (int i = 0; i < n; ++i) p[i + 1] = p[i] + (damped(i) ? 1 : 0);
            return p;
        };
        const std::vector<int> px = prefix(g.nnx, [&](int i) { return axf[i] != (T)1; });
        const std::vector<int> py = prefix(g.nny, [&](int i) { return ayf[i] != (T)1; });
        const std::vector<int> pz = prefix(g.nnz, [&](int i) { return azf[i] != (T)1; });
        // bytes per cell by the number of damped directions (VTI, w in the stress sweep)
        const double w1[4] = {48, 92, 92, 92}, w3[4] = {48, 56, 64, 68};
        const double* wc = sa.unsplit >= 2 ? w3 : w1;
        const Launch L = sa.l;
        const int tyr = c->ty[1];
        balance_bands(sa.l, [&](int tt) {
            const int tix = tt % L.ntx, r = tt / L.ntx, tiy = r % L.nty, seg = r / L.nty;
            const int x0 = std::max(tix * TX, MM), x1 = std::min(tix * TX + TX, g.nnx - MM);
            const int y0 = std::max(MM + tiy * tyr, MM), y1 = std::min(MM + tiy * tyr + tyr, g.nny - MM);
            const int z0 = L.zbeg + seg * L.zlen, z1 = std::min(z0 + L.zlen, L.zend);
            if (x1 <= x0 || y1 <= y0 || z1 <= z0) return 0.0;
            const int n[3][2] = {{x1 - x0 - (px[x1] - px[x0]), px[x1] - px[x0]},
                                 {y1 - y0 - (py[y1] - py[y0]), py[y1] - py[y0]},
                                 {z1 - z0 - (pz[z1] - pz[z0]), pz[z1] - pz[z0]}};
            double cost = 0;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    for (int e = 0; e < 2; ++e) cost += (double)n[0][a] * n[1][b] * n[2][e] * wc[a + b + e];
            return cost;
        });
    }
    // the stress sweep's (and the one-pass step's) interleaved model (VTI, float model),
    // built once per model
    // TTI (float model): the four rotation terms interleaved, one 16-byte load per point in
    // both sweeps; ISO (float model): (vp, rho), one 8-byte load in the stress sweep
    sa.m4 = nullptr;
    va.t4 = nullptr;
    if (op != A3D_OP_2ND && prec != PREC_FP64) {
        if (!c->vol[V_M4] && hipMalloc(&c->vol[V_M4], c->nvox * (op == A3D_OP_ISO ? 8 : 16)) != hipSuccess) {
            c->vol[V_M4] = nullptr;
            return fail(A3D_ENOMEM, "interleaved model: allocation failed");
        }
        const bool ang = kernel_op(c) == 4;
        if (!c->m4_ok || c->m4_ang != ang) {
            if (ang)
                HIP_TRY(launch_pack_angles((const float*)c->vol[V_CT], (const float*)c->vol[V_ST],
                                           (const float*)c->vol[V_CP], (const float*)c->vol[V_SP], c->vol[V_M4],
                                           c->nvox, c->st));
            else if (op == A3D_OP_VTI)
                HIP_TRY(launch_pack4((const float*)c->vol[V_VP], (const float*)c->vol[V_RHO],
                                     (const float*)c->vol[V_EPS], (const float*)c->vol[V_DEL], c->vol[V_M4], c->nvox,
                                     c->st));
            else if (op == A3D_OP_ISO)
                HIP_TRY(launch_pack2((const float*)c->vol[V_VP], (const float*)c->vol[V_RHO], c->vol[V_M4], c->nvox,
                                     c->st));
            else
                HIP_TRY(launch_pack4((const float*)c->vol[V_CT], (const float*)c->vol[V_ST],
                                     (const float*)c->vol[V_CP], (const float*)c->vol[V_SP], c->vol[V_M4], c->nvox,
                                     c->st));
            c->m4_ok = true;
            c->m4_ang = ang;
        }
        sa.m4 = c->vol[V_M4];
        if (op == A3D_OP_TTI) va.t4 = c->vol[V_M4];
    }
    // learned strip, coefficient head: the multipliers start at 1 (the plain ramp) until the
    // network first runs
    va.mfac = sa.mfac = nullptr;
    if (c->strip_net && c->strip_head == 2) {
        if (!c->strip_mfac && hipMalloc(&c->strip_mfac, c->nvox * sizeof(float)) != hipSuccess) {
            c->strip_mfac = nullptr;
            return fail(A3D_ENOMEM, "strip multiplier allocation failed");
        }
        HIP_TRY(launch_fill(c->strip_mfac, c->nvox, 1.0f, c->st));
        va.mfac = sa.mfac = c->strip_mfac;
    }
    va.s.s0 = s0;
    // The time grid is the reference's float accumulation t = dt, t += dt and the gate
    // t <= 2*tdelay is taken in float (iso:90, :521) in every precision.
    const float tdelay = (float)(1.0 / d->fm);
    const T pfac = (T)d->pfac;
    srcs.assign(nt, Src{-1, -1, 0.0, 0.0, 0.0, 0.0});
    float tcur = d->dt;
    for (int it = 0; it < nt; it++, tcur += d->dt) {
        // add_source (reference iso:523-530; vti:557; tti:571), applied on load
        if (tcur <= 2 * tdelay) {
            const T src = wavelet_value<T>(d->wtype, tcur, d->fm);
            if (d->op == A3D_OP_ISO) {
                const T mp = (T)(-1.0 * pfac);
                const T a0 = mp * src, a1 = pfac * src;
                srcs[it] = {s0, s0 + 2 * g.pl, (double)a0, (double)a1, (double)a0, (double)a1};
            } else {
                const T a0 = pfac * src;
                srcs[it] = {s0, -1, (double)a0, 0.0, (double)a0, 0.0};
            }
        }
    }
    va.s = Src{-1, -1, 0.0, 0.0, 0.0, 0.0};
    return A3D_OK;
}

StripArgs strip_args(const a3d_ctx* c, const a3d_shot_desc* d) {
    StripArgs a{};
    a.g = c->g;
    a.ids = c->strip_ids;
    a.meta = c->strip_meta;
    a.n = c->strip_n;
    a.P = (const float*)c->vol[V_P];
    a.Q = (const float*)c->vol[V_Q];
    a.u = (const float*)c->vol[V_U];
    a.v = (const float*)c->vol[V_V];
    a.w = (const float*)c->vol[V_W];
    a.vp = (const float*)c->vol[V_VP];
    a.rho = (const float*)c->vol[V_RHO];
    a.eps = (const float*)c->vol[V_EPS];
    a.del = (const float*)c->vol[V_DEL];
    a.cdt[0] = d->dt / d->dx;
    a.cdt[1] = d->dt / d->dy;
    a.cdt[2] = d->dt / d->dz;
    a.dirsplit = d->scheme >= 3;
    a.normal = c->strip_head == 1;
    a.mfac = c->strip_head == 2 ? c->strip_mfac : nullptr;
    a.vref = c->vel_d0;
    return a;
}

// Learned absorbing strip after the stress sweep of step it, on stream st: the damping
// heads (0, 1) evaluate the network and apply its factors; the coefficient head (2) writes
// the cells' damping multipliers, read by the next steps' sweeps, every strip_period steps.
int strip_step_on(a3d_ctx* c, const StripArgs& a, hipStream_t st, int it) {
    if (c->strip_head == 2) {
        if (c->strip_period > 0 ? it % c->strip_period == 0 : it == 0)
            HIP_TRY(launch_strip_infer(a, c->strip_net, c->strip_delta, c->strip_mfma, st));
        return A3D_OK;
    }
    HIP_TRY(launch_strip_infer(a, c->strip_net, c->strip_delta, c->strip_mfma, st));
    HIP_TRY(launch_strip_apply(a, c->strip_delta, (float*)c->vol[V_P], (float*)c->vol[V_Q], (float*)c->vol[V_PX],
                               (float*)c->vol[V_PY], (float*)c->vol[V_PZ], (float*)c->vol[V_QX], (float*)c->vol[V_QY],
                               (float*)c->vol[V_QZ], (float*)c->vol[V_U], (float*)c->vol[V_V], (float*)c->vol[V_W],
                               st));
    return A3D_OK;
}

int strip_step(a3d_ctx* c, const StripArgs& a, int it) { return strip_step_on(c, a, c->st, it); }

template <class T>
int run_shot_t(a3d_ctx* c, const a3d_shot_desc* d, float* shot_out, const SnapFn& snapfn, int* nsnap_taken,
               double* loop_ms) {
    const Geo g = c->g;
    const int prec = c->prec;
    const int nx = d->nx, ny = d->ny, nz = d->nz, nt = d->nt;
    const size_t plane_rec = (size_t)nx * ny;
    int rc;
    if ((rc = ensure(&c->rec, &c->rec_cap, plane_rec * nt))) return rc;
    const Scales sc = field_scales(c, d);

    std::vector<float> hsnap;
    const bool snaps = d->snap_interval > 0 && snapfn;
    if (snaps) {
        if (!c->snap_dev && hipMalloc(&c->snap_dev, (size_t)nx * ny * nz * sizeof(float)) != hipSuccess) {
            c->snap_dev = nullptr;
            return fail(A3D_ENOMEM, "snapshot buffer allocation failed");
        }
        hsnap.resize((size_t)nx * ny * nz);
    }
    auto want_snap = [&](int it) {
        return d->snap_interval > 0 && it % d->snap_interval == 0 && !(d->snap_skip0 && it == 0);
    };
    int nsnap = 0;
    auto take_snap = [&](int it, const void* src, int s0, double a0, int s1, double a1, bool sponge) -> int {
        ++nsnap;
        if (!snaps) return A3D_OK;
        const double* fx = sponge ? c->dtab : nullptr;
        const double* fy = sponge ? c->dtab + g.nnx : nullptr;
        const double* fz = sponge ? c->dtab + g.nnx + g.nny : nullptr;
        HIP_TRY(launch_snap(prec, src, g, c->off, nx, ny, nz, s0, a0, s1, a1, fx, fy, fz, sc.s[F_P], c->snap_dev,
                            c->st));
        HIP_TRY(hipMemcpyAsync(hsnap.data(), c->snap_dev, hsnap.size() * sizeof(float), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
        snapfn(nsnap - 1, it, hsnap.data());
        return A3D_OK;
    };
    const size_t wbytes = c->nvox * c->wsize;

    if (d->op == A3D_OP_2ND) {
        // -------- 2nd order (reference cuda_3dfd, Toa_3dfd_cuda.cu:366-508)
        if (c->vv_dt != d->dt) {
            HIP_TRY(launch_pad(c->stage, nx, ny, nz, c->off, g, c->vol[V_VV], prec == PREC_FP64, 1, d->dt, c->st));
            c->vv_dt = d->dt;
        }
        HIP_TRY(hipMemsetAsync(c->vol[V_PA], 0, wbytes, c->st));
        HIP_TRY(hipMemsetAsync(c->vol[V_PB], 0, wbytes, c->st));
        std::vector<T> wlt(nt);
        for (int it = 0; it < nt; ++it) wlt[it] = ricker_value<T>(it, d->fm, d->dt);
        Step2Args a{};
        a.g = g;
        a.l = make_launch(g, c->ty[0], c->zseg[0]);
        a.vv = c->vol[V_VV];
        a.fx = c->dtab;
        a.fy = c->dtab + g.nnx;
        a.fz = c->dtab + g.nnx + g.nny;
        a.dz2 = (T)(1.0 / ((T)d->dz * (T)d->dz));
        a.dx2 = (T)(1.0 / ((T)d->dx * (T)d->dx));
        a.dy2 = (T)(1.0 / ((T)d->dy * (T)d->dy));
        a.sc = sc;
        const int off = c->off;
        a.s0 = (d->sx + off) + g.ld * (d->sy + off) + g.pl * (d->sz + off);
        a.s1 = a.s0 + 2 * g.pl;
        a.zrec = off + d->rec_z;
        a.off = off;
        a.nx = nx;
        a.ny = ny;
        void* X = c->vol[V_PA];
        void* Y = c->vol[V_PB];
        HIP_TRY(mark(c->ev0, c->st));
        StepGraphs sg(c);
        for (int it = 0; it < nt; ++it) {
            if ((rc = sg.begin())) return rc;
            a.X = X;
            a.Y = Y;
            a.ya0 = wlt[it];
            a.ya1 = -wlt[it];
            a.xa0 = it > 0 ? (double)wlt[it - 1] : 0.0;
            a.xa1 = it > 0 ? -(double)wlt[it - 1] : 0.0;
            a.rec = c->rec + plane_rec * it;
            timing_step(c, it);
            timing_begin(c, 0);
            HIP_TRY(launch_step2(prec, c->ty[0], a, c->st));
            timing_end(c);
            if ((rc = sg.step_done(want_snap(it) || it == nt - 1))) return rc;
            if (want_snap(it)) {
                if ((rc = take_snap(it, Y, a.s0, a.ya0, a.s1, a.ya1, true))) return rc;
            }
            std::swap(X, Y);
        }
        HIP_TRY(mark(c->ev1, c->st));
    } else {
        // -------- 1st-order family
        VelArgs va;
        StressArgs sa;
        std::vector<Src> srcs;
        if ((rc = prepare_first<T>(c, d, sc, va, sa, srcs))) return rc;
        const int kop = kernel_op(c);
        const bool strip = c->strip_net != nullptr;
        const StripArgs stp = strip_args(c, d);
        if (d->scheme == 2) {
            // one-pass step: old and new stresses/velocities in separate buffers, swapped every step
            for (int v : {V_P2, V_Q2, V_U2, V_V2, V_W2}) {
                if (!c->vol[v] && hipMalloc(&c->vol[v], wbytes) != hipSuccess) {
                    c->vol[v] = nullptr;
                    return fail(A3D_ENOMEM, "one-pass step buffers: allocation failed");
                }
                HIP_TRY(hipMemsetAsync(c->vol[v], 0, wbytes, c->st));
            }
            FusedArgs fa{};
            fa.g = g;
            fa.l = make_launch(g, c->ty[1], c->zseg[1]);
            fa.c = sa.c;
            fa.tab = c->tab;
            fa.px = (float*)c->vol[V_PX];
            fa.py = (float*)c->vol[V_PY];
            fa.pz = (float*)c->vol[V_PZ];
            fa.qx = (float*)c->vol[V_QX];
            fa.qy = (float*)c->vol[V_QY];
            fa.qz = (float*)c->vol[V_QZ];
            fa.vp = (const float*)c->vol[V_VP];
            fa.rho = (const float*)c->vol[V_RHO];
            fa.eps = (const float*)c->vol[V_EPS];
            fa.del = (const float*)c->vol[V_DEL];
            fa.dtx = sa.dtx;
            fa.dty = sa.dty;
            fa.dtz = sa.dtz;
            fa.sv = sa.sv;
            fa.sxp = sa.sxp;
            fa.syp = sa.syp;
            fa.szp = sa.szp;
            fa.taper = sa.taper;
            fa.zrec = sa.zrec;
            fa.nb = sa.nb;
            fa.nx = sa.nx;
            fa.ny = sa.ny;
            fa.m4 = sa.m4;
            fa.zi0 = sa.zi0;
            fa.zi1 = sa.zi1;
            float* cur[5] = {(float*)c->vol[V_P], (float*)c->vol[V_Q], (float*)c->vol[V_U], (float*)c->vol[V_V],
                             (float*)c->vol[V_W]};
            float* nxt[5] = {(float*)c->vol[V_P2], (float*)c->vol[V_Q2], (float*)c->vol[V_U2], (float*)c->vol[V_V2],
                             (float*)c->vol[V_W2]};
            HIP_TRY(mark(c->ev0, c->st));
            StepGraphs sg(c);
            for (int it = 0; it < nt; it++) {
                if ((rc = sg.begin())) return rc;
                fa.P0 = cur[0];
                fa.Q0 = cur[1];
                fa.u0 = cur[2];
                fa.v0 = cur[3];
                fa.w0 = cur[4];
                fa.P1 = nxt[0];
                fa.Q1 = nxt[1];
                fa.u1 = nxt[2];
                fa.v1 = nxt[3];
                fa.w1 = nxt[4];
                fa.s = srcs[it];
                fa.rec = c->rec + plane_rec * it;
                timing_step(c, it);
                timing_begin(c, 1);
                HIP_TRY(launch_fused_vti(c->ty[1], fa, c->st));
                timing_end(c);
                if ((rc = sg.step_done(want_snap(it) || it == nt - 1))) return rc;
                for (int k = 0; k < 5; ++k) std::swap(cur[k], nxt[k]);
                if (want_snap(it)) {
                    if ((rc = take_snap(it, cur[0], -1, 0.0, -1, 0.0, false))) return rc;
                }
            }
            HIP_TRY(mark(c->ev1, c->st));
        } else {
        // z chunks: velocity sweep of chunk k+1 before the stress sweep of chunk k (its w window
        // reaches 3 planes into chunk k+1; the velocity sweep of chunk k+1 reads the stresses
        // of chunk k before they are updated), so the velocities a stress chunk reads were
        // written one launch earlier (the same arithmetic as whole-grid sweeps, bit for bit)
        std::vector<VelArgs> vch;
        std::vector<StressArgs> sch;
        for (int z0 = MM; c->zchunk > 0 && z0 < g.nnz - MM; z0 += c->zchunk) {
            const int z1 = std::min(z0 + c->zchunk, g.nnz - MM);
            vch.push_back(va);
            vch.back().l = make_launch(g, c->ty[0], c->zseg[0], z0, z1);
            sch.push_back(sa);
            sch.back().l = make_launch(g, c->ty[1], c->zseg[1], z0, z1);
        }
        const int nch = (int)vch.size();
        HIP_TRY(mark(c->ev0, c->st));
        StepGraphs sg(c);
        for (int it = 0; it < nt; it++) {
            if ((rc = sg.begin())) return rc;
            va.s = srcs[it];
            sa.s = srcs[it];
            sa.rec = c->rec + plane_rec * it;
            timing_step(c, it);
            if (nch > 0) {
                for (int k = 0; k <= nch; ++k) {
                    if (k < nch) {
                        vch[k].s = srcs[it];
                        timing_begin(c, 0);
                        HIP_TRY(launch_vel(prec, kop, c->ty[0], vch[k], c->st));
                        timing_end(c);
                    }
                    if (k >= 1) {
                        sch[k - 1].rec = sa.rec;
                        timing_begin(c, 1);
                        HIP_TRY(launch_stress(prec, kop, c->ty[1], sch[k - 1], c->st));
                        timing_end(c);
                    }
                }
            } else {
                timing_begin(c, 0);
                HIP_TRY(launch_vel(prec, kop, c->ty[0], va, c->st));
                timing_end(c);
                timing_begin(c, 1);
                HIP_TRY(launch_stress(prec, kop, c->ty[1], sa, c->st));
                timing_end(c);
            }
            if (strip && (rc = strip_step(c, stp, it))) return rc;
            if ((rc = sg.step_done(want_snap(it) || it == nt - 1))) return rc;
            if (want_snap(it)) {
                if ((rc = take_snap(it, c->vol[V_P], -1, 0.0, -1, 0.0, false))) return rc;
            }
        }
        HIP_TRY(mark(c->ev1, c->st));
        }
    }
    float ms = 0.f;
    if (c->ev1) {
        HIP_TRY(hipEventSynchronize(c->ev1));
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    }
    if (loop_ms) *loop_ms = ms;
    if (c->tm.on) timing_collect(c);
    if (nsnap_taken) *nsnap_taken = nsnap;

    // non-finite sentinel over the whole record
    HIP_TRY(hipMemsetAsync(c->flag, 0, sizeof(int), c->st));
    HIP_TRY(launch_check_finite(c->rec, plane_rec * nt, c->flag, c->st));
    int hflag = 0;
    HIP_TRY(hipMemcpyAsync(&hflag, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));

    if (shot_out || c->out_device) {
        // out_device: c->shot_dev is the caller's device buffer of plane_rec * nt floats
        if (!c->out_device && (rc = ensure(&c->shot_dev, &c->shot_cap, plane_rec * nt))) return rc;
        MuteArgs m{};
        m.mute = d->mute;
        m.dx = d->dx;
        m.dy = d->dy;
        m.dz = d->dz;
        m.dt = d->dt;
        m.fm = d->fm;
        m.sx = d->sx;
        m.sy = d->sy;
        m.sz = d->sz;
        m.tt = d->mute_tt;
        if (d->v_mute > 0) m.vcut = d->v_mute;
        else if (d->op == A3D_OP_2ND) m.vcut = c->vel_raw1;
        else if (d->op == A3D_OP_ISO) m.vcut = c->vel_first;
        else m.vcut = c->vel_first * sqrtf(1 + 2 * c->eps_first);
        HIP_TRY(launch_shot_transpose(c->rec, nt, nx, ny, m, c->shot_dev, c->st));
        if (shot_out)
            HIP_TRY(hipMemcpyAsync(shot_out, c->shot_dev, plane_rec * nt * sizeof(float), hipMemcpyDeviceToHost,
                                   c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
    }
    if (hflag) return fail(A3D_EUNSTABLE, "non-finite values in the shot record (unstable time step?)");
    return A3D_OK;
}

// Several 2nd-order shots on the context's model in one launch per step (k_step2_shots):
// shot k's source at src[3k..3k+2] (the descriptor's source ignored), records into
// shots_out + k * ny * nx * nt. No snapshots. Each shot's record equals its own run's.
template <class T>
int run_shots_2nd_t(a3d_ctx* c, const a3d_shot_desc* d, int n, const int* src, float* shots_out, double* loop_ms) {
    const Geo g = c->g;
    const int prec = c->prec;
    const int nx = d->nx, ny = d->ny, nz = d->nz, nt = d->nt;
    const size_t plane_rec = (size_t)nx * ny, rec_n = plane_rec * nt;
    const size_t wbytes = c->nvox * c->wsize;
    int rc;
    if (c->multi_cap < 2 * n * wbytes) {
        if (c->multi) (void)hipFree(c->multi);
        c->multi = nullptr;
        c->multi_cap = 0;
        if (hipMalloc(&c->multi, 2 * n * wbytes) != hipSuccess) {
            c->multi = nullptr;
            return fail(A3D_ENOMEM, "multi-shot levels: allocation of %zu bytes failed", 2 * n * wbytes);
        }
        c->multi_cap = 2 * n * wbytes;
    }
    if ((rc = ensure(&c->mrec, &c->mrec_cap, rec_n * n))) return rc;
    if (c->vv_dt != d->dt) {
        HIP_TRY(launch_pad(c->stage, nx, ny, nz, c->off, g, c->vol[V_VV], prec == PREC_FP64, 1, d->dt, c->st));
        c->vv_dt = d->dt;
    }
    unsigned char* X = (unsigned char*)c->multi;
    unsigned char* Y = X + n * wbytes;
    HIP_TRY(hipMemsetAsync(c->multi, 0, 2 * n * wbytes, c->st));
    std::vector<T> wlt(nt);
    for (int it = 0; it < nt; ++it) wlt[it] = ricker_value<T>(it, d->fm, d->dt);
    Step2Args a{};
    a.g = g;
    const int ty = n > 1 && c->mshape ? c->mty : c->ty[0];
    a.l = make_launch(g, ty, n > 1 && c->mshape ? c->mzseg : c->zseg[0]);
    a.l.ntiles *= n;
    uniform_bands(a.l);
    a.vv = c->vol[V_VV];
    a.fx = c->dtab;
    a.fy = c->dtab + g.nnx;
    a.fz = c->dtab + g.nnx + g.nny;
    a.dz2 = (T)(1.0 / ((T)d->dz * (T)d->dz));
    a.dx2 = (T)(1.0 / ((T)d->dx * (T)d->dx));
    a.dy2 = (T)(1.0 / ((T)d->dy * (T)d->dy));
    a.sc = field_scales(c, d);
    const int off = c->off;
    a.zrec = off + d->rec_z;
    a.off = off;
    a.nx = nx;
    a.ny = ny;
    Step2Shots sh{};
    sh.n = n;
    sh.stride = wbytes;
    sh.rstride = rec_n;
    for (int k = 0; k < n; ++k) {
        sh.s0[k] = (src[3 * k] + off) + g.ld * (src[3 * k + 1] + off) + g.pl * (src[3 * k + 2] + off);
        sh.s1[k] = sh.s0[k] + 2 * g.pl;
    }
    HIP_TRY(mark(c->ev0, c->st));
    for (int it = 0; it < nt; ++it) {
        a.X = X;
        a.Y = Y;
        a.ya0 = wlt[it];
        a.ya1 = -wlt[it];
        a.xa0 = it > 0 ? (double)wlt[it - 1] : 0.0;
        a.xa1 = it > 0 ? -(double)wlt[it - 1] : 0.0;
        a.rec = c->mrec + plane_rec * it;
        timing_step(c, it);
        timing_begin(c, 0);
        HIP_TRY(launch_step2_shots(prec, ty, a, sh, c->st));
        timing_end(c);
        std::swap(X, Y);
    }
    HIP_TRY(mark(c->ev1, c->st));
    float ms = 0.f;
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (loop_ms) *loop_ms = ms;
    if (c->tm.on) timing_collect(c);
    HIP_TRY(hipMemsetAsync(c->flag, 0, sizeof(int), c->st));
    HIP_TRY(launch_check_finite(c->mrec, rec_n * n, c->flag, c->st));
    int hflag = 0;
    HIP_TRY(hipMemcpyAsync(&hflag, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (shots_out) {
        if ((rc = ensure(&c->shot_dev, &c->shot_cap, rec_n))) return rc;
        for (int k = 0; k < n; ++k) {
            MuteArgs m{};
            m.mute = d->mute;
            m.dx = d->dx;
            m.dy = d->dy;
            m.dz = d->dz;
            m.dt = d->dt;
            m.fm = d->fm;
            m.sx = src[3 * k];
            m.sy = src[3 * k + 1];
            m.sz = src[3 * k + 2];
            m.tt = d->mute_tt;
            m.vcut = d->v_mute > 0 ? d->v_mute : c->vel_raw1;
            HIP_TRY(launch_shot_transpose(c->mrec + k * rec_n, nt, nx, ny, m, c->shot_dev, c->st));
            HIP_TRY(hipMemcpyAsync(shots_out + k * rec_n, c->shot_dev, rec_n * sizeof(float), hipMemcpyDeviceToHost,
                                   c->st));
            HIP_TRY(hipStreamSynchronize(c->st));
        }
    }
    if (hflag) return fail(A3D_EUNSTABLE, "non-finite values in a shot record (unstable time step?)");
    return A3D_OK;
}

int run_shot(a3d_ctx* c, const a3d_shot_desc* d, float* shot_out, const SnapFn& snapfn, int* nsnap_taken,
             double* loop_ms) {
    int rc = check_desc_shot(c, d);
    if (rc) return rc;
    if (!c->model_set) return fail(A3D_ESTATE, "model not set");
    HIP_TRY(hipSetDevice(c->device));
    if (c->prec == PREC_FP64) return run_shot_t<double>(c, d, shot_out, snapfn, nsnap_taken, loop_ms);
    return run_shot_t<float>(c, d, shot_out, snapfn, nsnap_taken, loop_ms);
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// Autotuner: per sweep, the tile rows TY and z segmentation with the lowest
// measured launch time on this grid, from a few real time steps per candidate
// (every candidate gives bit-identical results). Results are cached per
// process by (device, operator, padded grid).
// ---------------------------------------------------------------------------
struct TuneKey {
    int dev, op, prec, scheme, zchunk, nnx, nny, nnz;
    bool operator==(const TuneKey& o) const {
        return dev == o.dev && op == o.op && prec == o.prec && scheme == o.scheme && zchunk == o.zchunk &&
               nnx == o.nnx && nny == o.nny && nnz == o.nnz;
    }
};
struct TuneVal {
    int ty[2], zseg[2], wz;
};
std::vector<std::pair<TuneKey, TuneVal>> g_tune_cache;
std::mutex g_tune_mu;

int autotune(a3d_ctx* c, const a3d_shot_desc* d, int steps, double* best_ms) {
    const TuneKey key{c->device, kernel_op(c), c->prec, d->scheme * 4 + (c->wz_mode + 1), c->zchunk,
                      c->g.nnx, c->g.nny, c->g.nnz};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (auto& kv : g_tune_cache)
            if (kv.first == key) {
                for (int k = 0; k < 2; ++k) {
                    c->ty[k] = kv.second.ty[k];
                    c->zseg[k] = kv.second.zseg[k];
                }
                c->wz_tuned = kv.second.wz;
                c->tuned = true;
                return A3D_OK;
            }
    }
    a3d_shot_desc td = *d;
    td.nt = std::max(2, steps);
    td.snap_interval = 0;
    td.mute = 0;
    const Timing saved = c->tm;
    const int nzi = c->g.nnz - 2 * MM;
    // best[m][k]: fastest launch of sweep k with w updated by sweep m (0 velocity, 1 stress);
    // the two modes are compared on the sum of their two sweeps
    double best[2][2] = {{1e30, 1e30}, {1e30, 1e30}};
    int bty[2][2], bzs[2][2];
    for (int m = 0; m < 2; ++m)
        for (int k = 0; k < 2; ++k) {
            bty[m][k] = c->ty[k];
            bzs[m][k] = c->zseg[k];
        }
    const int tys[3] = {2, 4, 8};
    const int segs[9] = {1, 2, 3, 4, 6, 8, 12, 16, 24};
    const bool elig = wz_eligible(c, d);
    const int m0 = elig && c->wz_mode == 1 ? 1 : 0, m1 = elig && c->wz_mode != 0 ? 1 : 0;
    int rc = A3D_OK;
    c->tm = Timing();
    c->tm.pool = saved.pool;   // reuse the event pool
    c->tm.on = true;
    c->tm.skip = td.nt >= 6 ? 2 : 0;   // the first launches of a shot run on cold caches
    c->tuned = true;           // run_shot below must not re-enter the tuner
    const int wz_mode = c->wz_mode;
    for (int m = m0; m <= m1 && !rc; ++m) {
        c->wz_mode = m;
        // the 2nd-order step also tries segments of 4 and 2 planes (k_step2_blk: every load of
        // a segment at once), which pay on grids too small for plane-streaming tiles
        std::vector<int> cand(segs, segs + 9);
        if (c->d.op == A3D_OP_2ND)
            for (int zl : {4, 2}) cand.push_back((nzi + zl - 1) / zl);
        for (int ty : tys) {
            for (int ns : cand) {
                if (ns > 1 && nzi / ns < 8 && !(c->d.op == A3D_OP_2ND && (nzi + ns - 1) / ns <= 4)) continue;
                for (int k = 0; k < 2; ++k) {
                    c->ty[k] = ty;
                    c->zseg[k] = ns;
                }
                c->tm.ms[0] = c->tm.ms[1] = 0;
                c->tm.n[0] = c->tm.n[1] = 0;
                rc = run_shot(c, &td, nullptr, SnapFn(), nullptr, nullptr);
                if (rc && rc != A3D_EUNSTABLE) break;
                rc = A3D_OK;
                for (int k = 0; k < 2; ++k) {
                    if (c->tm.n[k] == 0) continue;
                    const double avg = c->tm.ms[k] / c->tm.n[k];
                    if (avg < best[m][k]) {
                        best[m][k] = avg;
                        bty[m][k] = ty;
                        bzs[m][k] = ns;
                    }
                }
            }
            if (rc) break;
        }
    }
    c->wz_mode = wz_mode;
    const int bm = m0 == m1 ? m0 : (best[1][0] + best[1][1] < best[0][0] + best[0][1] ? 1 : 0);
    std::vector<hipEvent_t> pool = c->tm.pool;
    c->tm = saved;
    c->tm.pool = pool;
    c->tm.used = 0;
    c->tm.marks.clear();
    for (int k = 0; k < 2; ++k) {
        c->ty[k] = bty[bm][k];
        c->zseg[k] = bzs[bm][k];
    }
    if (bm == 1) c->zseg[1] = 1;   // the stress sweep that updates w runs one z segment
    c->wz_tuned = bm;
    if (rc) {
        c->tuned = false;
        return rc;
    }
    if (best_ms) {
        best_ms[0] = best[bm][0] < 1e29 ? best[bm][0] : 0.0;
        best_ms[1] = best[bm][1] < 1e29 ? best[bm][1] : 0.0;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache.push_back({key, TuneVal{{c->ty[0], c->ty[1]}, {c->zseg[0], c->zseg[1]}, bm}});
    return A3D_OK;
}

// Tunes on the first shot of a grid of >= 1M updated points with nt >= 100
// unless A3D_AUTOTUNE=0 or a shape was set explicitly.
int maybe_autotune(a3d_ctx* c, const a3d_shot_desc* d) {
    if (c->tuned) return A3D_OK;
    const char* e = getenv("A3D_AUTOTUNE");
    if (e && e[0] == '0') return A3D_OK;
    const double pts = (double)(c->g.nnx - 2 * MM) * (c->g.nny - 2 * MM) * (c->g.nnz - 2 * MM);
    if (pts < 1.0e6 || d->nt < 100) return A3D_OK;
    if (check_desc_shot(c, d) || !c->model_set) return A3D_OK;   // run_shot reports the error
    return autotune(c, d, 10, nullptr);
}

}  // namespace

// ===========================================================================
// Handle API
// ===========================================================================
namespace {

// Model-derived scalars of a context whose interior is the window (w.o*, c->d.n*) of a
// z-fastest host model of size w.n*: the d0 sample, the mute velocity samples, the
// smallest impedance (binary16 scales).
// Interior cells (iz, ix, iy) of the model samples a shot's scalars come from. 1st order:
// the d0 sample, padded z-fastest linear index N/2 (reference get_d0, iso:194), and the
// mute-velocity sample, padded index 1 = (iz=1, ix=0, iy=0) (mute_directwave, iso:340,
// vti:357), both clamped into the interior (edge replication); 2nd order: the interior
// sample (iz=1, ix=0, iy=0) (Toa_3dfd_cuda.cu:403) in `first`.
struct Samples {
    int d0[3], first[3];
};

Samples model_samples(const a3d_ctx* c) {
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    Samples s{};
    if (c->d.op != A3D_OP_2ND) {
        const int nb = c->d.nb;
        const int nnx = nx + 2 * nb, nny = ny + 2 * nb, nnz = nz + 2 * nb;
        const long long i = (long long)nny * nnx * nnz / 2;
        const int iz = (int)(i % nnz), ix = (int)((i / nnz) % nnx), iy = (int)(i / ((long long)nnz * nnx));
        s.d0[0] = clampi(iz - nb, 0, nz - 1);
        s.d0[1] = clampi(ix - nb, 0, nx - 1);
        s.d0[2] = clampi(iy - nb, 0, ny - 1);
        s.first[0] = clampi(1 - nb, 0, nz - 1);
        s.first[1] = clampi(-nb, 0, nx - 1);
        s.first[2] = clampi(-nb, 0, ny - 1);
    } else {
        s.first[0] = nz > 1 ? 1 : 0;
    }
    return s;
}

// Model-derived scalars of a context whose interior is the window (w.o*, c->d.n*) of a
// z-fastest host model of size w.n*: the d0 sample, the mute velocity samples, the
// smallest impedance (binary16 scales).
void model_scalars(a3d_ctx* c, const float* vel, const float* rho, const float* eps, const Window& w) {
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    auto at = [&](const float* a, int iz, int ix, int iy) {
        return a[(size_t)(iz + w.oz) + (size_t)w.nz * (ix + w.ox) + (size_t)w.nz * w.nx * (iy + w.oy)];
    };
    const Samples sm = model_samples(c);
    if (c->d.op != A3D_OP_2ND) {
        c->vel_d0 = at(vel, sm.d0[0], sm.d0[1], sm.d0[2]);
        c->vel_first = at(vel, sm.first[0], sm.first[1], sm.first[2]);
        c->eps_first = eps ? at(eps, sm.first[0], sm.first[1], sm.first[2]) : 0.0f;
        // smallest impedance rho*v (TTI has no density: rho = 1)
        double zmin = 1e300;
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
                for (int iz = 0; iz < nz; ++iz)
                    zmin = std::min(zmin, (double)at(vel, iz, ix, iy) * (rho ? (double)at(rho, iz, ix, iy) : 1.0));
        c->zmin = (float)std::max(zmin, 1e-30);
    } else {
        c->vel_raw1 = at(vel, sm.first[0], sm.first[1], sm.first[2]);
    }
}

// Geometry, precision and element sizes of a context for *desc on `device`.
void init_geometry(a3d_ctx* c, const a3d_shot_desc* desc, int device) {
    c->d = *desc;
    c->device = device;
    const int op = desc->op;
    c->off = op == A3D_OP_2ND ? MM + desc->nb : desc->nb;
    const int ring = op == A3D_OP_2ND ? 2 * MM + 2 * desc->nb : 2 * desc->nb;
    c->g.nnx = desc->nx + ring;
    c->g.nny = desc->ny + ring;
    c->g.nnz = desc->nz + ring;
    c->g.ld = (c->g.nnx + TX - 1) / TX * TX;
    c->g.pl = c->g.ld * c->g.nny;
    c->nvox = (size_t)c->g.pl * c->g.nnz;
    c->prec = desc->storage == 1 ? PREC_FP16 : (desc->storage == 2 ? PREC_FP64 : PREC_FP32);
    c->wsize = c->prec == PREC_FP16 ? 2 : (c->prec == PREC_FP64 ? 8 : 4);
    c->msize = c->prec == PREC_FP64 ? 8 : 4;
    c->csize = c->msize;
    if (int ncu = 0; hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess)
        c->ncu = ncu;
}

// Elements of the PML table (arithmetic type) and of the double table (sponge, taper).
size_t table_count(const a3d_ctx* c) { return 6 * (size_t)(c->g.nnx + c->g.nny + c->g.nnz); }
size_t dtable_count(const a3d_ctx* c) { return (size_t)c->g.nnx + c->g.nny + c->g.nnz + 226; }

// Double tables: sponge factors (2nd order) and SV taper (VTI/TTI), uploaded to c->dtab.
int upload_dtab(a3d_ctx* c) {
    const a3d_shot_desc* desc = &c->d;
    std::vector<double> dt(dtable_count(c), 1.0);
    if (desc->op == A3D_OP_2ND) {
        for (int i = 0; i < c->g.nnx; ++i) dt[i] = sponge_value(i, desc->nx, desc->nb);
        for (int i = 0; i < c->g.nny; ++i) dt[c->g.nnx + i] = sponge_value(i, desc->ny, desc->nb);
        for (int i = 0; i < c->g.nnz; ++i) dt[c->g.nnx + c->g.nny + i] = sponge_value(i, desc->nz, desc->nb);
    }
    double* taper = dt.data() + c->g.nnx + c->g.nny + c->g.nnz;
    for (int r2 = 0; r2 < 226; ++r2) {
        // reference vti:163-166: 0.5*(1-cos(pi*((sqrtf(r2)-r)*4.0/(R*3.0-1)))), R=15, r=4;
        // the FP64 instance takes the square root in double
        const double rr = c->prec == PREC_FP64 ? std::sqrt((double)r2) - 4 : (double)(sqrtf((float)r2) - 4);
        taper[r2] = r2 <= 16 ? 0.0 : 0.5 * (1 - std::cos(kPi * (rr * 4.0 / (15 * 3.0 - 1))));
    }
    if (hipMemcpy(c->dtab, dt.data(), dt.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return fail(A3D_EHIP, "table upload failed");
    return A3D_OK;
}

}  // namespace

extern "C" {

const char* a3d_last_error(void) { return g_err.c_str(); }
const char* a3d_version(void) { return "acoustic3d-mi355x 0.1 (gfx950)"; }

int a3d_create(const a3d_shot_desc* desc, int device, a3d_ctx** out) {
    if (!out) return fail(A3D_EINVAL, "null output pointer");
    *out = nullptr;
    int rc = check_desc_geometry(desc);
    if (rc) return rc;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(A3D_EINVAL, "device %d not present (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    a3d_ctx* c = new a3d_ctx();
    init_geometry(c, desc, device);
    const int op = desc->op;
    auto cleanup = [&](int code) {
        a3d_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        c->st = nullptr;
        return cleanup(fail(A3D_EHIP, "stream creation failed"));
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return cleanup(fail(A3D_EHIP, "event creation failed"));
    for (int v : volumes_of(op)) {
        const size_t esz = is_split(v) ? std::max(c->wsize, (size_t)4) : (is_wavefield(v) ? c->wsize : c->msize);
        const size_t bytes = c->nvox * esz;
        if (hipMalloc(&c->vol[v], bytes) != hipSuccess) {
            c->vol[v] = nullptr;
            return cleanup(fail(A3D_ENOMEM, "device allocation of %zu MiB failed", bytes >> 20));
        }
    }
    const size_t ntab = table_count(c), ndtab = dtable_count(c);
    if (hipMalloc(&c->tab, ntab * c->csize) != hipSuccess) { c->tab = nullptr; return cleanup(fail(A3D_ENOMEM, "table allocation failed")); }
    if (hipMalloc(&c->dtab, ndtab * sizeof(double)) != hipSuccess) { c->dtab = nullptr; return cleanup(fail(A3D_ENOMEM, "table allocation failed")); }
    if (hipMalloc(&c->flag, sizeof(int)) != hipSuccess) { c->flag = nullptr; return cleanup(fail(A3D_ENOMEM, "flag allocation failed")); }
    if (int rc2 = upload_dtab(c)) return cleanup(rc2);
    default_shapes(c);
    *out = c;
    return A3D_OK;
}

int a3d_destroy(a3d_ctx* c) {
    if (!c) return A3D_OK;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (auto& p : c->vol)
        if (p) (void)hipFree(p);
    for (void* p : {(void*)c->stage, c->tab, (void*)c->rec, (void*)c->shot_dev, (void*)c->snap_dev})
        if (p) (void)hipFree(p);
    if (c->dtab) (void)hipFree(c->dtab);
    if (c->flag) (void)hipFree(c->flag);
    for (void* p : {(void*)c->strip_net, (void*)c->strip_ids, (void*)c->strip_meta, (void*)c->strip_delta,
                    (void*)c->strip_mfac, c->multi, (void*)c->mrec})
        if (p) (void)hipFree(p);
    for (auto e : c->tm.pool) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return A3D_OK;
}

int a3d_set_model(a3d_ctx* c, const float* vel, const float* rho, const float* eps, const float* del,
                  const float* theta, const float* phi) {
    if (!c) return fail(A3D_EINVAL, "null context");
    const int op = c->d.op;
    if (!vel) return fail(A3D_EINVAL, "velocity model required");
    if ((op == A3D_OP_ISO || op == A3D_OP_VTI) && !rho) return fail(A3D_EINVAL, "density model required");
    if ((op == A3D_OP_VTI || op == A3D_OP_TTI) && (!eps || !del)) return fail(A3D_EINVAL, "epsilon and delta required");
    if (op == A3D_OP_TTI && (!theta || !phi)) return fail(A3D_EINVAL, "theta and phi required");
    HIP_TRY(hipSetDevice(c->device));
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    const size_t n = (size_t)nx * ny * nz;
    float* stage = nullptr;
    if (hipMalloc(&stage, n * sizeof(float)) != hipSuccess) return fail(A3D_ENOMEM, "staging allocation failed");
    auto up = [&](const float* h, int v) -> int {
        HIP_TRY(hipMemcpy(stage, h, n * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(launch_pad(stage, nx, ny, nz, c->off, c->g, c->vol[v], c->prec == PREC_FP64, 0, 0.f, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
        return A3D_OK;
    };
    int rc = A3D_OK;
    if (op == A3D_OP_2ND) {
        HIP_TRY(hipMemcpy(stage, vel, n * sizeof(float), hipMemcpyHostToDevice));
        if (c->stage) (void)hipFree(c->stage);
        c->stage = stage;
        stage = nullptr;
        c->vv_dt = 0.0f;   // (v*dt)^2 is built at the first run
    } else {
        if (!rc) rc = up(vel, V_VP);
        if (!rc && (op == A3D_OP_ISO || op == A3D_OP_VTI)) rc = up(rho, V_RHO);
        if (!rc && op != A3D_OP_ISO) rc = up(eps, V_EPS);
        if (!rc && op != A3D_OP_ISO) rc = up(del, V_DEL);
        if (!rc && op == A3D_OP_TTI) rc = up(theta, V_CT);
        if (!rc && op == A3D_OP_TTI) rc = up(phi, V_CP);
        if (!rc && op == A3D_OP_TTI) {
            if (hipSuccess != launch_trig(c->vol[V_CT], c->vol[V_ST], c->vol[V_CP], c->vol[V_SP], c->nvox,
                                          c->prec == PREC_FP64, c->st) ||
                hipSuccess != hipStreamSynchronize(c->st))
                rc = fail(A3D_EHIP, "TTI rotation terms failed");
        }
        (void)hipFree(stage);
    }
    if (rc) return rc;
    model_scalars(c, vel, rho, eps, Window{0, 0, 0, nx, ny, nz});
    c->model_set = true;
    c->m4_ok = false;
    return A3D_OK;
}

int a3d_run(a3d_ctx* c, const a3d_shot_desc* d, float* shot_out, float* snaps_out, int nsnap_max,
            int* nsnap_taken, double* loop_ms) {
    if (!c || !d) return fail(A3D_EINVAL, "null argument");
    const size_t vol = (size_t)d->nx * d->ny * d->nz;
    if (int rc = maybe_autotune(c, d)) return rc;
    SnapFn fn;
    if (snaps_out && nsnap_max > 0)
        fn = [&](int k, int, const float* h) {
            if (k < nsnap_max) std::memcpy(snaps_out + (size_t)k * vol, h, vol * sizeof(float));
        };
    return run_shot(c, d, shot_out, fn, nsnap_taken, loop_ms);
}

int a3d_run_shots(a3d_ctx* c, const a3d_shot_desc* d, int nshot, const int* src, float* shots_out, double* loop_ms) {
    if (!c || !d || !src) return fail(A3D_EINVAL, "null argument");
    if (d->op != A3D_OP_2ND) return fail(A3D_EINVAL, "several shots in one launch: 2nd-order operator only");
    if (nshot < 1 || nshot > MAX_SHOTS) return fail(A3D_EINVAL, "shots per run must be in 1..%d, got %d", MAX_SHOTS, nshot);
    if (d->snap_interval != 0) return fail(A3D_EINVAL, "several shots in one launch take no snapshots");
    for (int k = 0; k < nshot; ++k) {
        a3d_shot_desc dk = *d;
        dk.sx = src[3 * k];
        dk.sy = src[3 * k + 1];
        dk.sz = src[3 * k + 2];
        if (int rc = check_desc_shot(c, &dk)) return rc;
    }
    if (!c->model_set) return fail(A3D_ESTATE, "model not set");
    a3d_shot_desc d0 = *d;   // the one-shot tuner runs the first shot (the descriptor's source is not used)
    d0.sx = src[0];
    d0.sy = src[1];
    d0.sz = src[2];
    d = &d0;
    if (int rc = maybe_autotune(c, d)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    auto run = [&](const a3d_shot_desc* dd, float* out, double* ms) {
        return c->prec == PREC_FP64 ? run_shots_2nd_t<double>(c, dd, nshot, src, out, ms)
                                    : run_shots_2nd_t<float>(c, dd, nshot, src, out, ms);
    };
    if (nshot > 1 && !c->mshape) {
        // several shots fill the GPU by themselves: their launch shape is timed apart from the
        // one-shot shape (24 steps per candidate: 4 or 8 tile rows, 12- or 4-plane segments)
        a3d_shot_desc td = *d;
        td.nt = std::min(d->nt, 24);
        const int nzi = c->g.nnz - 2 * MM;
        double best = 1e30;
        int bty = c->ty[0], bzs = c->zseg[0];
        c->mshape = true;
        for (int ty : {4, 8})
            for (int zl : {12, 4}) {
                c->mty = ty;
                c->mzseg = std::max(1, (nzi + zl - 1) / zl);
                double ms = 0;
                if (int rc = run(&td, nullptr, &ms)) {
                    c->mshape = false;
                    return rc;
                }
                if (ms < best) {
                    best = ms;
                    bty = c->mty;
                    bzs = c->mzseg;
                }
            }
        c->mty = bty;
        c->mzseg = bzs;
    }
    return run(d, shots_out, loop_ms);
}

int a3d_set_kernel_timing(a3d_ctx* c, int enable) {
    if (!c) return fail(A3D_EINVAL, "null context");
    if (enable < 0) return fail(A3D_EINVAL, "kernel timing: negative stride %d", enable);
    c->tm.on = enable != 0;
    c->tm.stride = enable > 1 ? enable : 1;
    c->tm.sample = true;
    return A3D_OK;
}

int a3d_kernel_times(a3d_ctx* c, double* ms, long long* launches, int reset) {
    if (!c) return fail(A3D_EINVAL, "null context");
    for (int k = 0; k < 2; ++k) {
        if (ms) ms[k] = c->tm.ms[k];
        if (launches) launches[k] = c->tm.n[k];
    }
    if (reset) {
        c->tm.ms[0] = c->tm.ms[1] = 0;
        c->tm.n[0] = c->tm.n[1] = 0;
    }
    return A3D_OK;
}

int a3d_set_launch(a3d_ctx* c, int ty_vel, int zseg_vel, int ty_stress, int zseg_stress) {
    if (!c) return fail(A3D_EINVAL, "null context");
    auto ok = [](int t) { return t == 2 || t == 4 || t == 8; };
    if (!ok(ty_vel) || !ok(ty_stress)) return fail(A3D_EINVAL, "tile rows must be 2, 4 or 8");
    if (zseg_vel < 0 || zseg_stress < 0) return fail(A3D_EINVAL, "z segments must be >= 0 (0: automatic)");
    c->ty[0] = ty_vel;
    c->ty[1] = ty_stress;
    c->zseg[0] = zseg_vel > 0 ? zseg_vel : single_wave_zseg(c, 0, ty_vel);
    c->mty = c->ty[0];   // a3d_run_shots takes the given shape too
    c->mzseg = c->zseg[0];
    c->mshape = true;
    c->zseg[1] = zseg_stress > 0 ? zseg_stress : single_wave_zseg(c, 1, ty_stress);
    c->tuned = true;   // an explicit shape is never overridden by the autotuner
    return A3D_OK;
}

int a3d_get_launch(a3d_ctx* c, int* out6) {
    if (!c || !out6) return fail(A3D_EINVAL, "null argument");
    for (int k = 0; k < 2; ++k) {
        const Launch l = make_launch(c->g, c->ty[k], c->zseg[k]);
        out6[3 * k + 0] = c->ty[k];
        out6[3 * k + 1] = l.nseg;
        out6[3 * k + 2] = l.grid;
    }
    return A3D_OK;
}

int a3d_autotune(a3d_ctx* c, const a3d_shot_desc* d, int steps, double* best_ms) {
    if (!c || !d) return fail(A3D_EINVAL, "null argument");
    return autotune(c, d, steps, best_ms);
}

int a3d_set_w_sweep(a3d_ctx* c, int mode) {
    if (!c) return fail(A3D_EINVAL, "null context");
    if (mode < -1 || mode > 1) return fail(A3D_EINVAL, "w sweep mode %d not in -1..1", mode);
    c->wz_mode = mode;
    return A3D_OK;
}

int a3d_get_w_sweep(a3d_ctx* c, const a3d_shot_desc* d, int* mode) {
    if (!c || !d || !mode) return fail(A3D_EINVAL, "null argument");
    *mode = wz_active(c, d) ? 1 : 0;
    return A3D_OK;
}

int a3d_probe_bandwidth(int device, size_t bytes, int reps, double* copy_gbs, double* read_gbs) {
    if (bytes < (1u << 20) || reps < 1) return fail(A3D_EINVAL, "probe needs >= 1 MiB and >= 1 repetition");
    HIP_TRY(hipSetDevice(device));
    bytes &= ~(size_t)15;
    float *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return fail(A3D_ENOMEM, "probe allocation failed");
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(A3D_ENOMEM, "probe allocation failed");
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = A3D_OK;
    auto run = [&](int ro, int grid, double* gbs) -> int {
        HIP_TRY(launch_probe(ro, a, b, bytes, grid, st));   // warm-up
        HIP_TRY(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) HIP_TRY(launch_probe(ro, a, b, bytes, grid, st));
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        const double moved = (ro ? 1.0 : 2.0) * (double)bytes * reps;
        if (gbs) *gbs = moved / (ms * 1e-3) / 1e9;
        return A3D_OK;
    };
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        rc = fail(A3D_EHIP, "probe stream/event creation failed");
    if (!rc) rc = hipMemsetAsync(a, 0, bytes, st) == hipSuccess ? A3D_OK : fail(A3D_EHIP, "memset failed");
    // the grid-stride copy (the byte counters' calibration dispatch), then the ceiling forms
    if (!rc) rc = run(0, 1024, nullptr);
    if (!rc) rc = run(0, 0, copy_gbs);
    if (!rc) rc = run(1, 0, read_gbs);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int a3d_set_strip_net(a3d_ctx* c, const float* params, int n) {
    if (!c) return fail(A3D_EINVAL, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (!params) {
        if (c->strip_net) (void)hipFree(c->strip_net);
        c->strip_net = nullptr;
        return A3D_OK;
    }
    if (n != LB_NPAR) return fail(A3D_EINVAL, "strip network: %d parameters expected, got %d", LB_NPAR, n);
    if (c->d.op != A3D_OP_VTI || c->prec != PREC_FP32)
        return fail(A3D_EINVAL, "the learned strip is built for the VTI operator with FP32 storage");
    if (c->d.nb != STRIP_NB) return fail(A3D_EINVAL, "the learned strip needs nb = %d (got %d)", STRIP_NB, c->d.nb);
    if (!c->strip_ids) {
        // strip cells: updated cells within STRIP_CELLS of the updated region's faces
        const Geo g = c->g;
        std::vector<int> ids, meta;
        const int n3[3] = {g.nnx, g.nny, g.nnz};
        for (int z = MM; z < g.nnz - MM; ++z)
            for (int y = MM; y < g.nny - MM; ++y)
                for (int x = MM; x < g.nnx - MM; ++x) {
                    const int p[3] = {x, y, z};
                    int best = 1 << 30, axis = 0, side = 0;
                    for (int a = 2; a >= 0; --a) {   // ties go to z, then y, then x
                        const int lo = p[a] - MM + 1, hi = n3[a] - MM - p[a];
                        if (lo < best) best = lo, axis = a, side = 0;
                        if (hi < best) best = hi, axis = a, side = 1;
                    }
                    if (best > STRIP_CELLS) continue;
                    // directions whose stress PML factor is not 1 (padded index below nb or
                    // above nn - nb, as pml_axis builds the tables)
                    int dirs = 0;
                    for (int a = 0; a < 3; ++a)
                        if (p[a] < STRIP_NB || p[a] > n3[a] - STRIP_NB) dirs |= 1 << a;
                    ids.push_back(x + g.ld * y + g.pl * z);
                    meta.push_back(axis | side << 2 | best << 3 | dirs << 8);
                }
        const size_t nb = ids.size() * sizeof(int);
        if (hipMalloc(&c->strip_ids, nb) != hipSuccess || hipMalloc(&c->strip_meta, nb) != hipSuccess ||
            hipMalloc(&c->strip_delta, ids.size() * sizeof(float2)) != hipSuccess) {
            for (void** q : {(void**)&c->strip_ids, (void**)&c->strip_meta, (void**)&c->strip_delta}) {
                if (*q) (void)hipFree(*q);
                *q = nullptr;
            }
            return fail(A3D_ENOMEM, "strip cell list allocation failed");
        }
        HIP_TRY(hipMemcpy(c->strip_ids, ids.data(), nb, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->strip_meta, meta.data(), nb, hipMemcpyHostToDevice));
        c->strip_n = (int)ids.size();
    }
    if (!c->strip_net && hipMalloc(&c->strip_net, LB_NPAR * sizeof(float)) != hipSuccess) {
        c->strip_net = nullptr;
        return fail(A3D_ENOMEM, "strip network allocation failed");
    }
    HIP_TRY(hipMemcpy(c->strip_net, params, LB_NPAR * sizeof(float), hipMemcpyHostToDevice));
    return A3D_OK;
}

int a3d_set_tti_angles(a3d_ctx* c, int on) {
    if (!c) return fail(A3D_EINVAL, "null context");
    if (on != 0 && on != 1) return fail(A3D_EINVAL, "TTI angle mode must be 0 or 1");
    if (on && (c->d.op != A3D_OP_TTI || c->prec == PREC_FP64))
        return fail(A3D_EINVAL, "rotation terms from the angles: TTI with FP32 or FP16 storage only");
    c->tti_angles = on;
    return A3D_OK;
}

int a3d_set_graph(a3d_ctx* c, int steps) {
    if (!c || steps < 0) return fail(A3D_EINVAL, "graph steps must be >= 0 (0: no graphs)");
    c->graph_steps = steps;
    return A3D_OK;
}

int a3d_set_zchunk(a3d_ctx* c, int planes) {
    if (!c || planes < 0 || (planes > 0 && planes < 8)) return fail(A3D_EINVAL, "z chunk must be 0 or >= 8 planes");
    c->zchunk = planes;
    c->tuned = false;
    return A3D_OK;
}

int a3d_set_strip_precision(a3d_ctx* c, int mfma) {
    if (!c || mfma < 0 || mfma > 1) return fail(A3D_EINVAL, "strip precision must be 0 or 1");
    c->strip_mfma = mfma;
    return A3D_OK;
}

int a3d_set_strip_head(a3d_ctx* c, int mode) {
    if (!c || mode < 0 || mode > 2) return fail(A3D_EINVAL, "strip head mode must be 0, 1 or 2");
    c->strip_head = mode;
    return A3D_OK;
}

int a3d_set_strip_period(a3d_ctx* c, int steps) {
    if (!c || steps < 0) return fail(A3D_EINVAL, "strip network period must be >= 0 steps");
    c->strip_period = steps;
    return A3D_OK;
}

int a3d_set_strip_profile(a3d_ctx* c, float amp) {
    if (!c || !(amp >= 0.0f) || !(amp <= 1000.0f)) return fail(A3D_EINVAL, "strip profile amplitude out of range");
    c->strip_amp = amp;
    return A3D_OK;
}

int a3d_strip_info(const a3d_ctx* c, int* ncells) {
    if (!c || !ncells) return fail(A3D_EINVAL, "null argument");
    *ncells = c->strip_n;
    return A3D_OK;
}

int a3d_strip_collect(a3d_ctx* t, a3d_ctx* s, const a3d_shot_desc* d, int every, int start, int horizon, float keep,
                      unsigned seed, float* X, float* T, float* S, int cap, int* count) {
    if (!t || !s || !d || !X || !T || !S || !count || cap <= 0 || every <= 0 || !(keep > 0))
        return fail(A3D_EINVAL, "bad argument");
    if (t->d.op != A3D_OP_VTI || s->d.op != A3D_OP_VTI || t->prec != PREC_FP32 || s->prec != PREC_FP32)
        return fail(A3D_EINVAL, "strip sampling: VTI contexts with FP32 storage");
    if (s->d.nx != t->d.nx || s->d.ny != t->d.ny || s->d.nz != t->d.nz || s->d.dx != t->d.dx ||
        s->d.dy != t->d.dy || s->d.dz != t->d.dz || t->d.nb <= s->d.nb)
        return fail(A3D_EINVAL, "strip sampling: same interior, teacher boundary wider than the student's");
    if (!t->model_set || !s->model_set) return fail(A3D_ESTATE, "model not set");
    if (!s->strip_ids) return fail(A3D_ESTATE, "student strip cells not built (set a strip network first)");
    if (s->strip_head == 2 && horizon > 1)
        return fail(A3D_EINVAL, "coefficient head: samples are taken teacher-forced only (horizon <= 1)");
    int rc = check_desc_shot(t, d);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(t->device));
    a3d_shot_desc ds = *d;
    ds.nb = s->d.nb;
    VelArgs vt, vs;
    StressArgs stt, sts;
    std::vector<Src> srcs, srcs_s;
    if ((rc = prepare_first<float>(t, d, field_scales(t, d), vt, stt, srcs))) return rc;
    if ((rc = prepare_first<float>(s, &ds, field_scales(s, &ds), vs, sts, srcs_s))) return rc;
    HIP_TRY(hipStreamSynchronize(s->st));
    stt.rec = nullptr;
    sts.rec = nullptr;
    const StripArgs a = strip_args(s, &ds);
    float *dX = nullptr, *dT = nullptr, *dS = nullptr;
    int* dn = nullptr;
    auto cleanup = [&](int r) {
        (void)hipStreamSynchronize(t->st);
        for (void* q : {(void*)dX, (void*)dT, (void*)dS, (void*)dn})
            if (q) (void)hipFree(q);
        return r;
    };
    if (hipMalloc(&dX, (size_t)cap * LB_NIN * sizeof(float)) != hipSuccess ||
        hipMalloc(&dT, (size_t)cap * LB_NOUT * sizeof(float)) != hipSuccess ||
        hipMalloc(&dS, (size_t)cap * sizeof(float)) != hipSuccess || hipMalloc(&dn, sizeof(int)) != hipSuccess)
        return cleanup(fail(A3D_ENOMEM, "sample buffer allocation failed"));
    HIP_TRY(hipMemsetAsync(dn, 0, sizeof(int), t->st));
    const int off = t->d.nb - s->d.nb;
    const unsigned keep_u = (unsigned)std::min(4294967295.0, (double)keep * 4294967296.0);
    const int fields[] = {V_U, V_V, V_W, V_P, V_Q, V_PX, V_PY, V_PZ, V_QX, V_QY, V_QZ};
    const bool free_run = horizon > 1;
    for (int it = 0; it < d->nt; ++it) {
        vt.s = srcs[it];
        HIP_TRY(launch_vel(PREC_FP32, A3D_OP_VTI, t->ty[0], vt, t->st));
        const bool live = it >= start;
        const bool samp = live && (it - start) % every == 0;
        const bool sync = live && (free_run ? (it - start) % horizon == 0 : samp);
        if (sync) {
            // the teacher's state (velocities at n+1/2, stresses at n) onto the strip grid
            for (int v : fields)
                HIP_TRY(launch_strip_copy((const float*)t->vol[v], t->g, (float*)s->vol[v], s->g, off, t->st));
        } else if (free_run && live) {
            vs.s = srcs_s[it];
            HIP_TRY(launch_vel(PREC_FP32, A3D_OP_VTI, s->ty[0], vs, t->st));
        }
        if (sync || (free_run && live)) HIP_TRY(launch_stress(PREC_FP32, A3D_OP_VTI, s->ty[1], sts, t->st));
        HIP_TRY(launch_stress(PREC_FP32, A3D_OP_VTI, t->ty[1], stt, t->st));
        if (samp)
            HIP_TRY(launch_strip_sample(a, (const float*)t->vol[V_P], (const float*)t->vol[V_Q], t->g, off, keep_u, seed,
                                        (unsigned)it, cap, dX, dT, dS, dn, t->st));
        if (free_run && live && s->strip_net) {
            // the strip grid runs on with its own network between synchronisations
            HIP_TRY(launch_strip_infer(a, s->strip_net, s->strip_delta, s->strip_mfma, t->st));
            HIP_TRY(launch_strip_apply(a, s->strip_delta, (float*)s->vol[V_P], (float*)s->vol[V_Q], (float*)s->vol[V_PX],
                                       (float*)s->vol[V_PY], (float*)s->vol[V_PZ], (float*)s->vol[V_QX],
                                       (float*)s->vol[V_QY], (float*)s->vol[V_QZ], (float*)s->vol[V_U],
                                       (float*)s->vol[V_V], (float*)s->vol[V_W], t->st));
        }
    }
    int hn = 0;
    HIP_TRY(hipMemcpyAsync(&hn, dn, sizeof(int), hipMemcpyDeviceToHost, t->st));
    HIP_TRY(hipStreamSynchronize(t->st));
    const int n = std::min(hn, cap);
    HIP_TRY(hipMemcpy(X, dX, (size_t)n * LB_NIN * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(T, dT, (size_t)n * LB_NOUT * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(S, dS, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    *count = n;
    return cleanup(A3D_OK);
}

int a3d_grid_info(const a3d_ctx* c, int* nnx, int* nny, int* nnz, long long* updated) {
    if (!c) return fail(A3D_EINVAL, "null context");
    if (nnx) *nnx = c->g.nnx;
    if (nny) *nny = c->g.nny;
    if (nnz) *nnz = c->g.nnz;
    if (updated)
        *updated = (long long)(c->g.nnx - 2 * MM) * (c->g.nny - 2 * MM) * (c->g.nnz - 2 * MM);
    return A3D_OK;
}

}  // extern "C"

// ===========================================================================
// Legacy entry points
// ===========================================================================
namespace {

int read_model(const char* path, size_t n, std::vector<float>& out) {
    if (!path) return fail(A3D_EIO, "null model file name");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(A3D_EIO, "error open <%s>!", path);
    out.resize(n);
    const size_t got = fread(out.data(), sizeof(float), n, f);
    fclose(f);
    if (got != n) return fail(A3D_EIO, "model file <%s> holds %zu of %zu floats", path, got, n);
    return A3D_OK;
}

// Decimated shot write (reference iso:591-597): fseek to is*nt*nx*ny floats,
// then [iy/jy][ix/jx][it/jt], time fastest.
int write_shot(FILE* f, int is, const float* shot, int nx, int ny, int nt, float dx, float dy,
               float dt, float dgx, float dgy, float dgt) {
    const int jx = int(dgx / dx), jy = int(dgy / dy), jt = int(dgt / dt);
    if (jx < 1 || jy < 1 || jt < 1)
        return fail(A3D_EINVAL, "receiver/time decimation dgx/dx=%d dgy/dy=%d dgt/dt=%d must be >= 1", jx, jy, jt);
    std::vector<float> buf;
    buf.reserve((size_t)((ny + jy - 1) / jy) * ((nx + jx - 1) / jx) * ((nt + jt - 1) / jt));
    for (int iy = 0; iy < ny; iy += jy)
        for (int ix = 0; ix < nx; ix += jx)
            for (int it = 0; it < nt; it += jt) buf.push_back(shot[(size_t)iy * nx * nt + (size_t)ix * nt + it]);
    if (fseek(f, (long)is * nt * nx * ny * (long)sizeof(float), SEEK_SET) != 0) return fail(A3D_EIO, "fseek failed");
    if (fwrite(buf.data(), sizeof(float), buf.size(), f) != buf.size()) return fail(A3D_EIO, "short write of shot file");
    return A3D_OK;
}

struct Legacy {
    const char* name;
    a3d_shot_desc d;
    const char* snapfmt;    // snapshot name after the FNsnap prefix
    int snap_is;            // shot index in the snapshot name
};

int legacy_run(Legacy& L, const char* FNsnap, const char* FNshot, FILE* fshot, int is_first, int ns,
               int sxbeg, int sybeg, int jsx, int jsy, int shot_is_offset, float dgx, float dgy, float dgt,
               const float* vel, const float* rho, const float* eps, const float* del, const float* th,
               const float* ph, int device) {
    (void)FNshot;
    a3d_ctx* c = nullptr;
    int rc = a3d_create(&L.d, device, &c);
    if (rc) return rc;
    rc = a3d_set_model(c, vel, rho, eps, del, th, ph);
    std::vector<float> shot((size_t)L.d.nx * L.d.ny * L.d.nt);
    for (int is = is_first; !rc && is < is_first + ns; ++is) {
        a3d_shot_desc d = L.d;
        if (d.op != A3D_OP_2ND) {
            d.sx = sxbeg + is % ns * jsx;   // reference vti:65-66 with nsx = ns
            d.sy = sybeg + is / ns * jsy;
        }
        const int snap_is = d.op == A3D_OP_2ND ? L.snap_is : is;
        SnapFn fn = [&](int, int it, const float* h) {
            char name[600];
            char idbuf[64];
            snprintf(idbuf, sizeof idbuf, L.snapfmt, snap_is, it);
            snprintf(name, sizeof name, "%s%s", FNsnap ? FNsnap : "", idbuf);
            FILE* f = fopen(name, "wb");
            if (!f) {
                printf("cannot write snapfile\n");
                return;
            }
            fwrite(h, sizeof(float), (size_t)d.nx * d.ny * d.nz, f);
            fclose(f);
        };
        int nsnap = 0;
        double ms = 0;
        rc = maybe_autotune(c, &d);
        if (!rc) rc = run_shot(c, &d, shot.data(), fn, &nsnap, &ms);
        if (rc == A3D_EUNSTABLE) {
            fprintf(stderr, "%s: %s\n", L.name, g_err.c_str());
            rc = A3D_OK;   // the shot is still written, as the reference would
        }
        if (!rc)
            rc = write_shot(fshot, shot_is_offset >= 0 ? shot_is_offset : is, shot.data(), d.nx, d.ny, d.nt,
                            d.dx, d.dy, d.dt, dgx, dgy, dgt);
    }
    a3d_destroy(c);
    return rc;
}

a3d_shot_desc base_desc(int op, int nx, int ny, int nz, float dx, float dy, float dz, int nb, int nt,
                        float dt, float fm) {
    a3d_shot_desc d{};
    // A3D_SCHEME=1 selects the unsplit-interior stress update for the legacy entry points,
    // A3D_SCHEME=2 the one-pass step for VTI (scheme 1 for ISO/TTI), A3D_SCHEME=3 split
    // fields for the damped directions only, A3D_SCHEME=4 one field per damped direction;
    // default 0: the reference's split fields everywhere, bit for bit.
    if (const char* e = getenv("A3D_SCHEME")) {
        if (op != A3D_OP_2ND && (e[0] == '1' || e[0] == '2')) d.scheme = 1;
        if (op == A3D_OP_VTI && e[0] == '2') d.scheme = 2;
        if (op != A3D_OP_2ND && e[0] == '3') d.scheme = 3;
        if (op != A3D_OP_2ND && e[0] == '4') d.scheme = 4;
    }
    d.op = op;
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.dx = dx;
    d.dy = dy;
    d.dz = dz;
    d.nb = nb;
    d.nt = nt;
    d.dt = dt;
    d.fm = fm;
    d.wtype = 1;
    d.mute_tt = 60;
    return d;
}

void report(const char* who, int rc) {
    if (rc) fprintf(stderr, "%s: error %d: %s\n", who, rc, g_err.c_str());
}

}  // namespace

extern "C" {

void cuda_3dfd(char* FNvel, char* FNsnap, char* FNshot, int is, int ns, int nx, int ny, int nz, float dx,
               float dy, float dz, int sxbeg, int sybeg, int szbeg, int jsx, int jsy, int jsz, float dgx,
               float dgy, float dgt, int nt, int kt, float dt, float fm, bool show_snapshot,
               bool cut_directwave, int snap_interval, int cudaDevicei) {
    (void)ns;
    (void)kt;
    (void)cut_directwave;   // mute is dead code in the reference (Toa_3dfd_cuda.cu:474-477)
    (void)snap_interval;    // cadence fixed at 500 (Toa_3dfd_cuda.cu:456)
    FILE* fshot = fopen(FNshot, "wb");
    if (!fshot) return report("cuda_3dfd", fail(A3D_EIO, "cannot open shot file <%s>", FNshot));
    std::vector<float> vel;
    int rc = read_model(FNvel, (size_t)nx * ny * nz, vel);
    if (!rc) {
        Legacy L{"cuda_3dfd", base_desc(A3D_OP_2ND, nx, ny, nz, dx, dy, dz, 50, nt, dt, fm), "ishot_%d_it_%d.bin", is};
        // cuda_set_s (Toa_3dfd_cuda.cu:51): shot `is` of the (jsx, jsy, jsz) line
        L.d.sx = sxbeg + is * jsx;
        L.d.sy = sybeg + is * jsy;
        L.d.sz = szbeg + is * jsz;
        L.d.rec_z = 1;
        L.d.snap_interval = show_snapshot ? 500 : 0;
        rc = legacy_run(L, FNsnap, FNshot, fshot, 0, 1, L.d.sx, L.d.sy, 0, 0, is, dgx, dgy, dgt, vel.data(),
                        nullptr, nullptr, nullptr, nullptr, nullptr, cudaDevicei);
    }
    fclose(fshot);
    report("cuda_3dfd", rc);
}

void cuda_3dfd_1order(char* FNvel, char* FNrho, char* FNsnap, char* FNshot, int is, int ns, int nx, int ny,
                      int nz, float dx, float dy, float dz, int sxbeg, int sybeg, int szbeg, int jsx, int jsy,
                      int jsz, float dgx, float dgy, float dgt, int nt, float dt, int wtype, float fm,
                      bool show_snapshot, bool cut_directwave, int snap_interval, int cudaDevicei) {
    (void)is;
    (void)jsz;
    FILE* fshot = fopen(FNshot, "wb");
    if (!fshot) return report("cuda_3dfd_1order", fail(A3D_EIO, "cannot open shot file <%s>", FNshot));
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> vel, rho;
    int rc = read_model(FNvel, n, vel);
    if (!rc) rc = read_model(FNrho, n, rho);
    if (!rc && ns < 1) rc = fail(A3D_EINVAL, "ns must be >= 1");
    if (!rc) {
        Legacy L{"cuda_3dfd_1order", base_desc(A3D_OP_ISO, nx, ny, nz, dx, dy, dz, 20, nt, dt, fm), "ishot_%dit_%d.bin", 0};
        L.d.sz = szbeg;
        L.d.wtype = wtype;
        L.d.pfac = 10000.0f;
        L.d.snap_interval = show_snapshot ? snap_interval : 0;
        L.d.mute = cut_directwave;
        rc = legacy_run(L, FNsnap, FNshot, fshot, 0, ns, sxbeg, sybeg, jsx, jsy, -1, dgx, dgy, dgt, vel.data(),
                        rho.data(), nullptr, nullptr, nullptr, nullptr, cudaDevicei);
    }
    fclose(fshot);
    report("cuda_3dfd_1order", rc);
}

void cuda_3dfd_vti(char* FNvel, char* FNrho, char* FNepsilon, char* FNdelta, char* FNsnap, char* FNshot, int is,
                   int ns, int nx, int ny, int nz, float dx, float dy, float dz, int sxbeg, int sybeg, int szbeg,
                   int jsx, int jsy, int jsz, float dgx, float dgy, float dgt, int nt, float dt, float fm,
                   bool show_snapshot, bool cut_directwave, int snap_interval, int cudaDevicei) {
    (void)is;
    (void)jsz;
    FILE* fshot = fopen(FNshot, "wb");
    if (!fshot) return report("cuda_3dfd_vti", fail(A3D_EIO, "cannot open shot file <%s>", FNshot));
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> vel, rho, eps, del;
    int rc = read_model(FNvel, n, vel);
    if (!rc) rc = read_model(FNepsilon, n, eps);
    if (!rc) rc = read_model(FNdelta, n, del);
    if (!rc) rc = read_model(FNrho, n, rho);
    if (!rc && ns < 1) rc = fail(A3D_EINVAL, "ns must be >= 1");
    if (!rc) {
        Legacy L{"cuda_3dfd_vti", base_desc(A3D_OP_VTI, nx, ny, nz, dx, dy, dz, 20, nt, dt, fm), "ishot_%d_it_%d.bin", 0};
        L.d.sz = szbeg;
        L.d.pfac = 10.0f;
        L.d.sv = 1;
        L.d.snap_interval = show_snapshot ? snap_interval : 0;
        L.d.mute = cut_directwave;
        rc = legacy_run(L, FNsnap, FNshot, fshot, 0, ns, sxbeg, sybeg, jsx, jsy, -1, dgx, dgy, dgt, vel.data(),
                        rho.data(), eps.data(), del.data(), nullptr, nullptr, cudaDevicei);
    }
    fclose(fshot);
    report("cuda_3dfd_vti", rc);
}

int cuda_3dfd_tti(char* FNvel, char* FNrho, char* FNepsilon, char* FNdelta, char* FNtheta, char* FNpsi,
                  char* FNsnap, char* FNshot, int is, int ns, int nx, int ny, int nz, float dx, float dy,
                  float dz, int sxbeg, int sybeg, int szbeg, int jsx, int jsy, int jsz, float dgx, float dgy,
                  float dgt, int nt, float dt, float fm, bool show_snapshot, bool cut_directwave,
                  int snap_interval, int cudaDevicei) {
    (void)FNrho;            // TTI has no density (reference tti:461)
    (void)is;
    (void)ns;               // forced to 1 (reference tti:439)
    (void)jsz;
    (void)show_snapshot;    // snapshots every 100 steps regardless (reference tti:593)
    (void)snap_interval;
    FILE* fshot = fopen(FNshot, "wb");
    if (!fshot) {
        printf("cannot write shot file");
        return 1;
    }
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> vel, eps, del, th, ph;
    int rc = read_model(FNvel, n, vel);
    if (!rc) rc = read_model(FNepsilon, n, eps);
    if (!rc) rc = read_model(FNdelta, n, del);
    if (!rc) rc = read_model(FNtheta, n, th);
    if (!rc) rc = read_model(FNpsi, n, ph);
    if (!rc) {
        Legacy L{"cuda_3dfd_tti", base_desc(A3D_OP_TTI, nx, ny, nz, dx, dy, dz, 40, nt, dt, fm), "ishot_%d_it_%d.bin", 0};
        L.d.sz = szbeg;
        L.d.pfac = 10.0f;
        L.d.sv = 1;
        L.d.snap_interval = 100;
        L.d.snap_skip0 = 1;
        L.d.mute = cut_directwave;
        rc = legacy_run(L, FNsnap, FNshot, fshot, 0, 1, sxbeg, sybeg, jsx, jsy, -1, dgx, dgy, dgt, vel.data(),
                        nullptr, eps.data(), del.data(), th.data(), ph.data(), cudaDevicei);
    }
    fclose(fshot);
    report("cuda_3dfd_tti", rc);
    return rc;
}

}  // extern "C"

// ===========================================================================
// Batched windows: many shots, each an aperture window cut from one model
// resident on the device, advanced together (one batched launch per sweep and
// tile-shape group per time step), optionally captured into one hipGraph.
// ===========================================================================
namespace {

struct BatchGroup {
    int ty = 4;
    std::vector<int> wins;          // window indices in this group
    void* args = nullptr;           // device array of VelArgs or StressArgs
    int* base = nullptr;            // device prefix sums of tile counts (nwin+1)
    Src* src = nullptr;             // velocity groups: per-step sources [nt][group windows]
    Launch lb{};
};

}  // namespace

struct a3d_batch {
    a3d_shot_desc d{};              // full-model geometry and the shared shot parameters
    int device = 0;
    std::vector<a3d_window> win;
    std::vector<a3d_ctx*> ctx;      // one context (buffers, tables) per window
    std::vector<a3d_shot_desc> wd;  // per-window shot descriptors
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool model_set = false;
    std::vector<int> ty_vel, ty_str; // tile rows per window and sweep
    int zs_vel = 1, zs_str = 1;      // z segments per tile column of every window, per sweep
    int wz_mode = -1;                // batched sweeps: w updated by 0 the velocity, 1 the stress
    int wz_tuned = 0;                //   sweep, -1 the tuner's choice (wz_tuned) (ISO, VTI)
    bool tuned = false;
    // state of the captured time loop, kept so later runs replay the same graph
    std::vector<BatchGroup> gv, gs;
    std::vector<hipGraph_t> graph;        // the time loop in chunks of graph_chunk steps
    std::vector<hipGraphExec_t> exec;
    int graph_chunk = 0;                  // 0: one graph for the whole loop
    long long graph_launches = 0;
    int persist_ty = 4;                   // mode 3: tile rows of the persistent kernel (4 or 8)
    int persist_steps = 100;              // mode 3: time steps per persistent launch
    int persist_group = 1;                // mode 3: consecutive tiles per ticket
};

namespace {

void free_groups(std::vector<BatchGroup>& gs) {
    for (auto& g : gs) {
        if (g.args) (void)hipFree(g.args);
        if (g.base) (void)hipFree(g.base);
        if (g.src) (void)hipFree(g.src);
        g.args = nullptr;
        g.base = nullptr;
        g.src = nullptr;
    }
}

// Per-step source table of a velocity group: src[it * n + k] for its k-th window.
int group_sources(BatchGroup& g, const std::vector<Src>& all, int nt, int nw) {
    const int n = (int)g.wins.size();
    std::vector<Src> t((size_t)nt * n);
    for (int it = 0; it < nt; ++it)
        for (int k = 0; k < n; ++k) t[(size_t)it * n + k] = all[(size_t)it * nw + g.wins[k]];
    if (hipMalloc(&g.src, t.size() * sizeof(Src)) != hipSuccess) {
        g.src = nullptr;
        return fail(A3D_ENOMEM, "batch source table allocation failed");
    }
    HIP_TRY(hipMemcpy(g.src, t.data(), t.size() * sizeof(Src), hipMemcpyHostToDevice));
    return A3D_OK;
}

// Groups windows by tile rows and uploads their arguments and tile prefix sums; every
// window's tile columns are cut into zs segments (one where a stress sweep updates w).
template <class A>
int build_groups(a3d_batch* b, const std::vector<A>& args, const std::vector<int>& tys, int zs, bool stress,
                 std::vector<BatchGroup>& out) {
    for (int ty : {2, 4, 8}) {
        BatchGroup g;
        g.ty = ty;
        for (size_t w = 0; w < args.size(); ++w)
            if (tys[w] == ty) g.wins.push_back((int)w);
        if (g.wins.empty()) continue;
        std::vector<A> ga;
        std::vector<int> base(1, 0);
        for (int w : g.wins) {
            A a = args[w];
            a.l = make_launch(a.g, ty, stress && a.wz ? 1 : zs);
            ga.push_back(a);
            base.push_back(base.back() + a.l.ntiles);
        }
        g.lb.ntiles = base.back();
        uniform_bands(g.lb);
        if (hipMalloc(&g.args, ga.size() * sizeof(A)) != hipSuccess ||
            hipMalloc(&g.base, base.size() * sizeof(int)) != hipSuccess) {
            out.push_back(g);
            return fail(A3D_ENOMEM, "batch argument allocation failed");
        }
        HIP_TRY(hipMemcpy(g.args, ga.data(), ga.size() * sizeof(A), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g.base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
        out.push_back(g);
    }
    return A3D_OK;
}

void drop_graph(a3d_batch* b) {
    free_groups(b->gv);
    free_groups(b->gs);
    b->gv.clear();
    b->gs.clear();
    for (auto e : b->exec) (void)hipGraphExecDestroy(e);
    for (auto g : b->graph) (void)hipGraphDestroy(g);
    b->exec.clear();
    b->graph.clear();
}

// Per-window sentinel, record transpose (+ mute) and copy to the caller's shots.
int batch_outputs(a3d_batch* b, float** shots) {
    const int nw = (int)b->ctx.size();
    const int nt = b->d.nt;
    const int op = b->d.op;
    int rc = A3D_OK;
    int unstable = 0;
    for (int w = 0; w < nw && !rc; ++w) {
        a3d_ctx* c = b->ctx[w];
        const a3d_shot_desc* d = &b->wd[w];
        const size_t n = (size_t)d->nx * d->ny * nt;
        HIP_TRY(hipMemsetAsync(c->flag, 0, sizeof(int), c->st));
        HIP_TRY(launch_check_finite(c->rec, n, c->flag, c->st));
        int hflag = 0;
        HIP_TRY(hipMemcpyAsync(&hflag, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
        unstable |= hflag;
        if (shots && shots[w]) {
            if ((rc = ensure(&c->shot_dev, &c->shot_cap, n))) return rc;
            MuteArgs m{};
            m.mute = d->mute;
            m.dx = d->dx;
            m.dy = d->dy;
            m.dz = d->dz;
            m.dt = d->dt;
            m.fm = d->fm;
            m.sx = d->sx;
            m.sy = d->sy;
            m.sz = d->sz;
            m.tt = d->mute_tt;
            m.vcut = d->v_mute > 0 ? d->v_mute
                                   : (op == A3D_OP_ISO ? c->vel_first : c->vel_first * sqrtf(1 + 2 * c->eps_first));
            HIP_TRY(launch_shot_transpose(c->rec, nt, d->nx, d->ny, m, c->shot_dev, c->st));
            HIP_TRY(hipMemcpyAsync(shots[w], c->shot_dev, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
            HIP_TRY(hipStreamSynchronize(c->st));
        }
    }
    if (!rc && unstable) return fail(A3D_EUNSTABLE, "non-finite values in a window's shot record");
    return rc;
}


// Mode 3: the time loop as persistent launches of persist_steps steps each (k_persist:
// resident grid, ticket counters, grid barriers), captured in one hipGraph and launched
// once (with A3D_PROGRESS set: launched one by one, a line after each). FP32 storage,
// PML boundary. Every window's results equal its plain run bit for bit: each tile is
// updated by the same code as in the other modes, and the barriers order the sweeps.
int persist_loop(a3d_batch* b, std::vector<VelArgs>& va, std::vector<StressArgs>& sa, const std::vector<Src>& src_all,
                 bool wz, double* loop_ms) {
    const int nw = (int)b->ctx.size(), nt = b->d.nt, op = kernel_op(b->ctx[0]), ty = b->persist_ty;
    const int dirs = dirs_mode(b->d.scheme);
    if (b->ctx[0]->prec != PREC_FP32) return fail(A3D_EINVAL, "the persistent batch (mode 3) runs FP32 storage");
    if (b->d.scheme == 2) return fail(A3D_EINVAL, "the persistent batch (mode 3) runs stress schemes 0, 1, 3 and 4");
    for (auto* c : b->ctx)
        if (c->strip_net) return fail(A3D_EINVAL, "the persistent batch (mode 3) runs the PML boundary");
    // At most 16 waves per CU (a workgroup is ty waves): the most seen resident together
    // on MI355X. The occupancy query said 7 workgroups of 4 waves for a 69-VGPR ISO
    // instance, and a grid of that size did not meet at its first barrier.
    const int per_cu = std::min(persist_blocks_per_cu(op, ty, dirs, wz), 16 / ty);
    if (per_cu < 1) return fail(A3D_EINVAL, "no persistent kernel for operator %d with %d tile rows", op, ty);
    const unsigned nblocks = (unsigned)(per_cu * b->ctx[0]->ncu) & ~7u;
    if (nblocks < 8) return fail(A3D_EINVAL, "persistent grid too small (%d workgroups per CU)", per_cu);
    // launches for the persistent kernel's tile rows; ticket order: the windows with the
    // longest tile columns first; each window's tiles in 8 contiguous bands, one per XCD
    for (int w = 0; w < nw; ++w) {
        va[w].l = make_launch(va[w].g, ty, b->zs_vel);
        sa[w].l = make_launch(sa[w].g, ty, wz ? 1 : b->zs_str);
    }
    std::vector<int> vord(nw), sord(nw);
    for (int w = 0; w < nw; ++w) vord[w] = sord[w] = w;
    std::stable_sort(vord.begin(), vord.end(), [&](int a, int c) { return va[a].l.zlen > va[c].l.zlen; });
    std::stable_sort(sord.begin(), sord.end(), [&](int a, int c) { return sa[a].l.zlen > sa[c].l.zlen; });
    // per band: prefix sums of ticket counts (a ticket covers `group` consecutive tiles of
    // one window) and each window's tile range in the band
    const int group = b->persist_group, n1 = nw + 1;
    std::vector<int> vb(8 * n1), sb(8 * n1), vl(16 * n1), sl(16 * n1);
    auto bands = [&](const std::vector<int>& ord, auto ntiles, std::vector<int>& base, std::vector<int>& lo) {
        for (int x = 0; x < 8; ++x) {
            base[x * n1] = 0;
            for (int j = 0; j < nw; ++j) {
                const int n = ntiles(ord[j]), per = (n + 7) / 8;
                const int l0 = std::min(x * per, n), l1 = std::min((x + 1) * per, n);
                lo[x * n1 + j] = l0;
                lo[8 * n1 + x * n1 + j] = l1;
                base[x * n1 + j + 1] = base[x * n1 + j] + (l1 - l0 + group - 1) / group;
            }
        }
    };
    bands(vord, [&](int w) { return va[w].l.ntiles; }, vb, vl);
    bands(sord, [&](int w) { return sa[w].l.ntiles; }, sb, sl);
    const int steps = std::max(1, std::min(b->persist_steps, nt));
    const size_t nsync = (size_t)PERSIST_LINE * (1 + 2 * (size_t)steps * 8);
    // one device block for everything the kernel reads besides the windows' own buffers
    std::vector<int> tabs;
    for (const std::vector<int>* v : {&vb, &sb, &vl, &sl, &vord, &sord}) tabs.insert(tabs.end(), v->begin(), v->end());
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    const size_t o_va = 0, o_sa = o_va + al(nw * sizeof(VelArgs)), o_tab = o_sa + al(nw * sizeof(StressArgs)),
                 o_src = o_tab + al(tabs.size() * sizeof(int)), o_sync = o_src + al(src_all.size() * sizeof(Src)),
                 o_err = o_sync + al(nsync * sizeof(unsigned)), total = o_err + al(sizeof(int));
    unsigned char* mem = nullptr;
    if (hipMalloc(&mem, total) != hipSuccess) return fail(A3D_ENOMEM, "persistent batch argument allocation failed");
    auto done = [&](int code) {
        (void)hipFree(mem);
        return code;
    };
#define P_TRY(x)                                                                                    \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return done(fail(A3D_EHIP, "%s: %s", #x, hipGetErrorString(e_)));    \
    } while (0)
    P_TRY(hipMemcpy(mem + o_va, va.data(), nw * sizeof(VelArgs), hipMemcpyHostToDevice));
    P_TRY(hipMemcpy(mem + o_sa, sa.data(), nw * sizeof(StressArgs), hipMemcpyHostToDevice));
    P_TRY(hipMemcpy(mem + o_tab, tabs.data(), tabs.size() * sizeof(int), hipMemcpyHostToDevice));
    P_TRY(hipMemcpy(mem + o_src, src_all.data(), src_all.size() * sizeof(Src), hipMemcpyHostToDevice));
    P_TRY(hipMemset(mem + o_err, 0, sizeof(int)));
    int khz = 0;
    P_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    PersistArgs p{};
    p.va = (const VelArgs*)(mem + o_va);
    p.sa = (const StressArgs*)(mem + o_sa);
    p.tabs = (const int*)(mem + o_tab);
    p.srcs = (const Src*)(mem + o_src);
    p.sync = (unsigned*)(mem + o_sync);
    p.err = (int*)(mem + o_err);
    p.nwin = nw;
    p.group = group;
    p.nblocks = nblocks;
    // a barrier waits at most 4 s (a sweep over every window takes milliseconds)
    p.wait_ticks = 4000LL * std::max(khz, 1);
    const bool progress = getenv("A3D_PROGRESS") != nullptr;
    auto enqueue = [&](int i0) -> hipError_t {
        hipError_t e = launch_zero_u32(p.sync, nsync, b->st);
        if (e == hipSuccess) e = launch_persist(op, ty, dirs, wz, p, i0, std::min(nt, i0 + steps), b->st);
        return e;
    };
    // the launches captured in one hipGraph; with A3D_PROGRESS, one graph per launch,
    // each followed by a synchronisation and a progress line
    std::vector<hipGraphExec_t> execs;
    auto capture = [&](int i0, int i1) -> hipError_t {
        hipError_t e = hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        for (int i = i0; i < i1 && e == hipSuccess; i += steps) e = enqueue(i);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->st, &g);
        if (e == hipSuccess) e = e2;
        hipGraphExec_t x = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (x) execs.push_back(x);
        return e;
    };
    auto run_all = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        if (progress) {
            int tv = 0, ts = 0;
            for (int x = 0; x < 8; ++x) {
                tv += vb[x * n1 + nw];
                ts += sb[x * n1 + nw];
            }
            fprintf(stderr,
                    "a3d: persistent grid of %u workgroups (%d per CU), %d tile rows, %d steps per launch, "
                    "%d tiles per ticket, %d + %d tickets per step\n",
                    nblocks, per_cu, ty, steps, group, tv, ts);
            for (int i0 = 0; i0 < nt && e == hipSuccess; i0 += steps) e = capture(i0, std::min(nt, i0 + steps));
        } else {
            e = capture(0, nt);
        }
        if (e == hipSuccess) e = hipEventRecord(b->ev0, b->st);
        for (size_t k = 0; k < execs.size() && e == hipSuccess; ++k) {
            e = hipGraphLaunch(execs[k], b->st);
            ++b->graph_launches;
            if (progress && e == hipSuccess) {
                e = hipStreamSynchronize(b->st);
                fprintf(stderr, "a3d: persistent launch %zu/%zu done\n", k + 1, execs.size());
            }
        }
        if (e == hipSuccess) e = hipEventRecord(b->ev1, b->st);
        if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
        else (void)hipStreamSynchronize(b->st);   // nothing of this run may still be in flight
        return e;
    };
    const hipError_t erun = run_all();
    for (auto x : execs) (void)hipGraphExecDestroy(x);
    P_TRY(erun);
    float ms = 0.f;
    P_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (loop_ms) *loop_ms = ms;
    int herr = 0;
    P_TRY(hipMemcpy(&herr, p.err, sizeof(int), hipMemcpyDeviceToHost));
#undef P_TRY
    if (herr) return done(fail(A3D_EHIP, "persistent batch: a grid barrier wait ran out (grid of %u workgroups)", nblocks));
    return done(A3D_OK);
}

template <class T>
int batch_run_t(a3d_batch* b, int mode, float** shots, double* loop_ms, double* group_ms) {
    const int nw = (int)b->ctx.size();
    const int nt = b->d.nt;
    const int op = b->d.op;
    int rc = A3D_OK;
    std::vector<VelArgs> va(nw);
    std::vector<StressArgs> sa(nw);
    std::vector<Src> src_all((size_t)nt * nw);
    for (int w = 0; w < nw; ++w) {
        a3d_ctx* c = b->ctx[w];
        const a3d_shot_desc* d = &b->wd[w];
        if ((rc = ensure(&c->rec, &c->rec_cap, (size_t)d->nx * d->ny * nt))) return rc;
        std::vector<Src> srcs;
        if ((rc = prepare_first<T>(c, d, field_scales(c, d), va[w], sa[w], srcs))) return rc;
        sa[w].rec = c->rec;
        for (int it = 0; it < nt; ++it) src_all[(size_t)it * nw + w] = srcs[it];
    }
    // the wavefield memsets of prepare_first ran on the windows' streams
    for (auto* c : b->ctx) HIP_TRY(hipStreamSynchronize(c->st));
    // the batched sweeps' w mode, the same for every window
    // learned strip with the coefficient head on the windows: w stays in the velocity sweep
    const bool sm = b->ctx[0]->strip_net && b->ctx[0]->strip_head == 2;
    const bool wz = (op == A3D_OP_ISO || op == A3D_OP_VTI) && b->d.scheme != 2 && !sm &&
                    (b->wz_mode == 1 || (b->wz_mode < 0 && b->wz_tuned));
    for (int w = 0; w < nw; ++w) va[w].wz = va[w].zrev = sa[w].wz = wz ? 1 : 0;
    if (mode == 3) {
        if ((rc = persist_loop(b, va, sa, src_all, wz, loop_ms))) return rc;
        return batch_outputs(b, shots);
    }
    const bool replay = mode == 2 && !b->exec.empty();
    const bool progress = getenv("A3D_PROGRESS") != nullptr;
    std::vector<BatchGroup> gv_local, gs_local;
    std::vector<BatchGroup>& gv = mode == 2 ? b->gv : gv_local;
    std::vector<BatchGroup>& gs = mode == 2 ? b->gs : gs_local;
    auto cleanup = [&](int code) {
        if (mode != 2) {
            free_groups(gv_local);
            free_groups(gs_local);
        } else if (code) {
            drop_graph(b);
        }
        return code;
    };
    if (!replay) {
        if (mode == 2) drop_graph(b);
        if ((rc = build_groups(b, va, b->ty_vel, b->zs_vel, false, gv)) ||
            (rc = build_groups(b, sa, b->ty_str, b->zs_str, true, gs)))
            return cleanup(rc);
        for (auto& g : gv)
            if ((rc = group_sources(g, src_all, nt, nw))) return cleanup(rc);
        if (wz)   // the stress sweep that updates w adds the step's source as the velocity sweep does
            for (auto& g : gs)
                if ((rc = group_sources(g, src_all, nt, nw))) return cleanup(rc);
    }
    // windows with a learned strip: their strip passes follow each step's stress sweeps
    std::vector<int> swin;
    std::vector<StripArgs> sargs;
    for (int w = 0; w < nw; ++w)
        if (b->ctx[w]->strip_net) {
            swin.push_back(w);
            sargs.push_back(strip_args(b->ctx[w], &b->wd[w]));
        }
    // per-group timing (used by the tuner): events around each group's launches
    std::vector<hipEvent_t> gev;
    const bool timed = group_ms != nullptr && mode != 2;
    auto step = [&](int it) -> int {
        for (size_t k = 0; k < gv.size(); ++k) {
            const BatchGroup& g = gv[k];
            if (timed) HIP_TRY(hipEventRecord(gev[2 * k], b->st));
            HIP_TRY(launch_vel_batch(b->ctx[0]->prec, kernel_op(b->ctx[0]), g.ty, (const VelArgs*)g.args, g.base, (int)g.wins.size(),
                                     g.lb, g.src, it, wz, b->st, sm));
            if (timed) HIP_TRY(hipEventRecord(gev[2 * k + 1], b->st));
        }
        for (size_t k = 0; k < gs.size(); ++k) {
            const BatchGroup& g = gs[k];
            const size_t e = 2 * (gv.size() + k);
            if (timed) HIP_TRY(hipEventRecord(gev[e], b->st));
            HIP_TRY(launch_stress_batch(b->ctx[0]->prec, kernel_op(b->ctx[0]), g.ty, (const StressArgs*)g.args, g.base,
                                        (int)g.wins.size(), g.lb, it, dirs_mode(b->d.scheme), wz, g.src, b->st, sm));
            if (timed) HIP_TRY(hipEventRecord(gev[e + 1], b->st));
        }
        for (size_t k = 0; k < swin.size(); ++k) {
            const int rc2 = strip_step_on(b->ctx[swin[k]], sargs[k], b->st, it);
            if (rc2) return rc2;
        }
        return A3D_OK;
    };
    if (timed) {
        gev.resize(2 * (gv.size() + gs.size()));
        for (auto& e : gev) (void)hipEventCreate(&e);
        for (size_t k = 0; k < gv.size() + gs.size(); ++k) group_ms[k] = 0.0;
    }
    if (mode == 2) {
        const int chunk = b->graph_chunk > 0 ? std::min(b->graph_chunk, nt) : nt;
        if (!replay) {
            // the time loop captured as graphs of `chunk` steps, kept and replayed by later runs
            for (int i0 = 0; i0 < nt && !rc; i0 += chunk) {
                hipGraph_t gr = nullptr;
                hipGraphExec_t ex = nullptr;
                HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
                for (int it = i0; it < std::min(nt, i0 + chunk) && !rc; ++it) rc = step(it);
                hipError_t e = hipStreamEndCapture(b->st, &gr);
                if (gr) b->graph.push_back(gr);
                if (rc) return cleanup(rc);
                if (e != hipSuccess) return cleanup(fail(A3D_EHIP, "graph capture failed: %s", hipGetErrorString(e)));
                if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess)
                    return cleanup(fail(A3D_EHIP, "graph instantiation failed"));
                b->exec.push_back(ex);
            }
        }
        HIP_TRY(hipEventRecord(b->ev0, b->st));
        for (size_t k = 0; k < b->exec.size(); ++k) {
            HIP_TRY(hipGraphLaunch(b->exec[k], b->st));
            ++b->graph_launches;
            if (progress) {
                HIP_TRY(hipStreamSynchronize(b->st));
                fprintf(stderr, "a3d: batch graph %zu/%zu done (%d steps)\n", k + 1, b->exec.size(),
                        (int)std::min((size_t)nt, (k + 1) * (size_t)chunk));
            }
        }
    } else {
        HIP_TRY(hipEventRecord(b->ev0, b->st));
        for (int it = 0; it < nt; ++it) {
            if ((rc = step(it))) return cleanup(rc);
            // the first two steps run on cold caches: left out of the timings when nt >= 6
            if (timed && (nt < 6 || it >= 2)) {
                HIP_TRY(hipEventSynchronize(gev.back()));
                for (size_t k = 0; k < gv.size() + gs.size(); ++k) {
                    float ms = 0.f;
                    if (hipEventElapsedTime(&ms, gev[2 * k], gev[2 * k + 1]) == hipSuccess) group_ms[k] += ms;
                }
            }
            if (progress && (it + 1) % 250 == 0) {
                HIP_TRY(hipStreamSynchronize(b->st));
                fprintf(stderr, "a3d: batch step %d/%d done\n", it + 1, nt);
            }
        }
    }
    HIP_TRY(hipEventRecord(b->ev1, b->st));
    HIP_TRY(hipEventSynchronize(b->ev1));
    for (auto& e : gev) (void)hipEventDestroy(e);
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (loop_ms) *loop_ms = ms;
    rc = cleanup(A3D_OK);
    if (rc) return rc;
    return batch_outputs(b, shots);
}

int batch_run(a3d_batch* b, int mode, float** shots, double* loop_ms, double* group_ms) {
    if (b->ctx[0]->prec == PREC_FP64) return batch_run_t<double>(b, mode, shots, loop_ms, group_ms);
    return batch_run_t<float>(b, mode, shots, loop_ms, group_ms);
}

}  // namespace

extern "C" {

int a3d_batch_create(const a3d_shot_desc* desc, int nwin, const a3d_window* wins, int device, a3d_batch** out) {
    if (!out || !desc || !wins || nwin < 1) return fail(A3D_EINVAL, "null argument or no windows");
    *out = nullptr;
    if (desc->op == A3D_OP_2ND) return fail(A3D_EINVAL, "batched windows cover the 1st-order operators");
    a3d_batch* b = new a3d_batch();
    b->d = *desc;
    b->device = device;
    auto cleanup = [&](int code) {
        a3d_batch_destroy(b);
        return code;
    };
    for (int w = 0; w < nwin; ++w) {
        const a3d_window& W = wins[w];
        if (W.ox < 0 || W.oy < 0 || W.oz < 0 || W.nx < 1 || W.ny < 1 || W.nz < 1 || W.ox + W.nx > desc->nx ||
            W.oy + W.ny > desc->ny || W.oz + W.nz > desc->nz)
            return cleanup(fail(A3D_EINVAL, "window %d outside the model", w));
        a3d_shot_desc d = *desc;
        d.nx = W.nx;
        d.ny = W.ny;
        d.nz = W.nz;
        d.sx = W.sx;
        d.sy = W.sy;
        d.sz = W.sz;
        d.snap_interval = 0;
        a3d_ctx* c = nullptr;
        int rc = a3d_create(&d, device, &c);
        if (rc) return cleanup(rc);
        b->ctx.push_back(c);
        b->wd.push_back(d);
        b->win.push_back(W);
        if ((rc = check_desc_shot(c, &d))) return cleanup(rc);
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return cleanup(fail(A3D_EHIP, "batch stream/event creation failed"));
    b->ty_vel.assign(nwin, 8);
    b->ty_str.assign(nwin, 4);
    *out = b;
    return A3D_OK;
}

int a3d_batch_destroy(a3d_batch* b) {
    if (!b) return A3D_OK;
    if (b->st) (void)hipStreamSynchronize(b->st);
    drop_graph(b);
    for (auto* c : b->ctx) a3d_destroy(c);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return A3D_OK;
}

int a3d_batch_set_model(a3d_batch* b, const float* vel, const float* rho, const float* eps, const float* del,
                        const float* theta, const float* phi) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    const int op = b->d.op;
    if (!vel || ((op == A3D_OP_ISO || op == A3D_OP_VTI) && !rho) || (op != A3D_OP_ISO && (!eps || !del)) ||
        (op == A3D_OP_TTI && (!theta || !phi)))
        return fail(A3D_EINVAL, "model arrays missing for the operator");
    HIP_TRY(hipSetDevice(b->device));
    drop_graph(b);
    const size_t n = (size_t)b->d.nx * b->d.ny * b->d.nz;
    float* stage = nullptr;
    if (hipMalloc(&stage, n * sizeof(float)) != hipSuccess) return fail(A3D_ENOMEM, "model staging allocation failed");
    int rc = A3D_OK;
    const Window full{0, 0, 0, b->d.nx, b->d.ny, b->d.nz};
    auto up = [&](const float* h, int v) -> int {
        HIP_TRY(hipMemcpy(stage, h, n * sizeof(float), hipMemcpyHostToDevice));
        for (size_t w = 0; w < b->ctx.size(); ++w) {
            a3d_ctx* c = b->ctx[w];
            const Window win{b->win[w].ox, b->win[w].oy, b->win[w].oz, full.nx, full.ny, full.nz};
            HIP_TRY(launch_pad(stage, c->d.nx, c->d.ny, c->d.nz, c->off, c->g, c->vol[v], c->prec == PREC_FP64, 0, 0.f,
                               b->st, &win));
        }
        HIP_TRY(hipStreamSynchronize(b->st));
        return A3D_OK;
    };
    if (!rc) rc = up(vel, V_VP);
    if (!rc && (op == A3D_OP_ISO || op == A3D_OP_VTI)) rc = up(rho, V_RHO);
    if (!rc && op != A3D_OP_ISO) rc = up(eps, V_EPS);
    if (!rc && op != A3D_OP_ISO) rc = up(del, V_DEL);
    if (!rc && op == A3D_OP_TTI) rc = up(theta, V_CT);
    if (!rc && op == A3D_OP_TTI) rc = up(phi, V_CP);
    for (size_t w = 0; w < b->ctx.size() && !rc && op == A3D_OP_TTI; ++w) {
        a3d_ctx* c = b->ctx[w];
        if (hipSuccess != launch_trig(c->vol[V_CT], c->vol[V_ST], c->vol[V_CP], c->vol[V_SP], c->nvox,
                                      c->prec == PREC_FP64, b->st) ||
            hipSuccess != hipStreamSynchronize(b->st))
            rc = fail(A3D_EHIP, "TTI rotation terms failed");
    }
    (void)hipFree(stage);
    if (rc) return rc;
    for (size_t w = 0; w < b->ctx.size(); ++w) {
        const Window win{b->win[w].ox, b->win[w].oy, b->win[w].oz, full.nx, full.ny, full.nz};
        model_scalars(b->ctx[w], vel, rho, eps, win);
        b->ctx[w]->model_set = true;
        b->ctx[w]->m4_ok = false;
    }
    b->model_set = true;
    return A3D_OK;
}

int a3d_batch_run(a3d_batch* b, int mode, float** shots, double* loop_ms) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    if (!b->model_set) return fail(A3D_ESTATE, "model not set");
    if (mode < 0 || mode > 3) return fail(A3D_EINVAL, "mode must be 0, 1, 2 or 3");
    HIP_TRY(hipSetDevice(b->device));
    if (mode == 0) {
        // baseline: the windows one after the other as plain single-shot runs
        double total = 0.0;
        for (size_t w = 0; w < b->ctx.size(); ++w) {
            double ms = 0.0;
            int rc = maybe_autotune(b->ctx[w], &b->wd[w]);
            if (!rc) rc = run_shot(b->ctx[w], &b->wd[w], shots ? shots[w] : nullptr, SnapFn(), nullptr, &ms);
            if (rc) return rc;
            total += ms;
            if (getenv("A3D_PROGRESS")) fprintf(stderr, "a3d: window %zu/%zu done (%.1f ms)\n", w + 1, b->ctx.size(), ms);
        }
        if (loop_ms) *loop_ms = total;
        return A3D_OK;
    }
    return batch_run(b, mode, shots, loop_ms, nullptr);
}

int a3d_batch_set_persist(a3d_batch* b, int tile_rows, int steps, int group) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    if (tile_rows != 4 && tile_rows != 8) return fail(A3D_EINVAL, "persistent tile rows must be 4 or 8");
    if (steps < 1) return fail(A3D_EINVAL, "steps per persistent launch must be positive");
    if (group < 1 || group > 64) return fail(A3D_EINVAL, "tiles per ticket must be in 1..64");
    b->persist_ty = tile_rows;
    b->persist_steps = steps;
    b->persist_group = group;
    return A3D_OK;
}

int a3d_batch_set_tti_angles(a3d_batch* b, int on) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    for (auto* c : b->ctx)
        if (int rc = a3d_set_tti_angles(c, on)) return rc;
    drop_graph(b);   // a kept graph holds the other kernels
    return A3D_OK;
}

int a3d_batch_tune(a3d_batch* b, int steps, int* nclasses) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    if (!b->model_set) return fail(A3D_ESTATE, "model not set");
    HIP_TRY(hipSetDevice(b->device));
    const int nw = (int)b->ctx.size();
    // size classes: windows with the same x-tile count and similar plane size
    std::vector<int> cls(nw);
    std::vector<long long> keys;
    for (int w = 0; w < nw; ++w) {
        const Geo& g = b->ctx[w]->g;
        const long long key = (long long)(g.ld / TX) * 1000 + (g.nny + 63) / 64;
        auto it = std::find(keys.begin(), keys.end(), key);
        if (it == keys.end()) {
            keys.push_back(key);
            cls[w] = (int)keys.size() - 1;
        } else {
            cls[w] = (int)(it - keys.begin());
        }
    }
    if (nclasses) *nclasses = (int)keys.size();
    drop_graph(b);
    // tile rows and z segments are tuned with w in the velocity sweep (a stress sweep that
    // updates w runs one z segment), unless the caller fixed the stress sweep for w
    if (b->wz_mode < 0) b->wz_tuned = 0;
    const int saved_nt = b->d.nt;
    const int tsteps = std::max(6, steps);   // timed from the third step on
    for (auto& d : b->wd) d.nt = tsteps;
    b->d.nt = tsteps;
    // time every class with each tile-row choice (all windows of the class in one group)
    std::vector<double> best_v(keys.size(), 1e30), best_s(keys.size(), 1e30);
    std::vector<int> tv(keys.size(), 8), ts(keys.size(), 4);
    int rc = A3D_OK;
    for (int ty : {2, 4, 8}) {
        b->ty_vel.assign(nw, ty);
        b->ty_str.assign(nw, ty);
        // one run per class so each group holds one class
        for (size_t k = 0; k < keys.size() && !rc; ++k) {
            std::vector<int> saved_v = b->ty_vel;
            // windows outside the class are moved to another group (a different tile-row value)
            const int other = ty == 2 ? 4 : 2;
            for (int w = 0; w < nw; ++w)
                if (cls[w] != (int)k) b->ty_vel[w] = b->ty_str[w] = other;
            double gms[6] = {0, 0, 0, 0, 0, 0};
            double ms = 0.0;
            rc = batch_run(b, 1, nullptr, &ms, gms);
            if (rc == A3D_EUNSTABLE) rc = A3D_OK;
            // group order: vel groups by ty (2,4,8 present), then stress groups likewise
            std::vector<int> present;
            for (int t : {2, 4, 8}) {
                bool any = false;
                for (int w = 0; w < nw; ++w) any |= b->ty_vel[w] == t;
                if (any) present.push_back(t);
            }
            const int ng = (int)present.size();
            for (int gi = 0; gi < ng; ++gi) {
                if (present[gi] != ty) continue;
                if (gms[gi] < best_v[k]) {
                    best_v[k] = gms[gi];
                    tv[k] = ty;
                }
                if (gms[ng + gi] < best_s[k]) {
                    best_s[k] = gms[ng + gi];
                    ts[k] = ty;
                }
            }
            b->ty_vel = saved_v;
            b->ty_str = saved_v;
        }
    }
    for (int w = 0; w < nw; ++w) {
        b->ty_vel[w] = tv[cls[w]];
        b->ty_str[w] = ts[cls[w]];
    }
    // then the z segments of every window's tile columns (all windows alike), per sweep,
    // on the tuned tile rows: the sum of each sweep's group times over the timed steps
    {
        auto distinct = [&](const std::vector<int>& v) {
            int n = 0;
            for (int t : {2, 4, 8}) n += std::find(v.begin(), v.end(), t) != v.end() ? 1 : 0;
            return n;
        };
        const int nv = distinct(b->ty_vel), ns = distinct(b->ty_str);
        double bv = 1e30, bs = 1e30;
        int zv = 1, zst = 1;
        for (int zs : {1, 2, 4, 8, 12, 16}) {
            if (rc) break;
            b->zs_vel = b->zs_str = zs;
            double gms[6] = {0, 0, 0, 0, 0, 0};
            double ms = 0.0;
            rc = batch_run(b, 1, nullptr, &ms, gms);
            if (rc == A3D_EUNSTABLE) rc = A3D_OK;
            if (rc) break;
            double sv = 0.0, ss = 0.0;
            for (int k = 0; k < nv; ++k) sv += gms[k];
            for (int k = 0; k < ns; ++k) ss += gms[nv + k];
            if (sv < bv) bv = sv, zv = zs;
            if (ss < bs) bs = ss, zst = zs;
        }
        b->zs_vel = zv;
        b->zs_str = zst;
    }
    // then which sweep updates w (ISO, VTI): both sweeps' group times summed, both ways
    // (the stress sweep that updates w keeps one z segment, so its tile rows are re-timed
    // for that form: the tuned ones and each uniform choice; the faster form is kept)
    const bool strip_coef = b->ctx[0]->strip_net && b->ctx[0]->strip_head == 2;
    if (!rc && (b->d.op == A3D_OP_ISO || b->d.op == A3D_OP_VTI) && b->d.scheme != 2 && b->wz_mode < 0 && !strip_coef) {
        auto timed = [&](double* out) -> int {
            double gms[6] = {0, 0, 0, 0, 0, 0};
            double loop = 0.0;
            int r = batch_run(b, 1, nullptr, &loop, gms);
            if (r == A3D_EUNSTABLE) r = A3D_OK;
            *out = gms[0] + gms[1] + gms[2] + gms[3] + gms[4] + gms[5];
            return r;
        };
        double ms0 = 1e30, ms1 = 1e30;
        b->wz_mode = 0;
        rc = timed(&ms0);
        const std::vector<int> ty_tuned = b->ty_str;
        std::vector<int> ty_best1 = ty_tuned;
        b->wz_mode = 1;
        for (int cand = 0; cand < 4 && !rc; ++cand) {
            if (cand > 0) b->ty_str.assign(nw, cand == 1 ? 2 : (cand == 2 ? 4 : 8));
            double t = 1e30;
            rc = timed(&t);
            if (t < ms1) {
                ms1 = t;
                ty_best1 = b->ty_str;
            }
        }
        b->wz_mode = -1;
        b->wz_tuned = ms1 < ms0 ? 1 : 0;
        b->ty_str = b->wz_tuned ? ty_best1 : ty_tuned;
    }
    for (auto& d : b->wd) d.nt = saved_nt;
    b->d.nt = saved_nt;
    b->tuned = rc == A3D_OK;
    return rc;
}

int a3d_batch_set_w_sweep(a3d_batch* b, int mode) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    if (mode < -1 || mode > 1) return fail(A3D_EINVAL, "w sweep mode %d not in -1..1", mode);
    if (mode != b->wz_mode) drop_graph(b);
    b->wz_mode = mode;
    return A3D_OK;
}

int a3d_batch_set_strip_net(a3d_batch* b, const float* params, int n) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    drop_graph(b);
    for (a3d_ctx* c : b->ctx) {
        const int rc = a3d_set_strip_net(c, params, n);
        if (rc) return rc;
    }
    return A3D_OK;
}

int a3d_batch_set_strip_head(a3d_batch* b, int mode) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    drop_graph(b);
    for (a3d_ctx* c : b->ctx) {
        const int rc = a3d_set_strip_head(c, mode);
        if (rc) return rc;
    }
    return A3D_OK;
}

int a3d_batch_set_strip_period(a3d_batch* b, int steps) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    drop_graph(b);
    for (a3d_ctx* c : b->ctx) {
        const int rc = a3d_set_strip_period(c, steps);
        if (rc) return rc;
    }
    return A3D_OK;
}

int a3d_batch_set_strip_profile(a3d_batch* b, float amp) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    drop_graph(b);
    for (a3d_ctx* c : b->ctx) {
        const int rc = a3d_set_strip_profile(c, amp);
        if (rc) return rc;
    }
    return A3D_OK;
}

int a3d_batch_set_zseg(a3d_batch* b, int zseg_vel, int zseg_stress) {
    if (!b || zseg_vel < 1 || zseg_vel > 64 || zseg_stress < 1 || zseg_stress > 64)
        return fail(A3D_EINVAL, "z segments must be in 1..64");
    if (zseg_vel != b->zs_vel || zseg_stress != b->zs_str) drop_graph(b);
    b->zs_vel = zseg_vel;
    b->zs_str = zseg_stress;
    return A3D_OK;
}

int a3d_batch_get_zseg(const a3d_batch* b, int* zseg_vel, int* zseg_stress) {
    if (!b || !zseg_vel || !zseg_stress) return fail(A3D_EINVAL, "null argument");
    *zseg_vel = b->zs_vel;
    *zseg_stress = b->zs_str;
    return A3D_OK;
}

int a3d_batch_set_graph_chunk(a3d_batch* b, int steps) {
    if (!b || steps < 0) return fail(A3D_EINVAL, "bad argument");
    if (steps != b->graph_chunk) drop_graph(b);
    b->graph_chunk = steps;
    return A3D_OK;
}

int a3d_batch_info(a3d_batch* b, long long* updated_points, int* nwin, int* tile_rows_out) {
    if (!b) return fail(A3D_EINVAL, "null batch");
    (void)b->graph_launches;
    long long pts = 0;
    for (auto* c : b->ctx) pts += (long long)(c->g.nnx - 2 * MM) * (c->g.nny - 2 * MM) * (c->g.nnz - 2 * MM);
    if (updated_points) *updated_points = pts;
    if (nwin) *nwin = (int)b->ctx.size();
    if (tile_rows_out)
        for (size_t w = 0; w < b->ctx.size(); ++w) {
            tile_rows_out[2 * w] = b->ty_vel[w];
            tile_rows_out[2 * w + 1] = b->ty_str[w];
        }
    return A3D_OK;
}

}  // extern "C"

// ===========================================================================
// Allocation-free shot: every buffer the shot needs beyond its inputs and its
// output lives in one file-scope device array, `scratch`, of fixed size. The
// runtime gives each device its own copy of the array; its address on a device
// is taken once with hipGetSymbolAddress, with that device current, and only
// that address is handed to kernels and copies (the array's name in host code
// is a host-side stand-in, not a device address).
// ===========================================================================

static __device__ __align__(256) unsigned char scratch[(size_t)A3D_STATIC_SCRATCH_MIB << 20];

namespace {

constexpr size_t kScratchBytes = (size_t)A3D_STATIC_SCRATCH_MIB << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device `device` (taken on first use).
int scratch_base(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(A3D_EINVAL, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(A3D_EHIP, "address of the scratch array on device %d: %s", device, hipGetErrorString(e));
        g_scratch[device] = (unsigned char*)p;
    }
    *out = g_scratch[device];
    return A3D_OK;
}

// Bump allocation of 256-byte aligned slices; base == nullptr only counts bytes.
struct Carver {
    unsigned char* base;
    size_t used = 0;
    void* take(size_t n) {
        used = (used + 255) & ~(size_t)255;
        void* p = base ? base + used : nullptr;
        used += n;
        return p;
    }
};

// Lays out every device buffer of a FP32 shot of context *c in cv (volumes of the operator,
// interleaved model, PML and double tables, sentinel flag, receiver record).
void carve(a3d_ctx* c, Carver& cv) {
    const int op = c->d.op;
    for (int v : volumes_of(op)) c->vol[v] = cv.take(c->nvox * c->wsize);
    if (op != A3D_OP_2ND) c->vol[V_M4] = cv.take(c->nvox * (op == A3D_OP_ISO ? 8 : 16));
    c->tab = cv.take(table_count(c) * c->csize);
    c->dtab = (double*)cv.take(dtable_count(c) * sizeof(double));
    c->flag = (int*)cv.take(sizeof(int));
    c->rec_cap = (size_t)c->d.nx * c->d.ny * c->d.nt;
    c->rec = (float*)cv.take(c->rec_cap * sizeof(float));
}

// One float of a device array (a model sample) to the host.
int sample(const float* a, size_t i, float* out) {
    HIP_TRY(hipMemcpy(out, a + i, sizeof(float), hipMemcpyDeviceToHost));
    return A3D_OK;
}

}  // namespace

extern "C" int a3d_run_static(const a3d_shot_desc* desc, int device, const float* vel, const float* rho,
                              const float* eps, const float* del, const float* theta, const float* phi, float* shot,
                              void* stream) {
    int rc = check_desc_geometry(desc);
    if (rc) return rc;
    const int op = desc->op;
    if (!vel || !shot) return fail(A3D_EINVAL, "velocity model and shot record are required");
    if ((op == A3D_OP_ISO || op == A3D_OP_VTI) && !rho) return fail(A3D_EINVAL, "density model required");
    if ((op == A3D_OP_VTI || op == A3D_OP_TTI) && (!eps || !del)) return fail(A3D_EINVAL, "epsilon and delta required");
    if (op == A3D_OP_TTI && (!theta || !phi)) return fail(A3D_EINVAL, "theta and phi required");
    if (desc->storage != 0) return fail(A3D_EINVAL, "the allocation-free shot runs FP32 storage only");
    if (desc->scheme == 2) return fail(A3D_EINVAL, "the allocation-free shot runs schemes 0, 1, 3 and 4");
    if (desc->snap_interval != 0) return fail(A3D_EINVAL, "the allocation-free shot takes no snapshots");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(A3D_EINVAL, "device %d not present (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    unsigned char* base = nullptr;
    if ((rc = scratch_base(device, &base))) return rc;

    a3d_ctx c;   // on the host stack: nothing of it is allocated
    init_geometry(&c, desc, device);
    if ((rc = check_desc_shot(&c, desc))) return rc;
    Carver count{nullptr};
    carve(&c, count);
    if (count.used > kScratchBytes)
        return fail(A3D_EINVAL, "shot needs %zu MiB of scratch, the allocation-free entry point has %zu MiB",
                    (count.used + (1u << 20) - 1) >> 20, kScratchBytes >> 20);
    Carver cv{base};
    carve(&c, cv);
    c.st = (hipStream_t)stream;
    c.out_device = true;
    c.shot_dev = shot;
    c.shot_cap = c.rec_cap;
    // the caller's work that produced the inputs, and an earlier call's, has finished
    HIP_TRY(hipStreamSynchronize(c.st));
    if ((rc = upload_dtab(&c))) return rc;

    const int nx = desc->nx, ny = desc->ny, nz = desc->nz;
    const Samples sm = model_samples(&c);
    auto lin = [&](const int* s) { return (size_t)s[0] + (size_t)nz * s[1] + (size_t)nz * nx * s[2]; };
    if (op == A3D_OP_2ND) {
        c.stage = const_cast<float*>(vel);   // run_shot pads (v*dt)^2 from it; never freed here
        c.vv_dt = 0.0f;
        if ((rc = sample(vel, lin(sm.first), &c.vel_raw1))) return rc;
    } else {
        auto pad = [&](const float* src, int v) -> int {
            HIP_TRY(launch_pad(src, nx, ny, nz, c.off, c.g, c.vol[v], 0, 0, 0.f, c.st));
            return A3D_OK;
        };
        if (!rc) rc = pad(vel, V_VP);
        if (!rc && (op == A3D_OP_ISO || op == A3D_OP_VTI)) rc = pad(rho, V_RHO);
        if (!rc && op != A3D_OP_ISO) rc = pad(eps, V_EPS);
        if (!rc && op != A3D_OP_ISO) rc = pad(del, V_DEL);
        if (!rc && op == A3D_OP_TTI) rc = pad(theta, V_CT);
        if (!rc && op == A3D_OP_TTI) rc = pad(phi, V_CP);
        if (!rc && op == A3D_OP_TTI)
            HIP_TRY(launch_trig(c.vol[V_CT], c.vol[V_ST], c.vol[V_CP], c.vol[V_SP], c.nvox, 0, c.st));
        if (!rc) rc = sample(vel, lin(sm.d0), &c.vel_d0);
        if (!rc) rc = sample(vel, lin(sm.first), &c.vel_first);
        if (!rc && eps) rc = sample(eps, lin(sm.first), &c.eps_first);
        if (rc) return rc;
    }
    c.model_set = true;
    c.m4_ok = false;
    default_shapes(&c);
    c.tuned = true;
    return run_shot(&c, desc, nullptr, SnapFn(), nullptr, nullptr);
}
