// a3d_internal.h -- shared between the HIP kernels (a3d_kernels.hip) and the
// host runtime (a3d_host.hip). Not part of the public ABI.
//
// Device storage of every padded volume: x fastest, then y, then z
// (index x + ld*y + ld*nny*z, ld a multiple of 64 floats). The time-stepping
// kernels stream along z: a workgroup owns an x-y tile (64 x TY points, one
// wavefront per tile row) and walks a segment of z planes, keeping the
// z-neighbourhood of its own column in registers and the x-y neighbourhood of
// the current plane in LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace a3d {

constexpr int MM = 4;    // half stencil length (8th order)

// Records "what: detail" as the library's last error message and returns code
// (defined with the rest of the error handling in a3d_host.hip).
int report_error(int code, const char* what, const char* detail);
constexpr int TX = 64;   // tile width: one wavefront per tile row

struct Geo {
    int nnx, nny, nnz;   // padded extents
    int ld;              // x pitch in elements
    int pl;              // plane stride ld*nny
};

// Work decomposition of one sweep: ntx*nty tiles per plane times nseg
// z segments of zlen planes covering the planes [zbeg, zend) (the whole
// updated range, or one chunk of it). grid = 8*ceil(ntiles/8) workgroups.
struct Launch {
    int ntx, nty, nseg, zlen, ntiles, grid;
    int zbeg, zend;
    int band[9];   // tiles [band[k], band[k+1]) go to XCD k (workgroups with blockIdx % 8 == k),
                   // in order; grid = 8 x the longest band
};

// Precision instances of the sweep kernels.
enum { PREC_FP32 = 0, PREC_FP16 = 1, PREC_FP64 = 2 };

// Wavefields of the 1st-order family (and the 2nd-order pressure, F_P).
enum { F_P, F_Q, F_U, F_V, F_W, F_PX, F_PY, F_PZ, F_QX, F_QY, F_QZ, NF };

// Power-of-two scale per wavefield for binary16 storage: value = half * s[f];
// stored half = value * inv[f]. Unused by the FP32 and FP64 instances.
struct Scales {
    float s[NF];
    float inv[NF];
};

// Per-axis split-PML profiles (1st-order operators), padded index: offsets
// into one table of the arithmetic type (float or double).
struct Pml {
    int cx1, cx2, cy1, cy2, cz1, cz2;   // velocity sweep (half-cell)
    int ax1, ax2, ay1, ay2, az1, az2;   // stress sweep
    // half the damping step dt*d of each index (0 outside the PML), velocity and stress
    // profiles: the learned strip's coefficient head recomputes c1 = 1/(1 + m h),
    // c2 = c1 (1 - m h) per cell from them with the cell's multiplier m
    int hvx, hvy, hvz, hsx, hsy, hsz;
};

// Learned strip, coefficient head: the damping multiplier m of a cell is
// STRIP_MMAX * sigmoid(y_0), so m = 1 (the plain ramp) at y_0 = -ln 2.
constexpr float STRIP_MMAX = 3.0f;

// Transient point source of the 1st-order operators: added to P (and Q) on
// every load of the velocity sweep, never stored (the stress sweep rewrites
// P and Q everywhere in the interior). Index -1: no source this step.
struct Src {
    int s0, s1;
    double p0, p1, q0, q1;
};

// Array pointers are untyped: each precision instance casts them to its
// storage types (wavefields: float / half / double; model: float / double).
struct VelArgs {
    Geo g;
    Launch l;
    Pml c;
    const void* tab;
    const void* P;
    const void* Q;
    const void* rho;
    const void *ct, *st, *cp, *sp;   // TTI: cos/sin of tilt and azimuth, per point
    const void* t4;            // TTI, float model: (cos, sin tilt, cos, sin azimuth) interleaved, or null
    void* u;
    void* v;
    void* w;
    double dtx, dty, dtz;
    Src s;
    Scales sc;
    int wz;                    // 1 (ISO, VTI): w is updated by the stress sweep instead
    int zrev;                  // with wz (no z window in this sweep): planes walked downward
    const float* mfac;         // learned strip, coefficient head: damping multiplier per padded point, or null
};

struct StressArgs {
    Geo g;
    Launch l;
    Pml c;
    const void* tab;
    const void* u;
    const void* v;
    const void* w;
    void *px, *py, *pz, *qx, *qy, *qz;
    const void *vp, *rho, *eps, *del;
    const void *ct, *st, *cp, *sp;   // TTI: cos/sin of tilt and azimuth, per point
    void* P;
    void* Q;
    double dtx, dty, dtz;
    int sv, sxp, syp, szp;     // SV taper: on/off and padded source cell
    const double* taper;       // taper factor by r^2, valid for 16 < r^2 <= 225
    float* rec;                // receiver plane of this step ([ny][nx]) or nullptr
    int zrec, nb, nx, ny;
    int unsplit;               // 1: P, Q updated directly where all stress PML factors are 1;
                               // 2: also split fields only for the damped directions of PML cells;
                               // 3: as 2 with one field per damped direction (VTI, TTI: the
                               //    split p and q fields are multiples of it)
    int zi0, zi1;              // planes whose z stress factor is exactly 1: zi0 <= z < zi1
    Scales sc;
    int wz;                    // 1 (ISO, VTI; one z segment): this sweep also updates w
    Src s;                     // wz: this step's source, added to P / Q as the velocity sweep reads them
    const void* m4;            // VTI, float model: (vp, rho, eps, delta) interleaved per point, or null;
                               // TTI, float model: (cos, sin tilt, cos, sin azimuth); ISO: (vp, rho)
    const float* mfac;         // learned strip, coefficient head: damping multiplier per padded point, or null
};

// One-pass VTI step (FP32, scheme-1 arithmetic): old fields in, new fields out
// (ping-pong buffers), split PML fields in place.
struct FusedArgs {
    Geo g;
    Launch l;
    Pml c;
    const void* tab;
    const float *P0, *Q0, *u0, *v0, *w0;   // stresses at step n, velocities at n-1/2
    float *P1, *Q1, *u1, *v1, *w1;         // stresses at n+1, velocities at n+1/2
    float *px, *py, *pz, *qx, *qy, *qz;
    const float *vp, *rho, *eps, *del;
    double dtx, dty, dtz;
    Src s;                                 // this step's source, added to P0, Q0 on load
    int sv, sxp, syp, szp;
    const double* taper;
    float* rec;
    int zrec, nb, nx, ny;
    const void* m4;                        // (vp, rho, eps, delta) interleaved per point
    int zi0, zi1;                          // planes whose z stress factor is exactly 1
};

struct Step2Args {
    Geo g;
    Launch l;
    void* X;                   // level n-1 in, level n+1 (damped once) out, in place
    const void* Y;             // level n (damped once, without this step's dipole)
    const void* vv;            // (v*dt)^2
    const double *fx, *fy, *fz;  // sponge factors per padded index (1.0: undamped)
    double dx2, dy2, dz2;
    int s0, s1;                // dipole cells
    double ya0, ya1;           // this step's dipole, added to Y on load
    double xa0, xa1;           // previous step's dipole, added to X on load
    float* rec;                // receiver plane of this step or nullptr
    int zrec, off, nx, ny;
    Scales sc;
};

hipError_t launch_vel(int prec, int op, int ty, const VelArgs& a, hipStream_t st);
hipError_t launch_fused_vti(int ty, const FusedArgs& a, hipStream_t st);
int fused_blocks_per_cu(int ty);
hipError_t launch_stress(int prec, int op, int ty, const StressArgs& a, hipStream_t st);
hipError_t launch_step2(int prec, int ty, const Step2Args& a, hipStream_t st);
// Several shots of the 2nd-order step on one model in one launch: shot k's levels at
// a.X / a.Y + k * stride bytes, its dipole cells s0[k], s1[k], its receiver plane at
// a.rec + k * rstride floats. a.l: the launch of one shot with ntiles = n x its tiles
// (tile t belongs to shot t % n), bands over that count.
constexpr int MAX_SHOTS = 16;
struct Step2Shots {
    int n;
    int s0[MAX_SHOTS], s1[MAX_SHOTS];
    size_t stride;             // bytes between two shots' volumes
    size_t rstride;            // floats between two shots' receiver records
};
hipError_t launch_step2_shots(int prec, int ty, const Step2Args& a, const Step2Shots& s, hipStream_t st);

// Batched sweeps over several windows (each an independent padded grid with its own
// arguments in device memory). base: nwin+1 prefix sums of the windows' tile counts;
// lb: the combined decomposition (ntiles = base[nwin], grid = 8*ceil(ntiles/8)).
// wz (ISO, VTI): w is updated by the stress sweep (all windows of a batch share it; the
// stress sweep then takes the per-step sources srcs[it * nwin + w] too, and its windows
// run one z segment per tile column).
// sm (VTI, FP32, w in the velocity sweep): the windows carry the learned strip's damping
// multipliers (mfac)
hipError_t launch_vel_batch(int prec, int op, int ty, const VelArgs* wins, const int* base, int nwin, const Launch& lb,
                            const Src* srcs, int it, bool wz, hipStream_t st, bool sm = false);
// dirs: the windows' stress args have unsplit == 2 (dirs 1) or 3 (dirs 2), else 0 (all
// windows of a batch share it)
hipError_t launch_stress_batch(int prec, int op, int ty, const StressArgs* wins, const int* base, int nwin,
                               const Launch& lb, int it, int dirs, bool wz, const Src* srcs, hipStream_t st,
                               bool sm = false);

// Persistent batched time loop: one launch runs steps it0 .. it1-1 of every window. A
// resident grid (nblocks workgroups, a multiple of 8, all held by the GPU at once) takes
// each sweep's tiles from ticket counters and meets at a grid barrier after each sweep.
// Every window's tiles are cut into 8 contiguous bands, band x for the workgroups with
// blockIdx % 8 == x (the hardware deals workgroup b to XCD b % 8), so each XCD gets an
// eighth of every window and neighbouring tiles share its L2; each XCD has its own
// counter per sweep and step.
constexpr int PERSIST_LINE = 16;   // unsigned counters per 64-byte line
struct PersistArgs {
    const VelArgs* va;       // per-window sweep arguments (device), launches for this TY
    const StressArgs* sa;
    const int* tabs;         // per XCD band x: vbase[8][nwin + 1], sbase[8][nwin + 1] (prefix sums of
                             // the windows' ticket counts in ticket order; a ticket covers `group`
                             // consecutive tiles), vlo[16][nwin + 1], slo[16][nwin + 1] (rows 0-7:
                             // each window's first tile in band x, rows 8-15: its end), then
                             // vwin[nwin], swin[nwin] (the window of each ticket-order slot)
    const Src* srcs;         // per-step sources [nt][nwin] by window index
    unsigned* sync;          // [0] barrier arrivals; the tickets of step it0 + k, sweep s and
                             // band x at PERSIST_LINE * (1 + (2k + s) * 8 + x) (zeroed before
                             // each launch)
    int* err;                // raised when a barrier wait ran out
    int nwin, group;
    unsigned nblocks;
    long long wait_ticks;    // barrier wait limit in wall-clock ticks
};
// FP32 storage; ty 4 or 8; dirs as launch_stress_batch; wz (ISO, VTI).
hipError_t launch_persist(int op, int ty, int dirs, bool wz, const PersistArgs& p, int it0, int it1, hipStream_t st);
// Resident workgroups per CU of the persistent kernel instance (0 if it does not exist).
int persist_blocks_per_cu(int op, int ty, int dirs, bool wz);
hipError_t launch_zero_u32(unsigned* a, size_t n, hipStream_t st);

// Resident workgroups per CU of a sweep kernel: kind 0 = velocity sweep (or
// the 2nd-order step for op 0), kind 1 = stress sweep.
int max_blocks_per_cu(int prec, int kind, int op, int ty);

// A sub-volume of a larger z-fastest model: origin (ox, oy, oz), full model size (nx, ny, nz).
struct Window {
    int ox, oy, oz, nx, ny, nz;
};

// Pads an interior z-fastest float model (nx,ny,nz) into the device layout
// (float, or double when model_double) by edge replication; the interior
// starts at padded index `off` on every axis. mode 0: copy; mode 1: (v*dt)^2.
// win (may be null): the interior is cut at origin win->o* from a model of size win->n*.
hipError_t launch_pad(const float* src, int nx, int ny, int nz, int off, const Geo& g, void* dst, int model_double,
                      int mode, float dt, hipStream_t st, const Window* win = nullptr);

// TTI angles to their cosines and sines, in place over n padded points:
// th -> cos(th) (in th), sin(th) (in st); ph -> cos(ph) (in ph), sin(ph) (in sp),
// each (M)cos((double)angle) with M float or double (model_double).
hipError_t launch_trig(void* th, void* st, void* ph, void* sp, size_t n, int model_double, hipStream_t s);
// (vp, rho, eps, delta) of a float model interleaved into one float4 per point.
// TTI: tilt and azimuth per point as one float2, atan2 of the float rotation terms in
// double (the angles to within the terms' rounding), for sweeps that evaluate the terms.
hipError_t launch_pack_angles(const float* ct, const float* st, const float* cp, const float* sp, void* out, size_t n,
                              hipStream_t s);
hipError_t launch_pack4(const float* vp, const float* rho, const float* eps, const float* del, void* out, size_t n,
                        hipStream_t s);
// (vp, rho) of a float model interleaved into one float2 per point.
hipError_t launch_pack2(const float* vp, const float* rho, void* out, size_t n, hipStream_t s);

// Interior crop of a padded wavefield into z-fastest float order ([ny][nx][nz]).
// 2nd-order snapshots: dipole (s0,a0),(s1,a1) added and the sponge applied
// (fx != nullptr); 1st-order: plain copy. scale: binary16 field scale.
hipError_t launch_snap(int prec, const void* src, const Geo& g, int off, int nx, int ny, int nz, int s0, double a0,
                       int s1, double a1, const double* fx, const double* fy, const double* fz, float scale,
                       float* dst, hipStream_t st);

// Record [nt][ny][nx] -> shot [ny][nx][nt], zeroing samples before the
// direct-wave cut when mute != 0.
struct MuteArgs {
    int mute;
    float dx, dy, dz, dt, fm, vcut;
    int sx, sy, sz, tt;
};
hipError_t launch_shot_transpose(const float* rec, int nt, int nx, int ny, const MuteArgs& m,
                                 float* shot, hipStream_t st);

// Calibration streams: float4 copy a -> b, or read-only reduction of a (b: 64 B scratch).
// grid > 0: the grid-stride float4 copy / read-only sum with that many workgroups of 256;
// grid 0: the ceiling forms (flat, non-temporal float4 streams)
hipError_t launch_probe(int read_only, const float* a, float* b, size_t bytes, int grid, hipStream_t st);

// Learned absorbing strip (VTI, FP32): the outer STRIP_CELLS updated cells of
// every face; the grid's boundary padding is STRIP_CELLS + MM.
constexpr int STRIP_CELLS = 6;
constexpr int STRIP_NB = STRIP_CELLS + MM;

// Strip cells of one grid: padded linear index and meta = axis | side << 2 | depth << 3 |
// dirs << 8 (axis of the nearest face, side 1 for the high face, depth 1 = outermost
// updated cell, dirs bit a set when the cell's stress PML factor along axis a is not 1).
struct StripArgs {
    Geo g;
    const int* ids;
    const int* meta;
    int n;
    const float *P, *Q, *u, *v, *w, *vp, *rho, *eps, *del;
    float cdt[3];   // dt/dx, dt/dy, dt/dz
    int dirsplit;   // the stress sweep keeps split fields only for a cell's damped directions
    int normal;     // head mode 1: damp the normal-axis split components and velocity only
    float* mfac;    // head mode 2: the network writes each strip cell's damping multiplier here
    float vref;     // head mode 2: the velocity sample the ramp's d0 is computed from (input 15 = vp / vref)
};

// delta[k] = the damping factors (g_P, g_Q) of every strip cell (after the stress sweep);
// mfma 1: the network on the matrix cores with binary16 operands, 0: FP32 per lane.
hipError_t launch_strip_infer(const StripArgs& a, const float* net, float2* delta, int mfma, hipStream_t st);
// Scales P, its split components, u and v by 1 - g_P; Q, its split components and w by 1 - g_Q.
hipError_t launch_strip_apply(const StripArgs& a, const float2* delta, float* P, float* Q, float* px, float* py,
                              float* pz, float* qx, float* qy, float* qz, float* u, float* v, float* w,
                              hipStream_t st);
// Appends teacher-forced samples (features, (teacher - strip state) / s, s).
hipError_t launch_strip_sample(const StripArgs& a, const float* Pt, const float* Qt, const Geo& gt, int off,
                               unsigned keep, unsigned seed, unsigned tag, int cap, float* X, float* T, float* S,
                               int* count, hipStream_t st);
// Copies a teacher field onto the student grid's updated region (index shift off).
hipError_t launch_strip_copy(const float* src, const Geo& gs, float* dst, const Geo& gd, int off, hipStream_t st);

// a[0..n) = v.
hipError_t launch_fill(float* a, size_t n, float v, hipStream_t st);

// Sets *flag to 1 if any of the n values is not finite.
hipError_t launch_check_finite(const float* a, size_t n, int* flag, hipStream_t st);

}  // namespace a3d
