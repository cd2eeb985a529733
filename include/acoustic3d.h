/*
 * acoustic3d.h -- C ABI of libacoustic3d.so, the MI355X (gfx950) 3D acoustic
 * finite-difference forward modeller.
 *
 * Two groups of entry points:
 *
 *  1. The four legacy entry points, with the exact parameter lists and
 *     semantics of the reference library symbols that readmodel.py binds with
 *     ctypes (one call = one shot: read model files, pad, time-step, record,
 *     optionally mute and write snapshots, write the shot file).
 *
 *  2. A handle API (a3d_*) for in-memory models and outputs, model data that
 *     stays resident in HBM across shots, per-kernel timing, and the
 *     parameters the legacy symbols hard-code (boundary width, receiver depth,
 *     mute velocity, snapshot cadence).
 *
 * Every a3d_* function returns A3D_OK (0) or a negative A3D_E* code;
 * a3d_last_error() returns a message for the calling thread's last failure.
 * The legacy void symbols print the message and return (the reference calls
 * exit(0) on a GPU error; this library never terminates the host process).
 *
 * Threading: contexts are independent; one context must not be used by two
 * host threads at once. All GPU work of a context runs on its own HIP stream.
 */
#ifndef ACOUSTIC3D_H
#define ACOUSTIC3D_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------- */
/* 1. Legacy entry points                                                   */
/* ----------------------------------------------------------------------- */

/* Replaces cuda_3dfd, reference src/Toa_3dfd_cuda.cu:366-370 (acoustic2order.so;
 * bound by readmodel.py:324-339). 2nd-order isotropic operator, 50-cell sponge,
 * receivers at interior z=1, snapshots every 500 steps when show_snapshot;
 * kt, cut_directwave and snap_interval are accepted and ignored, as in the
 * reference (readmodel.py passes the last two swapped). */
void cuda_3dfd(char* FNvel, char* FNsnap, char* FNshot, int is, int ns, int nx, int ny, int nz,
               float dx, float dy, float dz, int sxbeg, int sybeg, int szbeg, int jsx, int jsy,
               int jsz, float dgx, float dgy, float dgt, int nt, int kt, float dt, float fm,
               bool show_snapshot, bool cut_directwave, int snap_interval, int cudaDevicei);

/* Replaces cuda_3dfd_1order, reference src/Toa_gpu_3diso_fd_1orderfunciton.cu:372-377
 * (acoustic1order.so; bound by readmodel.py:228-244). Staggered-grid isotropic
 * operator with density, 20-cell split PML, z-dipole source (pfac 1e4), wavelet
 * types 1..6, receivers at interior z=0, snapshots when it % snap_interval == 0. */
void cuda_3dfd_1order(char* FNvel, char* FNrho, char* FNsnap, char* FNshot, int is, int ns,
                      int nx, int ny, int nz, float dx, float dy, float dz, int sxbeg, int sybeg,
                      int szbeg, int jsx, int jsy, int jsz, float dgx, float dgy, float dgt, int nt,
                      float dt, int wtype, float fm, bool show_snapshot, bool cut_directwave,
                      int snap_interval, int cudaDevicei);

/* Replaces cuda_3dfd_vti, reference src/Toa_gpu_3dvti_fd_1orderfunciton.cu:389-395
 * (acoustic1vti.so; bound by readmodel.py:278-296). VTI pseudo-acoustic, 20-cell
 * split PML, Ricker monopole into P and Q (pfac 10), SV taper around the source. */
void cuda_3dfd_vti(char* FNvel, char* FNrho, char* FNepsilon, char* FNdelta, char* FNsnap,
                   char* FNshot, int is, int ns, int nx, int ny, int nz, float dx, float dy,
                   float dz, int sxbeg, int sybeg, int szbeg, int jsx, int jsy, int jsz, float dgx,
                   float dgy, float dgt, int nt, float dt, float fm, bool show_snapshot,
                   bool cut_directwave, int snap_interval, int cudaDevicei);

/* Replaces cuda_3dfd_tti, reference src/Toa_gpu_3dtti_fd_1orderfunciton.cu:390-398
 * (acoustic1tti.so; bound by readmodel.py:366-386). TTI pseudo-acoustic without
 * density (FNrho is accepted and not read), 40-cell split PML, one shot (ns is
 * forced to 1), snapshots every 100 steps for it != 0 regardless of
 * show_snapshot. Returns 0, or 1 when the shot file cannot be opened; other
 * failures return a negative A3D_E* code (the reference returns 0 or 1 only). */
int cuda_3dfd_tti(char* FNvel, char* FNrho, char* FNepsilon, char* FNdelta, char* FNtheta,
                  char* FNpsi, char* FNsnap, char* FNshot, int is, int ns, int nx, int ny, int nz,
                  float dx, float dy, float dz, int sxbeg, int sybeg, int szbeg, int jsx, int jsy,
                  int jsz, float dgx, float dgy, float dgt, int nt, float dt, float fm,
                  bool show_snapshot, bool cut_directwave, int snap_interval, int cudaDevicei);

/* ----------------------------------------------------------------------- */
/* 2. Handle API (no counterpart in the reference: in-memory form of the    */
/*    entry points above, used by the Python driver, the tests and bench)   */
/* ----------------------------------------------------------------------- */

enum {
    A3D_OK = 0,
    A3D_EINVAL = -1,      /* bad argument */
    A3D_EHIP = -2,        /* HIP runtime error */
    A3D_ENOMEM = -3,      /* device or host allocation failed */
    A3D_EIO = -4,         /* file missing, short, or not writable */
    A3D_EUNSTABLE = -5,   /* non-finite wavefield detected */
    A3D_ESTATE = -6       /* call out of order (e.g. run before set_model) */
};

enum { A3D_OP_2ND = 0, A3D_OP_ISO = 1, A3D_OP_VTI = 2, A3D_OP_TTI = 3 };

/* One shot's description. Conventions follow the reference operator chosen:
 * op 2ND: source 0-based interior index, boundary = sponge width npd
 *         (plus the 4-cell ring), receivers default rec_z = 1;
 * op ISO/VTI/TTI: source 1-based interior index, boundary = npml,
 *         receivers default rec_z = 0. */
typedef struct a3d_shot_desc {
    int op;
    int nx, ny, nz;        /* interior model size */
    float dx, dy, dz;
    int nb;                /* boundary cells (reference: npd 50; npml 20, 20, 40) */
    int sx, sy, sz;        /* source position (see above) */
    int nt;
    float dt, fm;
    int wtype;             /* ISO: 1..6; VTI/TTI reference: 1; 2ND: Ricker table */
    float pfac;            /* 1st-order source factor (reference: 1e4 ISO, 10 VTI/TTI) */
    int rec_z;             /* receiver plane, interior z index */
    int sv;                /* VTI/TTI: SV-suppression taper around the source */
    int snap_interval;     /* 0: no snapshots */
    int snap_skip0;        /* 1: no snapshot at it == 0 */
    int mute;              /* 1: direct-wave mute of the shot record */
    float v_mute;          /* <= 0: reference velocity taken from the model */
    int mute_tt;           /* mute pad in samples (reference: 60) */
    int storage;           /* 0: FP32 (bitwise equal to the FP32 oracle); 1: wavefields stored as
                              binary16 times a power-of-two scale per field (split PML fields,
                              model and arithmetic FP32); 2: FP64 storage and arithmetic (error
                              reference). Outputs
                              (shot records, snapshots) are FP32 in every mode. Fixed per context. */
    int scheme;            /* 1st-order stress update. 0: split PML fields stored and updated
                              everywhere, the reference's arithmetic (FP32 bitwise equal to it);
                              1: split fields only in PML cells, P and Q updated directly where all
                              stress PML factors are 1 (same equations, different rounding: within
                              1e-5 relative L2 of scheme 0, 80 instead of 120 B/point for VTI);
                              2: VTI with FP32 storage only: the scheme-1 arithmetic in one
                              z-streaming pass per step (velocities made on the chip where the
                              stress update needs them; bitwise equal to scheme 1, 56 B/point
                              in the interior); 3: scheme 1, and a PML cell keeps the split
                              fields of its damped directions only (P, Q minus them carry the
                              undamped directions' updates): 64 instead of 92 B/point in the
                              faces' PML cells, within 1e-5 relative L2 of scheme 0; 4 (VTI,
                              TTI; ISO runs it as 3): scheme 3 with one field f per damped
                              direction instead of the pair (p, q): the reference's split p and q
                              of a direction obey one recurrence and differ only by the per-point
                              anisotropy factors of their source terms, so p_x,y = (1+2eps) f,
                              q_x,y = sqrt(1+2delta) f, p_z = sqrt(1+2delta) f, q_z = f (56
                              instead of 64 B/point in the faces' PML cells; not with the
                              learned strip's damping heads). Storage 1 (binary16) uses scheme 1,
                              3 or 4. */
} a3d_shot_desc;

typedef struct a3d_ctx a3d_ctx;

/* Allocates device memory for the geometry and operator of *desc on device
 * `device` and creates the context's stream. */
int a3d_create(const a3d_shot_desc* desc, int device, a3d_ctx** out);
int a3d_destroy(a3d_ctx* ctx);

/* Uploads the model: interior arrays of nx*ny*nz floats, z fastest
 * (iz + nz*ix + nz*nx*iy) as in the model files; NULL for arrays the operator
 * does not use (2ND: vel; ISO: vel, rho; VTI: vel, rho, eps, del; TTI: vel, eps,
 * del, theta, phi in radians). Pads by edge replication and builds the tables. */
int a3d_set_model(a3d_ctx* ctx, const float* vel, const float* rho, const float* eps,
                  const float* del, const float* theta, const float* phi);

/* Runs one shot. desc gives the source, time axis and output options; its
 * geometry fields (op, nx, ny, nz, dx, dy, dz, nb) must equal the context's.
 * shot_out: host [ny][nx][nt] (time fastest), undecimated; may be NULL (the
 * record stays on the device). snaps_out: host, up to nsnap_max interior
 * crops of nz*nx*ny (z fastest); may be NULL. *nsnap_taken receives the
 * number of snapshot steps. *loop_ms receives the time loop's duration
 * (HIP events on the context stream), excluding setup and host copies. */
int a3d_run(a3d_ctx* ctx, const a3d_shot_desc* desc, float* shot_out, float* snaps_out,
            int nsnap_max, int* nsnap_taken, double* loop_ms);

/* Per-kernel timing: when enabled, the launches of the two hot kernels are
 * bracketed by HIP events on the context stream. enable = 0 turns it off,
 * 1 times every time step, N > 1 times every N-th time step (it % N == 0,
 * fewer events in the stream). a3d_kernel_times returns the summed
 * milliseconds and launch counts of the timed launches since the last reset
 * (reset=1 clears after reading). Index 0: the 2nd-order step or 1st-order
 * velocity sweep; index 1: the 1st-order stress sweep (or the one-pass
 * step). A negative enable is A3D_EINVAL. */
int a3d_set_kernel_timing(a3d_ctx* ctx, int enable);
int a3d_kernel_times(a3d_ctx* ctx, double* ms, long long* launches, int reset);

/* Launch shape per sweep: tile rows TY (2, 4 or 8; a tile is 64 x TY points)
 * and the number of z segments each tile column is split into (0: one wave
 * of resident workgroups). Sweep 0 is the 1st-order velocity sweep or the
 * 2nd-order step; sweep 1 the stress sweep. Every shape gives bit-identical
 * results. An explicit shape disables the autotuner for the context. */
int a3d_set_launch(a3d_ctx* ctx, int ty_vel, int zseg_vel, int ty_stress, int zseg_stress);
/* out6 = {ty, z segments, workgroups} of sweep 0, then of sweep 1. */
int a3d_get_launch(a3d_ctx* ctx, int* out6);

/* Which sweep updates the z velocity w (ISO and VTI two-sweep step; no counterpart in
 * the reference, whose velocity kernel updates u, v and w): 0 the velocity sweep, as the
 * reference; 1 the stress sweep, from its own column's old P (ISO) or Q (VTI), one plane
 * ahead of where it needs w -- the same arithmetic, bit-identical results, 8 fewer
 * bytes per point and step, the stress sweep in one z segment; -1 (default) the
 * autotuner's choice (the velocity sweep when the tuner does not run).
 * a3d_get_w_sweep reports the mode a shot of *desc runs with. */
int a3d_set_w_sweep(a3d_ctx* ctx, int mode);

/* TTI rotation terms (no counterpart in the reference, whose TTI kernels recompute
 * cos/sin of the tilt and azimuth per point in double, Toa_gpu_3dtti_fd_1orderfunciton.cu:
 * 101-103,184-186): on = 0 (default) the sweeps read the four terms, (float)cos((double)
 * angle) built once per model (16 bytes per point per sweep, results bitwise the
 * oracle's); on = 1 they read the two angles (8 bytes per point) and evaluate the terms
 * in FP32 (abs. error <= 7e-8 per term). TTI with FP32 or FP16 storage only. */
int a3d_set_tti_angles(a3d_ctx* ctx, int on);
int a3d_get_w_sweep(a3d_ctx* ctx, const a3d_shot_desc* desc, int* mode);

/* Time loop as hipGraphs (no counterpart in the reference, which launches every kernel
 * from the host): steps > 0 captures the loop of each a3d_run into graphs of at most
 * `steps` time steps, cut before every snapshot, each launched as soon as it is
 * instantiated; 0 (default) launches every step from the host. Bit-identical results.
 * Not used while per-kernel timing is on. */
int a3d_set_graph(a3d_ctx* ctx, int steps);

/* Several 2nd-order shots on the context's model at once (no counterpart in the
 * reference, which runs one shot per cuda_3dfd call; Toa_3dfd_cuda.cu:366-508): nshot
 * (1..16) shots whose sources are src[3k], src[3k+1], src[3k+2] (sx, sy, sz as in
 * a3d_shot_desc; the descriptor's own source is not used) step together, one launch per
 * time step over every shot's tiles, so a small grid fills the GPU. shots_out (host,
 * may be NULL) receives nshot records of [ny][nx][nt] floats, each equal bit for bit to
 * a3d_run of that shot. No snapshots (desc->snap_interval must be 0); *loop_ms: the time
 * loop. Launch shape: a3d_set_launch's if given, else timed on the first run with several
 * shots (24 steps each of 4 or 8 tile rows, 12- or 4-plane segments) and kept. Allocates
 * 2 x nshot pressure volumes and the records on the device. */
int a3d_run_shots(a3d_ctx* ctx, const a3d_shot_desc* desc, int nshot, const int* src, float* shots_out,
                  double* loop_ms);

/* 1st-order two-sweep step in z chunks of `planes` planes (0: whole-grid sweeps, the
 * default; else >= 8): per step the velocity sweep of chunk k+1 runs before the stress
 * sweep of chunk k, so the velocities a stress chunk reads were written one launch
 * earlier and can still sit in the last-level cache. Bit-identical results. */
int a3d_set_zchunk(a3d_ctx* ctx, int planes);

/* Autotuner: times `steps` real time steps of *desc for each candidate
 * shape (TY in {2,4,8} x z segments in {1..24}; with w made by either sweep
 * where both can) and keeps the fastest per sweep; with steps >= 6 the first
 * two launches of each kernel (cold caches) are not counted. best_ms (may be
 * NULL) receives the two best average launch times. Results are cached per
 * process by (device, operator, precision, scheme, padded grid). a3d_run and
 * the legacy entry points call it (10 steps) on the first shot of a grid with
 * at least 1M updated points and nt >= 100 unless A3D_AUTOTUNE=0. */
int a3d_autotune(a3d_ctx* ctx, const a3d_shot_desc* desc, int steps, double* best_ms);

/* ----------------------------------------------------------------------- */
/* Batched windows (no counterpart in the reference): many shots, each an   */
/* aperture window cut from one model resident on the device.               */
/* ----------------------------------------------------------------------- */
typedef struct a3d_window {
    int ox, oy, oz;        /* window origin in the full model (0-based interior cells) */
    int nx, ny, nz;        /* window size */
    int sx, sy, sz;        /* source inside the window (operator convention: 1-based) */
} a3d_window;

typedef struct a3d_batch a3d_batch;

/* desc: the full model's geometry (op ISO/VTI/TTI, nx, ny, nz, spacing, nb) and
 * the shot parameters shared by all windows (nt, dt, fm, pfac, rec_z, sv, mute,
 * storage, scheme). Each window gets its own padded grid and receivers on its
 * full surface. */
int a3d_batch_create(const a3d_shot_desc* desc, int nwin, const a3d_window* wins, int device, a3d_batch** out);
int a3d_batch_destroy(a3d_batch* b);
/* Full-model arrays (nx*ny*nz floats, z fastest), uploaded once; every window's
 * padded model is cut from the device copy. */
int a3d_batch_set_model(a3d_batch* b, const float* vel, const float* rho, const float* eps, const float* del,
                        const float* theta, const float* phi);
/* Tile rows per window-size class (windows with equal x-tile and y-block counts), then
 * the z segments per sweep and (ISO, VTI) which sweep updates w, each candidate timed
 * over max(steps, 6) batched steps with the first two left out; *nclasses receives the
 * class count. */
int a3d_batch_tune(a3d_batch* b, int steps, int* nclasses);
/* mode 0: the windows one after another as plain single-shot runs (baseline);
 * mode 1: all windows advanced together, one batched launch per sweep and
 * tile-row group per step; mode 2: mode 1 with the whole time loop captured in
 * one hipGraph and launched once; mode 3 (FP32 storage, PML boundary): a
 * persistent kernel runs the time loop on the device -- a resident grid takes
 * the tiles of every window from a ticket counter per sweep (windows with the
 * longest tile columns first) and meets at a grid barrier after each sweep --
 * one launch per a3d_batch_set_persist steps, the launches captured in one
 * hipGraph (each barrier wait is bounded; a wait that runs out fails the run
 * with A3D_EHIP). shots: nwin host pointers ([ny][nx][nt] of each window) or
 * NULL; entries may be NULL. *loop_ms: time-loop duration. Results are bitwise
 * identical in all modes. */
int a3d_batch_run(a3d_batch* b, int mode, float** shots, double* loop_ms);
/* Mode 3: tile rows of the persistent kernel (4, the default, or 8; the z segments
 * per sweep are those of a3d_batch_set_zseg / a3d_batch_tune), time steps per
 * persistent launch (default 100) and consecutive tiles per ticket (1..64, default 1). */
int a3d_batch_set_persist(a3d_batch* b, int tile_rows, int steps, int group);
/* a3d_set_tti_angles for every window of the batch. */
int a3d_batch_set_tti_angles(a3d_batch* b, int on);
/* Mode 2 captures the time loop in graphs of `steps` time steps each (0, the
 * default: one graph for the whole loop); with the environment variable
 * A3D_PROGRESS set, a line is printed to stderr after each graph completes. */
int a3d_batch_set_graph_chunk(a3d_batch* b, int steps);
/* Updated points per step summed over windows; window count; tile rows per
 * window (2 ints each: velocity, stress), may be NULL. */
int a3d_batch_info(a3d_batch* b, long long* updated_points, int* nwin, int* tile_rows_out);
/* z segments per tile column of every window (1..64), per sweep; a stress sweep that
 * updates w keeps one. a3d_batch_tune picks them after the tile rows; all choices give
 * bitwise the same records. */
int a3d_batch_set_zseg(a3d_batch* b, int zseg_vel, int zseg_stress);
/* Which sweep updates w in the batched modes (ISO, VTI): 0 the velocity sweep, 1 the
 * stress sweep (one z segment per tile column), -1 (default) the choice of
 * a3d_batch_tune (velocity sweep until tuned); as a3d_set_w_sweep for one context.
 * Bitwise the same records either way. */
int a3d_batch_set_w_sweep(a3d_batch* b, int mode);
/* The learned absorbing strip on every window (VTI, FP32, nb = A3D_STRIP_NB): the
 * same network (params, a3d_mlp_nparams() floats, or NULL for the plain PML) and ramp
 * amplitude as a3d_set_strip_net / a3d_set_strip_profile give one context. The strip
 * passes of all windows follow each step's batched stress sweeps (and are captured
 * with them in mode 2); every mode gives bitwise the records of single runs. */
int a3d_batch_set_strip_net(a3d_batch* b, const float* params, int n);
int a3d_batch_set_strip_profile(a3d_batch* b, float amp);
/* The strip's head mode and network period on every window (as a3d_set_strip_head and
 * a3d_set_strip_period for one context). */
int a3d_batch_set_strip_head(a3d_batch* b, int mode);
int a3d_batch_set_strip_period(a3d_batch* b, int steps);
int a3d_batch_get_zseg(const a3d_batch* b, int* zseg_vel, int* zseg_stress);

/* ----------------------------------------------------------------------- */
/* Learned absorbing strip (no counterpart in the reference): VTI operator,   */
/* FP32 storage, nb = 10 (6 updated strip cells per face plus the 4-cell     */
/* zero ring, run with the PML update of that width). After every stress    */
/* sweep a network (16 inputs, two hidden tanh layers of 64, 2 outputs; 5378 */
/* FP32 parameters) gives each strip cell two damping factors g in          */
/* (0, 0.25): P, its split components, u and v are scaled by 1 - g_P; Q, its */
/* split components and w by 1 - g_Q.                                        */
/* ----------------------------------------------------------------------- */
#define A3D_STRIP_NB 10
/* params: a3d_mlp_nparams() floats, or NULL to run the plain PML of width nb.
 * Bit-identical to the plain PML when the output biases are <= -30 (1 - g rounds to 1). */
int a3d_set_strip_net(a3d_ctx* ctx, const float* params, int n);
/* Network evaluation: 1 (default) on the matrix cores, inputs, weights of the two
 * hidden layers and hidden activations rounded to binary16, sums in FP32; 0: FP32
 * per lane. */
int a3d_set_strip_precision(a3d_ctx* ctx, int mfma);
/* While a strip network is set, the strip's PML ramps over its 6 updated cells with
 * d0 = amp * v * ln(1e5) / (2 * 6 * d) (the reference's formula with 6 cells in place
 * of npml and amp in place of 10; default amp = 10). */
int a3d_set_strip_profile(a3d_ctx* ctx, float amp);
/* What the damping factors act on. 0 (default, "sponge"): P, its split components, u
 * and v are scaled by 1 - g_P; Q, its split components and w by 1 - g_Q. 1
 * ("normal"): only the split components along the cell's outward normal axis a
 * (nearest face) and the normal velocity: d = g_P * p_a; p_a -= d; P -= d (Q, q_a
 * with g_Q likewise), and the normal velocity (u, v for a = x, y with g_P; w with g_Q)
 * scaled by 1 - g -- a local increase of the PML's own damping along a. Cells whose
 * stress PML factor along a is 1 are left alone. */
int a3d_set_strip_head(a3d_ctx* ctx, int mode);
/* Head mode 2 ("coefficient"): the network's first output y_0 gives each strip cell the
 * multiplier m = 3 sigmoid(y_0) in (0, 3) of its PML damping (m = 1 at y_0 = -ln 2: the
 * plain ramp). The velocity and stress sweeps recompute the cell's split-PML coefficients
 * c1 = 1 / (1 + m h), c2 = c1 (1 - m h) from m and the profile's half damping step h =
 * dt d / 2 inside their update; nothing is applied after the sweeps. The network runs after
 * the stress sweep of every `steps`-th time step (a3d_set_strip_period, default 1; 0: after
 * the first step's only, for a network whose multipliers depend on the model and the strip
 * depth alone) and the multipliers hold until it runs again (1 before its first run). w may
 * be updated by either sweep (a3d_set_w_sweep). */
int a3d_set_strip_period(a3d_ctx* ctx, int steps);
/* Number of strip cells of the context's grid (0 before a network was set). */
int a3d_strip_info(const a3d_ctx* ctx, int* ncells);
/* Teacher-forced training samples: runs shot *desc on `teacher` (same interior,
 * wider PML) and copies the teacher's state onto `student` (nb = A3D_STRIP_NB,
 * network set) at steps start, start+h, start+2h, ... (h = horizon; horizon <= 1:
 * h = every). With horizon > 1 the student runs its own steps in between, its
 * network's strip corrections included. At steps start, start+every, ... a fraction
 * `keep` of the strip cells yields, after the student's stress sweep, features
 * X[16], target T[2] = (teacher P, Q - student P, Q) / s and scale S = s.
 * *count: samples written (<= cap). A student with the coefficient head (mode 2) is sampled
 * teacher-forced only (horizon <= 1), its input 15 filled as the head reads it. */
int a3d_strip_collect(a3d_ctx* teacher, a3d_ctx* student, const a3d_shot_desc* desc, int every, int start,
                      int horizon, float keep, unsigned seed, float* X, float* T, float* S, int cap, int* count);

/* The network on the GPU: forward, gradient and Adam training kernels. */
typedef struct a3d_mlp a3d_mlp;
int a3d_mlp_nparams(void);
int a3d_mlp_create(int device, a3d_mlp** out);
int a3d_mlp_destroy(a3d_mlp* m);
int a3d_mlp_set_params(a3d_mlp* m, const float* params, int n);
int a3d_mlp_get_params(a3d_mlp* m, float* params, int n);
/* What predict, loss_grad and train fit: 0 (default) the damping head's change below;
 * 2 the coefficient head's multiplier, Y = (m, 0) with m = 3 sigmoid(y_0) for every sample
 * (output 1 unused, no loss on it). */
int a3d_mlp_set_head(a3d_mlp* m, int head);
/* Y[n][2]: the damping head's change of the cell's P/s and Q/s,
 * (-X[.][0] g_P, -X[.][7] g_Q), for features X[n][16] (head 0; see a3d_mlp_set_head). */
int a3d_mlp_predict(a3d_mlp* m, const float* X, int n, float* Y);
/* Mean over samples of the summed squared error of Y against T and its gradient (FP32). */
int a3d_mlp_loss_grad(a3d_mlp* m, const float* X, const float* T, int n, double* loss, float* grad);
/* Adam over shuffled minibatches, learning rate cosine-decayed from lr to lr_end;
 * loss_per_epoch (may be NULL): mean training loss of each epoch. */
int a3d_mlp_train(a3d_mlp* m, const float* X, const float* T, int n, int epochs, int batch, float lr, float lr_end,
                  unsigned seed, double* loss_per_epoch);

/* ----------------------------------------------------------------------- */
/* Allocation-free shot. No counterpart in the reference: for callers that  */
/* cannot allocate device memory (tight serving loops, harnesses that        */
/* forbid hipMalloc).                                                        */
/* ----------------------------------------------------------------------- */
/* Size of the library's static scratch array, held by every device the library runs
 * the entry point below on, until the process ends. */
#define A3D_STATIC_SCRATCH_MIB 256
/* One shot of *desc, as a3d_run with FP32 storage (storage 0), stress scheme 0, 1, 3 or 4,
 * no snapshots, the default launch shape and w in the velocity sweep, calling no
 * hipMalloc/hipFree or other allocating routine: the padded model, wavefields, tables
 * and receiver record live in the static scratch array (A3D_STATIC_SCRATCH_MIB MiB per
 * device); a shot that needs more returns A3D_EINVAL with the size it would need.
 * vel..phi: DEVICE pointers to the interior model arrays (nx*ny*nz floats, z fastest, as
 * for a3d_set_model; NULL where the operator uses none). shot: DEVICE pointer to the
 * record [ny][nx][nt] (time fastest, muted if desc->mute). stream: a hipStream_t, or
 * NULL for the default stream. The call synchronizes with the stream on entry and
 * returns after the record is written (it cannot be captured into a graph). Results
 * equal a3d_run's bit for bit. All calls on one device share its scratch array: calls
 * that overlap on one device (several host threads) race and must be serialized by the
 * caller. */
int a3d_run_static(const a3d_shot_desc* desc, int device, const float* vel, const float* rho, const float* eps,
                   const float* del, const float* theta, const float* phi, float* shot, void* stream);

/* Padded grid actually updated per step: points = (nnx-8)*(nny-8)*(nnz-8). */
int a3d_grid_info(const a3d_ctx* ctx, int* nnx, int* nny, int* nnz, long long* updated_points);

/* Measurement support: achievable HBM bandwidth of a float4 copy (read+write
 * bytes / time; one float4 per thread, non-temporal) and of a float4 read-only
 * stream (four per thread, non-temporal) over `bytes` per buffer, `reps`
 * repetitions, on `device`. It first runs the grid-stride float4 copy
 * (k_probe_copy, 1024 x 256 threads), the calibration dispatch of the
 * FETCH_SIZE / WRITE_SIZE byte counters. */
int a3d_probe_bandwidth(int device, size_t bytes, int reps, double* copy_gbs, double* read_gbs);

/* Measurement support: the VTI stress sweep's memory access pattern replayed without
 * its arithmetic on a padded nnx x nny x nnz grid (20-cell PMLs), `reps` launches per
 * pattern on `device`. ms[0..4] and bytes[0..4] (algorithmic bytes per launch) for the
 * patterns: 0 its six read and three write streams over every lane (no leads, separate
 * outputs); 1 the sweep's interior pattern (in place, the z+3 / z+7 leads); 2 plus the
 * z-PML split pairs; 3 plus the y-PML pairs; 4 plus the x-PML pairs (the whole pattern). */
int a3d_probe_stress_pattern(int device, int nnx, int nny, int nnz, int reps, double* ms, double* bytes);

/* Measurement support: the FP32 sine and cosine the TTI sweeps use for the rotation terms
 * with a3d_set_tti_angles(ctx, 1), evaluated on `device` for the n angles t (host arrays). */
int a3d_probe_rot_sincos(int device, const float* t, float* s, float* c, int n);

const char* a3d_last_error(void);
const char* a3d_version(void);

#ifdef __cplusplus
}
#endif
#endif
