/*
 * oracle.h -- TEST INFRASTRUCTURE, not product code.
 *
 * CPU restatement of the reference's per-shot forward modeller arithmetic
 * (sh39o/Acoustic3d-fd, the four src/ CUDA files), used only as the checker by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg. The product
 * library never links, loads or calls anything under oracle/.
 *
 * Two instances of every operator: FP32 (the parity oracle: the reference's
 * float/double mix restated operation by operation) and FP64 (every value
 * and table in double: the error reference).
 */
#ifndef A3D_ORACLE_H
#define A3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OP_2ND = 0, ORC_OP_ISO = 1, ORC_OP_VTI = 2, ORC_OP_TTI = 3 };

typedef struct OracleShot {
    int op;              /* ORC_OP_* */
    int nx, ny, nz;      /* interior model size */
    float dx, dy, dz;
    int nb;              /* 2nd order: sponge width npd (ring mm=4 added); 1st order: npml */
    int sx, sy, sz;      /* source: 2nd order 0-based interior, 1st order 1-based interior */
    int nt;
    float dt, fm;
    int wtype;           /* 1st order wavelet code 1..6 (2nd order: Ricker table) */
    float pfac;          /* 1st order source factor (ref: 1e4 iso, 10 VTI/TTI) */
    int rec_z;           /* receiver plane, interior z index (ref: 1 for 2nd, 0 for 1st) */
    int sv;              /* VTI/TTI SV taper around the source */
    int snap_interval;   /* 0: no snapshots */
    int snap_skip0;      /* skip it == 0 */
    int mute;            /* apply the direct-wave mute to the shot */
    float v_mute;        /* <= 0: the reference's model-derived velocity */
    int mute_tt;         /* mute pad in samples (ref: 60) */
    int nthreads;        /* OpenMP threads (0: runtime default) */
} OracleShot;

/*
 * Runs one shot. prec = 32 (float outputs) or 64 (double outputs).
 * Model arrays: interior, z fastest (iz + nz*ix + nz*nx*iy), NULL when unused.
 * shot: [ny][nx][nt] (time fastest), undecimated.
 * snaps: nsnap_max interior crops of nz*nx*ny each (z fastest); may be NULL.
 * Returns the number of snapshots taken (>=0), or a negative error code.
 */
int oracle_run(int prec, const OracleShot* p,
               const float* vel, const float* rho, const float* eps,
               const float* del, const float* theta, const float* phi,
               void* shot, void* snaps, int nsnap_max, double* loop_seconds);

/* Learned absorbing strip of the FP32 VTI instance (grid padding nb = 10):
 * the network's 5378 parameters, or NULL to switch it off. Returns 0, or -1
 * for a wrong parameter count. */
int oracle_set_strip_net(const float* params, int n);

/* Strip network evaluation: 1 (default) binary16 operands of the hidden layers
 * (the library's matrix-core evaluation), 0 plain FP32 with fused multiply-adds. */
int oracle_set_strip_half(int on);
/* Nearest binary16 value of f (ties to even), as a float. */
float oracle_round_half(float f);

/* PML amplitude of the strip's 6-cell ramp while a strip network is set (default 10). */
int oracle_set_strip_profile(float amp);
/* What the damping factors act on: 0 (default) every field of the cell, 1 the split
 * components along the cell's normal axis and the normal velocity (cells whose stress
 * PML factor along the normal is 1 are left alone), 2 the coefficient head (below). */
int oracle_set_strip_head(int mode);
/* Head 2 ("coefficient"): the network's first output y_0 gives each strip cell the
 * multiplier m = 3 sigmoid(y_0) of its PML damping; the update rebuilds the cell's split-PML
 * coefficients c1 = 1/(1 + m h), c2 = c1 (1 - m h) from the profile's half damping step h.
 * The network runs after the stress update of every `steps`-th step (default 1); the
 * multipliers start at 1. Returns -1 for a bad value. */
int oracle_set_strip_period(int steps);

/* Plants a mistake in the TTI rotation so a test can show that it catches it
 * (0 = none, the default; 1..8 flip a velocity-sweep term's sign, 9..16 a
 * stress-sweep term's, 17/18 swap an operand; oracle.cpp lists them).
 * Returns -1 for an unknown code. */
int oracle_set_tti_fault(int code);

/* One strip correction on padded arrays (oracle layout: z fastest, then x, then y;
 * extents nnz, nnx, nny) with the network set above; -1 if none is set. */
int oracle_strip_apply(int nnx, int nny, int nnz, float dt, float dx, float dy, float dz, float* P, float* Q,
                       float* u, float* v, float* w, float* px, float* py, float* pz,
                       float* qx, float* qy, float* qz, const float* vp, const float* rho, const float* eps,
                       const float* del);

/* Time loop only, for the CPU baseline: runs `nsteps` steps and returns seconds. */
double oracle_time_steps(const OracleShot* p, const float* vel, const float* rho,
                         const float* eps, const float* del, const float* theta,
                         const float* phi, int nsteps);

/* Setup tables, exposed so tests can pin them against closed forms. */
int oracle_ricker_table(int prec, float fm, float dt, int nt, void* out);
int oracle_wavelet(int prec, int wtype, float fm, float dt, int nt, void* out);
int oracle_pml_tables(int prec, int n, int npml, float dt, float d0, void* coff1,
                      void* coff2, void* acoff1, void* acoff2);
double oracle_d0(const float* vp_padded, int nnx, int nny, int nnz, int npml,
                 float dx, float dy, float dz);
int oracle_pad(const float* in, int nx, int ny, int nz, int nb, float* out);
int oracle_sponge_factor(int n, int npd, double* out);
int oracle_taper(int r2, double* out);
int oracle_staggered_dx(const float* in, int n, int forward, float* out);
int oracle_laplacian_z(const float* in, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif
