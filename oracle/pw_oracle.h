/*
 * pw_oracle.h -- CPU restatement of the reference's 3-D Biot poroelastic
 * staggered-grid finite-difference path (Float/main.cu and Half/main.cu).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the parity oracle: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product library never links, loads or calls it.
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py):
 *   or_create_model_borehole  -- method.md:919-1022 closed forms (rho, mu, Kb, Ks,
 *                                alpha, M, C, H, H-2mu, C1, C2)
 *   or_precompute_*           -- closed forms (d0, staggered means) + fixtures below
 *   or_run_f32                -- the 11 committed Float/<name>.dat snapshots (192^3,
 *                                501..3500 steps), mirror symmetry, fluid vw == 0
 *   or_run_f64                -- or_run_f32 within FP32 rounding; textbook acoustic
 *                                limit (fluid medium: c = sqrt(Kf/rhof) to 0.4 %, 1/r
 *                                spreading, vw and shear stresses exactly 0);
 *                                linearity in the source (bit-exact)
 *   or_half_*  codec          -- IEEE-754 binary16 (exhaustive vs numpy float16),
 *                                residual bound |dec(enc(v)) - v| <= unit/2
 *   or_half_precompute        -- method.md:1073-1169, 1187-1307 (em=-23, Aw, Bp, Bt,
 *                                scaled moduli)
 *   or_run_half               -- no reference data (no Half output is committed in
 *                                the reference); pinned by invariants: within the
 *                                compensated format's error of or_run_f64, outputs
 *                                independent of the internal scaling (dynamic rescale
 *                                off / every step, es = 4), linear in the source.
 */
#ifndef PW_ORACLE_H
#define PW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Field ids shared by receivers, snapshots and final-state dumps. */
enum {
    OR_VUX = 0, OR_VUY, OR_VUZ, OR_VWX, OR_VWY, OR_VWZ,
    OR_TXX, OR_TYY, OR_TZZ, OR_TXY, OR_TXZ, OR_TYZ, OR_P, OR_NFIELDS
};

/* Base (cell-centred, interior) medium fields, in this order, each
 * (nz-2np) x (ny-2np) x (nx-2np), x fastest.  This is what create_model
 * (Float/main.cu:1590-1655) hands to the rest of the program. */
enum { OR_RHO = 0, OR_RHOF, OR_C1, OR_C2, OR_HH, OR_H2U, OR_C, OR_M, OR_MU, OR_NBASE };

typedef struct {
    int nx, ny, nz;          /* extended grid incl. PML (NX_ext, NY_ext, NZ_ext) */
    int np;                  /* PML thickness in cells (NP) */
    double h, dt;            /* H, DT */
    double rc, v0;           /* PML reflection coefficient RC and velocity v0 */
    int nt;                  /* steps to run */
    double f0;               /* Ricker peak frequency F0, T0 = 1.2/F0 */
    int sx, sy, sz;          /* source cell */
    double src_w[4];         /* weights for txx, tyy, tzz, p (reference: 1,1,1,0) */
    int nrec;                /* receivers: (field id, flat extended index) */
    const int* rec_field;
    const long long* rec_index;
    float* seis;             /* out: [nrec][nt] */
    int nsnap;               /* snapshots: y-slice at snap_y after step snap_step */
    const int* snap_step;
    const int* snap_field;
    int snap_y;
    float* snaps;            /* out: [nsnap][nz][nx] */
    float* final_state;      /* out (optional): [13][nz*ny*nx] */
    int pml_everywhere;      /* 1: run every CPML recursion on the whole grid as
                                written; 0: skip cells whose profile is zero on that
                                axis (bit-identical, the memory variable stays 0) */
    int nthreads;            /* OpenMP threads (0: runtime default) */
    /* Half path only (Half/main.cu:27-35) */
    int half_es;             /* WAVEFIELD_ES_INT */
    double half_dyn_end;     /* DYNAMIC_SCALE_END */
    int half_dyn_interval;   /* DYNAMIC_SCALE_INTERVAL */
    int half_em_out, half_epml_out;  /* out: chosen exponents */
    int order;               /* spatial order of the staggered differences: 2 (the
                                reference; 0 means 2) or 4 (FP32/FP64 runs only) */
    int pml_open;            /* faces without an absorbing layer, bit 2*axis (low
                                face) and 2*axis+1 (high face), axis 0 = x: the
                                interior profile (d = 0, e = 1) on that half of the
                                axis.  0: the reference (all six faces damped) */
} or_args;

/* create_model (Float/main.cu:1590-1655): the borehole-in-formation model.
 * out: OR_NBASE arrays of nz*ny*nx floats (interior sizes). */
void or_create_model_borehole(int nz, int ny, int nx, float* out);

/* Moduli-form Biot medium for one homogeneous cell (same formulas as
 * Float/main.cu:1635-1650, inputs given as moduli).  out[OR_NBASE]. */
void or_biot_from_moduli(double Ks, double Kb, double mu, double Kf, double rhos,
                         double rhof, double phi, double kappa, double eta,
                         double tortuosity, double pore_m, double* out);

/* 1-D CPML profiles (Float/main.cu:570-654) along one axis of n cells:
 * d, d2 (half-point), e = exp(-d dt), e2. */
void or_pml_profile_f32(int n, int np, float h, float dt, float rc, float v0,
                        float* d, float* d2, float* e, float* e2);
void or_pml_profile_f64(int n, int np, double h, double dt, double rc, double v0,
                        double* d, double* d2, double* e, double* e2);

/* Ricker source sample at step it (Float/main.cu:1050, 1059). */
float or_ricker_f32(int it, float dt, float f0);
double or_ricker_f64(int it, double dt, double f0);

/* Full runs.  base: OR_NBASE interior arrays (float for f32/half, double for f64).
 * Return 0 on success, nonzero on bad arguments / allocation failure. */
int or_run_f32(or_args* a, const float* base);
int or_run_f64(or_args* a, const double* base);
int or_run_half(or_args* a, const float* base);

/* CPU baseline timing: set up as or_run_f32, run `warm` untimed steps, then time
 * `steps` steps (source + FD_V + FD_T, slab skip per a->pml_everywhere) with the
 * OpenMP wall clock.  *seconds receives the timed total; *threads the team size. */
int or_time_f32(or_args* a, const float* base, int warm, int steps, double* seconds, int* threads);

/* Staggered medium precompute exposed for tests: 17 arrays of nz*ny*nx floats in
 * the order rho_tempx,y,z, rhof, C1x,y,z, C2x,y,z, C, M, HH, H2u, muxy, muxz, muyz. */
int or_precompute_f32(const or_args* a, const float* base, float* out17);

/* Compensated FP16 codec (Half/main.cu:49-100). */
unsigned short or_f32_to_f16(float v);
float or_f16_to_f32(unsigned short h);
float or_residual_unit(float hi_float);
void or_store_ch(float v, unsigned short* hi, signed char* lo);
float or_load_ch(unsigned short hi, signed char lo);
/* the two over arrays of n elements */
void or_load_ch_n(const unsigned short* hi, const signed char* lo, long n, float* out);
void or_store_ch_n(const float* v, long n, unsigned short* hi, signed char* lo);

/* Half exponent + coefficient precompute (Half/main.cu:745-1045) for a
 * homogeneous-by-cell medium: returns em, epml and writes the 22 FP16 coefficient
 * arrays (order: rho_inv_tilde x,y,z, rhof_rho_inv x,y,z, Aw x,y,z, Bp x,y,z,
 * Bt x,y,z, HH, H2u, C, M, muxy, muxz, muyz) when coef != 0. */
int or_half_precompute(const or_args* a, const float* base, int* em, int* epml,
                       unsigned short* coef22);

#ifdef __cplusplus
}
#endif
/* The staggered difference the update applies (order 2 or 4), FP64, on a 1-D
 * array: forward (point o + st/2) or backward (point o - st/2), undivided. */
double or_stagger_diff_f64(int order, int forward, const double* f, long o, long st);

#endif
