/*
 * porowave.h -- C ABI of the MI355X-native 3-D Biot poroelastic staggered-grid
 * finite-difference solver.
 *
 * The reference has no library interface: its operator surface is a monolithic
 * main() (Float/main.cu:300-1477, Half/main.cu:376-1850) whose parameters are
 * #defines and literals and whose outputs are .dat text files.  Each entry point
 * below replaces one part of that surface; the comment on each names what it
 * replaces.  Plain C types only: opaque handle, plain pointers and sizes, status
 * codes, thread-local error text.
 *
 * Conventions
 *  - Grid sizes nx, ny, nz are the TOTAL grid including the CPML pad (the
 *    reference's NX_ext = NX + 2*NP).  Flat index = ix + iy*nx + iz*nx*ny, x fastest
 *    (Float/main.cu:96).  The outermost layer of cells is never updated
 *    (Float/main.cu:103, 232).
 *  - Staggering: p and tau_ii at (i,j,k); vux/vwx at (i+1/2,j,k); vuy/vwy at
 *    (i,j+1/2,k); vuz/vwz at (i,j,k+1/2); txy (i+1/2,j+1/2,k); txz (i+1/2,j,k+1/2);
 *    tyz (i,j+1/2,k+1/2).
 *  - All outputs are in physical units (the compensated-FP16 exponent and dynamic
 *    scales are undone inside the library).
 *  - The handle owns its device memory; the caller owns host buffers.  One handle
 *    is not thread-safe; distinct handles may run concurrently on distinct streams.
 *  - pw_run() is asynchronous on the handle's stream; pw_get_* synchronise.
 */
#ifndef POROWAVE_H
#define POROWAVE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Raised whenever a struct below changes size or layout (3: pw_server_config
 * gained persistent and max_wait_s). */
#define PW_ABI_VERSION 3

typedef struct pw_solver pw_solver;

typedef enum {
    PW_OK = 0,
    PW_ERR_INVALID = 1,       /* bad argument or configuration */
    PW_ERR_HIP = 2,           /* HIP runtime error */
    PW_ERR_OUT_OF_MEMORY = 3, /* device allocation refused */
    PW_ERR_STATE = 4,         /* call out of order (e.g. run before set_medium) */
    PW_ERR_UNSTABLE = 5,      /* Courant condition violated by the medium */
    PW_ERR_UNSUPPORTED = 6,   /* valid request the library does not implement */
    PW_ERR_IO = 7,            /* file could not be written */
    PW_ERR_TIMEOUT = 8,       /* pw_server_wait: the shot did not finish in time */
    PW_ERR_CANCELLED = 9,     /* the shot was cancelled (pw_server_cancel) */
    PW_ERR_NONFINITE = 10     /* the shot's seismogram holds NaN or Inf (an unstable run) */
} pw_status;

/* Storage precision of wavefields and CPML memory variables.  Arithmetic is FP32
 * except for PW_FP64, which is the FP64 oracle run (everything in double). */
typedef enum {
    PW_FP32 = 0,   /* Float/main.cu storage */
    PW_COMP16 = 1, /* Half/main.cu: FP16 value + int8 residual, FP16 coefficients */
    PW_BF16 = 2,   /* BF16 wavefields + memory variables, FP16 coefficients */
    PW_FP64 = 3,   /* FP64 oracle run */
    PW_COMP16F = 4 /* the comp16 state (FP16 value + int8 residual, residual rounded to
                      nearest; stresses stored times 2^(es+6), velocities 2^(es+6-em) with
                      em of Half/main.cu:793-807) under an FP32 update of the Float
                      program's equations: cell-centred coefficients held as
                      compensated values (FP16 + int8, a power-of-two scale each),
                      FP32 memory variables, hardware reciprocals */
} pw_precision;

/* Wavefield ids for receivers, snapshots and pw_get_field. */
enum {
    PW_VUX = 0, PW_VUY, PW_VUZ, PW_VWX, PW_VWY, PW_VWZ,
    PW_TXX, PW_TYY, PW_TZZ, PW_TXY, PW_TXZ, PW_TYZ, PW_P,
    PW_NFIELDS
};

/* Cell-centred base medium fields handed to pw_set_medium, in this order: what
 * create_model (Float/main.cu:1590-1655) produces and the rest of the reference
 * program consumes. */
enum {
    PW_RHO = 0, /* bulk density rho = (1-phi) rho_s + phi rho_f */
    PW_RHOF,    /* fluid density rho_f */
    PW_C1,      /* viscous coupling eta / kappa */
    PW_C2,      /* inertial coupling (1 + 2/m) tau rho_f / phi */
    PW_HH,      /* H = alpha^2 M + Kb + 4/3 mu */
    PW_H2U,     /* H - 2 mu */
    PW_CC,      /* C = alpha M */
    PW_MM,      /* M (Biot modulus) */
    PW_MU,      /* shear modulus */
    PW_NBASE
};

typedef struct pw_config {
    int nx, ny, nz;         /* total grid incl. CPML (NX_ext, NY_ext, NZ_ext) */
    int np;                 /* CPML thickness in cells (NP) */
    double h;               /* grid spacing [m] (H) */
    double dt;              /* time step [s] (DT) */
    int nt;                 /* steps in the shot: sizes the seismogram buffer and
                               sets the comp16 dynamic-scale ramp (NT) */
    int precision;          /* pw_precision */
    double cpml_r;          /* CPML reflection coefficient (RC; reference 1e-6) */
    double cpml_v0;         /* CPML velocity v0 (reference 4000 m/s) */
    int half_es;            /* comp16/bf16: WAVEFIELD_ES_INT (reference 0); comp16f
                               stores stresses times 2^(half_es + 6) */
    double half_dyn_end;    /* comp16/bf16: DYNAMIC_SCALE_END (reference 16) */
    int half_dyn_interval;  /* comp16/bf16: DYNAMIC_SCALE_INTERVAL (reference 100) */
    int device;             /* HIP device ordinal */
    void* stream;           /* hipStream_t to run on; NULL: the library creates one */
    int check_cfl;          /* 1: pw_set_medium refuses media violating the Courant
                               limit (the reference only prints a warning,
                               Float/main.cu:940-942) */
    int space_order;        /* staggered differences: 2 (the reference, Float/main.cu:115-128,
                               234-249) or 4 (9/8, -1/24; two-cell halo, the outermost
                               two layers are not updated; Courant limit 6/7).  4 runs
                               the two-kernel step with one cell per lane. */
    int cpml_open;          /* faces without an absorbing layer, a bit mask
                               (PW_FACE_*; 0, the default, damps all six as the
                               reference does).  An open face keeps the interior
                               profile (no damping, no memory variables) and its
                               outermost plane is held like every edge plane: the
                               ghost plane of one rank's z-slab in a decomposed
                               grid.  No counterpart in the reference. */
    /* Kernel tiling of the two step kernels (no counterpart in the reference,
       whose kernels take one thread per cell).  0 selects the measured default
       of the precision; every setting gives bit-identical results. */
    int tile_cells;         /* cells of an x row per lane: 1, 2 or 4 (order 4: 1) */
    int tile_waves_x;       /* waves side by side along x in a tile: 1 or 2 */
    int tile_rows;          /* waves per block: 1, 2, 4 or 8 */
    int tile_planes;        /* z planes each block marches through, >= 1 */
} pw_config;

/* pw_config.cpml_open bits */
#define PW_FACE_X_LO 1
#define PW_FACE_X_HI 2
#define PW_FACE_Y_LO 4
#define PW_FACE_Y_HI 8
#define PW_FACE_Z_LO 16
#define PW_FACE_Z_HI 32

/* Fills a config with the reference's values (Float/main.cu:20-23, 305-322):
 * 192^3 total, NP 32, H 0.01 m, DT 0.9 us, NT 3500, RC 1e-6, v0 4000, FP32. */
void pw_config_default(pw_config* cfg);

/* Replaces main()'s setup: cudaSetDevice + the 93 cudaMalloc calls
 * (Float/main.cu:814-935).  One pooled device allocation per handle. */
pw_status pw_create(const pw_config* cfg, pw_solver** out);

/* Replaces the frees at Float/main.cu:1225-1474. NULL is accepted. */
void pw_destroy(pw_solver* s);

/* Text of the last error raised on the calling thread ("" if none). */
const char* pw_last_error(void);

/* Replaces extmodel (Float/main.cu:387-399, 1662-1794), the staggered averaging
 * (:473-513), the Half exponent/quantisation precompute (Half/main.cu:745-1045)
 * and the 93 H2D copies (Float/main.cu:948-1040).  base: PW_NBASE pointers to
 * cell-centred interior fields of nxi*nyi*nzi doubles (nxi = nx - 2*np, ...),
 * x fastest.  The pad is filled by edge replication, staggered coefficients are
 * computed in FP64 on the GPU and stored in the handle's precision.  Resets the
 * wavefields to zero and the step counter to 0.  PW_HH and PW_H2U are H and
 * H - 2 mu (Float/main.cu:1644-1645): when they agree with PW_MU to 1e-6 of H,
 * as the reference's formulas make them, FP32 forms H as (H - 2 mu) + 2 mu in
 * its stress update; comp16f always does, and refuses a medium where they do
 * not agree (PW_ERR_UNSUPPORTED). */
pw_status pw_set_medium(pw_solver* s, const double* const* base, int nxi, int nyi, int nzi);

/* pw_set_medium for base fields already in device memory, allocating nothing.
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  d_base: PW_NBASE device pointers, on the handle's device,
 * to the cell-centred interior fields (nxi*nyi*nzi doubles each, x fastest), as
 * pw_set_medium takes them in host memory; the installed medium is
 * bit-identical.  The medium's maxima (stability check, Half exponents, comp16f
 * coefficient scales) are reduced into a file-scope device array of 64 doubles
 * instead of an allocation; the reduction needs 15 of them whatever the grid
 * size.  Every call
 * on a device shares that device's array and clears it first; calls in one
 * process are serialised on it (calls that overlapped on one device would
 * race).  Refuses a handle with cpml_v0 <= 0 (PW_ERR_UNSUPPORTED): its CPML
 * tables are built from the medium on the host and uploaded. */
pw_status pw_set_medium_device(pw_solver* s, const double* const* d_base, int nxi, int nyi, int nzi);

/* Reference Ricker source (Float/main.cu:310-313, 321-322, 1059-1060): wavelet
 * -(1 - 2 pi^2 f0^2 tau^2) exp(-pi^2 f0^2 tau^2), tau = it*dt - t0, t0 = 1.2/f0,
 * added to txx, tyy, tzz and p at cell (sx, sy, sz) with the given weights
 * (reference: {1, 1, 1, 0}; NULL means that). */
pw_status pw_set_source(pw_solver* s, int sx, int sy, int sz, double f0, const double* weights4);

/* Replaces the receiver gather (Float/main.cu:1083-1093): record field[r] at
 * flat index[r] after every step into a device buffer of [n][nt] floats. */
pw_status pw_set_receivers(pw_solver* s, int n, const int* field, const long long* index);

/* The reference's receivers: every depth iz of the vertical line through (sx, sy)
 * for the 5 channels sisx/sisz/sisvu/sisvw/sisp = txx/tzz/vux/vwx/p
 * (Float/main.cu:1085-1093).  Channel c, depth iz is receiver c*nz + iz. */
pw_status pw_set_receivers_reference(pw_solver* s, int sx, int sy);

/* Replaces the snapshot copies (Float/main.cu:1094-1163): after step step[k],
 * copy the y = plane_y slice of field[k] ([nz][nx] floats). */
pw_status pw_set_snapshots(pw_solver* s, int n, const int* step, const int* field, int plane_y);

/* Replaces the time loop (Float/main.cu:1048-1164): enqueue nsteps steps
 * (source -> velocity kernel -> stress kernel -> receivers/snapshots) on the
 * handle's stream.  Asynchronous.  Steps continue from the current counter and
 * may not exceed cfg.nt in total. */
pw_status pw_run(pw_solver* s, int nsteps);

/* Blocks until all work enqueued on the handle's stream has finished. */
pw_status pw_synchronize(pw_solver* s);

/* Steps completed (enqueued) so far. */
int pw_current_step(const pw_solver* s);

/* Zero the wavefields and memory variables and reset the step counter; keeps the
 * medium, source, receivers and snapshots (a new shot on the same handle). */
pw_status pw_reset(pw_solver* s);

/* Copies out the seismograms: [nrec][nt] floats (steps not yet run are 0). */
pw_status pw_get_seismogram(pw_solver* s, float* host);

/* Copies out the snapshots: [nsnap][nz][nx] floats. */
pw_status pw_get_snapshots(pw_solver* s, float* host);

/* Copies out one full wavefield decoded to physical FP32: nz*ny*nx floats. */
pw_status pw_get_field(pw_solver* s, int field, float* host);

/* Full 3-D snapshots without stalling the run (no reference counterpart: the
 * reference copies whole arrays to the host synchronously inside its loop,
 * Float/main.cu:1075-1081, and keeps y-slices only).  After step step[k] the
 * field is decoded to physical FP32 on the handle's stream into one of two
 * device staging volumes, and copied to library-owned pinned host memory on a
 * second stream while the following steps run.  Replaces any earlier set; n = 0
 * clears it.  Holds n * nz*ny*nx floats of pinned host memory and two staging
 * volumes on the device. */
pw_status pw_set_volume_snapshots(pw_solver* s, int n, const int* step, const int* field);
/* Waits for volume snapshot k's copy and copies it out: nz*ny*nx floats.
 * PW_ERR_STATE if its step has not run yet. */
pw_status pw_get_volume_snapshot(pw_solver* s, int k, float* host);

/* Replaces wfile / wfile3d (Float/main.cu:1501-1516, 1563-1588): writes nr rows
 * of nc values as "%e " text, each row followed by "\n" and a blank line. */
pw_status pw_write_dat(const char* path, const float* data, int nr, int nc);

/* Binary seismogram / snapshot file (no reference counterpart: the reference
 * writes text only).  An 80-byte header, then nr*nc little-endian FP32 values,
 * row-major.  Values times `scale` are in physical units (the library writes 1). */
enum { PW_BIN_SEISMOGRAM = 0, PW_BIN_SNAPSHOT = 1 };
typedef struct pw_bin_header {
    char magic[4];          /* "PWB1" (set by pw_write_bin) */
    int version;            /* 1 (set by pw_write_bin) */
    int kind;               /* PW_BIN_SEISMOGRAM ([receiver][step]) or PW_BIN_SNAPSHOT ([z][x]) */
    int nr, nc;             /* rows, columns */
    int field;              /* pw field id, or -1 for a mixed receiver set */
    int step;               /* snapshot step, or -1 */
    int reserved;
    double dt, h;           /* time step [s] and grid step [m] of the run */
    double scale;
    char label[24];         /* NUL-terminated name, e.g. "sisx", "txx500" */
} pw_bin_header;
pw_status pw_write_bin(const char* path, const pw_bin_header* hdr, const float* data);
/* Reads the header, and nr*nc values into data unless data is NULL. */
pw_status pw_read_bin(const char* path, pw_bin_header* hdr, float* data);

/* Writes the reference program's output set from a handle into directory dir
 * (Float/main.cu:1187-1222): with pw_set_receivers_reference, sisx, sisz, sisvu,
 * sisvw, sisp ([nz][nt] each); otherwise one "seismogram" file ([nrec][nt]); and
 * one file per snapshot named field + step ("txx500", ...).  format: PW_FORMAT_DAT
 * (the reference's text, ".dat") or PW_FORMAT_BIN (".bin"). */
enum { PW_FORMAT_DAT = 0, PW_FORMAT_BIN = 1 };
pw_status pw_write_outputs(pw_solver* s, const char* dir, int format);

typedef struct pw_info {
    int precision;
    int em, epml;               /* comp16/bf16 exponents (Half/main.cu:793-843) */
    double bytes_per_cell_step; /* algorithmic HBM bytes per cell-step, this grid */
    double bytes_velocity;      /* ... of which the velocity kernel (per cell) */
    double bytes_stress;        /* ... of which the stress kernel (per cell) */
    double vp_max;              /* fastest P velocity of the medium [m/s] */
    double courant;             /* vp_max * dt / h * sqrt(3) */
    long long device_bytes;     /* device memory held by the handle */
    int slab_lo[3], slab_hi[3]; /* CPML slab widths per axis */
} pw_info;

pw_status pw_get_info(pw_solver* s, pw_info* out);

/* Per-kernel timing: with profiling enabled, pw_run brackets every velocity and
 * stress launch with HIP events on the handle's stream.  pw_get_kernel_times
 * synchronises, returns the summed milliseconds of each kernel and the number of
 * profiled steps since the last call, and clears them.  No counterpart in the reference, which times only the
 * whole loop (Float/main.cu:1043-1046, 1176-1184). */
pw_status pw_set_profiling(pw_solver* s, int enable);
pw_status pw_get_kernel_times(pw_solver* s, double* ms_velocity, double* ms_stress, long long* launches);

/* The compensated FP16 codec of the comp16 and comp16f step kernels, run alone
 * on the GPU so that it can be checked value by value.  No counterpart in the
 * reference (its codec, Half/main.cu:49-100, is inline device code).  All
 * pointers are device memory of n elements on the current device; stream is a
 * hipStream_t (NULL: the null stream); both calls wait for the stream.
 * pw_codec_decode: out[i] = hi[i] + lo[i] * u(hi[i]), hi FP16 bit patterns, u
 *   the residual unit 2^(max(e,1) - 33) for hi's biased exponent e.
 * pw_codec_encode: hi[i] = in[i] rounded to FP16 (nearest, ties to even), lo[i]
 *   = (in[i] - hi[i]) / u(hi[i]) rounded as the precision's kernels round it and
 *   clamped to [-127, 127]: PW_COMP16 half away from zero (the reference's
 *   store_ch), PW_COMP16F to nearest, ties to even. */
pw_status pw_codec_decode(const unsigned short* hi, const signed char* lo, float* out, long long n, void* stream);
pw_status pw_codec_encode(const float* in, unsigned short* hi, signed char* lo, long long n, int precision,
                          void* stream);

/* ---------------------------------------------------------------------------
 * Medium front-ends (host, FP64).  Each writes PW_NBASE arrays of nxi*nyi*nzi
 * doubles into out (contiguous, field-major).
 * ------------------------------------------------------------------------- */

/* create_model (Float/main.cu:1590-1655): fluid borehole of radius 10 cells on
 * the vertical axis through the interior centre, in the reference formation.
 * Values are the reference's float results widened to double. */
pw_status pw_medium_borehole(int nxi, int nyi, int nzi, double* out);

typedef struct pw_rock {
    double Ks, Kb, mu, Kf;    /* grain, dry-frame, shear, fluid moduli [Pa] */
    double rhos, rhof;        /* grain and fluid densities [kg/m^3] */
    double phi;               /* porosity */
    double kappa, eta;        /* permeability [m^2], viscosity [Pa s] */
    double tortuosity, pore_m;/* tortuosity, pore-shape factor m (reference 8) */
} pw_rock;

/* Biot coefficients of one rock (the formulas of Float/main.cu:1635-1650 with
 * moduli given directly): out[PW_NBASE]. */
pw_status pw_biot_from_rock(const pw_rock* r, double* out);

#define PW_MAX_LAYERS 16
typedef struct pw_layered {
    int nlayers;                 /* 1..PW_MAX_LAYERS */
    int z_top[PW_MAX_LAYERS];    /* first total-grid z index of each layer, ascending,
                                    z_top[0] == 0 */
    pw_rock rock[PW_MAX_LAYERS];
    double perturb;              /* relative amplitude of the Kb/mu/phi perturbation
                                    (0.05 = 5 %; 0 disables) */
    double corr_cells;           /* Gaussian smoothing sigma in cells */
    double clamp;                /* clamp of the unit-variance field (3) */
    unsigned long long seed;
} pw_layered;

/* Horizontally layered medium with a smooth random perturbation, generated on
 * the handle's GPU in FP64 and installed as with pw_set_medium.  The perturbation
 * of Kb, mu and phi is white noise (splitmix64 counter hash, Box-Muller)
 * smoothed by a separable Gaussian, scaled to unit variance, clamped and applied
 * as a factor (1 + perturb * n). */
pw_status pw_set_medium_layered(pw_solver* s, const pw_layered* spec);

/* Copies the cell-centred base fields installed by the last pw_set_medium* call:
 * PW_NBASE arrays of nxi*nyi*nzi doubles (for feeding the same medium to other
 * tools). */
pw_status pw_get_base_fields(pw_solver* s, double* out);

/* ---------------------------------------------------------------------------
 * Shot server.
 *
 * No counterpart in the reference, which runs one shot per process: main()
 * sets up, steps and writes .dat files for a single source
 * (Float/main.cu:300-1477).  The server replaces a queue of such runs.  It
 * admits shot requests first come, first served while the device-memory budget
 * and the slot count allow; it keeps every admitted shot resident; and it
 * advances all of them together, in rounds of steps:
 *  - one launch per step kernel covers every active shot of a precision;
 *  - a round that repeats is captured once as a hipGraph and replayed;
 *  - or (persistent) one cooperative launch per precision runs the whole round.
 * Shots join and leave between rounds.  A shot's seismogram is bit-identical
 * to the same shot run alone through pw_run, because the per-cell code is the
 * same.  All calls are thread-safe; one worker thread owns the device.
 * ------------------------------------------------------------------------- */
typedef struct pw_server pw_server;

typedef struct pw_server_config {
    int device;
    void* stream;           /* hipStream_t for all device work; NULL: the server creates one */
    long long hbm_budget;   /* device bytes the admitted shots may hold; 0: 90 % of the memory
                               free at creation, less the medium */
    int max_active;         /* shots advanced together (slots), >= 1 */
    int round_steps;        /* steps per scheduling round, >= 1 */
    int use_graphs;         /* 1: a round that repeats (same shots, same length) is captured
                               once as a hipGraph and replayed */
    int batched;            /* 1: one launch per step kernel for all active shots of a
                               precision; 0: each shot stepped by pw_run in turn */
    long long fill_cells;   /* admission stops adding shots once the admitted shots hold
                               this many grid cells (nx*ny*nz summed; a shot is always
                               admitted when none is active).  Shots large enough to
                               fill the GPU then run one or two at a time, first come
                               first served, which keeps latency at the sequential
                               level under load, while small shots still share the
                               device.  0: no limit besides max_active and the budget */
    int persistent;         /* 1: each round of a batch runs as ONE persistent cooperative
                               launch (as many blocks as the device holds at once, grid
                               barriers between the phases of a step) instead of one launch
                               per step kernel; results are bit-identical.  Needs batched */
    double max_wait_s;      /* admission order.  0: first come first served.  > 0: the queued
                               shot with the least work (cells x steps) first, except that
                               shots queued longer than max_wait_s seconds go first, oldest
                               first: short shots pass long ones, which wait at most about
                               that long before they are next.  No counterpart in the
                               reference */
} pw_server_config;

typedef struct pw_shot {
    int nx, ny, nz, np;     /* total grid and CPML width, as in pw_config */
    double h, dt;
    int nt;
    int precision;          /* pw_precision */
    double cpml_r, cpml_v0;
    int crop[3];            /* origin (x, y, z) of the shot's interior inside the server's
                               medium volume, in interior cells */
    int sx, sy, sz;         /* source cell, total grid */
    double f0;              /* Ricker peak frequency [Hz] */
    double w[4];            /* source weights into txx, tyy, tzz, p */
    int nrec;
    const int* rec_field;   /* nrec field ids; copied by pw_server_submit */
    const long long* rec_index; /* nrec flat total-grid indices; copied */
} pw_shot;

typedef enum {
    PW_SHOT_QUEUED = 0,
    PW_SHOT_RUNNING = 1,
    PW_SHOT_DONE = 2,
    PW_SHOT_FAILED = 3
} pw_shot_state;

typedef struct pw_shot_status {
    int state;                 /* pw_shot_state */
    int precision, nt, nrec;
    long long cells;           /* nx*ny*nz */
    long long device_bytes;    /* device memory the shot holds while admitted */
    double t_submit, t_admit, t_done; /* seconds on the server clock (-1: not yet) */
    pw_status error;           /* PW_OK unless state == PW_SHOT_FAILED */
    char error_text[256];
} pw_shot_status;

typedef struct pw_server_stats {
    long long submitted, completed, failed;
    long long rounds, graphs_built, launches;
    double busy_s;             /* time spent in rounds (device work issued and waited for) */
    double setup_s;            /* time spent admitting shots (allocation, medium, tables) */
    double graph_s;            /* time spent capturing and instantiating round graphs */
    double retire_s;           /* time spent copying results out and freeing finished shots */
    double cell_steps;         /* sum of nx*ny*nz*nt over completed shots */
    long long peak_bytes;      /* largest device memory held by admitted shots */
    int peak_active;           /* largest number of shots advanced together */
} pw_server_stats;

/* max_active 16, round_steps 32, graphs on, batched, budget 0 (automatic),
 * fill_cells 2^23 (half a 256^3 grid: a 256^3 shot runs alone, smaller shots
 * share the device), persistent off. */
void pw_server_config_default(pw_server_config* cfg);
pw_status pw_server_create(const pw_server_config* cfg, pw_server** out);
/* Finishes the admitted shots, fails the queued ones, stops the worker. */
void pw_server_destroy(pw_server* srv);
/* The medium volume shots are cropped from: PW_NBASE host arrays of nxi*nyi*nzi
 * doubles (interior cells, x fastest), kept on the device.  Only while no shot
 * is queued or admitted. */
pw_status pw_server_set_medium(pw_server* srv, const double* const* base, int nxi, int nyi, int nzi);
/* The same from a layered specification (pw_set_medium_layered's generator),
 * generated for an interior of nxi*nyi*nzi cells of a grid with CPML width np. */
pw_status pw_server_set_medium_layered(pw_server* srv, const pw_layered* spec, int nxi, int nyi, int nzi, int np);
/* Queues a shot (validated here); returns its id.  Non-blocking. */
pw_status pw_server_submit(pw_server* srv, const pw_shot* shot, long long* id);
/* Blocks until the shot is done or failed (its error), or timeout_s passes
 * (PW_ERR_TIMEOUT; timeout_s < 0 waits without limit). */
pw_status pw_server_wait(pw_server* srv, long long id, double timeout_s);
pw_status pw_server_status(pw_server* srv, long long id, pw_shot_status* out);
/* Copies a finished shot's seismogram, [nrec][nt] floats in physical units. */
pw_status pw_server_result(pw_server* srv, long long id, float* seis);
/* Forgets a finished or failed shot and its host-side result (PW_ERR_STATE while
 * the shot is queued or running, or while another thread waits on it). */
pw_status pw_server_release(pw_server* srv, long long id);
/* Cancels a shot.  A queued shot fails at once; a running one leaves the device
 * at the end of the current round (its memory is freed then).  Either way it
 * ends in PW_SHOT_FAILED with PW_ERR_CANCELLED, and the other shots are not
 * disturbed.  PW_ERR_STATE for a shot that has already finished. */
pw_status pw_server_cancel(pw_server* srv, long long id);
pw_status pw_server_get_stats(pw_server* srv, pw_server_stats* out);
/* Seconds since pw_server_create (the clock of pw_shot_status). */
double pw_server_clock(pw_server* srv);

#ifdef __cplusplus
}
#endif
#endif
