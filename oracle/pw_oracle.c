/*
 * pw_oracle.c -- CPU restatement of the reference's Biot staggered-grid FD path
 * (Float/main.cu, Half/main.cu).  TEST INFRASTRUCTURE ONLY: the parity oracle of the
 * HIP product, loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; never by the product.  Pinning status per function is listed
 * in pw_oracle.h.
 *
 * Expressions keep the reference's operand types so that float/double promotion
 * happens where the reference's does (e.g. the literal 0.5 in the CPML lines turns
 * those into double arithmetic).  The one thing not reproducible on the host is
 * nvcc's FMA contraction of device code, which moves results by last-place ulps.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "pw_oracle.h"

#define OR_PI 3.1415926535 /* Float/main.cu:11 */

/* ------------------------------------------------------------------------- */
/* Medium: create_model, Float/main.cu:1590-1655                              */
/* ------------------------------------------------------------------------- */
void or_create_model_borehole(int nr, int ny, int nc, float* out)
{
    const size_t n = (size_t)nr * ny * nc;
    float *rho = out + OR_RHO * n, *rhof = out + OR_RHOF * n, *C1 = out + OR_C1 * n,
          *C2 = out + OR_C2 * n, *HH = out + OR_HH * n, *H2u = out + OR_H2U * n,
          *C = out + OR_C * n, *M = out + OR_M * n, *mu = out + OR_MU * n;
    double Ks1, Kb1, Kf1, a1, tao, eta, porousm, perm, por;
    int ix, iy, iz;
    porousm = 8;
    eta = 1.0 * pow(10.0, -3);
    for (iz = 0; iz < nr; iz++)
        for (iy = 0; iy < ny; iy++)
            for (ix = 0; ix < nc; ix++) {
                const size_t o = (size_t)ix + (size_t)iy * nc + (size_t)iz * nc * ny;
                float vp, vs, rhos, vf;
                if (((ix - nc / 2) * (ix - nc / 2) + (iy - ny / 2) * (iy - ny / 2)) <= 100) {
                    /* fluid-filled borehole, Float/main.cu:1608-1626 */
                    vp = 1500.0f; vs = 0.0f; vf = 1500.0f; rhos = 1000.0f;
                    rhof[o] = 1000.0f; rho[o] = 1000.0f;
                    por = 1.0;
                    mu[o] = 0.0f;
                    Kb1 = 0.0;
                    Ks1 = rhof[o] * vp * vp;
                    Kf1 = rhof[o] * vf * vf;
                    C[o] = Kf1; M[o] = Kf1; HH[o] = Kf1; H2u[o] = Kf1;
                    C1[o] = 0.0f;
                    C2[o] = rhof[o];
                    (void)vs; (void)rhos; (void)por; (void)Kb1; (void)Ks1;
                } else {
                    /* formation, Float/main.cu:1628-1651 */
                    vp = 4000.0f; vs = 2700.0f; rhos = 2650.0f; rhof[o] = 1000.0f; vf = 1500.0f;
                    por = 0.1;
                    rho[o] = (1 - por) * rhos + por * rhof[o];
                    perm = 2 * pow(10.0, -12);
                    mu[o] = (1 - por) * rhos * vs * vs;
                    Kb1 = rhos * (1 - por) * (vp * vp - vs * vs * 4.0 / 3.0);
                    Ks1 = rhos * (vp * vp - vs * vs * 4.0 / 3.0);
                    Kf1 = rhof[o] * vf * vf;
                    tao = 3.0;
                    a1 = 1 - Kb1 / Ks1;
                    M[o] = Kf1 * Ks1 / (por * Ks1 + (a1 - por) * Kf1);
                    HH[o] = a1 * a1 * M[o] + Kb1 + mu[o] * 4.0 / 3.0;
                    H2u[o] = HH[o] - 2 * mu[o];
                    C[o] = M[o] * a1;
                    C1[o] = eta / perm;
                    C2[o] = (1 + 2 / porousm) * tao * rhof[o] / por;
                }
            }
}

/* The same Biot relations (Float/main.cu:1635-1650) for moduli-given input. */
void or_biot_from_moduli(double Ks, double Kb, double mu, double Kf, double rhos,
                         double rhof, double phi, double kappa, double eta,
                         double tortuosity, double pore_m, double* out)
{
    const double a1 = 1 - Kb / Ks;
    const double M = Kf * Ks / (phi * Ks + (a1 - phi) * Kf);
    const double HH = a1 * a1 * M + Kb + mu * 4.0 / 3.0;
    out[OR_RHO] = (1 - phi) * rhos + phi * rhof;
    out[OR_RHOF] = rhof;
    out[OR_C1] = eta / kappa;
    out[OR_C2] = (1 + 2 / pore_m) * tortuosity * rhof / phi;
    out[OR_HH] = HH;
    out[OR_H2U] = HH - 2 * mu;
    out[OR_C] = M * a1;
    out[OR_M] = M;
    out[OR_MU] = mu;
}

/* ------------------------------------------------------------------------- */
/* CPML profiles, Float/main.cu:570-654, per axis                             */
/* ------------------------------------------------------------------------- */
void or_pml_profile_f32(int n, int np, float H, float DT, float RC, float v0_in,
                        float* d, float* d2, float* e, float* e2)
{
    const float DP = np * H;
    const float oleft = DP;
    const float oright = (n - 1) * H - DP;
    int i;
    for (i = 0; i < n; i++) { d[i] = 0.0f; d2[i] = 0.0f; e[i] = 1.0f; e2[i] = 1.0f; }
    for (i = 0; i < n; i++) {
        const float x = i * H;
        if (x >= 0 && x < oleft) {
            const float v0 = v0_in;
            const float d0 = 3.0 * v0 * log(1.0 / RC) / (2.0 * DP);
            d[i] = d0 * pow(((oleft - x) / DP), 2);
            d2[i] = d0 * pow(((oleft - x - 0.5 * H) / DP), 2);
            e[i] = expf(-(d[i]) * DT);
            e2[i] = expf(-(d2[i]) * DT);
        }
        if (x >= oright && x < n * H) {
            const float v0 = v0_in;
            const float d0 = 3.0 * v0 * log(1.0 / RC) / (2.0 * DP);
            d[i] = d0 * pow(((x - oright) / DP), 2);
            d2[i] = d0 * pow(((x + 0.5 * H - oright) / DP), 2);
            e[i] = expf(-(d[i]) * DT);
            e2[i] = expf(-(d2[i]) * DT);
        }
    }
}

void or_pml_profile_f64(int n, int np, double H, double DT, double RC, double v0,
                        double* d, double* d2, double* e, double* e2)
{
    const double DP = np * H;
    const double oleft = DP;
    const double oright = (n - 1) * H - DP;
    const double d0 = 3.0 * v0 * log(1.0 / RC) / (2.0 * DP);
    int i;
    for (i = 0; i < n; i++) {
        /* compare on the integer index: the same cells as the float program */
        d[i] = 0.0; d2[i] = 0.0; e[i] = 1.0; e2[i] = 1.0;
        if (i < np) {
            d[i] = d0 * pow((oleft - i * H) / DP, 2);
            d2[i] = d0 * pow((oleft - i * H - 0.5 * H) / DP, 2);
        }
        if (i >= n - 1 - np) {
            d[i] = d0 * pow((i * H - oright) / DP, 2);
            d2[i] = d0 * pow((i * H + 0.5 * H - oright) / DP, 2);
        }
        e[i] = exp(-d[i] * DT);
        e2[i] = exp(-d2[i] * DT);
    }
}

/* Faces listed in or_args.pml_open: the interior profile on that half of the
 * axis, i.e. what the reference profile gives every cell outside its layers. */
static void open_faces_f32(int mask, int axis, int n, float* d, float* d2, float* e, float* e2)
{
    const int lo = (mask >> (2 * axis)) & 1, hi = (mask >> (2 * axis + 1)) & 1;
    int i;
    for (i = 0; i < n; i++)
        if (i < n / 2 ? lo : hi) { d[i] = 0.0f; d2[i] = 0.0f; e[i] = 1.0f; e2[i] = 1.0f; }
}
static void open_faces_f64(int mask, int axis, int n, double* d, double* d2, double* e, double* e2)
{
    const int lo = (mask >> (2 * axis)) & 1, hi = (mask >> (2 * axis + 1)) & 1;
    int i;
    for (i = 0; i < n; i++)
        if (i < n / 2 ? lo : hi) { d[i] = 0.0; d2[i] = 0.0; e[i] = 1.0; e2[i] = 1.0; }
}

/* ------------------------------------------------------------------------- */
/* Ricker, Float/main.cu:1050, 1059                                           */
/* ------------------------------------------------------------------------- */
float or_ricker_f32(int it, float DT, float F0)
{
    const float T0 = 1.2 / F0;
    const float tt = it * DT;
    float I_sou;
    I_sou = -(1 - 2 * OR_PI * OR_PI * F0 * F0 * (tt - T0) * (tt - T0)) *
            exp(-OR_PI * OR_PI * F0 * F0 * (tt - T0) * (tt - T0));
    return I_sou;
}

double or_ricker_f64(int it, double DT, double F0)
{
    const double T0 = 1.2 / F0;
    const double tt = it * DT;
    return -(1 - 2 * OR_PI * OR_PI * F0 * F0 * (tt - T0) * (tt - T0)) *
           exp(-OR_PI * OR_PI * F0 * F0 * (tt - T0) * (tt - T0));
}

/* ------------------------------------------------------------------------- */
/* Core step loops (float and double instantiations)                          */
/* ------------------------------------------------------------------------- */
#define FN(x) f32_##x
#define REAL float
#include "pw_oracle_core.inc"
#undef FN
#undef REAL
#define FN(x) f64_##x
#define REAL double
#include "pw_oracle_core.inc"
#undef FN
#undef REAL

static int check_args(const or_args* a)
{
    if (!a || a->nx < 4 || a->ny < 4 || a->nz < 4 || a->np < 1) return 1;
    if (a->nx - 2 * a->np < 1 || a->ny - 2 * a->np < 1 || a->nz - 2 * a->np < 1) return 1;
    if (a->nt < 0 || a->h <= 0 || a->dt <= 0) return 1;
    if (a->order != 0 && a->order != 2 && a->order != 4) return 1;
    if (a->pml_open < 0 || a->pml_open > 63) return 1;
    if (a->sx < 0 || a->sx >= a->nx || a->sy < 0 || a->sy >= a->ny || a->sz < 0 || a->sz >= a->nz) return 1;
    if (a->nrec < 0 || (a->nrec > 0 && (!a->rec_field || !a->rec_index || !a->seis))) return 1;
    if (a->nsnap < 0 || (a->nsnap > 0 && (!a->snap_step || !a->snap_field || !a->snaps))) return 1;
    if (a->nsnap > 0 && (a->snap_y < 0 || a->snap_y >= a->ny)) return 1;
    {
        int r;
        const long long ncell = (long long)a->nx * a->ny * a->nz;
        for (r = 0; r < a->nrec; r++)
            if (a->rec_field[r] < 0 || a->rec_field[r] >= OR_NFIELDS || a->rec_index[r] < 0 ||
                a->rec_index[r] >= ncell)
                return 1;
        for (r = 0; r < a->nsnap; r++)
            if (a->snap_field[r] < 0 || a->snap_field[r] >= OR_NFIELDS) return 1;
    }
    return 0;
}

static void set_threads(const or_args* a)
{
#ifdef _OPENMP
    if (a->nthreads > 0) omp_set_num_threads(a->nthreads);
#else
    (void)a;
#endif
}

/* extmodel (Float/main.cu:1662-1794): the 27 loop blocks replicate the nearest
 * interior cell into the pad, i.e. clamp every index into the interior. */
static inline size_t ext_src(int iz, int iy, int ix, int nzi, int nyi, int nxi, int np)
{
    int kz = iz - np, ky = iy - np, kx = ix - np;
    kz = kz < 0 ? 0 : (kz >= nzi ? nzi - 1 : kz);
    ky = ky < 0 ? 0 : (ky >= nyi ? nyi - 1 : ky);
    kx = kx < 0 ? 0 : (kx >= nxi ? nxi - 1 : kx);
    return (size_t)kx + (size_t)ky * nxi + (size_t)kz * nxi * nyi;
}

/* Float medium precompute: extmodel + staggered averages, Float/main.cu:387-513.
 * Writes into the 17 medium arrays of a float state; mu_ext optionally returned. */
static void precompute_f32_into(const or_args* a, const float* base, f32_state* s, float* mu_ext)
{
    const int nx = a->nx, ny = a->ny, nz = a->nz, np = a->np;
    const int nxi = nx - 2 * np, nyi = ny - 2 * np, nzi = nz - 2 * np;
    const size_t ni = (size_t)nxi * nyi * nzi;
    const long sxy = (long)nx * ny;
    const float *rho = base + OR_RHO * ni, *rhof = base + OR_RHOF * ni, *C1 = base + OR_C1 * ni,
                *C2 = base + OR_C2 * ni, *mu = base + OR_MU * ni;
    int iz;
#pragma omp parallel for schedule(static)
    for (iz = 0; iz < nz; iz++) {
        int iy, ix;
        for (iy = 0; iy < ny; iy++)
            for (ix = 0; ix < nx; ix++) {
                const long o = ix + (long)iy * nx + iz * sxy;
                const size_t q = ext_src(iz, iy, ix, nzi, nyi, nxi, np);
                s->rhof[o] = rhof[q];
                s->HH[o] = base[OR_HH * ni + q];
                s->H2u[o] = base[OR_H2U * ni + q];
                s->C[o] = base[OR_C * ni + q];
                s->M[o] = base[OR_M * ni + q];
                if (mu_ext) mu_ext[o] = mu[q];
            }
    }
#pragma omp parallel for schedule(static)
    for (iz = 1; iz < nz - 1; iz++) {
        int iy, ix;
        for (iy = 1; iy < ny - 1; iy++)
            for (ix = 1; ix < nx - 1; ix++) {
                const long o = ix + (long)iy * nx + iz * sxy;
#define MU(dz, dy, dx) mu[ext_src(iz + (dz), iy + (dy), ix + (dx), nzi, nyi, nxi, np)]
#define RH(dz, dy, dx) rho[ext_src(iz + (dz), iy + (dy), ix + (dx), nzi, nyi, nxi, np)]
#define CA(arr, dz, dy, dx) arr[ext_src(iz + (dz), iy + (dy), ix + (dx), nzi, nyi, nxi, np)]
                /* Float/main.cu:480-491 */
                if (MU(0, 0, 0) == 0 || MU(0, 0, 1) == 0 || MU(1, 0, 0) == 0 || MU(1, 0, 1) == 0)
                    s->muxz[o] = 0.0;
                else
                    s->muxz[o] = 1 / (0.25 * (1 / MU(0, 0, 0) + 1 / MU(0, 0, 1) + 1 / MU(1, 0, 0) + 1 / MU(1, 0, 1)));
                if (MU(0, 0, 0) == 0 || MU(0, 1, 0) == 0 || MU(1, 0, 0) == 0 || MU(1, 1, 0) == 0)
                    s->muyz[o] = 0.0;
                else
                    s->muyz[o] = 1 / (0.25 * (1 / MU(0, 0, 0) + 1 / MU(0, 1, 0) + 1 / MU(1, 0, 0) + 1 / MU(1, 1, 0)));
                if (MU(0, 0, 0) == 0 || MU(0, 0, 1) == 0 || MU(0, 1, 0) == 0 || MU(0, 1, 1) == 0)
                    s->muxy[o] = 0.0;
                else
                    s->muxy[o] = 1 / (0.25 * (1 / MU(0, 0, 0) + 1 / MU(0, 0, 1) + 1 / MU(0, 1, 0) + 1 / MU(0, 1, 1)));
                /* Float/main.cu:499-509 */
                s->rhox[o] = 2 / (RH(0, 0, 0) + RH(0, 0, 1));
                s->rhoy[o] = 2 / (RH(0, 0, 0) + RH(0, 1, 0));
                s->rhoz[o] = 2 / (RH(0, 0, 0) + RH(1, 0, 0));
                s->C1x[o] = 0.5 * (CA(C1, 0, 0, 0) + CA(C1, 0, 0, 1));
                s->C1y[o] = 0.5 * (CA(C1, 0, 0, 0) + CA(C1, 0, 1, 0));
                s->C1z[o] = 0.5 * (CA(C1, 0, 0, 0) + CA(C1, 1, 0, 0));
                s->C2x[o] = 0.5 * (CA(C2, 0, 0, 0) + CA(C2, 0, 0, 1));
                s->C2y[o] = 0.5 * (CA(C2, 0, 0, 0) + CA(C2, 0, 1, 0));
                s->C2z[o] = 0.5 * (CA(C2, 0, 0, 0) + CA(C2, 1, 0, 0));
#undef MU
#undef RH
#undef CA
            }
    }
    or_pml_profile_f32(nx, np, (float)a->h, (float)a->dt, (float)a->rc, (float)a->v0, s->dxi, s->dxi2, s->exi, s->exi2);
    or_pml_profile_f32(ny, np, (float)a->h, (float)a->dt, (float)a->rc, (float)a->v0, s->dyj, s->dyj2, s->eyj, s->eyj2);
    or_pml_profile_f32(nz, np, (float)a->h, (float)a->dt, (float)a->rc, (float)a->v0, s->dzk, s->dzk2, s->ezk, s->ezk2);
    open_faces_f32(a->pml_open, 0, nx, s->dxi, s->dxi2, s->exi, s->exi2);
    open_faces_f32(a->pml_open, 1, ny, s->dyj, s->dyj2, s->eyj, s->eyj2);
    open_faces_f32(a->pml_open, 2, nz, s->dzk, s->dzk2, s->ezk, s->ezk2);
}

int or_precompute_f32(const or_args* a, const float* base, float* out17)
{
    f32_state s;
    size_t n, k;
    float** med;
    if (check_args(a) || !base || !out17) return 1;
    set_threads(a);
    if (f32_state_alloc(&s, a->nx, a->ny, a->nz)) { f32_state_free(&s); return 2; }
    precompute_f32_into(a, base, &s, 0);
    n = (size_t)a->nx * a->ny * a->nz;
    med = &s.rhox;
    for (k = 0; k < OR_NMED_ARRAYS; k++) memcpy(out17 + k * n, med[k], n * sizeof(float));
    f32_state_free(&s);
    return 0;
}

int or_run_f32(or_args* a, const float* base)
{
    f32_state s;
    float* stf;
    int it;
    if (check_args(a) || !base) return 1;
    set_threads(a);
    if (f32_state_alloc(&s, a->nx, a->ny, a->nz)) { f32_state_free(&s); return 2; }
    precompute_f32_into(a, base, &s, 0);
    stf = (float*)malloc(sizeof(float) * (a->nt > 0 ? a->nt : 1));
    for (it = 0; it < a->nt; it++) stf[it] = or_ricker_f32(it, (float)a->dt, (float)a->f0);
    f32_time_loop(&s, a, (float)a->h, (float)a->dt, stf);
    free(stf);
    f32_state_free(&s);
    return 0;
}

int or_time_f32(or_args* a, const float* base, int warm, int steps, double* seconds, int* threads)
{
    f32_state s;
    const long sxy = (long)a->nx * a->ny;
    const long sn = a->sx + (long)a->sy * a->nx + (long)a->sz * sxy;
    double t0 = 0.0, t1 = 0.0;
    int it;
    if (check_args(a) || !base || warm < 0 || steps < 1 || !seconds) return 1;
    set_threads(a);
    if (f32_state_alloc(&s, a->nx, a->ny, a->nz)) { f32_state_free(&s); return 2; }
    precompute_f32_into(a, base, &s, 0);
    for (it = 0; it < warm + steps; it++) {
        const float I = or_ricker_f32(it, (float)a->dt, (float)a->f0);
        if (it == warm) {
#ifdef _OPENMP
            t0 = omp_get_wtime();
#endif
        }
        s.txx[sn] += (float)a->src_w[0] * I;
        s.tyy[sn] += (float)a->src_w[1] * I;
        s.tzz[sn] += (float)a->src_w[2] * I;
        s.ss[sn] += (float)a->src_w[3] * I;
        f32_fd_v(&s, a->nx, a->ny, a->nz, (float)a->h, (float)a->dt, a->pml_everywhere, a->order);
        f32_fd_t(&s, a->nx, a->ny, a->nz, (float)a->h, (float)a->dt, a->pml_everywhere, a->order);
    }
#ifdef _OPENMP
    t1 = omp_get_wtime();
    if (threads) *threads = omp_get_max_threads();
#else
    if (threads) *threads = 1;
#endif
    *seconds = t1 - t0;
    f32_state_free(&s);
    return 0;
}

/* FP64 oracle: the same algorithm with every quantity in double. */
int or_run_f64(or_args* a, const double* base)
{
    const int nx = a->nx, ny = a->ny, nz = a->nz, np = a->np;
    const int nxi = nx - 2 * np, nyi = ny - 2 * np, nzi = nz - 2 * np;
    const size_t ni = (size_t)nxi * nyi * nzi;
    const long sxy = (long)nx * ny;
    f64_state s;
    double* stf;
    int it, iz;
    if (check_args(a) || !base) return 1;
    set_threads(a);
    if (f64_state_alloc(&s, nx, ny, nz)) { f64_state_free(&s); return 2; }
    {
        const double *rho = base + OR_RHO * ni, *C1 = base + OR_C1 * ni, *C2 = base + OR_C2 * ni,
                     *mu = base + OR_MU * ni;
#pragma omp parallel for schedule(static)
        for (iz = 0; iz < nz; iz++) {
            int iy, ix;
            for (iy = 0; iy < ny; iy++)
                for (ix = 0; ix < nx; ix++) {
                    const long o = ix + (long)iy * nx + iz * sxy;
                    const size_t q = ext_src(iz, iy, ix, nzi, nyi, nxi, np);
                    s.rhof[o] = base[OR_RHOF * ni + q];
                    s.HH[o] = base[OR_HH * ni + q];
                    s.H2u[o] = base[OR_H2U * ni + q];
                    s.C[o] = base[OR_C * ni + q];
                    s.M[o] = base[OR_M * ni + q];
                    if (iz < 1 || iy < 1 || ix < 1 || iz > nz - 2 || iy > ny - 2 || ix > nx - 2) continue;
#define MU(dz, dy, dx) mu[ext_src(iz + (dz), iy + (dy), ix + (dx), nzi, nyi, nxi, np)]
#define RH(dz, dy, dx) rho[ext_src(iz + (dz), iy + (dy), ix + (dx), nzi, nyi, nxi, np)]
#define CA(arr, dz, dy, dx) arr[ext_src(iz + (dz), iy + (dy), ix + (dx), nzi, nyi, nxi, np)]
                    if (MU(0, 0, 0) == 0 || MU(0, 0, 1) == 0 || MU(1, 0, 0) == 0 || MU(1, 0, 1) == 0) s.muxz[o] = 0;
                    else s.muxz[o] = 1 / (0.25 * (1 / MU(0, 0, 0) + 1 / MU(0, 0, 1) + 1 / MU(1, 0, 0) + 1 / MU(1, 0, 1)));
                    if (MU(0, 0, 0) == 0 || MU(0, 1, 0) == 0 || MU(1, 0, 0) == 0 || MU(1, 1, 0) == 0) s.muyz[o] = 0;
                    else s.muyz[o] = 1 / (0.25 * (1 / MU(0, 0, 0) + 1 / MU(0, 1, 0) + 1 / MU(1, 0, 0) + 1 / MU(1, 1, 0)));
                    if (MU(0, 0, 0) == 0 || MU(0, 0, 1) == 0 || MU(0, 1, 0) == 0 || MU(0, 1, 1) == 0) s.muxy[o] = 0;
                    else s.muxy[o] = 1 / (0.25 * (1 / MU(0, 0, 0) + 1 / MU(0, 0, 1) + 1 / MU(0, 1, 0) + 1 / MU(0, 1, 1)));
                    s.rhox[o] = 2 / (RH(0, 0, 0) + RH(0, 0, 1));
                    s.rhoy[o] = 2 / (RH(0, 0, 0) + RH(0, 1, 0));
                    s.rhoz[o] = 2 / (RH(0, 0, 0) + RH(1, 0, 0));
                    s.C1x[o] = 0.5 * (CA(C1, 0, 0, 0) + CA(C1, 0, 0, 1));
                    s.C1y[o] = 0.5 * (CA(C1, 0, 0, 0) + CA(C1, 0, 1, 0));
                    s.C1z[o] = 0.5 * (CA(C1, 0, 0, 0) + CA(C1, 1, 0, 0));
                    s.C2x[o] = 0.5 * (CA(C2, 0, 0, 0) + CA(C2, 0, 0, 1));
                    s.C2y[o] = 0.5 * (CA(C2, 0, 0, 0) + CA(C2, 0, 1, 0));
                    s.C2z[o] = 0.5 * (CA(C2, 0, 0, 0) + CA(C2, 1, 0, 0));
#undef MU
#undef RH
#undef CA
                }
        }
    }
    or_pml_profile_f64(nx, np, a->h, a->dt, a->rc, a->v0, s.dxi, s.dxi2, s.exi, s.exi2);
    or_pml_profile_f64(ny, np, a->h, a->dt, a->rc, a->v0, s.dyj, s.dyj2, s.eyj, s.eyj2);
    or_pml_profile_f64(nz, np, a->h, a->dt, a->rc, a->v0, s.dzk, s.dzk2, s.ezk, s.ezk2);
    open_faces_f64(a->pml_open, 0, nx, s.dxi, s.dxi2, s.exi, s.exi2);
    open_faces_f64(a->pml_open, 1, ny, s.dyj, s.dyj2, s.eyj, s.eyj2);
    open_faces_f64(a->pml_open, 2, nz, s.dzk, s.dzk2, s.ezk, s.ezk2);
    stf = (double*)malloc(sizeof(double) * (a->nt > 0 ? a->nt : 1));
    for (it = 0; it < a->nt; it++) stf[it] = or_ricker_f64(it, a->dt, a->f0);
    f64_time_loop(&s, a, a->h, a->dt, stf);
    free(stf);
    f64_state_free(&s);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Half path: compensated FP16 codec, Half/main.cu:49-100                     */
/* IEEE-754 binary16 conversions written from the standard (round to nearest, */
/* ties to even), as __float2half / __half2float define them.                 */
/* ------------------------------------------------------------------------- */
unsigned short or_f32_to_f16(float f)
{
    uint32_t x, sign, absx, e, m, q, rem;
    memcpy(&x, &f, 4);
    sign = (x >> 16) & 0x8000u;
    absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) /* inf / nan */
        return (unsigned short)(sign | 0x7c00u | (absx > 0x7f800000u ? (0x200u | ((absx >> 13) & 0x3ffu)) : 0u));
    if (absx >= 0x477ff000u) return (unsigned short)(sign | 0x7c00u); /* >= 65520 rounds to inf */
    if (absx < 0x38800000u) {                                      /* below 2^-14: subnormal */
        int shift;
        uint32_t half;
        if (absx <= 0x33000000u) return (unsigned short)sign; /* <= 2^-25: ties to even -> 0 */
        e = absx >> 23;
        m = (absx & 0x7fffffu) | 0x800000u;
        shift = 126 - (int)e; /* 14..24 */
        q = m >> shift;
        rem = m & ((1u << shift) - 1u);
        half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (unsigned short)(sign | q);
    }
    e = (absx >> 23) - 127 + 15;
    m = absx & 0x7fffffu;
    q = (e << 10) | (m >> 13);
    rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return (unsigned short)(sign | q);
}

float or_f16_to_f32(unsigned short h)
{
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    float f;
    if (e == 0) {
        f = ldexpf((float)m, -24);
        return sign ? -f : f;
    }
    if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else x = sign | ((e - 15 + 127) << 23) | (m << 13);
    memcpy(&f, &x, 4);
    return f;
}

static float g_h2f[65536];
static int g_h2f_ready = 0;
static void h2f_init(void)
{
    int i;
    if (g_h2f_ready) return;
    for (i = 0; i < 65536; i++) g_h2f[i] = or_f16_to_f32((unsigned short)i);
    g_h2f_ready = 1;
}
#define H2F(h) (g_h2f[(h)])

/* residual_unit_from_half_float, Half/main.cu:59-67 */
float or_residual_unit(float hi_float)
{
    const float a = fabsf(hi_float);
    int ex = 0;
    if (a < 6.103515625e-5f) return 2.3283064365386963e-10f;
    frexpf(a, &ex);
    return ldexpf(1.0f, ex - 19);
}

/* quantize_residual + store_ch, Half/main.cu:85-100 */
void or_store_ch(float value, unsigned short* hi, signed char* lo)
{
    const unsigned short h = or_f32_to_f16(value);
    const float hf = or_f16_to_f32(h);
    const float unit = or_residual_unit(hf);
    const float scaled = (value - hf) / unit;
    int q = (int)(scaled + (scaled >= 0.0f ? 0.5f : -0.5f));
    if (q > 127) q = 127;
    if (q < -127) q = -127;
    *hi = h;
    *lo = (signed char)q;
}

/* load_cr_value, Half/main.cu:69-73 */
float or_load_ch(unsigned short hi, signed char lo)
{
    const float hf = or_f16_to_f32(hi);
    return hf + (float)lo * or_residual_unit(hf);
}

/* or_load_ch / or_store_ch over arrays (codec tests) */
void or_load_ch_n(const unsigned short* hi, const signed char* lo, long n, float* out)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) out[i] = or_load_ch(hi[i], lo[i]);
}
void or_store_ch_n(const float* v, long n, unsigned short* hi, signed char* lo)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) or_store_ch(v[i], hi + i, lo + i);
}

static inline float ldch(const unsigned short* hi, const signed char* lo, long o)
{
    const float hf = H2F(hi[o]);
    return hf + (float)lo[o] * or_residual_unit(hf);
}
static inline void stch(unsigned short* hi, signed char* lo, long o, float v) { or_store_ch(v, hi + o, lo + o); }
#define STH(arr, o, v) ((arr)[o] = or_f32_to_f16(v))

/* Half state: 16 compensated arrays, 48 FP16 memory arrays, 22 FP16 coefficient
 * arrays, 12 FP16 PML tables (1-D in content, Half/main.cu:952-964). */
enum { HS_VUX = 0, HS_VUY, HS_VUZ, HS_VWX, HS_VWY, HS_VWZ, HS_VWX2, HS_VWY2, HS_VWZ2,
       HS_TXX, HS_TYY, HS_TZZ, HS_TXZ, HS_TXY, HS_TYZ, HS_SS, HS_N };
enum { HC_RIX = 0, HC_RIY, HC_RIZ, HC_FRX, HC_FRY, HC_FRZ, HC_AWX, HC_AWY, HC_AWZ, HC_BPX, HC_BPY, HC_BPZ,
       HC_BTX, HC_BTY, HC_BTZ, HC_HH, HC_H2U, HC_C, HC_M, HC_MUXY, HC_MUXZ, HC_MUYZ, HC_N };
/* memory arrays: V x-group (pmlxSxx, pmlySxy, pmlzSxz, pmlxss, SXxx, SXxy, SXxz, SXss),
 * y-group, z-group, then T: (pmlxVux, pmlyVuy, pmlzVuz, Vuxx, Vuyy, Vuzz),
 * (pmlxVwx, pmlyVwy, pmlzVwz, Vwxx, Vwyy, Vwzz), (pmlxVuy, pmlyVux, Vuxy, Vuyx),
 * (pmlxVuz, pmlzVux, Vuxz, Vuzx), (pmlyVuz, pmlzVuy, Vuyz, Vuzy) */
enum { HM_N = 48 };

typedef struct {
    unsigned short* hi[HS_N];
    signed char* lo[HS_N];
    unsigned short* mem[HM_N];
    unsigned short* coef[HC_N];
    unsigned short *hdx, *hdx2, *hex, *hex2, *hdy, *hdy2, *hey, *hey2, *hdz, *hdz2, *hez, *hez2;
    float inv_scale_pml;
} half_state;

static void half_free(half_state* h)
{
    int k;
    for (k = 0; k < HS_N; k++) { free(h->hi[k]); free(h->lo[k]); }
    for (k = 0; k < HM_N; k++) free(h->mem[k]);
    for (k = 0; k < HC_N; k++) free(h->coef[k]);
    free(h->hdx); free(h->hdy); free(h->hdz);
}

/* Exponent choice and FP16 coefficient quantisation, Half/main.cu:745-1045.
 * fs: a float state holding the Float precompute; mu_ext: padded mu. */
static void half_scales(const or_args* a, f32_state* fs, const float* mu_ext, int* em_out, int* epml_out,
                        float* scale_m, float* inv_scale_m, float* scale_pml, float* inv_scale_pml)
{
    const int nx = a->nx, ny = a->ny, nz = a->nz;
    const float DT = (float)a->dt;
    const size_t n = (size_t)nx * ny * nz;
    float Kmax = 0.0f, Rmax = 0.0f, pml_max = 0.0f;
    const float Q_EPS = 1.0e-12f;
    size_t o;
    int em_int = -23, epml_int = 0, k;
    for (o = 0; o < n; o++) {
        float rhof0, rho_inv_x, rho_inv_y, rho_inv_z, Qx, Qy, Qz;
        Kmax = fmaxf(Kmax, fabsf(fs->HH[o]));
        Kmax = fmaxf(Kmax, fabsf(fs->H2u[o]));
        Kmax = fmaxf(Kmax, fabsf(fs->C[o]));
        Kmax = fmaxf(Kmax, fabsf(fs->M[o]));
        Kmax = fmaxf(Kmax, fabsf(mu_ext[o]));
        Kmax = fmaxf(Kmax, fabsf(fs->muxy[o]));
        Kmax = fmaxf(Kmax, fabsf(fs->muxz[o]));
        Kmax = fmaxf(Kmax, fabsf(fs->muyz[o]));
        rhof0 = fs->rhof[o];
        rho_inv_x = fs->rhox[o]; rho_inv_y = fs->rhoy[o]; rho_inv_z = fs->rhoz[o];
        if (rho_inv_x > 0.0f) Rmax = fmaxf(Rmax, fabsf(rho_inv_x));
        if (rho_inv_y > 0.0f) Rmax = fmaxf(Rmax, fabsf(rho_inv_y));
        if (rho_inv_z > 0.0f) Rmax = fmaxf(Rmax, fabsf(rho_inv_z));
        Qx = fs->C2x[o] - rhof0 * rhof0 * rho_inv_x;
        Qy = fs->C2y[o] - rhof0 * rhof0 * rho_inv_y;
        Qz = fs->C2z[o] - rhof0 * rhof0 * rho_inv_z;
        if (fs->C1x[o] != 0.0f && fabsf(Qx) > Q_EPS) {
            Rmax = fmaxf(Rmax, 1.0f / fabsf(Qx));
            Rmax = fmaxf(Rmax, fabsf(rhof0 * rho_inv_x / Qx));
        }
        if (fs->C1y[o] != 0.0f && fabsf(Qy) > Q_EPS) {
            Rmax = fmaxf(Rmax, 1.0f / fabsf(Qy));
            Rmax = fmaxf(Rmax, fabsf(rhof0 * rho_inv_y / Qy));
        }
        if (fs->C1z[o] != 0.0f && fabsf(Qz) > Q_EPS) {
            Rmax = fmaxf(Rmax, 1.0f / fabsf(Qz));
            Rmax = fmaxf(Rmax, fabsf(rhof0 * rho_inv_z / Qz));
        }
    }
    if (Kmax > 0.0f && Rmax > 0.0f) {
        const float er = 1.0f / sqrtf(DT * DT * Kmax * Rmax);
        em_int = (int)lrintf(-log2f(er * DT * Kmax));
    }
    /* pml_max over the 12 profiles: 1-D in content, so scan the tables */
    {
        const float* dtab[6] = {fs->dxi, fs->dxi2, fs->dyj, fs->dyj2, fs->dzk, fs->dzk2};
        const float* etab[6] = {fs->exi, fs->exi2, fs->eyj, fs->eyj2, fs->ezk, fs->ezk2};
        const int len[6] = {nx, nx, ny, ny, nz, nz};
        for (k = 0; k < 6; k++) {
            int i;
            for (i = 0; i < len[k]; i++) {
                pml_max = fmaxf(pml_max, fabsf(DT * dtab[k][i]));
                pml_max = fmaxf(pml_max, fabsf(1.0f - etab[k][i]));
            }
        }
    }
    if (pml_max > 0.0f) {
        epml_int = (int)floorf(log2f(2048.0f / pml_max));
        if (epml_int > 20) epml_int = 20;
        if (epml_int < -20) epml_int = -20;
    }
    *em_out = em_int;
    *epml_out = epml_int;
    *scale_m = ldexpf(1.0f, em_int);
    *inv_scale_m = ldexpf(1.0f, -em_int);
    *scale_pml = ldexpf(1.0f, epml_int);
    *inv_scale_pml = ldexpf(1.0f, -epml_int);
}

static void half_coefficients(const or_args* a, f32_state* fs, float scale_m, float inv_scale_m,
                              float scale_pml, half_state* h)
{
    const int nx = a->nx, ny = a->ny, nz = a->nz;
    const float DT = (float)a->dt;
    const size_t n = (size_t)nx * ny * nz;
    const float Q_EPS = 1.0e-12f;
    long o;
#pragma omp parallel for schedule(static)
    for (o = 0; o < (long)n; o++) {
        const float rhof0 = fs->rhof[o];
        const float rix = fs->rhox[o], riy = fs->rhoy[o], riz = fs->rhoz[o];
        const float Qx = fs->C2x[o] - rhof0 * rhof0 * rix;
        const float Qy = fs->C2y[o] - rhof0 * rhof0 * riy;
        const float Qz = fs->C2z[o] - rhof0 * rhof0 * riz;
        const float denom_x = Qx + 0.5f * fs->C1x[o] * DT;
        const float denom_y = Qy + 0.5f * fs->C1y[o] * DT;
        const float denom_z = Qz + 0.5f * fs->C1z[o] * DT;
        float awx = 0, awy = 0, awz = 0, bpx = 0, bpy = 0, bpz = 0, btx = 0, bty = 0, btz = 0;
        h->coef[HC_RIX][o] = or_f32_to_f16(inv_scale_m * DT * rix);
        h->coef[HC_RIY][o] = or_f32_to_f16(inv_scale_m * DT * riy);
        h->coef[HC_RIZ][o] = or_f32_to_f16(inv_scale_m * DT * riz);
        h->coef[HC_FRX][o] = or_f32_to_f16(rhof0 * rix);
        h->coef[HC_FRY][o] = or_f32_to_f16(rhof0 * riy);
        h->coef[HC_FRZ][o] = or_f32_to_f16(rhof0 * riz);
        if (fs->C1x[o] != 0.0f && fabsf(Qx) > Q_EPS && fabsf(denom_x) > Q_EPS) {
            awx = (Qx - 0.5f * fs->C1x[o] * DT) / denom_x;
            bpx = inv_scale_m * DT / denom_x;
            btx = inv_scale_m * DT * rhof0 * rix / denom_x;
        }
        if (fs->C1y[o] != 0.0f && fabsf(Qy) > Q_EPS && fabsf(denom_y) > Q_EPS) {
            awy = (Qy - 0.5f * fs->C1y[o] * DT) / denom_y;
            bpy = inv_scale_m * DT / denom_y;
            bty = inv_scale_m * DT * rhof0 * riy / denom_y;
        }
        if (fs->C1z[o] != 0.0f && fabsf(Qz) > Q_EPS && fabsf(denom_z) > Q_EPS) {
            awz = (Qz - 0.5f * fs->C1z[o] * DT) / denom_z;
            bpz = inv_scale_m * DT / denom_z;
            btz = inv_scale_m * DT * rhof0 * riz / denom_z;
        }
        h->coef[HC_AWX][o] = or_f32_to_f16(awx);
        h->coef[HC_AWY][o] = or_f32_to_f16(awy);
        h->coef[HC_AWZ][o] = or_f32_to_f16(awz);
        h->coef[HC_BPX][o] = or_f32_to_f16(bpx);
        h->coef[HC_BPY][o] = or_f32_to_f16(bpy);
        h->coef[HC_BPZ][o] = or_f32_to_f16(bpz);
        h->coef[HC_BTX][o] = or_f32_to_f16(btx);
        h->coef[HC_BTY][o] = or_f32_to_f16(bty);
        h->coef[HC_BTZ][o] = or_f32_to_f16(btz);
        h->coef[HC_HH][o] = or_f32_to_f16(scale_m * DT * fs->HH[o]);
        h->coef[HC_H2U][o] = or_f32_to_f16(scale_m * DT * fs->H2u[o]);
        h->coef[HC_C][o] = or_f32_to_f16(scale_m * DT * fs->C[o]);
        h->coef[HC_M][o] = or_f32_to_f16(scale_m * DT * fs->M[o]);
        h->coef[HC_MUXZ][o] = or_f32_to_f16(scale_m * DT * fs->muxz[o]);
        h->coef[HC_MUXY][o] = or_f32_to_f16(scale_m * DT * fs->muxy[o]);
        h->coef[HC_MUYZ][o] = or_f32_to_f16(scale_m * DT * fs->muyz[o]);
    }
    {
        int i;
        for (i = 0; i < nx; i++) {
            h->hdx[i] = or_f32_to_f16(scale_pml * DT * fs->dxi[i]);
            h->hdx2[i] = or_f32_to_f16(scale_pml * DT * fs->dxi2[i]);
            h->hex[i] = or_f32_to_f16(scale_pml * (1.0f - fs->exi[i]));
            h->hex2[i] = or_f32_to_f16(scale_pml * (1.0f - fs->exi2[i]));
        }
        for (i = 0; i < ny; i++) {
            h->hdy[i] = or_f32_to_f16(scale_pml * DT * fs->dyj[i]);
            h->hdy2[i] = or_f32_to_f16(scale_pml * DT * fs->dyj2[i]);
            h->hey[i] = or_f32_to_f16(scale_pml * (1.0f - fs->eyj[i]));
            h->hey2[i] = or_f32_to_f16(scale_pml * (1.0f - fs->eyj2[i]));
        }
        for (i = 0; i < nz; i++) {
            h->hdz[i] = or_f32_to_f16(scale_pml * DT * fs->dzk[i]);
            h->hdz2[i] = or_f32_to_f16(scale_pml * DT * fs->dzk2[i]);
            h->hez[i] = or_f32_to_f16(scale_pml * (1.0f - fs->ezk[i]));
            h->hez2[i] = or_f32_to_f16(scale_pml * (1.0f - fs->ezk2[i]));
        }
    }
}

/* Builds the Float precompute, then the Half scales/coefficients. */
static int half_setup(or_args* a, const float* base, half_state* h, int* em, int* epml)
{
    const int nx = a->nx, ny = a->ny, nz = a->nz;
    const size_t n = (size_t)nx * ny * nz;
    f32_state fs;
    float* mu_ext;
    float scale_m, inv_scale_m, scale_pml, inv_scale_pml;
    int k;
    memset(h, 0, sizeof(*h));
    h2f_init();
    if (f32_state_alloc(&fs, nx, ny, nz)) { f32_state_free(&fs); return 2; }
    mu_ext = (float*)calloc(n, sizeof(float));
    if (!mu_ext) { f32_state_free(&fs); return 2; }
    precompute_f32_into(a, base, &fs, mu_ext);
    half_scales(a, &fs, mu_ext, em, epml, &scale_m, &inv_scale_m, &scale_pml, &inv_scale_pml);
    for (k = 0; k < HC_N; k++) {
        h->coef[k] = (unsigned short*)calloc(n, sizeof(unsigned short));
        if (!h->coef[k]) { free(mu_ext); f32_state_free(&fs); return 2; }
    }
    h->hdx = (unsigned short*)calloc((size_t)nx * 4, 2);
    h->hdy = (unsigned short*)calloc((size_t)ny * 4, 2);
    h->hdz = (unsigned short*)calloc((size_t)nz * 4, 2);
    if (!h->hdx || !h->hdy || !h->hdz) { free(mu_ext); f32_state_free(&fs); return 2; }
    h->hdx2 = h->hdx + nx; h->hex = h->hdx + 2 * nx; h->hex2 = h->hdx + 3 * nx;
    h->hdy2 = h->hdy + ny; h->hey = h->hdy + 2 * ny; h->hey2 = h->hdy + 3 * ny;
    h->hdz2 = h->hdz + nz; h->hez = h->hdz + 2 * nz; h->hez2 = h->hdz + 3 * nz;
    half_coefficients(a, &fs, scale_m, inv_scale_m, scale_pml, h);
    h->inv_scale_pml = inv_scale_pml;
    free(mu_ext);
    f32_state_free(&fs);
    return 0;
}

int or_half_precompute(const or_args* a, const float* base, int* em, int* epml, unsigned short* coef22)
{
    half_state h;
    int rc, k;
    if (check_args(a) || !base || !em || !epml) return 1;
    set_threads(a);
    rc = half_setup((or_args*)a, base, &h, em, epml);
    if (rc == 0 && coef22) {
        const size_t n = (size_t)a->nx * a->ny * a->nz;
        for (k = 0; k < HC_N; k++) memcpy(coef22 + k * n, h.coef[k], n * 2);
    }
    half_free(&h);
    return rc;
}

/* One Half CPML recursion, Half/main.cu:201 pattern. */
#define HPML(psi, prev, hd, he, der)                                                        \
    do {                                                                                   \
        if (all || (hd) != 0) {                                                            \
            const float pe = 1.0f - isp * H2F(he);                                          \
            const float pdd = isp * H2F(hd);                                               \
            const float nw = H2F(psi[o]) * pe + (-pdd * 0.5f) * (pe * H2F(prev[o]) + der); \
            STH(psi, o, nw);                                                               \
            STH(prev, o, der);                                                             \
            der = der + nw;                                                                \
        }                                                                                  \
    } while (0)

/* Half FD_V, Half/main.cu:156-268 */
static void half_fd_v(half_state* h, int nx, int ny, int nz, float H, int all)
{
    const long sxy = (long)nx * ny;
    const float isp = h->inv_scale_pml;
    unsigned short** hi = h->hi;
    signed char** lo = h->lo;
    unsigned short** m = h->mem;
    unsigned short** c = h->coef;
    int iz;
#pragma omp parallel for schedule(static)
    for (iz = 1; iz < nz - 1; iz++) {
        int iy, ix;
        for (iy = 1; iy < ny - 1; iy++)
            for (ix = 1; ix < nx - 1; ix++) {
                const long o = ix + (long)iy * nx + iz * sxy;
                const long ob = o - sxy, orr = o + 1, oh = o - nx, oq = o + nx, ol = o - 1, ou = o + sxy;
                float x1, x2, x3, y1, y2, y3, z1, z2, z3, s1, s2, s3;
                float dtx, dty, dtz, vwxo, vwyo, vwzo, vwxp, vwyp, vwzp;
                float awx, bpx, btx, awy, bpy, bty, awz, bpz, btz, vwxn, vwyn, vwzn, vuxn, vuyn, vuzn;
                x1 = (ldch(hi[HS_TXX], lo[HS_TXX], orr) - ldch(hi[HS_TXX], lo[HS_TXX], o)) / H;
                x2 = (ldch(hi[HS_TXY], lo[HS_TXY], o) - ldch(hi[HS_TXY], lo[HS_TXY], oh)) / H;
                x3 = (ldch(hi[HS_TXZ], lo[HS_TXZ], o) - ldch(hi[HS_TXZ], lo[HS_TXZ], ob)) / H;
                s1 = (ldch(hi[HS_SS], lo[HS_SS], orr) - ldch(hi[HS_SS], lo[HS_SS], o)) / H;
                y1 = (ldch(hi[HS_TYY], lo[HS_TYY], oq) - ldch(hi[HS_TYY], lo[HS_TYY], o)) / H;
                y2 = (ldch(hi[HS_TXY], lo[HS_TXY], o) - ldch(hi[HS_TXY], lo[HS_TXY], ol)) / H;
                y3 = (ldch(hi[HS_TYZ], lo[HS_TYZ], o) - ldch(hi[HS_TYZ], lo[HS_TYZ], ob)) / H;
                s2 = (ldch(hi[HS_SS], lo[HS_SS], oq) - ldch(hi[HS_SS], lo[HS_SS], o)) / H;
                z1 = (ldch(hi[HS_TZZ], lo[HS_TZZ], ou) - ldch(hi[HS_TZZ], lo[HS_TZZ], o)) / H;
                z2 = (ldch(hi[HS_TXZ], lo[HS_TXZ], o) - ldch(hi[HS_TXZ], lo[HS_TXZ], ol)) / H;
                z3 = (ldch(hi[HS_TYZ], lo[HS_TYZ], o) - ldch(hi[HS_TYZ], lo[HS_TYZ], oh)) / H;
                s3 = (ldch(hi[HS_SS], lo[HS_SS], ou) - ldch(hi[HS_SS], lo[HS_SS], o)) / H;

                HPML(m[0], m[4], h->hdx2[ix], h->hex2[ix], x1);  /* pmlxSxx / SXxx */
                HPML(m[1], m[5], h->hdy[iy], h->hey[iy], x2);    /* pmlySxy / SXxy */
                HPML(m[2], m[6], h->hdz[iz], h->hez[iz], x3);    /* pmlzSxz / SXxz */
                HPML(m[3], m[7], h->hdx2[ix], h->hex2[ix], s1);  /* pmlxss / SXss */
                HPML(m[8], m[12], h->hdx[ix], h->hex[ix], y2);   /* pmlxSxy / SYxy */
                HPML(m[9], m[13], h->hdy2[iy], h->hey2[iy], y1); /* pmlySyy / SYyy */
                HPML(m[10], m[14], h->hdz[iz], h->hez[iz], y3);  /* pmlzSyz / SYyz */
                HPML(m[11], m[15], h->hdy2[iy], h->hey2[iy], s2);/* pmlyss / SYss */
                HPML(m[16], m[20], h->hdx[ix], h->hex[ix], z2);  /* pmlxSxz / SZxz */
                HPML(m[17], m[21], h->hdy[iy], h->hey[iy], z3);  /* pmlySyz / SZyz */
                HPML(m[18], m[22], h->hdz2[iz], h->hez2[iz], z1);/* pmlzSzz / SZzz */
                HPML(m[19], m[23], h->hdz2[iz], h->hez2[iz], s3);/* pmlzss / SZss */

                dtx = x1 + x2 + x3; dty = y1 + y2 + y3; dtz = z1 + z2 + z3;
                vwxo = ldch(hi[HS_VWX], lo[HS_VWX], o);
                vwyo = ldch(hi[HS_VWY], lo[HS_VWY], o);
                vwzo = ldch(hi[HS_VWZ], lo[HS_VWZ], o);
                vwxp = ldch(hi[HS_VWX2], lo[HS_VWX2], o);
                vwyp = ldch(hi[HS_VWY2], lo[HS_VWY2], o);
                vwzp = ldch(hi[HS_VWZ2], lo[HS_VWZ2], o);
                awx = H2F(c[HC_AWX][o]); bpx = H2F(c[HC_BPX][o]); btx = H2F(c[HC_BTX][o]);
                awy = H2F(c[HC_AWY][o]); bpy = H2F(c[HC_BPY][o]); bty = H2F(c[HC_BTY][o]);
                awz = H2F(c[HC_AWZ][o]); bpz = H2F(c[HC_BPZ][o]); btz = H2F(c[HC_BTZ][o]);
                vwxn = (awx == 0.0f && bpx == 0.0f && btx == 0.0f) ? 0.0f : awx * vwxo - btx * dtx - bpx * s1;
                vwyn = (awy == 0.0f && bpy == 0.0f && bty == 0.0f) ? 0.0f : awy * vwyo - bty * dty - bpy * s2;
                vwzn = (awz == 0.0f && bpz == 0.0f && btz == 0.0f) ? 0.0f : awz * vwzo - btz * dtz - bpz * s3;
                vuxn = ldch(hi[HS_VUX], lo[HS_VUX], o) + H2F(c[HC_RIX][o]) * dtx - H2F(c[HC_FRX][o]) * (vwxn - vwxp);
                vuyn = ldch(hi[HS_VUY], lo[HS_VUY], o) + H2F(c[HC_RIY][o]) * dty - H2F(c[HC_FRY][o]) * (vwyn - vwyp);
                vuzn = ldch(hi[HS_VUZ], lo[HS_VUZ], o) + H2F(c[HC_RIZ][o]) * dtz - H2F(c[HC_FRZ][o]) * (vwzn - vwzp);
                stch(hi[HS_VWX], lo[HS_VWX], o, vwxn); stch(hi[HS_VWY], lo[HS_VWY], o, vwyn); stch(hi[HS_VWZ], lo[HS_VWZ], o, vwzn);
                stch(hi[HS_VUX], lo[HS_VUX], o, vuxn); stch(hi[HS_VUY], lo[HS_VUY], o, vuyn); stch(hi[HS_VUZ], lo[HS_VUZ], o, vuzn);
                stch(hi[HS_VWX2], lo[HS_VWX2], o, vwxn); stch(hi[HS_VWY2], lo[HS_VWY2], o, vwyn); stch(hi[HS_VWZ2], lo[HS_VWZ2], o, vwzn);
            }
    }
}

/* Half FD_T, Half/main.cu:270-371 */
static void half_fd_t(half_state* h, int nx, int ny, int nz, float H, int all)
{
    const long sxy = (long)nx * ny;
    const float isp = h->inv_scale_pml;
    unsigned short** hi = h->hi;
    signed char** lo = h->lo;
    unsigned short** m = h->mem;
    unsigned short** c = h->coef;
    int iz;
#pragma omp parallel for schedule(static)
    for (iz = 1; iz < nz - 1; iz++) {
        int iy, ix;
        for (iy = 1; iy < ny - 1; iy++)
            for (ix = 1; ix < nx - 1; ix++) {
                const long o = ix + (long)iy * nx + iz * sxy;
                const long ob = o - sxy, orr = o + 1, oh = o - nx, oq = o + nx, ol = o - 1, ou = o + sxy;
                float uxx, uyy, uzz, uxy, uxz, uyx, uyz, uzx, uzy, wx, wy, wz, div_u, div_w;
                uxx = (ldch(hi[HS_VUX], lo[HS_VUX], o) - ldch(hi[HS_VUX], lo[HS_VUX], ol)) / H;
                uyy = (ldch(hi[HS_VUY], lo[HS_VUY], o) - ldch(hi[HS_VUY], lo[HS_VUY], oh)) / H;
                uzz = (ldch(hi[HS_VUZ], lo[HS_VUZ], o) - ldch(hi[HS_VUZ], lo[HS_VUZ], ob)) / H;
                wx = (ldch(hi[HS_VWX], lo[HS_VWX], o) - ldch(hi[HS_VWX], lo[HS_VWX], ol)) / H;
                wy = (ldch(hi[HS_VWY], lo[HS_VWY], o) - ldch(hi[HS_VWY], lo[HS_VWY], oh)) / H;
                wz = (ldch(hi[HS_VWZ], lo[HS_VWZ], o) - ldch(hi[HS_VWZ], lo[HS_VWZ], ob)) / H;
                uxy = (ldch(hi[HS_VUX], lo[HS_VUX], oq) - ldch(hi[HS_VUX], lo[HS_VUX], o)) / H;
                uyx = (ldch(hi[HS_VUY], lo[HS_VUY], orr) - ldch(hi[HS_VUY], lo[HS_VUY], o)) / H;
                uxz = (ldch(hi[HS_VUX], lo[HS_VUX], ou) - ldch(hi[HS_VUX], lo[HS_VUX], o)) / H;
                uzx = (ldch(hi[HS_VUZ], lo[HS_VUZ], orr) - ldch(hi[HS_VUZ], lo[HS_VUZ], o)) / H;
                uyz = (ldch(hi[HS_VUY], lo[HS_VUY], ou) - ldch(hi[HS_VUY], lo[HS_VUY], o)) / H;
                uzy = (ldch(hi[HS_VUZ], lo[HS_VUZ], oq) - ldch(hi[HS_VUZ], lo[HS_VUZ], o)) / H;

                HPML(m[24], m[27], h->hdx[ix], h->hex[ix], uxx);   /* pmlxVux / Vuxx */
                HPML(m[25], m[28], h->hdy[iy], h->hey[iy], uyy);   /* pmlyVuy / Vuyy */
                HPML(m[26], m[29], h->hdz[iz], h->hez[iz], uzz);   /* pmlzVuz / Vuzz */
                HPML(m[30], m[33], h->hdx[ix], h->hex[ix], wx);    /* pmlxVwx / Vwxx */
                HPML(m[31], m[34], h->hdy[iy], h->hey[iy], wy);    /* pmlyVwy / Vwyy */
                HPML(m[32], m[35], h->hdz[iz], h->hez[iz], wz);    /* pmlzVwz / Vwzz */
                HPML(m[36], m[38], h->hdy2[iy], h->hey2[iy], uxy); /* pmlxVuy / Vuxy */
                HPML(m[37], m[39], h->hdx2[ix], h->hex2[ix], uyx); /* pmlyVux / Vuyx */
                HPML(m[40], m[42], h->hdz2[iz], h->hez2[iz], uxz); /* pmlxVuz / Vuxz */
                HPML(m[41], m[43], h->hdx2[ix], h->hex2[ix], uzx); /* pmlzVux / Vuzx */
                HPML(m[44], m[46], h->hdz2[iz], h->hez2[iz], uyz); /* pmlyVuz / Vuyz */
                HPML(m[45], m[47], h->hdy2[iy], h->hey2[iy], uzy); /* pmlzVuy / Vuzy */

                div_u = uxx + uyy + uzz;
                div_w = wx + wy + wz;
                stch(hi[HS_SS], lo[HS_SS], o, ldch(hi[HS_SS], lo[HS_SS], o) - H2F(c[HC_C][o]) * div_u - H2F(c[HC_M][o]) * div_w);
                stch(hi[HS_TXX], lo[HS_TXX], o, ldch(hi[HS_TXX], lo[HS_TXX], o) + H2F(c[HC_H2U][o]) * (uyy + uzz) + H2F(c[HC_HH][o]) * uxx + H2F(c[HC_C][o]) * div_w);
                stch(hi[HS_TYY], lo[HS_TYY], o, ldch(hi[HS_TYY], lo[HS_TYY], o) + H2F(c[HC_H2U][o]) * (uxx + uzz) + H2F(c[HC_HH][o]) * uyy + H2F(c[HC_C][o]) * div_w);
                stch(hi[HS_TZZ], lo[HS_TZZ], o, ldch(hi[HS_TZZ], lo[HS_TZZ], o) + H2F(c[HC_H2U][o]) * (uxx + uyy) + H2F(c[HC_HH][o]) * uzz + H2F(c[HC_C][o]) * div_w);
                stch(hi[HS_TXY], lo[HS_TXY], o, ldch(hi[HS_TXY], lo[HS_TXY], o) + H2F(c[HC_MUXY][o]) * (uxy + uyx));
                stch(hi[HS_TYZ], lo[HS_TYZ], o, ldch(hi[HS_TYZ], lo[HS_TYZ], o) + H2F(c[HC_MUYZ][o]) * (uyz + uzy));
                stch(hi[HS_TXZ], lo[HS_TXZ], o, ldch(hi[HS_TXZ], lo[HS_TXZ], o) + H2F(c[HC_MUXZ][o]) * (uxz + uzx));
            }
    }
}
#undef HPML

/* dynamic_scale_for_step, Half/main.cu:116-122 */
static float half_dynamic_scale(int it, int nt, float end)
{
    float alpha;
    if (nt <= 1) return end;
    alpha = (float)it / (float)(nt - 1);
    return 1.0f + (end - 1.0f) * alpha;
}

static int half_field_slot(int f, int* is_velocity)
{
    static const int map[OR_NFIELDS] = {HS_VUX, HS_VUY, HS_VUZ, HS_VWX, HS_VWY, HS_VWZ,
                                        HS_TXX, HS_TYY, HS_TZZ, HS_TXY, HS_TXZ, HS_TYZ, HS_SS};
    *is_velocity = f <= OR_VWZ;
    return map[f];
}

int or_run_half(or_args* a, const float* base)
{
    const int nx = a->nx, ny = a->ny, nz = a->nz;
    const size_t n = (size_t)nx * ny * nz;
    const long sxy = (long)nx * ny;
    const long sn = a->sx + (long)a->sy * nx + (long)a->sz * sxy;
    const float H = (float)a->h, DT = (float)a->dt;
    const float dyn_end = a->half_dyn_end > 0 ? (float)a->half_dyn_end : 16.0f;
    const int interval = a->half_dyn_interval;
    half_state h;
    int em, epml, k, it, r;
    float source_scale, stress_base, velocity_base, dynamic_scale, stress_out, velocity_out;
    if (check_args(a) || !base || a->order == 4) return 1;  /* the Half program is 2nd order */
    set_threads(a);
    if (half_setup(a, base, &h, &em, &epml)) { half_free(&h); return 2; }
    a->half_em_out = em;
    a->half_epml_out = epml;
    for (k = 0; k < HS_N; k++) {
        h.hi[k] = (unsigned short*)calloc(n, 2);
        h.lo[k] = (signed char*)calloc(n, 1);
        if (!h.hi[k] || !h.lo[k]) { half_free(&h); return 2; }
    }
    for (k = 0; k < HM_N; k++) {
        h.mem[k] = (unsigned short*)calloc(n, 2);
        if (!h.mem[k]) { half_free(&h); return 2; }
    }
    /* Half/main.cu:800-807 */
    source_scale = ldexpf(1.0f, a->half_es);
    stress_base = ldexpf(1.0f, -a->half_es);
    velocity_base = ldexpf(1.0f, em - a->half_es);
    dynamic_scale = 1.0f;
    stress_out = stress_base / dynamic_scale;
    velocity_out = velocity_base / dynamic_scale;
    for (it = 0; it < a->nt; it++) {
        /* dynamic rescale, Half/main.cu:1326-1364 */
        if (it == 0 || (interval > 0 && it % interval == 0)) {
            const float target = half_dynamic_scale(it, a->nt, dyn_end);
            const float ratio = target / dynamic_scale;
            if (fabsf(ratio - 1.0f) > 1.0e-6f) {
                long o;
#pragma omp parallel for schedule(static)
                for (o = 0; o < (long)n; o++) {
                    int q;
                    for (q = 0; q < HS_N; q++) stch(h.hi[q], h.lo[q], o, ldch(h.hi[q], h.lo[q], o) * ratio);
                    for (q = 0; q < HM_N; q++) h.mem[q][o] = or_f32_to_f16(H2F(h.mem[q][o]) * ratio);
                }
                dynamic_scale = target;
                stress_out = stress_base / dynamic_scale;
                velocity_out = velocity_base / dynamic_scale;
            }
        }
        {
            /* Source, Half/main.cu:135-151, 1372-1374 */
            const float I = dynamic_scale * source_scale * or_ricker_f32(it, DT, (float)a->f0);
            const int slot[4] = {HS_TXX, HS_TYY, HS_TZZ, HS_SS};
            for (k = 0; k < 4; k++)
                if (a->src_w[k] != 0)
                    stch(h.hi[slot[k]], h.lo[slot[k]], sn, ldch(h.hi[slot[k]], h.lo[slot[k]], sn) + (float)a->src_w[k] * I);
        }
        half_fd_v(&h, nx, ny, nz, H, a->pml_everywhere);
        half_fd_t(&h, nx, ny, nz, H, a->pml_everywhere);
        for (r = 0; r < a->nrec; r++) {
            int isv;
            const int q = half_field_slot(a->rec_field[r], &isv);
            a->seis[(size_t)r * a->nt + it] = (isv ? velocity_out : stress_out) * ldch(h.hi[q], h.lo[q], (long)a->rec_index[r]);
        }
        for (k = 0; k < a->nsnap; k++) {
            if (a->snap_step[k] == it) {
                int isv, iz, ix;
                const int q = half_field_slot(a->snap_field[k], &isv);
                float* dst = a->snaps + (size_t)k * nz * nx;
                for (iz = 0; iz < nz; iz++)
                    for (ix = 0; ix < nx; ix++)
                        dst[(size_t)iz * nx + ix] = (isv ? velocity_out : stress_out) *
                                                    ldch(h.hi[q], h.lo[q], ix + (long)a->snap_y * nx + iz * sxy);
            }
        }
    }
    if (a->final_state) {
        int f;
        for (f = 0; f < OR_NFIELDS; f++) {
            int isv;
            const int q = half_field_slot(f, &isv);
            size_t i;
            for (i = 0; i < n; i++)
                a->final_state[(size_t)f * n + i] = (isv ? velocity_out : stress_out) * ldch(h.hi[q], h.lo[q], (long)i);
        }
    }
    half_free(&h);
    return 0;
}

/* The staggered difference of the update (order 2 or 4) on a 1-D array, FP64. */
double or_stagger_diff_f64(int order, int forward, const double* f, long o, long st)
{
    return f64_stagger_diff(order, forward, f, o, st);
}
