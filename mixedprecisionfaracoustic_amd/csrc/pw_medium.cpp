// pw_medium.cpp -- host-side front-ends: the reference's borehole model, Biot
// relations, 1-D CPML profiles, Ricker wavelet and the .dat writer.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "pw_internal.h"

namespace pw {

static const double kPi = 3.1415926535;  // the reference's PI (Float/main.cu:11)

// CPML profile of one axis with the reference's float/double mix
// (Float/main.cu:570-654): positions and damping held in float, d0 and the
// squared ratio formed in double, e = expf(-d*dt).  1-D: the reference's 3-D
// profile arrays depend only on their own axis index.
void cpml_profile_reference_f32(int n, int np, float H, float DT, float RC, float v0, float* d, float* d2, float* e,
                                float* e2)
{
    const float DP = np * H;
    const float left = DP;
    const float right = (n - 1) * H - DP;
    const float d0 = 3.0 * v0 * std::log(1.0 / RC) / (2.0 * DP);
    for (int i = 0; i < n; ++i) {
        d[i] = 0.0f; d2[i] = 0.0f; e[i] = 1.0f; e2[i] = 1.0f;
        const float x = i * H;
        bool hit = false;
        if (x >= 0 && x < left) {
            const double r = (left - x) / DP, r2 = (left - x - 0.5 * H) / DP;
            d[i] = d0 * (r * r);
            d2[i] = d0 * (r2 * r2);
            hit = true;
        }
        if (x >= right && x < n * H) {
            const double r = (x - right) / DP, r2 = (x + 0.5 * H - right) / DP;
            d[i] = d0 * (r * r);
            d2[i] = d0 * (r2 * r2);
            hit = true;
        }
        if (hit) {
            e[i] = std::exp(-(d[i]) * DT);
            e2[i] = std::exp(-(d2[i]) * DT);
        }
    }
}

// The same profile evaluated in double throughout (FP64 runs), on the same cells.
void cpml_profile_f64(int n, int np, double H, double DT, double RC, double v0, double* d, double* d2, double* e,
                      double* e2)
{
    const double DP = np * H, right = (n - 1) * H - DP;
    const double d0 = 3.0 * v0 * std::log(1.0 / RC) / (2.0 * DP);
    for (int i = 0; i < n; ++i) {
        d[i] = 0.0; d2[i] = 0.0;
        if (i < np) {
            const double r = (DP - i * H) / DP, r2 = (DP - i * H - 0.5 * H) / DP;
            d[i] = d0 * r * r;
            d2[i] = d0 * r2 * r2;
        }
        if (i >= n - 1 - np) {
            const double r = (i * H - right) / DP, r2 = (i * H + 0.5 * H - right) / DP;
            d[i] = d0 * r * r;
            d2[i] = d0 * r2 * r2;
        }
        e[i] = std::exp(-d[i] * DT);
        e2[i] = std::exp(-d2[i] * DT);
    }
}

// Ricker sample at step it in the reference's float/double mix (Float/main.cu:1050, 1059).
float ricker_f32(int it, float DT, float F0)
{
    const float T0 = 1.2 / F0;
    const float tt = it * DT;
    const float tau = tt - T0;
    return (float)(-(1 - 2 * kPi * kPi * F0 * F0 * tau * tau) * std::exp(-kPi * kPi * F0 * F0 * tau * tau));
}

double ricker_f64(int it, double DT, double F0)
{
    const double tau = it * DT - 1.2 / F0;
    return -(1 - 2 * kPi * kPi * F0 * F0 * tau * tau) * std::exp(-kPi * kPi * F0 * F0 * tau * tau);
}

}  // namespace pw

using namespace pw;

extern "C" pw_status pw_write_dat(const char* path, const float* data, int nr, int nc)
{
    if (!path || (!data && nr * nc > 0) || nr < 0 || nc < 0) return fail(PW_ERR_INVALID, "pw_write_dat: bad arguments");
    FILE* fp = std::fopen(path, "wt");
    if (!fp) return fail(PW_ERR_IO, std::string("pw_write_dat: cannot open ") + path);
    for (int i = 0; i < nr; ++i) {
        for (int j = 0; j < nc; ++j) std::fprintf(fp, "%e ", data[(size_t)i * nc + j]);
        // wfile: "\n" after the row's last value, then one more "\n"
        if (nc > 0) std::fputc('\n', fp);
        std::fputc('\n', fp);
    }
    const bool bad = std::ferror(fp) != 0;
    if (std::fclose(fp) != 0 || bad) return fail(PW_ERR_IO, std::string("pw_write_dat: write failed: ") + path);
    return PW_OK;
}

// create_model (Float/main.cu:1590-1655).  The reference stores every field in
// float and forms Kb, Ks, M, H, C in double from float inputs; the results are
// reproduced in float and widened to double.
extern "C" pw_status pw_medium_borehole(int nxi, int nyi, int nzi, double* out)
{
    if (nxi < 1 || nyi < 1 || nzi < 1 || !out) return fail(PW_ERR_INVALID, "pw_medium_borehole: bad arguments");
    const size_t n = (size_t)nxi * nyi * nzi;
    // Formation cell (Float/main.cu:1629-1650).
    const float vp = 4000.0f, vs = 2700.0f, rhos = 2650.0f, rhof = 1000.0f, vf = 1500.0f;
    const double por = 0.1, perm = 2 * std::pow(10.0, -12), eta = 1.0 * std::pow(10.0, -3), tao = 3.0, pm = 8;
    const float rho_f = (1 - por) * rhos + por * rhof;
    const float mu_f = (1 - por) * rhos * vs * vs;
    const double Kb = rhos * (1 - por) * (vp * vp - vs * vs * 4.0 / 3.0);
    const double Ks = rhos * (vp * vp - vs * vs * 4.0 / 3.0);
    const double Kf = rhof * vf * vf;
    const double a1 = 1 - Kb / Ks;
    const float M_f = Kf * Ks / (por * Ks + (a1 - por) * Kf);
    const float HH_f = a1 * a1 * M_f + Kb + mu_f * 4.0 / 3.0;
    const float H2u_f = HH_f - 2 * mu_f;
    const float C_f = M_f * a1;
    const float C1_f = eta / perm;
    const float C2_f = (1 + 2 / pm) * tao * rhof / por;
    const double form[PW_NBASE] = {rho_f, rhof, C1_f, C2_f, HH_f, H2u_f, C_f, M_f, mu_f};
    // Borehole fluid cell (Float/main.cu:1609-1625): C = M = H = H-2mu = Kf, C2 = rhof.
    const float kf_b = 1000.0f * 1500.0f * 1500.0f;
    const double fluid[PW_NBASE] = {1000.0, 1000.0, 0.0, 1000.0, kf_b, kf_b, kf_b, kf_b, 0.0};
    for (int iz = 0; iz < nzi; ++iz)
        for (int iy = 0; iy < nyi; ++iy)
            for (int ix = 0; ix < nxi; ++ix) {
                const int dx = ix - nxi / 2, dy = iy - nyi / 2;
                const double* src = (dx * dx + dy * dy <= 100) ? fluid : form;
                const size_t o = (size_t)ix + (size_t)iy * nxi + (size_t)iz * nxi * nyi;
                for (int f = 0; f < PW_NBASE; ++f) out[f * n + o] = src[f];
            }
    return PW_OK;
}

extern "C" pw_status pw_biot_from_rock(const pw_rock* r, double* out)
{
    if (!r || !out) return fail(PW_ERR_INVALID, "pw_biot_from_rock: null argument");
    if (!(r->Ks > 0 && r->Kf > 0 && r->phi > 0 && r->phi < 1 && r->kappa > 0 && r->pore_m > 0 && r->rhos > 0 &&
          r->rhof > 0))
        return fail(PW_ERR_INVALID, "pw_biot_from_rock: non-physical rock parameters");
    const double a1 = 1 - r->Kb / r->Ks;
    const double M = r->Kf * r->Ks / (r->phi * r->Ks + (a1 - r->phi) * r->Kf);
    const double HH = a1 * a1 * M + r->Kb + r->mu * 4.0 / 3.0;
    out[PW_RHO] = (1 - r->phi) * r->rhos + r->phi * r->rhof;
    out[PW_RHOF] = r->rhof;
    out[PW_C1] = r->eta / r->kappa;
    out[PW_C2] = (1 + 2 / r->pore_m) * r->tortuosity * r->rhof / r->phi;
    out[PW_HH] = HH;
    out[PW_H2U] = HH - 2 * r->mu;
    out[PW_CC] = M * a1;
    out[PW_MM] = M;
    out[PW_MU] = r->mu;
    return PW_OK;
}
