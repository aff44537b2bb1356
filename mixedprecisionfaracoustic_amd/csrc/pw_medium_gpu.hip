// pw_medium_gpu.hip -- FP64 medium precompute on the GPU.
//
// Replaces the host precompute of the reference: extmodel edge replication
// (Float/main.cu:1662-1794), staggered averaging (:473-513), the Crank-Nicolson
// coefficient pair of FD_V (:106-113) and the Half exponent scan and FP16
// coefficient quantisation (Half/main.cu:745-1045).  Everything is evaluated in
// FP64 from the cell-centred base fields and rounded once to the storage type.
// Also holds the layered synthetic-medium generator.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "pw_device.h"

namespace pw {

typedef _Float16 f16;

struct BaseView {
    const double* b[PW_NBASE];
    int nxi, nyi, nzi, np;
    long long rs, ps;  // element strides of a row and a plane of the base arrays
    // padded (extmodel) value of field f at total-grid cell (ix, iy, iz)
    __device__ __forceinline__ double at(int f, int ix, int iy, int iz) const
    {
        int kx = ix - np, ky = iy - np, kz = iz - np;
        kx = kx < 0 ? 0 : (kx >= nxi ? nxi - 1 : kx);
        ky = ky < 0 ? 0 : (ky >= nyi ? nyi - 1 : ky);
        kz = kz < 0 ? 0 : (kz >= nzi ? nzi - 1 : kz);
        return b[f][(long long)kx + (long long)ky * rs + (long long)kz * ps];
    }
};

// Staggered quantities of one cell, Float/main.cu:473-513 (zero on the outer
// layer, which the reference never computes).
struct Stag {
    double rix, riy, riz, c1x, c1y, c1z, c2x, c2y, c2z, muxy, muxz, muyz;
};

__device__ __forceinline__ double harm4(double a, double b, double c, double d)
{
    if (a == 0.0 || b == 0.0 || c == 0.0 || d == 0.0) return 0.0;
    return 1.0 / (0.25 * (1.0 / a + 1.0 / b + 1.0 / c + 1.0 / d));
}

__device__ Stag staggered(const BaseView& v, int ix, int iy, int iz, int nx, int ny, int nz)
{
    Stag s = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (ix < 1 || iy < 1 || iz < 1 || ix > nx - 2 || iy > ny - 2 || iz > nz - 2) return s;
    const double r0 = v.at(PW_RHO, ix, iy, iz);
    s.rix = 2.0 / (r0 + v.at(PW_RHO, ix + 1, iy, iz));
    s.riy = 2.0 / (r0 + v.at(PW_RHO, ix, iy + 1, iz));
    s.riz = 2.0 / (r0 + v.at(PW_RHO, ix, iy, iz + 1));
    const double a0 = v.at(PW_C1, ix, iy, iz), b0 = v.at(PW_C2, ix, iy, iz);
    s.c1x = 0.5 * (a0 + v.at(PW_C1, ix + 1, iy, iz));
    s.c1y = 0.5 * (a0 + v.at(PW_C1, ix, iy + 1, iz));
    s.c1z = 0.5 * (a0 + v.at(PW_C1, ix, iy, iz + 1));
    s.c2x = 0.5 * (b0 + v.at(PW_C2, ix + 1, iy, iz));
    s.c2y = 0.5 * (b0 + v.at(PW_C2, ix, iy + 1, iz));
    s.c2z = 0.5 * (b0 + v.at(PW_C2, ix, iy, iz + 1));
    const double m000 = v.at(PW_MU, ix, iy, iz);
    const double m100 = v.at(PW_MU, ix + 1, iy, iz), m010 = v.at(PW_MU, ix, iy + 1, iz);
    const double m001 = v.at(PW_MU, ix, iy, iz + 1), m110 = v.at(PW_MU, ix + 1, iy + 1, iz);
    const double m101 = v.at(PW_MU, ix + 1, iy, iz + 1), m011 = v.at(PW_MU, ix, iy + 1, iz + 1);
    s.muxz = harm4(m000, m100, m001, m101);
    s.muyz = harm4(m000, m010, m001, m011);
    s.muxy = harm4(m000, m100, m010, m110);
    return s;
}

// Largest squared P velocity of the cell: the high-frequency Biot fast wave,
// the larger root of det([[H, C], [C, M]] - v^2 [[rho, rhof], [rhof, m]]) = 0
// with m = C2; a cell without viscous coupling (the borehole fluid) uses H/rho.
__device__ double vp2_cell(double rho, double rhof, double c1, double m, double H, double C, double M)
{
    const double a = rho * m - rhof * rhof;
    if (c1 == 0.0 || a <= 1e-12 * rho * m) return H / rho;
    const double b = H * m + M * rho - 2.0 * C * rhof;
    const double c = H * M - C * C;
    const double disc = fmax(b * b - 4.0 * a * c, 0.0);
    return (b + sqrt(disc)) / (2.0 * a);
}

__device__ __forceinline__ void atomic_max_pos(double* addr, double v)
{
    // Non-negative doubles order like their bit patterns.
    atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_medium_reduce(BaseView v, int nx, int ny, int nz, double dt, double* res)
{
    __shared__ double sh[kReduceSlots][256];
    const long long n = (long long)nx * ny * nz;
    double kmax = 0.0, rmax = 0.0, vmax = 0.0, wmax = 0.0;
    double fmx[PW_NBASE] = {}, hmis = 0.0, rspread = 0.0;
    const double rf0 = v.at(PW_RHOF, 0, 0, 0);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((long long)nx * ny));
        const Stag s = staggered(v, ix, iy, iz, nx, ny, nz);
        const double rhof = v.at(PW_RHOF, ix, iy, iz);
        const double HH = v.at(PW_HH, ix, iy, iz), H2u = v.at(PW_H2U, ix, iy, iz);
        const double C = v.at(PW_CC, ix, iy, iz), M = v.at(PW_MM, ix, iy, iz), mu = v.at(PW_MU, ix, iy, iz);
        // Kmax, Half/main.cu:754-761
        kmax = fmax(kmax, fmax(fmax(fabs(HH), fabs(H2u)), fmax(fabs(C), fabs(M))));
        kmax = fmax(kmax, fmax(fabs(mu), fmax(fabs(s.muxy), fmax(fabs(s.muxz), fabs(s.muyz)))));
        // Rmax, Half/main.cu:763-788
        const double ri[3] = {s.rix, s.riy, s.riz}, c1[3] = {s.c1x, s.c1y, s.c1z}, c2[3] = {s.c2x, s.c2y, s.c2z};
        for (int d = 0; d < 3; ++d) {
            if (ri[d] > 0.0) rmax = fmax(rmax, ri[d]);
            const double Q = c2[d] - rhof * rhof * ri[d];
            if (c1[d] != 0.0 && fabs(Q) > 1.0e-12) {
                rmax = fmax(rmax, 1.0 / fabs(Q));
                rmax = fmax(rmax, fabs(rhof * ri[d] / Q));
            }
            // relative fluid velocity per unit solid velocity in one step: the
            // vw forcing rhof*rho^-1*S/CC1 against the vu forcing rho^-1*S
            const double cc1 = 0.5 * c1[d] * dt + Q;
            if (c1[d] != 0.0 && cc1 != 0.0) wmax = fmax(wmax, fabs(rhof / cc1));
        }
        vmax = fmax(vmax, vp2_cell(v.at(PW_RHO, ix, iy, iz), rhof, v.at(PW_C1, ix, iy, iz), v.at(PW_C2, ix, iy, iz),
                                   HH, C, M));
        // largest magnitude of every base field (comp16f's coefficient scales)
        for (int f = 0; f < PW_NBASE; ++f) fmx[f] = fmax(fmx[f], fabs(v.at(f, ix, iy, iz)));
        // how far H is from (H - 2 mu) + 2 mu (NaN counts as far)
        const double dh = fabs(HH - H2u - 2.0 * mu);
        hmis = fmax(hmis, dh <= 0.0 ? 0.0 : (dh < fabs(HH) ? dh / fabs(HH) : 1.0));
        // one fluid throughout? (NaN counts as not)
        const double dr = fabs(rhof - rf0);
        rspread = fmax(rspread, rf0 > 0.0 && dr <= 0.0 ? 0.0 : (dr > 0.0 ? dr : 1.0));
    }
    sh[0][threadIdx.x] = kmax; sh[1][threadIdx.x] = rmax; sh[2][threadIdx.x] = vmax; sh[3][threadIdx.x] = wmax;
    for (int f = 0; f < PW_NBASE; ++f) sh[4 + f][threadIdx.x] = fmx[f];
    sh[kSlotHMismatch][threadIdx.x] = hmis;
    sh[kSlotRhofSpread][threadIdx.x] = rspread;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < kReduceSlots; ++k) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < kReduceSlots) atomic_max_pos(res + threadIdx.x, sh[threadIdx.x][0]);
}

template <class T> __device__ __forceinline__ T to_store(double v) { return (T)v; }

struct CoefOut {
    void* c[kMaxCoef];
    int cstride;  // elements between a coefficient's consecutive cells (scaled form)
    int pitch;
    long long plane;
};

// Physical form (FP32 / FP64 storage): the nine cell-centred fields on the padded
// grid (extmodel, Float/main.cu:1662-1794), rounded once to the storage type.
template <class T>
__global__ void k_coef_physical(BaseView v, int nx, int ny, int nz, double dt, CoefOut out)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= nx) return;
    const size_t o = (size_t)ix + (size_t)iy * out.pitch + (size_t)iz * out.plane;
    T** c = (T**)out.c;
    c[PC_RHO][o] = to_store<T>(v.at(PW_RHO, ix, iy, iz));
    c[PC_RHOF][o] = to_store<T>(v.at(PW_RHOF, ix, iy, iz));
    c[PC_C1][o] = to_store<T>(v.at(PW_C1, ix, iy, iz));
    c[PC_C2][o] = to_store<T>(v.at(PW_C2, ix, iy, iz));
    c[PC_C][o] = to_store<T>(v.at(PW_CC, ix, iy, iz));
    c[PC_M][o] = to_store<T>(v.at(PW_MM, ix, iy, iz));
    if (c[PC_HH]) c[PC_HH][o] = to_store<T>(v.at(PW_HH, ix, iy, iz));  // null: comp16f
    c[PC_H2U][o] = to_store<T>(v.at(PW_H2U, ix, iy, iz));
    c[PC_MU][o] = to_store<T>(v.at(PW_MU, ix, iy, iz));
    (void)ny; (void)nz; (void)dt;
}

// comp16f: the two coefficient records of each cell (rec_base order), each
// coefficient rounded to FP32, times its power-of-two scale (exact), then
// encoded as FP16 value + residual with comp16f's rounding.
struct RecOut {
    f16* hi[kRecords];
    int8_t* lo[kRecords];
    float enc[kRecords][4];
    int pitch;
    long long plane;
};
__global__ void k_coef_records(BaseView v, int nx, RecOut out)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= nx) return;
    const size_t o = (size_t)ix + (size_t)iy * out.pitch + (size_t)iz * out.plane;
    for (int r = 0; r < kRecords; ++r) {
        const int w = rec_width(r);
        for (int q = 0; q < w; ++q) {
            f16 h;
            int8_t l;
            c16_encode_rn((float)v.at(rec_base(r, q), ix, iy, iz) * out.enc[r][q], h, l, 0u);
            out.hi[r][w * o + q] = h;
            out.lo[r][w * o + q] = l;
        }
    }
}

// Scaled form (comp16 / bf16), Half/main.cu:967-1044: FP16 arrays prescaled by
// dt and the velocity exponent 2^em.  Evaluated with the reference's own
// arithmetic (float staggered averages of Float/main.cu:480-509 with their double
// steps, float Half formulas, no contraction) so that every FP16 coefficient is
// the one the reference stores: in stiff media Aw is close to -1 and one FP16 ulp
// of Aw is a large part of 1 + Aw.
__device__ __forceinline__ float harm4_ref(float a, float b, float c, float d)
{
#pragma clang fp contract(off)
    if (a == 0 || b == 0 || c == 0 || d == 0) return 0.0f;
    return (float)(1 / (0.25 * (double)(1 / a + 1 / b + 1 / c + 1 / d)));
}

__global__ void k_coef_scaled(BaseView v, int nx, int ny, int nz, float DT, float sm, float ism, CoefOut out,
                              unsigned z)
{
#pragma clang fp contract(off)
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= nx) return;
    const size_t o = (size_t)ix + (size_t)iy * out.pitch + (size_t)iz * out.plane;
    // comp16: the cell's record, c[k] pointing at coefficient k's slot in it;
    // BF16: one array per coefficient
    const size_t r = o * (size_t)out.cstride;
    f16** c = (f16**)out.c;
    auto F = [&](int f, int dx, int dy, int dz) { return (float)v.at(f, ix + dx, iy + dy, iz + dz); };
    const float rhof0 = F(PW_RHOF, 0, 0, 0);
    float ri[3] = {0, 0, 0}, c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0}, muxz = 0, muyz = 0, muxy = 0;
    if (ix >= 1 && iy >= 1 && iz >= 1 && ix <= nx - 2 && iy <= ny - 2 && iz <= nz - 2) {
        const float r0 = F(PW_RHO, 0, 0, 0), a0 = F(PW_C1, 0, 0, 0), b0 = F(PW_C2, 0, 0, 0);
        ri[0] = 2 / (r0 + F(PW_RHO, 1, 0, 0));
        ri[1] = 2 / (r0 + F(PW_RHO, 0, 1, 0));
        ri[2] = 2 / (r0 + F(PW_RHO, 0, 0, 1));
        c1[0] = (float)(0.5 * (double)(a0 + F(PW_C1, 1, 0, 0)));
        c1[1] = (float)(0.5 * (double)(a0 + F(PW_C1, 0, 1, 0)));
        c1[2] = (float)(0.5 * (double)(a0 + F(PW_C1, 0, 0, 1)));
        c2[0] = (float)(0.5 * (double)(b0 + F(PW_C2, 1, 0, 0)));
        c2[1] = (float)(0.5 * (double)(b0 + F(PW_C2, 0, 1, 0)));
        c2[2] = (float)(0.5 * (double)(b0 + F(PW_C2, 0, 0, 1)));
        const float m0 = F(PW_MU, 0, 0, 0);
        muxz = harm4_ref(m0, F(PW_MU, 1, 0, 0), F(PW_MU, 0, 0, 1), F(PW_MU, 1, 0, 1));
        muyz = harm4_ref(m0, F(PW_MU, 0, 1, 0), F(PW_MU, 0, 0, 1), F(PW_MU, 0, 1, 1));
        muxy = harm4_ref(m0, F(PW_MU, 1, 0, 0), F(PW_MU, 0, 1, 0), F(PW_MU, 1, 1, 0));
    }
    const float Q_EPS = 1.0e-12f;
    for (int d = 0; d < 3; ++d) {
        const float Q = c2[d] - rhof0 * rhof0 * ri[d];
        const float den = Q + 0.5f * c1[d] * DT;
        float aw = 0.0f, bp = 0.0f, bt = 0.0f;
        if (c1[d] != 0.0f && fabsf(Q) > Q_EPS && fabsf(den) > Q_EPS) {
            aw = (Q - 0.5f * c1[d] * DT) / den;
            bp = ism * DT / den;
            bt = ism * DT * rhof0 * ri[d] / den;
        }
        c[SC_RITX + d][r] = (f16)round32(ism * DT * ri[d], z);
        c[SC_FRX + d][r] = (f16)round32(rhof0 * ri[d], z);
        c[SC_AWX + d][r] = (f16)round32(aw, z);
        c[SC_BPX + d][r] = (f16)round32(bp, z);
        c[SC_BTX + d][r] = (f16)round32(bt, z);
    }
    c[SC_HH][r] = (f16)round32(sm * DT * F(PW_HH, 0, 0, 0), z);
    c[SC_H2U][r] = (f16)round32(sm * DT * F(PW_H2U, 0, 0, 0), z);
    c[SC_C][r] = (f16)round32(sm * DT * F(PW_CC, 0, 0, 0), z);
    c[SC_M][r] = (f16)round32(sm * DT * F(PW_MM, 0, 0, 0), z);
    c[SC_MUXZ][r] = (f16)round32(sm * DT * muxz, z);
    c[SC_MUXY][r] = (f16)round32(sm * DT * muxy, z);
    c[SC_MUYZ][r] = (f16)round32(sm * DT * muyz, z);
    if (out.cstride == 4) {  // the unused fourth slots of records 3 and 5
        c[SC_FRX][r + 3] = (f16)0.0f;
        c[SC_MUXY][r + 3] = (f16)0.0f;
    }
}

static BaseView make_view(const MediumArgs& m)
{
    BaseView v;
    for (int k = 0; k < PW_NBASE; ++k) v.b[k] = m.base[k];
    v.nxi = m.nxi; v.nyi = m.nyi; v.nzi = m.nzi; v.np = m.np;
    v.rs = m.row_stride > 0 ? m.row_stride : m.nxi;
    v.ps = m.plane_stride > 0 ? m.plane_stride : (long long)m.nxi * m.nyi;
    return v;
}

__global__ void k_fill_f64(double* p, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t medium_reduce(const MediumArgs& m, double* d_res, hipStream_t st)
{
    // the maxima start at zero (a fill kernel, not a memset: the same code serves
    // the scratch array, which keeps the previous call's values)
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, st, d_res, kReduceSlots, 0.0);
    hipLaunchKernelGGL(k_medium_reduce, dim3(2048), dim3(256), 0, st, make_view(m), m.nx, m.ny, m.nz, m.dt, d_res);
    return hipGetLastError();
}

// Scratch of pw_set_medium_device: the medium reduction's maxima go here
// instead of into an allocation.  Each device holds its own copy until the
// process ends.  Host code never passes the name to a kernel: it takes the
// array's device address once per device, with that device current.
static __device__ __align__(256) double scratch[kScratchDoubles];
static_assert(kReduceSlots <= kScratchDoubles, "the reduction must fit the scratch array");

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

std::mutex& scratch_mutex() { return g_scratch_mu; }

// Caller holds scratch_mutex() and has made dev current.
hipError_t scratch_for_device(int dev, double** out)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[dev]) {
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

hipError_t medium_coefficients(const MediumArgs& m, const DevBuffers& b, double scale_m, double inv_scale_m,
                               hipStream_t st)
{
    CoefOut out;
    for (int k = 0; k < kMaxCoef; ++k) out.c[k] = b.coef[k];
    out.cstride = has_srec(b.prec) ? 4 : 1;
    out.pitch = b.g.pitch;
    out.plane = b.g.plane;
    const dim3 grid((unsigned)((m.nx + 127) / 128), (unsigned)m.ny, (unsigned)m.nz), block(128);
    const BaseView v = make_view(m);
    switch (b.prec) {
    case PW_FP32: hipLaunchKernelGGL(k_coef_physical<float>, grid, block, 0, st, v, m.nx, m.ny, m.nz, m.dt, out); break;
    case PW_FP64: hipLaunchKernelGGL(k_coef_physical<double>, grid, block, 0, st, v, m.nx, m.ny, m.nz, m.dt, out); break;
    case PW_COMP16F: {
        RecOut ro;
        for (int r = 0; r < kRecords; ++r) {
            ro.hi[r] = (f16*)b.rec_hi[r];
            ro.lo[r] = (int8_t*)b.rec_lo[r];
            for (int q = 0; q < 4; ++q) ro.enc[r][q] = 1.0f / b.rec_scale[r][q];
        }
        ro.pitch = b.g.pitch;
        ro.plane = b.g.plane;
        hipLaunchKernelGGL(k_coef_records, grid, block, 0, st, v, m.nx, ro);
        break;
    }
    case PW_COMP16:
    case PW_BF16:
        hipLaunchKernelGGL(k_coef_scaled, grid, block, 0, st, v, m.nx, m.ny, m.nz, (float)m.dt, (float)scale_m,
                           (float)inv_scale_m, out, 0u);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// Layered synthetic medium (FP64)
// --------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Standard normal white noise, one independent stream per field.
__global__ void k_noise(double* out, long long n, unsigned long long seed, int stream)
{
    const unsigned long long key = seed ^ ((unsigned long long)(stream + 1) * 0xD1B54A32D192ED03ull);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const unsigned long long h1 = splitmix64(key + 2ull * (unsigned long long)i);
        const unsigned long long h2 = splitmix64(key + 2ull * (unsigned long long)i + 1ull);
        const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;   // (0, 1]
        const double u2 = (double)(h2 >> 11) * 0x1.0p-53;           // [0, 1)
        out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

struct GaussW {
    double w[97];
    int radius;
};

// One separable Gaussian pass along axis (0 = x, 1 = y, 2 = z), edge-clamped.
__global__ void k_smooth(const double* __restrict__ in, double* __restrict__ out, int nx, int ny, int nz, int axis,
                         GaussW g)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= nx) return;
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const int i0 = axis == 0 ? ix : (axis == 1 ? iy : iz);
    const long long stride = axis == 0 ? 1 : (axis == 1 ? nx : (long long)nx * ny);
    const long long base = (long long)ix + (long long)iy * nx + (long long)iz * nx * ny - (long long)i0 * stride;
    double acc = 0.0;
    for (int k = -g.radius; k <= g.radius; ++k) {
        int j = i0 + k;
        j = j < 0 ? 0 : (j >= n ? n - 1 : j);
        acc += g.w[k + g.radius] * in[base + (long long)j * stride];
    }
    out[(long long)ix + (long long)iy * nx + (long long)iz * nx * ny] = acc;
}

// Per-block sums (fixed grid, fixed order) for a deterministic mean / variance.
__global__ void k_partial_sums(const double* __restrict__ in, long long n, double* __restrict__ part)
{
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double v = in[i];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a; s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s1[threadIdx.x] += s1[threadIdx.x + w]; s2[threadIdx.x] += s2[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[blockIdx.x] = s1[0]; part[gridDim.x + blockIdx.x] = s2[0]; }
}

struct LayerSpec {
    int nlayers;
    int z_top[PW_MAX_LAYERS];
    pw_rock rock[PW_MAX_LAYERS];
    double amp, clamp;
    double mean[3], inv_std[3];
    int use_noise;
};

// Biot coefficients per interior cell (formulas of Float/main.cu:1635-1650).
__global__ void k_layered(double* base, const double* __restrict__ n_kb, const double* __restrict__ n_mu,
                          const double* __restrict__ n_phi, int nxi, int nyi, int nzi, int np, LayerSpec L)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= nxi) return;
    const long long n = (long long)nxi * nyi * nzi;
    const long long i = (long long)ix + (long long)iy * nxi + (long long)iz * nxi * nyi;
    const int zt = iz + np;
    int l = 0;
    while (l + 1 < L.nlayers && zt >= L.z_top[l + 1]) ++l;
    const pw_rock r = L.rock[l];
    double fkb = 1.0, fmu = 1.0, fphi = 1.0;
    if (L.use_noise) {
        const double a = fmin(fmax((n_kb[i] - L.mean[0]) * L.inv_std[0], -L.clamp), L.clamp);
        const double b = fmin(fmax((n_mu[i] - L.mean[1]) * L.inv_std[1], -L.clamp), L.clamp);
        const double c = fmin(fmax((n_phi[i] - L.mean[2]) * L.inv_std[2], -L.clamp), L.clamp);
        fkb = 1.0 + L.amp * a;
        fmu = 1.0 + L.amp * b;
        fphi = 1.0 + L.amp * c;
    }
    const double Kb = r.Kb * fkb, mu = r.mu * fmu, phi = r.phi * fphi;
    const double a1 = 1.0 - Kb / r.Ks;
    const double M = r.Kf * r.Ks / (phi * r.Ks + (a1 - phi) * r.Kf);
    const double HH = a1 * a1 * M + Kb + mu * 4.0 / 3.0;
    base[PW_RHO * n + i] = (1.0 - phi) * r.rhos + phi * r.rhof;
    base[PW_RHOF * n + i] = r.rhof;
    base[PW_C1 * n + i] = r.eta / r.kappa;
    base[PW_C2 * n + i] = (1.0 + 2.0 / r.pore_m) * r.tortuosity * r.rhof / phi;
    base[PW_HH * n + i] = HH;
    base[PW_H2U * n + i] = HH - 2.0 * mu;
    base[PW_CC * n + i] = M * a1;
    base[PW_MM * n + i] = M;
    base[PW_MU * n + i] = mu;
}

hipError_t generate_layered(const pw_layered& spec, int nxi, int nyi, int nzi, int np, double* d_base,
                            double* d_work, double* d_partial, hipStream_t st)
{
    const long long n = (long long)nxi * nyi * nzi;
    LayerSpec L;
    L.nlayers = spec.nlayers;
    for (int k = 0; k < PW_MAX_LAYERS; ++k) {
        L.z_top[k] = spec.z_top[k];
        L.rock[k] = spec.rock[k];
    }
    L.amp = spec.perturb;
    L.clamp = spec.clamp > 0 ? spec.clamp : 3.0;
    L.use_noise = spec.perturb != 0.0;
    const dim3 g3((unsigned)((nxi + 127) / 128), (unsigned)nyi, (unsigned)nzi);
    if (L.use_noise) {
        GaussW gw;
        const double sig = spec.corr_cells > 0 ? spec.corr_cells : 8.0;
        gw.radius = (int)std::ceil(3.0 * sig);
        if (gw.radius > 48) gw.radius = 48;
        double sum = 0.0;
        for (int k = -gw.radius; k <= gw.radius; ++k) sum += std::exp(-0.5 * k * k / (sig * sig));
        for (int k = -gw.radius; k <= gw.radius; ++k) gw.w[k + gw.radius] = std::exp(-0.5 * k * k / (sig * sig)) / sum;
        double* tmp = d_work + 3 * n;
        const int nblk = 1024;
        std::vector<double> part(2 * nblk);
        for (int f = 0; f < 3; ++f) {
            double* fld = d_work + f * n;
            hipLaunchKernelGGL(k_noise, dim3(4096), dim3(256), 0, st, fld, n, spec.seed, f);
            hipLaunchKernelGGL(k_smooth, g3, dim3(128), 0, st, fld, tmp, nxi, nyi, nzi, 0, gw);
            hipLaunchKernelGGL(k_smooth, g3, dim3(128), 0, st, tmp, fld, nxi, nyi, nzi, 1, gw);
            hipLaunchKernelGGL(k_smooth, g3, dim3(128), 0, st, fld, tmp, nxi, nyi, nzi, 2, gw);
            hipError_t e = hipMemcpyAsync(fld, tmp, n * sizeof(double), hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_partial_sums, dim3(nblk), dim3(256), 0, st, fld, n, d_partial);
            e = hipMemcpyAsync(part.data(), d_partial, 2 * nblk * sizeof(double), hipMemcpyDeviceToHost, st);
            if (e != hipSuccess) return e;
            e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            double s = 0.0, s2 = 0.0;
            for (int b = 0; b < nblk; ++b) { s += part[b]; s2 += part[nblk + b]; }
            const double mean = s / (double)n;
            const double var = s2 / (double)n - mean * mean;
            L.mean[f] = mean;
            L.inv_std[f] = var > 0 ? 1.0 / std::sqrt(var) : 0.0;
        }
    } else {
        for (int f = 0; f < 3; ++f) { L.mean[f] = 0; L.inv_std[f] = 0; }
    }
    hipLaunchKernelGGL(k_layered, g3, dim3(128), 0, st, d_base, d_work, d_work + n, d_work + 2 * n, nxi, nyi, nzi, np,
                       L);
    return hipGetLastError();
}

}  // namespace pw
