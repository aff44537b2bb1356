// pw_kernels.hip -- CDNA4 (gfx950) kernels of the Biot staggered-grid step.
//
// Per step (the order of Float/main.cu:1055-1093):
//   k_source   : add the source sample at one cell (Float/main.cu:45-59)
//   k_velocity : fused solid + relative-fluid velocity update with the
//                Crank-Nicolson viscous coupling (Float/main.cu:64-192)
//   k_stress   : fused stress + pore-pressure update (Float/main.cu:194-295)
//   k_gather   : receivers -> device seismogram buffer (Float/main.cu:1083-1093)
//
// Layout: one padded SoA array per wavefield (row pitch = nx rounded to 64
// elements).  A block is a 64 x 4 (x, y) column tile that marches up a chunk of
// z; each wave owns one 64-cell x row, so every load is a coalesced 256 B (FP32)
// row segment, x/y neighbours hit L1/L2, and z neighbours are carried in
// registers from the previous iteration (each value is loaded once per block).
//
// CPML: the memory variables exist only on the boundary slabs of each axis (in
// the interior the reference's psi is identically zero).  Each term keeps ONE
// memory variable phi instead of the reference's (psi, previous derivative)
// pair: with c = dt*d/2,
//   psi_n = e*psi_{n-1} - c*(e*der_{n-1} + der_n)         (Float/main.cu:130)
// equals psi_n = phi_{n-1} - c*der_n with phi_n = e*(psi_n - c*der_n), so
// phi carries e*psi_n - c*e*der_n and the per-term traffic halves.
//
// Storage policies: FP32 and FP64 use the reference Float formulas; comp16
// (FP16 hi + int8 residual, Half/main.cu:49-100) and BF16 use the reference
// Half formulas with FP16 prescaled coefficients.  All arithmetic is FP32
// except the FP64 policy.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pw_device.h"

namespace pw {

typedef _Float16 f16;
typedef __bf16 bf16;

// --------------------------------------------------------------------------
// Compensated FP16 codec (Half/main.cu:59-100).
// unit = 2^(e-19) where |hi| = f * 2^e, f in [1/2, 1): for a normal float with
// biased exponent E that is 2^(E-145); below 2^-14 the unit is 2^-32.
// --------------------------------------------------------------------------
__device__ __forceinline__ float c16_unit(float hf)
{
    const float a = __builtin_fabsf(hf);
    if (a < 6.103515625e-5f) return 2.3283064365386963e-10f;
    const unsigned E = (__float_as_uint(a) >> 23) & 0xffu;
    return __uint_as_float((E - 18u) << 23);
}
__device__ __forceinline__ float c16_inv_unit(float hf)
{
    const float a = __builtin_fabsf(hf);
    if (a < 6.103515625e-5f) return 4294967296.0f;
    const unsigned E = (__float_as_uint(a) >> 23) & 0xffu;
    return __uint_as_float((272u - E) << 23);   // 2^(145-E)
}
__device__ __forceinline__ float c16_decode(f16 hi, int8_t lo)
{
    const float hf = (float)hi;
    return hf + (float)lo * c16_unit(hf);
}
__device__ __forceinline__ void c16_encode(float v, f16& hi, int8_t& lo)
{
    const f16 h = (f16)v;
    const float hf = (float)h;
    const float scaled = (v - hf) * c16_inv_unit(hf);   // exact: power-of-two scale
    int q = (int)(scaled + (scaled >= 0.0f ? 0.5f : -0.5f));
    q = q > 127 ? 127 : (q < -127 ? -127 : q);
    hi = h;
    lo = (int8_t)q;
}

// --------------------------------------------------------------------------
// Storage policies
// --------------------------------------------------------------------------
template <class R> struct Tab4 { R cf, ef, ch, eh; };

struct PolF32 {
    using R = float; using Coef = float; using Mem = float;
    static constexpr bool scaled = false;
    struct Field { float* p; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return f.p[o]; }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v) { f.p[o] = v; }
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v) { m[o] = v; }
};
struct PolF64 {
    using R = double; using Coef = double; using Mem = double;
    static constexpr bool scaled = false;
    struct Field { double* p; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return f.p[o]; }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v) { f.p[o] = v; }
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v) { m[o] = v; }
};
struct PolC16 {
    using R = float; using Coef = f16; using Mem = f16;
    static constexpr bool scaled = true;
    struct Field { f16* hi; int8_t* lo; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return c16_decode(f.hi[o], f.lo[o]); }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v)
    {
        f16 h;
        int8_t l;
        c16_encode(v, h, l);
        f.hi[o] = h;
        f.lo[o] = l;
    }
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return (float)c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return (float)m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v) { m[o] = (f16)v; }
};
struct PolB16 {
    using R = float; using Coef = f16; using Mem = bf16;
    static constexpr bool scaled = true;
    struct Field { bf16* p; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return (float)f.p[o]; }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v) { f.p[o] = (bf16)v; }
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return (float)c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return (float)m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v) { m[o] = (bf16)v; }
};

template <class P> struct StepArgs {
    typename P::Field f[PW_NFIELDS];
    const typename P::Coef* c[kMaxCoef];
    typename P::Mem* m[3][kMemPerAxis];
    const Tab4<typename P::R>* tab[3];
    int nx, ny, nz, pitch;
    unsigned plane;
    int lo[3], hi[3];
    int zc;
    int gx, gy, nblocks;   // tile grid (x tiles, y tiles) and total blocks of the 1-D launch
    int zb, ze;            // planes [zb, ze) updated by this launch
    int xcd_map;           // 1: XCD-aware block -> tile mapping
    typename P::R dt, h;
};

template <class P> __host__ typename P::Field make_field(const DevBuffers& b, int k);
template <> __host__ PolF32::Field make_field<PolF32>(const DevBuffers& b, int k) { return {(float*)b.f[k]}; }
template <> __host__ PolF64::Field make_field<PolF64>(const DevBuffers& b, int k) { return {(double*)b.f[k]}; }
template <> __host__ PolC16::Field make_field<PolC16>(const DevBuffers& b, int k) { return {(f16*)b.f[k], b.lo[k]}; }
template <> __host__ PolB16::Field make_field<PolB16>(const DevBuffers& b, int k) { return {(bf16*)b.f[k]}; }

template <class P> __host__ StepArgs<P> make_args(const DevBuffers& b)
{
    using R = typename P::R;
    StepArgs<P> a;
    for (int k = 0; k < PW_NFIELDS; ++k) a.f[k] = make_field<P>(b, k);
    for (int k = 0; k < kMaxCoef; ++k) a.c[k] = (const typename P::Coef*)b.coef[k];
    for (int ax = 0; ax < 3; ++ax) {
        for (int k = 0; k < kMemPerAxis; ++k) a.m[ax][k] = (typename P::Mem*)b.mem[ax][k];
        a.tab[ax] = (const Tab4<R>*)b.tab[ax];
        a.lo[ax] = b.g.lo[ax];
        a.hi[ax] = b.g.hi[ax];
    }
    a.nx = b.g.nx; a.ny = b.g.ny; a.nz = b.g.nz; a.pitch = b.g.pitch;
    a.plane = (unsigned)b.g.plane;
    a.zc = b.zc;
    a.gx = (b.g.nx + 63) / 64;
    a.gy = (b.g.ny + 3) / 4;
    a.zb = 1;
    a.ze = b.g.nz - 1;
    a.nblocks = a.gx * a.gy * ((b.g.nz - 2 + b.zc - 1) / b.zc);
    a.xcd_map = b.xcd_map;
    a.dt = (R)b.dt;
    a.h = (R)b.h;
    return a;
}

// Slab-local index along one axis, -1 outside the CPML slabs.
__device__ __forceinline__ int slab_index(int i, int n, int lo, int hi)
{
    if (i < lo) return i;
    if (i >= n - hi) return i - (n - hi) + lo;
    return -1;
}

// Block -> (x tile, y tile, z chunk).  The hardware deals workgroups to the 8
// XCDs round-robin (block L runs on XCD L % 8), each XCD with its own L2.  With
// the XCD map, XCD k gets one contiguous run of tiles ordered x fastest, then y,
// then z, so the x and y neighbour tiles whose halo rows a tile re-reads were
// just fetched into the same L2.
struct Tile { int bx, by, bz; };
__device__ __forceinline__ Tile tile_of_block(int gx, int gy, int nblocks, int xcd_map)
{
    const unsigned L = blockIdx.x;
    unsigned t = L;
    if (xcd_map) {
        const unsigned q = (unsigned)nblocks / 8u, r = (unsigned)nblocks % 8u;
        const unsigned k = L % 8u, j = L / 8u;
        t = k * q + (k < r ? k : r) + j;
    }
    Tile tl;
    tl.bx = (int)(t % (unsigned)gx);
    const unsigned row = t / (unsigned)gx;
    tl.by = (int)(row % (unsigned)gy);
    tl.bz = (int)(row / (unsigned)gy);
    return tl;
}

// One CPML term in single-variable form (see header comment).
template <class P>
__device__ __forceinline__ typename P::R cpml(typename P::Mem* phi, unsigned k, typename P::R c, typename P::R e,
                                              typename P::R der)
{
    using R = typename P::R;
    const R psi = P::ldm(phi, k) - c * der;
    P::stm(phi, k, e * (psi - c * der));
    return der + psi;
}

// --------------------------------------------------------------------------
// Reference arithmetic of the staggered medium (Float/main.cu:473-513) and of
// the Crank-Nicolson pair (Float/main.cu:106-113), formed on the fly from the
// cell-centred fields so that FP32 runs see the reference's exact coefficients.
// --------------------------------------------------------------------------
// rho_temp = 2 / (rho_i + rho_j)  (float in the reference)
template <class R> __device__ __forceinline__ R inv_mean(R a, R b) { return (R)2 / (a + b); }
// C1x = 0.5 * (a + b): the sum is formed in the storage type, the halving is exact
template <class R> __device__ __forceinline__ R half_sum(R a, R b) { return (a + b) * (R)0.5; }
// mu_xy = 1 / (0.25 * (1/a + 1/b + 1/c + 1/d)), 0 if any is 0: reciprocals and
// their sum in the storage type, the last two operations in double
template <class R> __device__ __forceinline__ R harm4(R a, R b, R c, R d)
{
    if (a == (R)0 || b == (R)0 || c == (R)0 || d == (R)0) return (R)0;
    const R s = (R)1 / a + (R)1 / b + (R)1 / c + (R)1 / d;
    return (R)(1.0 / (0.25 * (double)s));
}
// CC1 = 0.5*C1*DT + C2 - rhof^2 rho^-1, CC2 = C2 - rhof^2 rho^-1 - 0.5*C1*DT:
// the reference's double literal 0.5 makes these double expressions
template <class R>
__device__ __forceinline__ void cn_pair(R c1, R c2, R rf, R ri, R dt, R& cc1, R& cc2)
{
    const double q = (double)(rf * rf * ri);
    const double h = 0.5 * (double)c1 * (double)dt;
    cc1 = (R)(h + (double)c2 - q);
    cc2 = (R)((double)c2 - q - h);
}

// --------------------------------------------------------------------------
// Velocity kernel: Float/main.cu:64-192 (physical) / Half/main.cu:156-268 (scaled)
// --------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(256) void k_velocity(const StepArgs<P> a)
{
    using R = typename P::R;
    const Tile tl = tile_of_block(a.gx, a.gy, a.nblocks, a.xcd_map);
    const int ix = tl.bx * 64 + threadIdx.x;
    const int iy = tl.by * 4 + threadIdx.y;
    if (ix < 1 || ix > a.nx - 2 || iy < 1 || iy > a.ny - 2) return;
    const int z0 = a.zb + tl.bz * a.zc;
    const int z1 = min(z0 + a.zc, a.ze);
    if (z0 >= z1) return;
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)z0 * PL;
    const R H = a.h, dt = a.dt;

    const int xs = slab_index(ix, a.nx, a.lo[0], a.hi[0]);
    const int ys = slab_index(iy, a.ny, a.lo[1], a.hi[1]);
    const int XW = a.lo[0] + a.hi[0], YW = a.lo[1] + a.hi[1];
    Tab4<R> tx = {0, 1, 0, 1}, ty = {0, 1, 0, 1};
    if (xs >= 0) tx = a.tab[0][ix];
    if (ys >= 0) ty = a.tab[1][iy];

    R txz_m = P::ld(a.f[PW_TXZ], o - PL);
    R tyz_m = P::ld(a.f[PW_TYZ], o - PL);
    R tzz_c = P::ld(a.f[PW_TZZ], o);
    R p_c = P::ld(a.f[PW_P], o);
    // physical form: cell-centred rho, C1, C2 of this plane, carried up the column
    R rho_c = 0, c1_c = 0, c2_c = 0;
    if constexpr (!P::scaled) {
        rho_c = P::ldc(a.c[PC_RHO], o);
        c1_c = P::ldc(a.c[PC_C1], o);
        c2_c = P::ldc(a.c[PC_C2], o);
    }

    for (int iz = z0; iz < z1; ++iz, o += PL) {
        const R txz_c = P::ld(a.f[PW_TXZ], o);
        const R tyz_c = P::ld(a.f[PW_TYZ], o);
        const R tzz_n = P::ld(a.f[PW_TZZ], o + PL);
        const R p_n = P::ld(a.f[PW_P], o + PL);
        const R txx_c = P::ld(a.f[PW_TXX], o), txx_r = P::ld(a.f[PW_TXX], o + 1);
        const R tyy_c = P::ld(a.f[PW_TYY], o), tyy_q = P::ld(a.f[PW_TYY], o + W);
        const R txy_c = P::ld(a.f[PW_TXY], o), txy_l = P::ld(a.f[PW_TXY], o - 1), txy_h = P::ld(a.f[PW_TXY], o - W);
        const R txz_l = P::ld(a.f[PW_TXZ], o - 1);
        const R tyz_h = P::ld(a.f[PW_TYZ], o - W);
        const R p_r = P::ld(a.f[PW_P], o + 1), p_q = P::ld(a.f[PW_P], o + W);

        // Float/main.cu:115-128
        R x1 = (txx_r - txx_c) / H, x2 = (txy_c - txy_h) / H, x3 = (txz_c - txz_m) / H, s1 = (p_r - p_c) / H;
        R y1 = (tyy_q - tyy_c) / H, y2 = (txy_c - txy_l) / H, y3 = (tyz_c - tyz_m) / H, s2 = (p_q - p_c) / H;
        R z1 = (tzz_n - tzz_c) / H, z2 = (txz_c - txz_l) / H, z3 = (tyz_c - tyz_h) / H, s3 = (p_n - p_c) / H;

        if (xs >= 0) {
            const unsigned k = ((unsigned)iz * a.ny + iy) * XW + xs;
            x1 = cpml<P>(a.m[0][0], k, tx.ch, tx.eh, x1);
            s1 = cpml<P>(a.m[0][1], k, tx.ch, tx.eh, s1);
            y2 = cpml<P>(a.m[0][2], k, tx.cf, tx.ef, y2);
            z2 = cpml<P>(a.m[0][3], k, tx.cf, tx.ef, z2);
        }
        if (ys >= 0) {
            const unsigned k = ((unsigned)iz * YW + ys) * W + ix;
            y1 = cpml<P>(a.m[1][0], k, ty.ch, ty.eh, y1);
            s2 = cpml<P>(a.m[1][1], k, ty.ch, ty.eh, s2);
            x2 = cpml<P>(a.m[1][2], k, ty.cf, ty.ef, x2);
            z3 = cpml<P>(a.m[1][3], k, ty.cf, ty.ef, z3);
        }
        const int zs = slab_index(iz, a.nz, a.lo[2], a.hi[2]);
        if (zs >= 0) {
            const Tab4<R> tz = a.tab[2][iz];
            const unsigned k = ((unsigned)zs * a.ny + iy) * W + ix;
            z1 = cpml<P>(a.m[2][0], k, tz.ch, tz.eh, z1);
            s3 = cpml<P>(a.m[2][1], k, tz.ch, tz.eh, s3);
            x3 = cpml<P>(a.m[2][2], k, tz.cf, tz.ef, x3);
            y3 = cpml<P>(a.m[2][3], k, tz.cf, tz.ef, y3);
        }

        const R sx = x1 + x2 + x3, sy = y1 + y2 + y3, sz = z1 + z2 + z3;
        const R wx0 = P::ld(a.f[PW_VWX], o), wy0 = P::ld(a.f[PW_VWY], o), wz0 = P::ld(a.f[PW_VWZ], o);
        R wx1, wy1, wz1, ux1, uy1, uz1;
        if constexpr (!P::scaled) {
            const R rho_n = P::ldc(a.c[PC_RHO], o + PL), rho_x = P::ldc(a.c[PC_RHO], o + 1),
                    rho_y = P::ldc(a.c[PC_RHO], o + W);
            const R c1_n = P::ldc(a.c[PC_C1], o + PL), c1_x = P::ldc(a.c[PC_C1], o + 1),
                    c1_y = P::ldc(a.c[PC_C1], o + W);
            const R c2_n = P::ldc(a.c[PC_C2], o + PL), c2_x = P::ldc(a.c[PC_C2], o + 1),
                    c2_y = P::ldc(a.c[PC_C2], o + W);
            const R rf = P::ldc(a.c[PC_RHOF], o);
            const R rix = inv_mean(rho_c, rho_x), riy = inv_mean(rho_c, rho_y), riz = inv_mean(rho_c, rho_n);
            const R c1x = half_sum(c1_c, c1_x), c1y = half_sum(c1_c, c1_y), c1z = half_sum(c1_c, c1_n);
            const R c2x = half_sum(c2_c, c2_x), c2y = half_sum(c2_c, c2_y), c2z = half_sum(c2_c, c2_n);
            R cc1x, cc2x, cc1y, cc2y, cc1z, cc2z;
            cn_pair(c1x, c2x, rf, rix, dt, cc1x, cc2x);
            cn_pair(c1y, c2y, rf, riy, dt, cc1y, cc2y);
            cn_pair(c1z, c2z, rf, riz, dt, cc1z, cc2z);
            // Float/main.cu:160-189
            wx1 = c1x == (R)0 ? (R)0 : (wx0 * cc2x - dt * (rf * rix * sx + s1)) / cc1x;
            wy1 = c1y == (R)0 ? (R)0 : (wy0 * cc2y - dt * (rf * riy * sy + s2)) / cc1y;
            wz1 = c1z == (R)0 ? (R)0 : (wz0 * cc2z - dt * (rf * riz * sz + s3)) / cc1z;
            ux1 = P::ld(a.f[PW_VUX], o) + dt * rix * sx - rf * rix * (wx1 - wx0);
            uy1 = P::ld(a.f[PW_VUY], o) + dt * riy * sy - rf * riy * (wy1 - wy0);
            uz1 = P::ld(a.f[PW_VUZ], o) + dt * riz * sz - rf * riz * (wz1 - wz0);
            rho_c = rho_n; c1_c = c1_n; c2_c = c2_n;
        } else {
            // Half/main.cu:243-262: vw = Aw*vw - Bt*S - Bp*dp (0 where Aw = Bp = Bt = 0);
            // vu += rho~^-1*S - rhof*rho^-1*(vw_new - vw_old)
            const R awx = P::ldc(a.c[SC_AWX], o), bpx = P::ldc(a.c[SC_BPX], o), btx = P::ldc(a.c[SC_BTX], o);
            const R awy = P::ldc(a.c[SC_AWY], o), bpy = P::ldc(a.c[SC_BPY], o), bty = P::ldc(a.c[SC_BTY], o);
            const R awz = P::ldc(a.c[SC_AWZ], o), bpz = P::ldc(a.c[SC_BPZ], o), btz = P::ldc(a.c[SC_BTZ], o);
            wx1 = (awx == 0.0f && bpx == 0.0f && btx == 0.0f) ? 0.0f : awx * wx0 - btx * sx - bpx * s1;
            wy1 = (awy == 0.0f && bpy == 0.0f && bty == 0.0f) ? 0.0f : awy * wy0 - bty * sy - bpy * s2;
            wz1 = (awz == 0.0f && bpz == 0.0f && btz == 0.0f) ? 0.0f : awz * wz0 - btz * sz - bpz * s3;
            ux1 = P::ld(a.f[PW_VUX], o) + P::ldc(a.c[SC_RITX], o) * sx - P::ldc(a.c[SC_FRX], o) * (wx1 - wx0);
            uy1 = P::ld(a.f[PW_VUY], o) + P::ldc(a.c[SC_RITY], o) * sy - P::ldc(a.c[SC_FRY], o) * (wy1 - wy0);
            uz1 = P::ld(a.f[PW_VUZ], o) + P::ldc(a.c[SC_RITZ], o) * sz - P::ldc(a.c[SC_FRZ], o) * (wz1 - wz0);
        }
        P::st(a.f[PW_VWX], o, wx1); P::st(a.f[PW_VWY], o, wy1); P::st(a.f[PW_VWZ], o, wz1);
        P::st(a.f[PW_VUX], o, ux1); P::st(a.f[PW_VUY], o, uy1); P::st(a.f[PW_VUZ], o, uz1);

        txz_m = txz_c; tyz_m = tyz_c; tzz_c = tzz_n; p_c = p_n;
    }
}

// --------------------------------------------------------------------------
// Stress kernel: Float/main.cu:194-295 (physical) / Half/main.cu:270-371 (scaled)
// --------------------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(256) void k_stress(const StepArgs<P> a)
{
    using R = typename P::R;
    const Tile tl = tile_of_block(a.gx, a.gy, a.nblocks, a.xcd_map);
    const int ix = tl.bx * 64 + threadIdx.x;
    const int iy = tl.by * 4 + threadIdx.y;
    if (ix < 1 || ix > a.nx - 2 || iy < 1 || iy > a.ny - 2) return;
    const int z0 = a.zb + tl.bz * a.zc;
    const int z1 = min(z0 + a.zc, a.ze);
    if (z0 >= z1) return;
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)z0 * PL;
    const R H = a.h, dt = a.dt;

    const int xs = slab_index(ix, a.nx, a.lo[0], a.hi[0]);
    const int ys = slab_index(iy, a.ny, a.lo[1], a.hi[1]);
    const int XW = a.lo[0] + a.hi[0], YW = a.lo[1] + a.hi[1];
    Tab4<R> tx = {0, 1, 0, 1}, ty = {0, 1, 0, 1};
    if (xs >= 0) tx = a.tab[0][ix];
    if (ys >= 0) ty = a.tab[1][iy];

    R vuz_m = P::ld(a.f[PW_VUZ], o - PL);
    R vwz_m = P::ld(a.f[PW_VWZ], o - PL);
    R vux_c = P::ld(a.f[PW_VUX], o);
    R vuy_c = P::ld(a.f[PW_VUY], o);
    // physical form: cell-centred mu at (x, y), (x+1, y), (x, y+1) of this plane
    R mu_c = 0, mu_x = 0, mu_y = 0;
    if constexpr (!P::scaled) {
        mu_c = P::ldc(a.c[PC_MU], o);
        mu_x = P::ldc(a.c[PC_MU], o + 1);
        mu_y = P::ldc(a.c[PC_MU], o + W);
    }

    for (int iz = z0; iz < z1; ++iz, o += PL) {
        const R vux_n = P::ld(a.f[PW_VUX], o + PL), vuy_n = P::ld(a.f[PW_VUY], o + PL);
        const R vuz_c = P::ld(a.f[PW_VUZ], o), vwz_c = P::ld(a.f[PW_VWZ], o);
        const R vux_l = P::ld(a.f[PW_VUX], o - 1), vux_q = P::ld(a.f[PW_VUX], o + W);
        const R vuy_h = P::ld(a.f[PW_VUY], o - W), vuy_r = P::ld(a.f[PW_VUY], o + 1);
        const R vuz_r = P::ld(a.f[PW_VUZ], o + 1), vuz_q = P::ld(a.f[PW_VUZ], o + W);
        const R vwx_c = P::ld(a.f[PW_VWX], o), vwx_l = P::ld(a.f[PW_VWX], o - 1);
        const R vwy_c = P::ld(a.f[PW_VWY], o), vwy_h = P::ld(a.f[PW_VWY], o - W);

        // Float/main.cu:234-249
        R uxx = (vux_c - vux_l) / H, uyy = (vuy_c - vuy_h) / H, uzz = (vuz_c - vuz_m) / H;
        R wx = (vwx_c - vwx_l) / H, wy = (vwy_c - vwy_h) / H, wz = (vwz_c - vwz_m) / H;
        R uxy = (vux_q - vux_c) / H, uyx = (vuy_r - vuy_c) / H;
        R uxz = (vux_n - vux_c) / H, uzx = (vuz_r - vuz_c) / H;
        R uyz = (vuy_n - vuy_c) / H, uzy = (vuz_q - vuz_c) / H;

        if (xs >= 0) {
            const unsigned k = ((unsigned)iz * a.ny + iy) * XW + xs;
            uxx = cpml<P>(a.m[0][4], k, tx.cf, tx.ef, uxx);
            wx = cpml<P>(a.m[0][5], k, tx.cf, tx.ef, wx);
            uyx = cpml<P>(a.m[0][6], k, tx.ch, tx.eh, uyx);
            uzx = cpml<P>(a.m[0][7], k, tx.ch, tx.eh, uzx);
        }
        if (ys >= 0) {
            const unsigned k = ((unsigned)iz * YW + ys) * W + ix;
            uyy = cpml<P>(a.m[1][4], k, ty.cf, ty.ef, uyy);
            wy = cpml<P>(a.m[1][5], k, ty.cf, ty.ef, wy);
            uxy = cpml<P>(a.m[1][6], k, ty.ch, ty.eh, uxy);
            uzy = cpml<P>(a.m[1][7], k, ty.ch, ty.eh, uzy);
        }
        const int zs = slab_index(iz, a.nz, a.lo[2], a.hi[2]);
        if (zs >= 0) {
            const Tab4<R> tz = a.tab[2][iz];
            const unsigned k = ((unsigned)zs * a.ny + iy) * W + ix;
            uzz = cpml<P>(a.m[2][4], k, tz.cf, tz.ef, uzz);
            wz = cpml<P>(a.m[2][5], k, tz.cf, tz.ef, wz);
            uxz = cpml<P>(a.m[2][6], k, tz.ch, tz.eh, uxz);
            uyz = cpml<P>(a.m[2][7], k, tz.ch, tz.eh, uyz);
        }

        const R du = uxx + uyy + uzz, dw = wx + wy + wz;
        R p1, xx1, yy1, zz1, xy1, yz1, xz1;
        if constexpr (!P::scaled) {
            const R mu_n = P::ldc(a.c[PC_MU], o + PL), mu_xn = P::ldc(a.c[PC_MU], o + PL + 1),
                    mu_yn = P::ldc(a.c[PC_MU], o + PL + W), mu_xy = P::ldc(a.c[PC_MU], o + 1 + W);
            // Float/main.cu:480-491
            const R muxz = harm4(mu_c, mu_x, mu_n, mu_xn);
            const R muyz = harm4(mu_c, mu_y, mu_n, mu_yn);
            const R muxy = harm4(mu_c, mu_x, mu_y, mu_xy);
            const R C = P::ldc(a.c[PC_C], o), M = P::ldc(a.c[PC_M], o);
            const R HH = P::ldc(a.c[PC_HH], o), H2u = P::ldc(a.c[PC_H2U], o);
            // Float/main.cu:287-293
            p1 = P::ld(a.f[PW_P], o) - dt * (C * du + M * dw);
            xx1 = P::ld(a.f[PW_TXX], o) + dt * (H2u * (uyy + uzz) + HH * uxx + C * dw);
            yy1 = P::ld(a.f[PW_TYY], o) + dt * (H2u * (uxx + uzz) + HH * uyy + C * dw);
            zz1 = P::ld(a.f[PW_TZZ], o) + dt * (H2u * (uxx + uyy) + HH * uzz + C * dw);
            xy1 = P::ld(a.f[PW_TXY], o) + muxy * dt * (uxy + uyx);
            yz1 = P::ld(a.f[PW_TYZ], o) + muyz * dt * (uyz + uzy);
            xz1 = P::ld(a.f[PW_TXZ], o) + muxz * dt * (uxz + uzx);
            mu_c = mu_n; mu_x = mu_xn; mu_y = mu_yn;
        } else {
            // Half/main.cu:362-370
            const R C = P::ldc(a.c[SC_C], o), M = P::ldc(a.c[SC_M], o);
            const R HH = P::ldc(a.c[SC_HH], o), H2u = P::ldc(a.c[SC_H2U], o);
            p1 = P::ld(a.f[PW_P], o) - C * du - M * dw;
            xx1 = P::ld(a.f[PW_TXX], o) + H2u * (uyy + uzz) + HH * uxx + C * dw;
            yy1 = P::ld(a.f[PW_TYY], o) + H2u * (uxx + uzz) + HH * uyy + C * dw;
            zz1 = P::ld(a.f[PW_TZZ], o) + H2u * (uxx + uyy) + HH * uzz + C * dw;
            xy1 = P::ld(a.f[PW_TXY], o) + P::ldc(a.c[SC_MUXY], o) * (uxy + uyx);
            yz1 = P::ld(a.f[PW_TYZ], o) + P::ldc(a.c[SC_MUYZ], o) * (uyz + uzy);
            xz1 = P::ld(a.f[PW_TXZ], o) + P::ldc(a.c[SC_MUXZ], o) * (uxz + uzx);
        }
        P::st(a.f[PW_P], o, p1);
        P::st(a.f[PW_TXX], o, xx1); P::st(a.f[PW_TYY], o, yy1); P::st(a.f[PW_TZZ], o, zz1);
        P::st(a.f[PW_TXY], o, xy1); P::st(a.f[PW_TYZ], o, yz1); P::st(a.f[PW_TXZ], o, xz1);

        vuz_m = vuz_c; vwz_m = vwz_c; vux_c = vux_n; vuy_c = vuy_n;
    }
}

// --------------------------------------------------------------------------
// Side kernels
// --------------------------------------------------------------------------
// Source: Float/main.cu:45-59 (Half: codec RMW, Half/main.cu:135-151).  One
// thread per channel; the four channels are distinct arrays.
template <class P>
__global__ void k_source(const StepArgs<P> a, unsigned o, typename P::R w0, typename P::R w1, typename P::R w2,
                         typename P::R w3, typename P::R amp)
{
    using R = typename P::R;
    const int t = threadIdx.x;
    const int fid[4] = {PW_TXX, PW_TYY, PW_TZZ, PW_P};
    const R w[4] = {w0, w1, w2, w3};
    if (t < 4 && w[t] != (R)0) {
        const typename P::Field f = a.f[fid[t]];
        P::st(f, o, P::ld(f, o) + w[t] * amp);
    }
}

template <class P> __device__ __forceinline__ typename P::R ld_any(const StepArgs<P>& a, int fid, unsigned o)
{
    switch (fid) {
    case PW_VUX: return P::ld(a.f[PW_VUX], o);
    case PW_VUY: return P::ld(a.f[PW_VUY], o);
    case PW_VUZ: return P::ld(a.f[PW_VUZ], o);
    case PW_VWX: return P::ld(a.f[PW_VWX], o);
    case PW_VWY: return P::ld(a.f[PW_VWY], o);
    case PW_VWZ: return P::ld(a.f[PW_VWZ], o);
    case PW_TXX: return P::ld(a.f[PW_TXX], o);
    case PW_TYY: return P::ld(a.f[PW_TYY], o);
    case PW_TZZ: return P::ld(a.f[PW_TZZ], o);
    case PW_TXY: return P::ld(a.f[PW_TXY], o);
    case PW_TXZ: return P::ld(a.f[PW_TXZ], o);
    case PW_TYZ: return P::ld(a.f[PW_TYZ], o);
    default: return P::ld(a.f[PW_P], o);
    }
}

// Receivers: seis[r][it] = scale * field(rfield[r])[roff[r]]  (Float/main.cu:1085-1093)
template <class P>
__global__ void k_gather(const StepArgs<P> a, const int* __restrict__ rfield, const unsigned* __restrict__ roff,
                         int nrec, float* __restrict__ seis, int nt, int it, float s_stress, float s_vel)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int fid = rfield[r];
    const float v = (float)ld_any<P>(a, fid, roff[r]);
    seis[(size_t)r * nt + it] = (fid <= PW_VWZ ? s_vel : s_stress) * v;
}

// y-slice of one field: dst[iz][ix] (Float/main.cu:1094-1108, wfile3d slice)
template <class P>
__global__ void k_snapshot(const StepArgs<P> a, int fid, int y, float* __restrict__ dst, float scale)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (ix >= a.nx) return;
    const unsigned o = (unsigned)ix + (unsigned)y * a.pitch + (unsigned)iz * a.plane;
    dst[(size_t)iz * a.nx + ix] = scale * (float)ld_any<P>(a, fid, o);
}

// Full field decoded to dense FP32 (x fastest, no pitch padding).
template <class P>
__global__ void k_decode(const StepArgs<P> a, int fid, float* __restrict__ dst, float scale)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= a.nx) return;
    const unsigned o = (unsigned)ix + (unsigned)iy * a.pitch + (unsigned)iz * a.plane;
    dst[(size_t)ix + (size_t)iy * a.nx + (size_t)iz * a.nx * a.ny] = scale * (float)ld_any<P>(a, fid, o);
}

// Dynamic rescale of every stored wavefield and memory variable
// (Half/main.cu:102-114, 1326-1364) in one launch: blockIdx.y selects the array.
template <class P>
__global__ void k_rescale(const StepArgs<P> a, unsigned long long ncell, unsigned long long nmem0,
                          unsigned long long nmem1, unsigned long long nmem2, float ratio)
{
    const int arr = blockIdx.y;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (arr < PW_NFIELDS) {
        const typename P::Field f = a.f[arr];
        for (; i < ncell; i += stride) P::st(f, (unsigned)i, P::ld(f, (unsigned)i) * ratio);
    } else {
        const int k = arr - PW_NFIELDS;
        const int ax = k / kMemPerAxis, t = k % kMemPerAxis;
        const unsigned long long n = ax == 0 ? nmem0 : (ax == 1 ? nmem1 : nmem2);
        typename P::Mem* m = a.m[ax][t];
        if (!m) return;
        for (; i < n; i += stride) P::stm(m, (unsigned)i, P::ldm(m, (unsigned)i) * ratio);
    }
}

__global__ void k_zero(uint4* __restrict__ p, unsigned long long n16)
{
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// --------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------
// Restricts a step launch to planes [zb, ze).
template <class P> static StepArgs<P> range_args(const DevBuffers& b, int zb, int ze)
{
    StepArgs<P> a = make_args<P>(b);
    a.zb = zb;
    a.ze = ze;
    a.nblocks = a.gx * a.gy * ((ze - zb + b.zc - 1) / b.zc);
    return a;
}

template <class F> static hipError_t dispatch(int prec, F&& f)
{
    switch (prec) {
    case PW_FP32: return f(PolF32{});
    case PW_FP64: return f(PolF64{});
    case PW_COMP16: return f(PolC16{});
    case PW_BF16: return f(PolB16{});
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_velocity(const DevBuffers& b, hipStream_t st, int zb, int ze)
{
    if (ze <= zb) return hipSuccess;
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        const StepArgs<P> a = range_args<P>(b, zb, ze);
        k_velocity<P><<<a.nblocks, dim3(64, 4), 0, st>>>(a);
        return hipGetLastError();
    });
}

hipError_t launch_stress(const DevBuffers& b, hipStream_t st, int zb, int ze)
{
    if (ze <= zb) return hipSuccess;
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        const StepArgs<P> a = range_args<P>(b, zb, ze);
        k_stress<P><<<a.nblocks, dim3(64, 4), 0, st>>>(a);
        return hipGetLastError();
    });
}

hipError_t launch_source(const DevBuffers& b, hipStream_t st, unsigned off, const double w[4], double amp)
{
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        using R = typename P::R;
        k_source<P><<<1, 64, 0, st>>>(make_args<P>(b), off, (R)w[0], (R)w[1], (R)w[2], (R)w[3], (R)amp);
        return hipGetLastError();
    });
}

hipError_t launch_gather(const DevBuffers& b, hipStream_t st, const int* rfield, const unsigned* roff, int nrec,
                         float* seis, int nt, int it, float s_stress, float s_vel)
{
    if (nrec <= 0) return hipSuccess;
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        k_gather<P><<<(nrec + 255) / 256, 256, 0, st>>>(make_args<P>(b), rfield, roff, nrec, seis, nt, it, s_stress,
                                                        s_vel);
        return hipGetLastError();
    });
}

hipError_t launch_snapshot(const DevBuffers& b, hipStream_t st, int field, int y, float* dst, float scale)
{
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        k_snapshot<P><<<dim3((unsigned)((b.g.nx + 255) / 256), (unsigned)b.g.nz), 256, 0, st>>>(make_args<P>(b), field,
                                                                                               y, dst, scale);
        return hipGetLastError();
    });
}

hipError_t launch_decode(const DevBuffers& b, hipStream_t st, int field, float* dst, float scale)
{
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        k_decode<P><<<dim3((unsigned)((b.g.nx + 255) / 256), (unsigned)b.g.ny, (unsigned)b.g.nz), 256, 0, st>>>(
            make_args<P>(b), field, dst, scale);
        return hipGetLastError();
    });
}

hipError_t launch_rescale(const DevBuffers& b, hipStream_t st, float ratio)
{
    if (!is_scaled(b.prec)) return hipSuccess;
    const unsigned long long ncell = (unsigned long long)b.g.cells;
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        k_rescale<P><<<dim3(2048, PW_NFIELDS + 3 * kMemPerAxis), 256, 0, st>>>(
            make_args<P>(b), ncell, (unsigned long long)b.g.slab_cells[0], (unsigned long long)b.g.slab_cells[1],
            (unsigned long long)b.g.slab_cells[2], ratio);
        return hipGetLastError();
    });
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t st)
{
    if (bytes == 0) return hipSuccess;
    if (bytes % 16 != 0 || ((uintptr_t)p) % 16 != 0) return hipErrorInvalidValue;
    const unsigned long long n16 = bytes / 16;
    unsigned blocks = (unsigned)((n16 + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, st, (uint4*)p, n16);
    return hipGetLastError();
}

}  // namespace pw
