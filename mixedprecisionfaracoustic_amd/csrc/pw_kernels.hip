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

#include <algorithm>
#include <cstdint>
#include <vector>

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
// The same unit from the FP16 bit pattern: biased FP16 exponent e5 (0 for
// zero/subnormal) gives 2^(max(e5, 1) - 32), i.e. float bits (max(e5,1)+95) << 23
// with the 127 bias: e5 = 1 -> 2^-32, the subnormal value.
__device__ __forceinline__ float c16_unit_bits(unsigned short hb)
{
    const unsigned e5 = (hb >> 10) & 31u;
    return __uint_as_float(((e5 > 1u ? e5 : 1u) + 94u) << 23);
}
__device__ __forceinline__ float c16_inv_unit_bits(unsigned short hb)
{
    const unsigned e5 = (hb >> 10) & 31u;
    return __uint_as_float((160u - (e5 > 1u ? e5 : 1u)) << 23);
}
__device__ __forceinline__ float c16_decode(f16 hi, int8_t lo)
{
    return (float)hi + (float)lo * c16_unit_bits(__builtin_bit_cast(unsigned short, hi));
}
__device__ __forceinline__ void c16_encode(float v, f16& hi, int8_t& lo)
{
    const f16 h = (f16)v;
    const float hf = (float)h;
    // exact: the unit is a power of two
    const float scaled = (v - hf) * c16_inv_unit_bits(__builtin_bit_cast(unsigned short, h));
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
    static __device__ __forceinline__ R rt(R v) { return v; }  // value as stored
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
    static __device__ __forceinline__ R rt(R v) { return v; }
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
    static __device__ __forceinline__ R rt(R v)
    {
        f16 h;
        int8_t l;
        c16_encode(v, h, l);
        return c16_decode(h, l);
    }
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
    static __device__ __forceinline__ R rt(R v) { return (float)(bf16)v; }
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
__device__ __forceinline__ unsigned xcd_remap(unsigned L, unsigned nblocks, int xcd_map)
{
    if (!xcd_map) return L;
    const unsigned q = nblocks / 8u, r = nblocks % 8u;
    const unsigned k = L % 8u, j = L / 8u;
    return k * q + (k < r ? k : r) + j;
}
__device__ __forceinline__ Tile tile_of_index(unsigned t, int gx, int gy)
{
    Tile tl;
    tl.bx = (int)(t % (unsigned)gx);
    const unsigned row = t / (unsigned)gx;
    tl.by = (int)(row % (unsigned)gy);
    tl.bz = (int)(row / (unsigned)gy);
    return tl;
}
__device__ __forceinline__ Tile tile_of_block(int gx, int gy, int nblocks, int xcd_map)
{
    return tile_of_index(xcd_remap(blockIdx.x, (unsigned)nblocks, xcd_map), gx, gy);
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

// Staggered difference divided by the grid step.  The FP32/FP64 forms divide
// as the reference does; the FP16-stored forms multiply by 1/h (the rounding
// difference is far below FP16 storage resolution).
template <class P> struct Deriv {
    typename P::R h, ih;
    __device__ __forceinline__ explicit Deriv(typename P::R h_) : h(h_), ih((typename P::R)1 / h_) {}
    __device__ __forceinline__ typename P::R operator()(typename P::R d) const
    {
        if constexpr (P::scaled) return d * ih;
        else return d / h;
    }
};

// --------------------------------------------------------------------------
// Velocity kernel: Float/main.cu:64-192 (physical), Half/main.cu:156-268 (scaled)
// --------------------------------------------------------------------------
template <class P>
__device__ __forceinline__ void velocity_tile(const StepArgs<P>& a, const Tile tl)
{
    using R = typename P::R;
    const int ix = tl.bx * 64 + threadIdx.x;
    const int iy = tl.by * 4 + threadIdx.y;
    if (ix < 1 || ix > a.nx - 2 || iy < 1 || iy > a.ny - 2) return;
    const int z0 = a.zb + tl.bz * a.zc;
    const int z1 = min(z0 + a.zc, a.ze);
    if (z0 >= z1) return;
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)z0 * PL;
    const Deriv<P> D(a.h);
    const R dt = a.dt;

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
        R x1 = D(txx_r - txx_c), x2 = D(txy_c - txy_h), x3 = D(txz_c - txz_m), s1 = D(p_r - p_c);
        R y1 = D(tyy_q - tyy_c), y2 = D(txy_c - txy_l), y3 = D(tyz_c - tyz_m), s2 = D(p_q - p_c);
        R z1 = D(tzz_n - tzz_c), z2 = D(txz_c - txz_l), z3 = D(tyz_c - tyz_h), s3 = D(p_n - p_c);

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
// Stress kernel: Float/main.cu:194-295 (physical), Half/main.cu:270-371 (scaled)
// --------------------------------------------------------------------------
template <class P>
__device__ __forceinline__ void stress_tile(const StepArgs<P>& a, const Tile tl)
{
    using R = typename P::R;
    const int ix = tl.bx * 64 + threadIdx.x;
    const int iy = tl.by * 4 + threadIdx.y;
    if (ix < 1 || ix > a.nx - 2 || iy < 1 || iy > a.ny - 2) return;
    const int z0 = a.zb + tl.bz * a.zc;
    const int z1 = min(z0 + a.zc, a.ze);
    if (z0 >= z1) return;
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)z0 * PL;
    const Deriv<P> D(a.h);
    const R dt = a.dt;

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
        R uxx = D(vux_c - vux_l), uyy = D(vuy_c - vuy_h), uzz = D(vuz_c - vuz_m);
        R wx = D(vwx_c - vwx_l), wy = D(vwy_c - vwy_h), wz = D(vwz_c - vwz_m);
        R uxy = D(vux_q - vux_c), uyx = D(vuy_r - vuy_c);
        R uxz = D(vux_n - vux_c), uzx = D(vuz_r - vuz_c);
        R uyz = D(vuy_n - vuy_c), uzy = D(vuz_q - vuz_c);

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

template <class P> __global__ __launch_bounds__(256) void k_velocity(const StepArgs<P> a)
{
    velocity_tile<P>(a, tile_of_block(a.gx, a.gy, a.nblocks, a.xcd_map));
}
template <class P> __global__ __launch_bounds__(256) void k_stress(const StepArgs<P> a)
{
    stress_tile<P>(a, tile_of_block(a.gx, a.gy, a.nblocks, a.xcd_map));
}

// --------------------------------------------------------------------------
// Fused step kernel: velocity and stress update of one time step in one launch.
//
// The state is double-buffered (read time n from s.f, write time n+1 to g), so
// each wavefield is read once and written once per step instead of read twice
// (39 -> 26 touches).  A block owns a 64 x 4 column tile and marches up its z
// chunk: at plane k it computes the new velocities of plane k+1 (tile cells: all
// three axes, stored; the one-cell ring around the tile: only the axis
// components the stress stencil reads, not stored), then the stress update of
// plane k from registers (own column, planes k-1..k+1) and LDS (x/y neighbours).
// Ring cells belong to neighbouring tiles and are recomputed here from the same
// time-n inputs, so they come out bit-identical to the owner's values.  The CPML
// memory variables are double-buffered likewise (read m, write mo).
// --------------------------------------------------------------------------
template <class P> struct FusedArgs {
    StepArgs<P> s;                        // time n (read): fields, coefficients, memory variables, geometry
    typename P::Field g[PW_NFIELDS];      // time n+1 (write)
    typename P::Mem* mo[3][kMemPerAxis];  // memory variables written
};

template <class P>
__device__ __forceinline__ typename P::R cpml2(const typename P::Mem* in, typename P::Mem* out, unsigned k,
                                               typename P::R c, typename P::R e, typename P::R der, bool write)
{
    using R = typename P::R;
    const R psi = P::ldm(in, k) - c * der;
    if (write) P::stm(out, k, e * (psi - c * der));
    return der + psi;
}

// Velocity update of one axis (AX 0: vux/vwx, 1: vuy/vwy, 2: vuz/vwz) at an
// updated cell; own: this block owns the cell and writes its memory variables.
template <class P, int AX>
__device__ __forceinline__ void v_axis(const FusedArgs<P>& A, int ix, int iy, int iz, bool own, typename P::R& vu1,
                                       typename P::R& vw1)
{
    using R = typename P::R;
    const StepArgs<P>& a = A.s;
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    const unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)iz * PL;
    const Deriv<P> D(a.h);
    const R dt = a.dt;
    const unsigned st = AX == 0 ? 1u : (AX == 1 ? W : PL);
    // d1: normal stress along AX (forward), d2/d3: shear stresses (backward),
    // ds: pressure along AX (forward)  (Float/main.cu:115-128)
    R d1, d2, d3;
    const R ds = D(P::ld(a.f[PW_P], o + st) - P::ld(a.f[PW_P], o));
    if (AX == 0) {
        d1 = D(P::ld(a.f[PW_TXX], o + 1) - P::ld(a.f[PW_TXX], o));
        d2 = D(P::ld(a.f[PW_TXY], o) - P::ld(a.f[PW_TXY], o - W));
        d3 = D(P::ld(a.f[PW_TXZ], o) - P::ld(a.f[PW_TXZ], o - PL));
    } else if (AX == 1) {
        d1 = D(P::ld(a.f[PW_TYY], o + W) - P::ld(a.f[PW_TYY], o));
        d2 = D(P::ld(a.f[PW_TXY], o) - P::ld(a.f[PW_TXY], o - 1));
        d3 = D(P::ld(a.f[PW_TYZ], o) - P::ld(a.f[PW_TYZ], o - PL));
    } else {
        d1 = D(P::ld(a.f[PW_TZZ], o + PL) - P::ld(a.f[PW_TZZ], o));
        d2 = D(P::ld(a.f[PW_TXZ], o) - P::ld(a.f[PW_TXZ], o - 1));
        d3 = D(P::ld(a.f[PW_TYZ], o) - P::ld(a.f[PW_TYZ], o - W));
    }
    R dsv = ds;
    // CPML: d1 and ds are damped along AX (half-point profile, terms 0 and 1);
    // d2 along x (AX 1, 2) or y (AX 0); d3 along z (AX 0, 1) or y (AX 2).
    const int xs = slab_index(ix, a.nx, a.lo[0], a.hi[0]);
    const int ys = slab_index(iy, a.ny, a.lo[1], a.hi[1]);
    const int zs = slab_index(iz, a.nz, a.lo[2], a.hi[2]);
    const unsigned kx = xs >= 0 ? ((unsigned)iz * a.ny + iy) * (a.lo[0] + a.hi[0]) + xs : 0u;
    const unsigned ky = ys >= 0 ? ((unsigned)iz * (a.lo[1] + a.hi[1]) + ys) * W + ix : 0u;
    const unsigned kz = zs >= 0 ? ((unsigned)zs * a.ny + iy) * W + ix : 0u;
    const int sa = AX == 0 ? xs : (AX == 1 ? ys : zs);
    if (sa >= 0) {
        const Tab4<R> t = a.tab[AX][AX == 0 ? ix : (AX == 1 ? iy : iz)];
        const unsigned k = AX == 0 ? kx : (AX == 1 ? ky : kz);
        d1 = cpml2<P>(a.m[AX][0], A.mo[AX][0], k, t.ch, t.eh, d1, own);
        dsv = cpml2<P>(a.m[AX][1], A.mo[AX][1], k, t.ch, t.eh, dsv, own);
    }
    if (AX == 0) {
        if (ys >= 0) { const Tab4<R> t = a.tab[1][iy]; d2 = cpml2<P>(a.m[1][2], A.mo[1][2], ky, t.cf, t.ef, d2, own); }
        if (zs >= 0) { const Tab4<R> t = a.tab[2][iz]; d3 = cpml2<P>(a.m[2][2], A.mo[2][2], kz, t.cf, t.ef, d3, own); }
    } else if (AX == 1) {
        if (xs >= 0) { const Tab4<R> t = a.tab[0][ix]; d2 = cpml2<P>(a.m[0][2], A.mo[0][2], kx, t.cf, t.ef, d2, own); }
        if (zs >= 0) { const Tab4<R> t = a.tab[2][iz]; d3 = cpml2<P>(a.m[2][3], A.mo[2][3], kz, t.cf, t.ef, d3, own); }
    } else {
        if (xs >= 0) { const Tab4<R> t = a.tab[0][ix]; d2 = cpml2<P>(a.m[0][3], A.mo[0][3], kx, t.cf, t.ef, d2, own); }
        if (ys >= 0) { const Tab4<R> t = a.tab[1][iy]; d3 = cpml2<P>(a.m[1][3], A.mo[1][3], ky, t.cf, t.ef, d3, own); }
    }
    const R S = d1 + d2 + d3;
    const R vw0 = P::ld(a.f[PW_VWX + AX], o), vu0 = P::ld(a.f[PW_VUX + AX], o);
    if constexpr (!P::scaled) {
        const R ri = inv_mean(P::ldc(a.c[PC_RHO], o), P::ldc(a.c[PC_RHO], o + st));
        const R c1 = half_sum(P::ldc(a.c[PC_C1], o), P::ldc(a.c[PC_C1], o + st));
        const R c2 = half_sum(P::ldc(a.c[PC_C2], o), P::ldc(a.c[PC_C2], o + st));
        const R rf = P::ldc(a.c[PC_RHOF], o);
        R cc1, cc2;
        cn_pair(c1, c2, rf, ri, dt, cc1, cc2);
        vw1 = c1 == (R)0 ? (R)0 : (vw0 * cc2 - dt * (rf * ri * S + dsv)) / cc1;
        vu1 = vu0 + dt * ri * S - rf * ri * (vw1 - vw0);
    } else {
        const R aw = P::ldc(a.c[SC_AWX + AX], o), bp = P::ldc(a.c[SC_BPX + AX], o), bt = P::ldc(a.c[SC_BTX + AX], o);
        vw1 = (aw == 0.0f && bp == 0.0f && bt == 0.0f) ? 0.0f : aw * vw0 - bt * S - bp * dsv;
        vu1 = vu0 + P::ldc(a.c[SC_RITX + AX], o) * S - P::ldc(a.c[SC_FRX + AX], o) * (vw1 - vw0);
    }
}

template <class P>
__global__ __launch_bounds__(256) void k_fused(const FusedArgs<P> A)
{
    using R = typename P::R;
    const StepArgs<P>& a = A.s;
    // velocity of the current plane for the tile and its ring: vux, vuy, vuz, vwx, vwy
    __shared__ R lds[2][5][6 * 66];
    const Tile tl = tile_of_block(a.gx, a.gy, a.nblocks, a.xcd_map);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
    const int ix = tl.bx * 64 + tx, iy = tl.by * 4 + ty;
    const int z0 = a.zb + tl.bz * a.zc;
    const int z1 = min(z0 + a.zc, a.ze);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    // no early return below: every thread takes part in the barriers
    const bool in_xy = ix <= nx - 1 && iy <= ny - 1;
    const bool up_xy = ix >= 1 && ix <= nx - 2 && iy >= 1 && iy <= ny - 2;
    const int lt = (ty + 1) * 66 + (tx + 1);
    // ring cell of this thread: 0 = row y-1 (vuy, vwy), 1 = row y+4 (vux, vuz),
    // 2 = column x-1 (vux, vwx), 3 = column x+64 (vuy, vuz)
    int hk = -1, hx = 0, hy = 0;
    if (tid < 64) { hk = 0; hx = tl.bx * 64 + tid; hy = tl.by * 4 - 1; }
    else if (tid < 128) { hk = 1; hx = tl.bx * 64 + tid - 64; hy = tl.by * 4 + 4; }
    else if (tid < 132) { hk = 2; hx = tl.bx * 64 - 1; hy = tl.by * 4 + tid - 128; }
    else if (tid < 136) { hk = 3; hx = tl.bx * 64 + 64; hy = tl.by * 4 + tid - 132; }
    const bool h_in = hk >= 0 && hx >= 0 && hx <= nx - 1 && hy >= 0 && hy <= ny - 1;
    const bool h_up = hk >= 0 && hx >= 1 && hx <= nx - 2 && hy >= 1 && hy <= ny - 2;
    const int lh = hk >= 0 ? (hy - tl.by * 4 + 1) * 66 + (hx - tl.bx * 64 + 1) : 0;

    // new velocities of the tile cell at plane z for the axes in mask (1 x, 2 y, 4 z);
    // cells that are never updated keep their stored value
    auto v_tile = [&](int z, bool own, int mask, R* u) {
        if (z0 >= z1) return;
        if (up_xy && z >= 1 && z <= nz - 2) {
            if (mask & 1) v_axis<P, 0>(A, ix, iy, z, own, u[0], u[3]);
            if (mask & 2) v_axis<P, 1>(A, ix, iy, z, own, u[1], u[4]);
            if (mask & 4) v_axis<P, 2>(A, ix, iy, z, own, u[2], u[5]);
            // the stress update reads the velocities as stored (Half: encoded)
            for (int c = 0; c < 6; ++c) u[c] = P::rt(u[c]);
        } else if (in_xy && z >= 0 && z <= nz - 1) {
            const unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)z * PL;
            for (int c = 0; c < 3; ++c)
                if (mask & (1 << c)) { u[c] = P::ld(a.f[PW_VUX + c], o); u[3 + c] = P::ld(a.f[PW_VWX + c], o); }
        }
    };
    auto store_tile = [&](int z, const R* u) {
        if (z0 >= z1 || !up_xy || z < 1 || z > nz - 2) return;
        const unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)z * PL;
        for (int c = 0; c < 6; ++c) P::st(A.g[PW_VUX + c], o, u[c]);
    };
    auto put_tile = [&](int buf, const R* u) {
        lds[buf][0][lt] = u[0]; lds[buf][1][lt] = u[1]; lds[buf][2][lt] = u[2];
        lds[buf][3][lt] = u[3]; lds[buf][4][lt] = u[4];
    };
    auto ring = [&](int z, int buf) {
        if (hk < 0 || z0 >= z1) return;
        R vu[3] = {0, 0, 0}, vw[3] = {0, 0, 0};
        const int need = hk == 0 ? 2 : (hk == 1 ? 5 : (hk == 2 ? 1 : 6));
        if (h_up && z >= 1 && z <= nz - 2) {
            if (need & 1) v_axis<P, 0>(A, hx, hy, z, false, vu[0], vw[0]);
            if (need & 2) v_axis<P, 1>(A, hx, hy, z, false, vu[1], vw[1]);
            if (need & 4) v_axis<P, 2>(A, hx, hy, z, false, vu[2], vw[2]);
            for (int c = 0; c < 3; ++c) { vu[c] = P::rt(vu[c]); vw[c] = P::rt(vw[c]); }
        } else if (h_in && z >= 0 && z <= nz - 1) {
            const unsigned o = (unsigned)hx + (unsigned)hy * W + (unsigned)z * PL;
            for (int c = 0; c < 3; ++c)
                if (need & (1 << c)) { vu[c] = P::ld(a.f[PW_VUX + c], o); vw[c] = P::ld(a.f[PW_VWX + c], o); }
        }
        if (hk == 0) { lds[buf][1][lh] = vu[1]; lds[buf][4][lh] = vw[1]; }
        else if (hk == 1) { lds[buf][0][lh] = vu[0]; lds[buf][2][lh] = vu[2]; }
        else if (hk == 2) { lds[buf][0][lh] = vu[0]; lds[buf][3][lh] = vw[0]; }
        else { lds[buf][1][lh] = vu[1]; lds[buf][2][lh] = vu[2]; }
    };

    R prev[6] = {0, 0, 0, 0, 0, 0}, cur[6] = {0, 0, 0, 0, 0, 0}, nxt[6] = {0, 0, 0, 0, 0, 0};
    v_tile(z0 - 1, false, 4, prev);  // vuz, vwz below the chunk (owned by the block below)
    v_tile(z0, true, 7, cur);
    store_tile(z0, cur);
    put_tile(0, cur);
    ring(z0, 0);
    __syncthreads();

    for (int k = z0; k < z1; ++k) {
        const int cb = (k - z0) & 1, nb = cb ^ 1;
        if (k + 1 < z1) {
            v_tile(k + 1, true, 7, nxt);
            store_tile(k + 1, nxt);
            put_tile(nb, nxt);
            ring(k + 1, nb);
        } else {
            v_tile(k + 1, false, 3, nxt);  // vux, vuy above the chunk (owned by the block above)
        }
        // stress update of plane k (Float/main.cu:234-293 / Half/main.cu:297-370)
        if (up_xy) {
            const unsigned o = (unsigned)ix + (unsigned)iy * W + (unsigned)k * PL;
            const Deriv<P> D(a.h);
    const R dt = a.dt;
            const R vux_l = lds[cb][0][lt - 1], vux_q = lds[cb][0][lt + 66];
            const R vuy_h = lds[cb][1][lt - 66], vuy_r = lds[cb][1][lt + 1];
            const R vuz_r = lds[cb][2][lt + 1], vuz_q = lds[cb][2][lt + 66];
            const R vwx_l = lds[cb][3][lt - 1], vwy_h = lds[cb][4][lt - 66];
            R uxx = D(cur[0] - vux_l), uyy = D(cur[1] - vuy_h), uzz = D(cur[2] - prev[2]);
            R wx = D(cur[3] - vwx_l), wy = D(cur[4] - vwy_h), wz = D(cur[5] - prev[5]);
            R uxy = D(vux_q - cur[0]), uyx = D(vuy_r - cur[1]);
            R uxz = D(nxt[0] - cur[0]), uzx = D(vuz_r - cur[2]);
            R uyz = D(nxt[1] - cur[1]), uzy = D(vuz_q - cur[2]);
            const int xs = slab_index(ix, nx, a.lo[0], a.hi[0]);
            const int ys = slab_index(iy, ny, a.lo[1], a.hi[1]);
            const int zs = slab_index(k, nz, a.lo[2], a.hi[2]);
            if (xs >= 0) {
                const Tab4<R> t = a.tab[0][ix];
                const unsigned q = ((unsigned)k * ny + iy) * (a.lo[0] + a.hi[0]) + xs;
                uxx = cpml2<P>(a.m[0][4], A.mo[0][4], q, t.cf, t.ef, uxx, true);
                wx = cpml2<P>(a.m[0][5], A.mo[0][5], q, t.cf, t.ef, wx, true);
                uyx = cpml2<P>(a.m[0][6], A.mo[0][6], q, t.ch, t.eh, uyx, true);
                uzx = cpml2<P>(a.m[0][7], A.mo[0][7], q, t.ch, t.eh, uzx, true);
            }
            if (ys >= 0) {
                const Tab4<R> t = a.tab[1][iy];
                const unsigned q = ((unsigned)k * (a.lo[1] + a.hi[1]) + ys) * W + ix;
                uyy = cpml2<P>(a.m[1][4], A.mo[1][4], q, t.cf, t.ef, uyy, true);
                wy = cpml2<P>(a.m[1][5], A.mo[1][5], q, t.cf, t.ef, wy, true);
                uxy = cpml2<P>(a.m[1][6], A.mo[1][6], q, t.ch, t.eh, uxy, true);
                uzy = cpml2<P>(a.m[1][7], A.mo[1][7], q, t.ch, t.eh, uzy, true);
            }
            if (zs >= 0) {
                const Tab4<R> t = a.tab[2][k];
                const unsigned q = ((unsigned)zs * ny + iy) * W + ix;
                uzz = cpml2<P>(a.m[2][4], A.mo[2][4], q, t.cf, t.ef, uzz, true);
                wz = cpml2<P>(a.m[2][5], A.mo[2][5], q, t.cf, t.ef, wz, true);
                uxz = cpml2<P>(a.m[2][6], A.mo[2][6], q, t.ch, t.eh, uxz, true);
                uyz = cpml2<P>(a.m[2][7], A.mo[2][7], q, t.ch, t.eh, uyz, true);
            }
            const R du = uxx + uyy + uzz, dw = wx + wy + wz;
            R p1, xx1, yy1, zz1, xy1, yz1, xz1;
            if constexpr (!P::scaled) {
                const R mu_c = P::ldc(a.c[PC_MU], o), mu_x = P::ldc(a.c[PC_MU], o + 1),
                        mu_y = P::ldc(a.c[PC_MU], o + W), mu_n = P::ldc(a.c[PC_MU], o + PL),
                        mu_xn = P::ldc(a.c[PC_MU], o + PL + 1), mu_yn = P::ldc(a.c[PC_MU], o + PL + W),
                        mu_xy = P::ldc(a.c[PC_MU], o + 1 + W);
                const R muxz = harm4(mu_c, mu_x, mu_n, mu_xn);
                const R muyz = harm4(mu_c, mu_y, mu_n, mu_yn);
                const R muxy = harm4(mu_c, mu_x, mu_y, mu_xy);
                const R C = P::ldc(a.c[PC_C], o), M = P::ldc(a.c[PC_M], o);
                const R HH = P::ldc(a.c[PC_HH], o), H2u = P::ldc(a.c[PC_H2U], o);
                p1 = P::ld(a.f[PW_P], o) - dt * (C * du + M * dw);
                xx1 = P::ld(a.f[PW_TXX], o) + dt * (H2u * (uyy + uzz) + HH * uxx + C * dw);
                yy1 = P::ld(a.f[PW_TYY], o) + dt * (H2u * (uxx + uzz) + HH * uyy + C * dw);
                zz1 = P::ld(a.f[PW_TZZ], o) + dt * (H2u * (uxx + uyy) + HH * uzz + C * dw);
                xy1 = P::ld(a.f[PW_TXY], o) + muxy * dt * (uxy + uyx);
                yz1 = P::ld(a.f[PW_TYZ], o) + muyz * dt * (uyz + uzy);
                xz1 = P::ld(a.f[PW_TXZ], o) + muxz * dt * (uxz + uzx);
            } else {
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
            P::st(A.g[PW_P], o, p1);
            P::st(A.g[PW_TXX], o, xx1); P::st(A.g[PW_TYY], o, yy1); P::st(A.g[PW_TZZ], o, zz1);
            P::st(A.g[PW_TXY], o, xy1); P::st(A.g[PW_TYZ], o, yz1); P::st(A.g[PW_TXZ], o, xz1);
        }
        __syncthreads();
        for (int c = 0; c < 6; ++c) { prev[c] = cur[c]; cur[c] = nxt[c]; }
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
// Multi-shot step: every shot of one precision advanced by one launch per
// kernel.  A descriptor table holds each shot's arguments; the launch's blocks
// are the concatenation of the shots' tile grids (block prefix bstart), dealt to
// the XCDs as one grid.  Each shot's current step is it[s] + k (k = step within
// the round), read on the device, so a round of launches can be captured once
// and replayed; a shot past its last step skips the work.  The per-cell code is
// the single-shot code (velocity_tile / stress_tile): results are bit-identical
// to a shot run alone.
// --------------------------------------------------------------------------
template <class P> struct ShotDesc {
    StepArgs<P> a;
    const typename P::R* amp;
    const float* oscale;
    const float* ratio;
    const int* rfield;
    const unsigned* roff;
    float* seis;
    int nrec, nt;
    unsigned src_off;
    int has_src;
    typename P::R w[4];
    unsigned long long ncell, nmem[3];
};

struct BatchArgs {
    const void* descs;
    const unsigned* bstart;  // [n+1] block prefix (without the XCD map)
    const unsigned* xstart;  // [8][n+1] per-XCD block prefix (with the XCD map)
    const unsigned* rstart;
    int* it;
    int n;
    unsigned nblocks, nrec;
    int xcd_map;
};

// largest s with start[s] <= t (start ascending, start[0] = 0)
__device__ __forceinline__ int find_shot(const unsigned* __restrict__ start, int n, unsigned t)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Block -> (shot, tile index within the shot).  With the XCD map, XCD x (the
// blocks L with L % 8 == x) gets the x-th eighth of EVERY shot's tiles, each
// eighth contiguous in the shot's x-y-z tile order: the XCDs share the work
// evenly whatever the mix of grid sizes (blocks of large shots march more
// planes), and each XCD's L2 still sees neighbouring tiles of one shot in turn.
// Returns false for the spare blocks of the XCDs with fewer tiles.
template <class P> __device__ __forceinline__ bool batch_block(const BatchArgs& c, int& s, unsigned& local)
{
    const ShotDesc<P>* D = (const ShotDesc<P>*)c.descs;
    if (!c.xcd_map) {
        const unsigned t = blockIdx.x;
        s = find_shot(c.bstart, c.n, t);
        local = t - c.bstart[s];
        return true;
    }
    const unsigned x = blockIdx.x % 8u, j = blockIdx.x / 8u;
    const unsigned* xs = c.xstart + x * (unsigned)(c.n + 1);
    if (j >= xs[c.n]) return false;
    s = find_shot(xs, c.n, j);
    const unsigned nb = (unsigned)D[s].a.nblocks, q = nb / 8u, r = nb % 8u;
    local = x * q + (x < r ? x : r) + (j - xs[s]);
    return true;
}

template <class P> __global__ __launch_bounds__(256) void k_velocity_batch(const BatchArgs c, int k)
{
    int s;
    unsigned t;
    if (!batch_block<P>(c, s, t)) return;
    const ShotDesc<P>* d = (const ShotDesc<P>*)c.descs + s;
    if (c.it[s] + k >= d->nt) return;
    const StepArgs<P> a = d->a;
    velocity_tile<P>(a, tile_of_index(t, a.gx, a.gy));
}

template <class P> __global__ __launch_bounds__(256) void k_stress_batch(const BatchArgs c, int k)
{
    int s;
    unsigned t;
    if (!batch_block<P>(c, s, t)) return;
    const ShotDesc<P>* d = (const ShotDesc<P>*)c.descs + s;
    if (c.it[s] + k >= d->nt) return;
    const StepArgs<P> a = d->a;
    stress_tile<P>(a, tile_of_index(t, a.gx, a.gy));
}

// one thread per (shot, source channel); the arithmetic of k_source
template <class P> __global__ void k_source_batch(const BatchArgs c, int k)
{
    using R = typename P::R;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = g >> 2, ch = g & 3;
    if (s >= c.n) return;
    const ShotDesc<P>& d = ((const ShotDesc<P>*)c.descs)[s];
    const int it = c.it[s] + k;
    if (it >= d.nt || !d.has_src || d.w[ch] == (R)0) return;
    const int fid[4] = {PW_TXX, PW_TYY, PW_TZZ, PW_P};
    const typename P::Field f = d.a.f[fid[ch]];
    const R amp = d.amp[it];
    P::st(f, d.src_off, P::ld(f, d.src_off) + d.w[ch] * amp);
}

// Multiplies array arr (0..12 wavefields, then the memory variables) of one
// shot by ratio over the indices i0, i0 + stride, ...  (k_rescale's arithmetic)
template <class P>
__device__ __forceinline__ void rescale_array(const StepArgs<P>& a, int arr, unsigned long long i0,
                                              unsigned long long stride, unsigned long long ncell,
                                              const unsigned long long nmem[3], float ratio)
{
    if (arr < PW_NFIELDS) {
        const typename P::Field f = a.f[arr];
        for (unsigned long long i = i0; i < ncell; i += stride) P::st(f, (unsigned)i, P::ld(f, (unsigned)i) * ratio);
    } else {
        const int q = arr - PW_NFIELDS;
        const int ax = q / kMemPerAxis, t = q % kMemPerAxis;
        typename P::Mem* m = a.m[ax][t];
        if (!m) return;
        for (unsigned long long i = i0; i < nmem[ax]; i += stride) P::stm(m, (unsigned)i, P::ldm(m, (unsigned)i) * ratio);
    }
}

// grid (blocks per shot, shots); idle unless the shot rescales at this step
template <class P> __global__ void k_rescale_batch(const BatchArgs c, int k)
{
    const int s = blockIdx.y;
    const ShotDesc<P>& d = ((const ShotDesc<P>*)c.descs)[s];
    const int it = c.it[s] + k;
    if (it >= d.nt || !d.ratio) return;
    const float ratio = d.ratio[it];
    if (ratio == 0.0f) return;
    const StepArgs<P> a = d.a;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    const unsigned long long i0 = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int arr = 0; arr < PW_NFIELDS + 3 * kMemPerAxis; ++arr) rescale_array<P>(a, arr, i0, stride, d.ncell, d.nmem, ratio);
}

// one thread per receiver of the batch; the arithmetic of k_gather
template <class P> __global__ void k_gather_batch(const BatchArgs c, int k)
{
    const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= c.nrec) return;
    const int s = find_shot(c.rstart, c.n, r);
    const ShotDesc<P>& d = ((const ShotDesc<P>*)c.descs)[s];
    const int it = c.it[s] + k;
    if (it >= d.nt) return;
    const unsigned q = r - c.rstart[s];
    const int fid = d.rfield[q];
    const float v = (float)ld_any<P>(d.a, fid, d.roff[q]);
    const float s_stress = d.oscale[2 * it], s_vel = d.oscale[2 * it + 1];
    d.seis[(size_t)q * d.nt + it] = (fid <= PW_VWZ ? s_vel : s_stress) * v;
}

__global__ void k_advance(int* it, const int* __restrict__ nt, int n, int K)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) it[s] = min(it[s] + K, nt[s]);
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

hipError_t launch_fused(const DevBuffers& b, hipStream_t st)
{
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        FusedArgs<P> A;
        A.s = make_args<P>(b);
        for (int k = 0; k < PW_NFIELDS; ++k) {
            DevBuffers w = b;
            w.f[k] = b.fo[k];
            w.lo[k] = b.loo[k];
            A.g[k] = make_field<P>(w, k);
        }
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k < kMemPerAxis; ++k) A.mo[ax][k] = (typename P::Mem*)b.memo[ax][k];
        k_fused<P><<<A.s.nblocks, dim3(64, 4), 0, st>>>(A);
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

// ---- multi-shot launchers ----
static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

void batch_free(BatchGroup& g)
{
    if (g.d_descs) (void)hipFree(g.d_descs);
    g = BatchGroup{};
}

hipError_t batch_build(BatchGroup& g, int prec, const BatchShot* shots, int n, const int* it, hipStream_t st)
{
    if (n <= 0) {
        g.n = 0;
        g.prec = prec;
        return hipSuccess;
    }
    return dispatch(prec, [&](auto pol) -> hipError_t {
        using P = decltype(pol);
        using R = typename P::R;
        std::vector<ShotDesc<P>> d((size_t)n);
        std::vector<unsigned> bstart((size_t)n + 1, 0u), rstart((size_t)n + 1, 0u);
        std::vector<int> ntv((size_t)n), itv(it, it + n);
        for (int s = 0; s < n; ++s) {
            const BatchShot& b = shots[s];
            if (b.b->prec != prec || b.b->fused) return hipErrorInvalidValue;
            ShotDesc<P>& q = d[(size_t)s];
            q.a = make_args<P>(*b.b);
            q.amp = (const R*)b.amp;
            q.oscale = b.oscale;
            q.ratio = b.ratio;
            q.rfield = b.rfield;
            q.roff = b.roff;
            q.seis = b.seis;
            q.nrec = b.nrec;
            q.nt = b.nt;
            q.src_off = b.src_off;
            q.has_src = b.has_src;
            for (int c = 0; c < 4; ++c) q.w[c] = (R)b.w[c];
            q.ncell = (unsigned long long)b.b->g.cells;
            for (int ax = 0; ax < 3; ++ax) q.nmem[ax] = (unsigned long long)b.b->g.slab_cells[ax];
            bstart[(size_t)s + 1] = bstart[(size_t)s] + (unsigned)q.a.nblocks;
            rstart[(size_t)s + 1] = rstart[(size_t)s] + (unsigned)b.nrec;
            ntv[(size_t)s] = b.nt;
        }
        // per-XCD prefix: XCD x holds q + (x < r) tiles of a shot with q*8 + r tiles
        std::vector<unsigned> xstart((size_t)8 * (n + 1), 0u);
        unsigned xmax = 0;
        for (int x = 0; x < 8; ++x) {
            unsigned* xs = xstart.data() + (size_t)x * (n + 1);
            for (int s = 0; s < n; ++s) {
                const unsigned nb = (unsigned)d[(size_t)s].a.nblocks;
                xs[s + 1] = xs[s] + nb / 8u + ((unsigned)x < nb % 8u ? 1u : 0u);
            }
            xmax = std::max(xmax, xs[n]);
        }
        const size_t o_desc = 0, o_b = al256(sizeof(ShotDesc<P>) * n), o_x = o_b + al256(4 * (n + 1)),
                     o_r = o_x + al256(4 * 8 * (n + 1)), o_it = o_r + al256(4 * (n + 1)), o_nt = o_it + al256(4 * n),
                     total = o_nt + al256(4 * n);
        if (g.bytes < total) {
            batch_free(g);
            hipError_t e = hipMalloc(&g.d_descs, total);
            if (e != hipSuccess) return e;
            g.bytes = total;
        }
        char* base = (char*)g.d_descs;
        g.prec = prec;
        g.n = n;
        g.d_bstart = (unsigned*)(base + o_b);
        g.d_xstart = (unsigned*)(base + o_x);
        g.d_rstart = (unsigned*)(base + o_r);
        g.d_it = (int*)(base + o_it);
        g.d_nt = (int*)(base + o_nt);
        g.xcd_map = shots[0].b->xcd_map;
        g.nblocks = g.xcd_map ? 8u * xmax : bstart[(size_t)n];
        g.nrec = rstart[(size_t)n];
        hipError_t e = hipMemcpyAsync(base + o_desc, d.data(), sizeof(ShotDesc<P>) * n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(g.d_bstart, bstart.data(), 4 * (n + 1), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(g.d_xstart, xstart.data(), 4 * 8 * (n + 1), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(g.d_rstart, rstart.data(), 4 * (n + 1), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(g.d_it, itv.data(), 4 * n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(g.d_nt, ntv.data(), 4 * n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);  // the host tables go out of scope
        return e;
    });
}

static BatchArgs batch_args(const BatchGroup& g)
{
    BatchArgs c;
    c.descs = g.d_descs;
    c.bstart = g.d_bstart;
    c.xstart = g.d_xstart;
    c.rstart = g.d_rstart;
    c.it = g.d_it;
    c.n = g.n;
    c.nblocks = g.nblocks;
    c.nrec = g.nrec;
    c.xcd_map = g.xcd_map;
    return c;
}

hipError_t batch_step(const BatchGroup& g, hipStream_t st, int k)
{
    if (g.n <= 0) return hipSuccess;
    const BatchArgs c = batch_args(g);
    return dispatch(g.prec, [&](auto pol) -> hipError_t {
        using P = decltype(pol);
        if (P::scaled) k_rescale_batch<P><<<dim3(64, (unsigned)g.n), 256, 0, st>>>(c, k);
        k_source_batch<P><<<(unsigned)((4 * g.n + 255) / 256), 256, 0, st>>>(c, k);
        k_velocity_batch<P><<<g.nblocks, dim3(64, 4), 0, st>>>(c, k);
        k_stress_batch<P><<<g.nblocks, dim3(64, 4), 0, st>>>(c, k);
        if (g.nrec > 0) k_gather_batch<P><<<(g.nrec + 255) / 256, 256, 0, st>>>(c, k);
        return hipGetLastError();
    });
}

hipError_t batch_advance(const BatchGroup& g, hipStream_t st, int K)
{
    if (g.n <= 0) return hipSuccess;
    k_advance<<<(unsigned)((g.n + 255) / 256), 256, 0, st>>>(g.d_it, g.d_nt, g.n, K);
    return hipGetLastError();
}

}  // namespace pw
