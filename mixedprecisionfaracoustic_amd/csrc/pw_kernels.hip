// pw_kernels.hip -- CDNA4 (gfx950) kernels of the Biot staggered-grid step.
//
// Per step (the order of Float/main.cu:1055-1093):
//   k_source   : add the source sample at one cell (Float/main.cu:45-59)
//   k_velocity : fused solid + relative-fluid velocity update with the
//                Crank-Nicolson viscous coupling (Float/main.cu:64-192)
//   k_stress   : fused stress + pore-pressure update (Float/main.cu:194-295)
//   k_gather   : receivers -> device seismogram buffer (Float/main.cu:1083-1093)
// (Half: k_rescale every 100 steps before the source, Half/main.cu:1326-1364)
//
// Layout: one padded SoA array per wavefield (row pitch = nx rounded to 64
// elements).  A block is a 64 x 4 (x, y) column tile that marches up a chunk of
// z (8 planes from 256 planes up); each wave owns one 64-cell x row, so every
// load is a coalesced 256 B (FP32) row segment.  (Tunable per kernel: N cells
// per lane, wx waves side by side along x, rows and planes per block; by
// default comp16's velocity kernel uses N = 2 and FP32's two kernels and BF16's
// velocity kernel wx = 2, where those wider segments fill the rows, comp16f
// runs 64 x 1 tiles, 16 planes from nz >= 512, and FP32's 4th-order velocity
// kernel 64 x 1 tiles of 16 planes from nz >= 512.)  x neighbours come from
// the adjacent lane (one DPP move; only the edge lane loads), y neighbours hit
// L1/L2, and z neighbours are carried in registers from the previous iteration
// (each value is loaded once per block).
// A plane's loads are issued ahead of its CPML memory-variable stores, which
// the compiler cannot prove disjoint from them.  Each field puts one pointer
// into the kernels' scalar registers (FA below).
//
// Also here: the batched kernels of the shot server (one launch per kernel over
// many shots, or one persistent cooperative launch per round), the rescale,
// snapshot, decode and fill side kernels.
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
// Half formulas with FP16 prescaled coefficients; comp16f keeps comp16's state
// under the Float equations in FP32 (stored units, compensated coefficient
// records).  All arithmetic is FP32 except the FP64 policy.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "pw_device.h"

namespace pw {

typedef _Float16 f16;
typedef __bf16 bf16;

// --------------------------------------------------------------------------
// Storage policies
//   ld / st        one value
//   ldv / stv      N consecutive values (offset a multiple of N) as one wide access
//   ldc / ldcv     coefficients;  ldm / stm  CPML memory variables
//   rt             a value as it reads back after a store
// --------------------------------------------------------------------------
template <class R> struct Tab4 { R cf, ef, ch, eh; };

template <class F, class = void> struct requires_zero : std::false_type {};
template <class F> struct requires_zero<F, std::void_t<decltype(F::z)>> : std::true_type {};

// N consecutive elements, aligned to their total size: one global load/store
template <class T, int N> struct alignas(sizeof(T) * N) Pack { T v[N]; };

// Update arithmetic of a policy:
//  kFormRef   the reference Float program's (Float/main.cu:64-295): staggered
//             coefficients formed on the fly with its float/double steps,
//             derivatives divided by h
//  kFormHalf  the reference Half program's (Half/main.cu:156-371): FP16
//             coefficients prescaled by dt and 2^em
//  kFormFast  this build's FP32 form of the same update: undivided differences
//             with dt/h and the storage scales folded into launch constants,
//             hardware reciprocals instead of IEEE divides
enum { kFormRef = 0, kFormHalf = 1, kFormFast = 2 };
template <int N, class T> __device__ __forceinline__ Pack<T, N> ld_pack(const T* p, unsigned o)
{
    return *(const Pack<T, N>*)(p + o);
}
template <int N, class T> __device__ __forceinline__ void st_pack(T* p, unsigned o, const Pack<T, N>& v)
{
    *(Pack<T, N>*)(p + o) = v;
}
// comp16 / BF16 coefficients: the 4-value FP16 records of cells o .. o+N-1 in
// one access (p: the record array, i.e. its slot-0 coefficient); v[q][j] is
// slot q of cell o + j
template <int N, class R, class C>
__device__ __forceinline__ void ld_srec(const C* p, unsigned o, R (&v)[4][N])
{
    const Pack<C, 4 * N> t = *(const Pack<C, 4 * N>*)(p + 4 * (size_t)o);
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q][j] = (R)t.v[4 * j + q];
}

// the run-time zero of a field (round32); plain FP32/FP64 fields carry none
template <class F> __device__ __forceinline__ unsigned field_zero(const F& f)
{
    if constexpr (requires_zero<F>::value) return f.z;
    else return 0u;
}
// R -> storage type T; the BF16 conversion starts from a rounded FP32 value.
template <class T, class R> __device__ __forceinline__ T narrow(R v, unsigned) { return (T)v; }
template <> __device__ __forceinline__ bf16 narrow<bf16, float>(float v, unsigned z) { return (bf16)round32(v, z); }

// Wide access helpers shared by the plain policies (storage type = T).
#define PW_PLAIN_VEC(T)                                                                           \
    template <int N> static __device__ __forceinline__ void ldv(const Field& f, unsigned o, R (&r)[N]) \
    {                                                                                             \
        const Pack<T, N> t = ld_pack<N>(f.p, o);                                                  \
        _Pragma("unroll") for (int j = 0; j < N; ++j) r[j] = (R)t.v[j];                           \
    }                                                                                             \
    template <int N> static __device__ __forceinline__ void stv(const Field& f, unsigned o, const R (&r)[N]) \
    {                                                                                             \
        Pack<T, N> t;                                                                             \
        _Pragma("unroll") for (int j = 0; j < N; ++j) t.v[j] = narrow<T>(r[j], field_zero(f));     \
        st_pack<N>(f.p, o, t);                                                                    \
    }                                                                                             \
    template <int N> static __device__ __forceinline__ void ldcv(const Coef* c, unsigned o, R (&r)[N]) \
    {                                                                                             \
        const Pack<Coef, N> t = ld_pack<N>(c, o);                                                 \
        _Pragma("unroll") for (int j = 0; j < N; ++j) r[j] = (R)t.v[j];                           \
    }

struct PolF32 {
    using R = float; using Coef = float; using Mem = float;
    static constexpr int form = kFormRef;
    static constexpr bool scaled = false;
    static constexpr bool pair_cpml = false;
    struct Field { float* p; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return f.p[o]; }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v) { f.p[o] = v; }
    PW_PLAIN_VEC(float)
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v, unsigned) { m[o] = v; }
};
struct PolF64 {
    using R = double; using Coef = double; using Mem = double;
    static constexpr int form = kFormRef;
    static constexpr bool scaled = false;
    static constexpr bool pair_cpml = false;
    struct Field { double* p; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return f.p[o]; }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v) { f.p[o] = v; }
    PW_PLAIN_VEC(double)
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v, unsigned) { m[o] = v; }
};
struct PolC16 {
    using R = float; using Coef = f16; using Mem = f16;
    static constexpr int form = kFormHalf;
    static constexpr bool scaled = true;
    static constexpr bool srec = true;  // coefficients in 4-value records
    // the reference Half program's CPML: (psi, previous derivative) FP16 pairs
    static constexpr bool pair_cpml = true;
    struct Field { f16* hi; int8_t* lo; unsigned z; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return c16_decode(f.hi[o], f.lo[o]); }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v)
    {
        f16 h;
        int8_t l;
        c16_encode(v, h, l, f.z);
        f.hi[o] = h;
        f.lo[o] = l;
    }
    template <int N> static __device__ __forceinline__ void ldv(const Field& f, unsigned o, R (&r)[N])
    {
        const Pack<f16, N> h = ld_pack<N>(f.hi, o);
        const Pack<int8_t, N> l = ld_pack<N>(f.lo, o);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = c16_decode(h.v[j], l.v[j]);
    }
    template <int N> static __device__ __forceinline__ void stv(const Field& f, unsigned o, const R (&r)[N])
    {
        Pack<f16, N> h;
        Pack<int8_t, N> l;
#pragma unroll
        for (int j = 0; j < N; ++j) c16_encode(r[j], h.v[j], l.v[j], f.z);
        st_pack<N>(f.hi, o, h);
        st_pack<N>(f.lo, o, l);
    }
    template <int N> static __device__ __forceinline__ void ldcv(const Coef* c, unsigned o, R (&r)[N])
    {
        const Pack<Coef, N> t = ld_pack<N>(c, o);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = (R)t.v[j];
    }
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return (float)c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return (float)m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v, unsigned z) { m[o] = (f16)round32(v, z); }
};
// comp16f: the comp16 state (FP16 value + int8 residual, residual rounded to
// nearest) under this build's FP32 update (kFormFast).  Stored values are
// physical values times a power-of-two scale per field class; coefficients are
// the cell-centred FP32 medium and memory variables FP32.  ld / st take and
// return physical values (side kernels); the step kernels work in stored units
// through FA<PolC16F>.
struct PolC16F {
    using R = float; using Coef = float; using Mem = float;
    static constexpr int form = kFormFast;
    static constexpr bool scaled = false;
    static constexpr bool pair_cpml = false;
    // st: 2 for a field stored in a pair (its values and residuals alternate
    // with the partner's), 1 for a field stored alone
    struct Field { f16* hi; int8_t* lo; unsigned z; float sc, isc; unsigned st; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o)
    {
        return c16_decode(f.hi[o * f.st], f.lo[o * f.st]) * f.isc;
    }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v)
    {
        f16 h;
        int8_t l;
        c16_encode_rn(v * f.sc, h, l, f.z);
        f.hi[o * f.st] = h;
        f.lo[o * f.st] = l;
    }
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return c[o]; }
    template <int N> static __device__ __forceinline__ void ldcv(const Coef* c, unsigned o, R (&r)[N])
    {
        const Pack<Coef, N> t = ld_pack<N>(c, o);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = t.v[j];
    }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v, unsigned) { m[o] = v; }
};

struct PolB16 {
    using R = float; using Coef = f16; using Mem = bf16;
    static constexpr int form = kFormHalf;
    static constexpr bool scaled = true;
    static constexpr bool srec = false;
    static constexpr bool pair_cpml = false;
    struct Field { bf16* p; unsigned z; };
    static __device__ __forceinline__ R ld(const Field& f, unsigned o) { return (float)f.p[o]; }
    static __device__ __forceinline__ void st(const Field& f, unsigned o, R v) { f.p[o] = (bf16)round32(v, f.z); }
    PW_PLAIN_VEC(bf16)
    static __device__ __forceinline__ R ldc(const Coef* c, unsigned o) { return (float)c[o]; }
    static __device__ __forceinline__ R ldm(const Mem* m, unsigned o) { return (float)m[o]; }
    static __device__ __forceinline__ void stm(Mem* m, unsigned o, R v, unsigned z) { m[o] = (bf16)round32(v, z); }
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
    int ty;                // waves per block (block = 64 x ty threads)
    int wx;                // waves side by side along x: a tile is 64*vx*wx x ty/wx cells
    int vx;                // cells per lane along x (tile = 64*vx x ty cells)
    unsigned zero;         // 0, opaque to the compiler (round32)
    int zb, ze;            // planes [zb, ze) updated by this launch
    int xcd_map;           // 1: XCD-aware block -> tile mapping
    typename P::R dt, h;
    // compensated fields: bytes from a field's FP16 values to its residual bytes
    // (the same for every field; comp16f: for the slot-0 field of every pair,
    // and loff1 for the field stored alone; 64-bit: comp16f's pairs put their
    // residuals 4 bytes per cell past the values)
    unsigned long long loff, loff1;
    // kFormFast launch constants (KF_*): dt/h, the factors of the staggered
    // means and the class storage scales folded together
    float kf[9];
    // comp16f coefficient records (rec_base order): FP16 values, residuals, and
    // the factor from a decoded value to the physical coefficient
    const f16* rh[kRecords];
    const int8_t* rl[kRecords];
    float rs[kRecords][4];
    typename P::R rhof;  // the medium's one rhof where c[PC_RHOF] is null (FP32, FP64)
};

// kFormFast constants; g = dt/h, s_u, s_w, s_t = storage scales of the solid
// velocity, relative fluid velocity and stress classes
enum {
    KF_DT2 = 0,  // dt/2
    KF_WS,       // 4 g s_w/s_t
    KF_WP,       // 2 g s_w/s_t
    KF_U,        // 2 g s_u/s_t
    KF_UW,       // 2 s_u/s_w
    KF_G,        // g s_t/s_u
    KF_G2,       // 2 g s_t/s_u
    KF_G4,       // 4 g s_t/s_u
    KF_W         // s_u/s_w
};

template <class P> __host__ typename P::Field make_field(const DevBuffers& b, int k);
template <> __host__ PolF32::Field make_field<PolF32>(const DevBuffers& b, int k) { return {(float*)b.f[k]}; }
template <> __host__ PolF64::Field make_field<PolF64>(const DevBuffers& b, int k) { return {(double*)b.f[k]}; }
template <> __host__ PolC16::Field make_field<PolC16>(const DevBuffers& b, int k) { return {(f16*)b.f[k], b.lo[k], 0u}; }
template <> __host__ PolB16::Field make_field<PolB16>(const DevBuffers& b, int k) { return {(bf16*)b.f[k], 0u}; }
template <> __host__ PolC16F::Field make_field<PolC16F>(const DevBuffers& b, int k)
{
    const float sc = k <= PW_VUZ ? b.sc_vel : (k <= PW_VWZ ? b.sc_vw : b.sc_stress);
    return {(f16*)b.f[k], b.lo[k], 0u, sc, 1.0f / sc, c16f_pair(k) >= 0 ? 2u : 1u};
}

template <class P> __host__ StepArgs<P> make_args(const DevBuffers& b)
{
    using R = typename P::R;
    StepArgs<P> a;
    for (int k = 0; k < PW_NFIELDS; ++k) a.f[k] = make_field<P>(b, k);
    for (int k = 0; k < kMaxCoef; ++k) a.c[k] = (const typename P::Coef*)b.coef[k];
    if (b.derive_h) a.c[PC_HH] = nullptr;  // FP32, consistent medium: the stress kernel forms H
    a.rhof = (R)b.rhof_uniform;
    if (b.rhof_uniform > 0) a.c[PC_RHOF] = nullptr;  // FP32 and FP64, one fluid: a scalar
    for (int ax = 0; ax < 3; ++ax) {
        for (int k = 0; k < kMemPerAxis; ++k) a.m[ax][k] = (typename P::Mem*)b.mem[ax][k];
        a.tab[ax] = (const Tab4<R>*)b.tab[ax];
        a.lo[ax] = b.g.lo[ax];
        a.hi[ax] = b.g.hi[ax];
    }
    a.nx = b.g.nx; a.ny = b.g.ny; a.nz = b.g.nz; a.pitch = b.g.pitch;
    a.plane = (unsigned)b.g.plane;
    a.zc = b.zc;
    a.ty = b.ty > 0 ? b.ty : 4;
    a.vx = (b.vx > 0 && b.order != 4) ? b.vx : 1;  // 4th order: one cell per lane
    a.zero = 0u;
    a.wx = (b.wx == 2 && a.ty % 2 == 0) ? 2 : 1;
    a.gx = (b.g.nx + 64 * a.vx * a.wx - 1) / (64 * a.vx * a.wx);
    a.gy = (b.g.ny + a.ty / a.wx - 1) / (a.ty / a.wx);
    const int halo = b.order == 4 ? 2 : 1;  // the outermost layers that are never updated
    a.zb = halo;
    a.ze = b.g.nz - halo;
    a.nblocks = a.gx * a.gy * ((a.ze - a.zb + b.zc - 1) / b.zc);
    a.xcd_map = b.xcd_map;
    a.dt = (R)b.dt;
    a.h = (R)b.h;
    // the pool stores each compensated field's residual bytes right after its
    // FP16 values, at one offset for all fields
    a.loff = b.lo[0] ? (unsigned long long)((const char*)b.lo[0] - (const char*)b.f[0]) : 0ull;
    a.loff1 = b.lo[kC16FSingle] ? (unsigned long long)((const char*)b.lo[kC16FSingle] - (const char*)b.f[kC16FSingle])
                                : 0ull;
    const double g = b.dt / b.h, su = b.sc_vel, sw = b.sc_vw, stt = b.sc_stress;
    a.kf[KF_DT2] = (float)(0.5 * b.dt);
    a.kf[KF_WS] = (float)(4.0 * g * sw / stt);
    a.kf[KF_WP] = (float)(2.0 * g * sw / stt);
    a.kf[KF_U] = (float)(2.0 * g * su / stt);
    a.kf[KF_UW] = (float)(2.0 * su / sw);
    a.kf[KF_G] = (float)(g * stt / su);
    a.kf[KF_G2] = (float)(2.0 * g * stt / su);
    a.kf[KF_G4] = (float)(4.0 * g * stt / su);
    a.kf[KF_W] = (float)(su / sw);
    for (int r = 0; r < kRecords; ++r) {
        a.rh[r] = (const f16*)b.rec_hi[r];
        a.rl[r] = (const int8_t*)b.rec_lo[r];
        for (int q = 0; q < 4; ++q) a.rs[r][q] = b.rec_scale[r][q];
    }
    return a;
}

// --------------------------------------------------------------------------
// Field access in the step kernels.  Each field puts one pointer, its value
// array, into the kernels' scalar registers.  The compensated policies find a
// field's residual bytes a.loff bytes after its FP16 values (one offset for all
// fields), and the zero of round32 and comp16f's scale for the field's class
// come from the launch, not from each field: otherwise 13 fields hold up to two
// pointers, a zero and two scales each, more than the scalar register file,
// and the overflow is reloaded lane by lane.  The values are exactly those of
// the policies' own ld / st.
// --------------------------------------------------------------------------
template <class P> struct FA {
    using R = typename P::R;
    static __device__ __forceinline__ R ld(const StepArgs<P>& a, int fid, unsigned o) { return P::ld(a.f[fid], o); }
    template <int N> static __device__ __forceinline__ void ldv(const StepArgs<P>& a, int fid, unsigned o, R (&r)[N])
    {
        P::template ldv<N>(a.f[fid], o, r);
    }
    static __device__ __forceinline__ void st(const StepArgs<P>& a, int fid, unsigned o, R v) { P::st(a.f[fid], o, v); }
    template <int N>
    static __device__ __forceinline__ void stv(const StepArgs<P>& a, int fid, unsigned o, const R (&r)[N])
    {
        P::template stv<N>(a.f[fid], o, r);
    }
};
template <> struct FA<PolB16> {
    using R = float;
    static __device__ __forceinline__ R ld(const StepArgs<PolB16>& a, int fid, unsigned o) { return (float)a.f[fid].p[o]; }
    template <int N> static __device__ __forceinline__ void ldv(const StepArgs<PolB16>& a, int fid, unsigned o, R (&r)[N])
    {
        const Pack<bf16, N> t = ld_pack<N>(a.f[fid].p, o);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = (float)t.v[j];
    }
    static __device__ __forceinline__ void st(const StepArgs<PolB16>& a, int fid, unsigned o, R v)
    {
        a.f[fid].p[o] = (bf16)round32(v, a.zero);
    }
    template <int N>
    static __device__ __forceinline__ void stv(const StepArgs<PolB16>& a, int fid, unsigned o, const R (&r)[N])
    {
        Pack<bf16, N> t;
#pragma unroll
        for (int j = 0; j < N; ++j) t.v[j] = narrow<bf16>(r[j], a.zero);
        st_pack<N>(a.f[fid].p, o, t);
    }
};
// comp16: the field's FP16 values at f[fid].hi, its residual bytes a.loff bytes after them
template <> struct FA<PolC16> {
    using R = float;
    using P = PolC16;
    static __device__ __forceinline__ R ld(const StepArgs<P>& a, int fid, unsigned o)
    {
        const f16* h = a.f[fid].hi;
        return c16_decode(h[o], ((const int8_t*)h)[(unsigned)a.loff + o]);
    }
    template <int N> static __device__ __forceinline__ void ldv(const StepArgs<P>& a, int fid, unsigned o, R (&r)[N])
    {
        const f16* h = a.f[fid].hi;
        const Pack<f16, N> hv = ld_pack<N>(h, o);
        const Pack<int8_t, N> lv = ld_pack<N>((const int8_t*)h, (unsigned)a.loff + o);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = c16_decode(hv.v[j], lv.v[j]);
    }
    static __device__ __forceinline__ void st(const StepArgs<P>& a, int fid, unsigned o, R v)
    {
        f16* h = a.f[fid].hi;
        f16 hh;
        int8_t l;
        c16_encode(v, hh, l, a.zero);
        h[o] = hh;
        ((int8_t*)h)[(unsigned)a.loff + o] = l;
    }
    template <int N>
    static __device__ __forceinline__ void stv(const StepArgs<P>& a, int fid, unsigned o, const R (&r)[N])
    {
        f16* h = a.f[fid].hi;
        Pack<f16, N> hv;
        Pack<int8_t, N> lv;
#pragma unroll
        for (int j = 0; j < N; ++j) c16_encode(r[j], hv.v[j], lv.v[j], a.zero);
        st_pack<N>(h, o, hv);
        st_pack<N>((int8_t*)h, (unsigned)a.loff + o, lv);
    }
};
// comp16f, in stored units (the class scales are folded into the launch
// constants), with the round-to-nearest encode.  A paired field in slot q has
// f[fid].hi = pair values + q (in FP16 units); the cell's values lie at 2o + q
// and its residual byte a.loff - q + 2o bytes after f[fid].hi.  The field
// stored alone has comp16's addressing with a.loff1.
template <> struct FA<PolC16F> {
    using R = float;
    using P = PolC16F;
    static __device__ __forceinline__ constexpr bool paired(int fid) { return c16f_pair(fid) >= 0; }
    static __device__ __forceinline__ constexpr int slot(int fid) { return c16f_slot(fid); }
    static __device__ __forceinline__ R ld(const StepArgs<P>& a, int fid, unsigned o)
    {
        const f16* h = a.f[fid].hi;
        if (paired(fid)) return c16_decode(h[2 * o], ((const int8_t*)h + a.loff - slot(fid))[2 * o]);
        return c16_decode(h[o], ((const int8_t*)h + a.loff1)[o]);
    }
    template <int N> static __device__ __forceinline__ void ldv(const StepArgs<P>& a, int fid, unsigned o, R (&r)[N])
    {
        const f16* h = a.f[fid].hi;
        if (paired(fid)) {
            // both fields' values of the N cells (one access each for values and
            // residuals), this field's slot taken
            const int q = slot(fid);
            const Pack<f16, 2 * N> hv = ld_pack<2 * N>(h - q, 2 * o);
            const Pack<int8_t, 2 * N> lv = ld_pack<2 * N>((const int8_t*)(h - q) + a.loff, 2 * o);
#pragma unroll
            for (int j = 0; j < N; ++j) r[j] = c16_decode(hv.v[2 * j + q], lv.v[2 * j + q]);
            return;
        }
        const Pack<f16, N> hv = ld_pack<N>(h, o);
        const Pack<int8_t, N> lv = ld_pack<N>((const int8_t*)h + a.loff1, o);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = c16_decode(hv.v[j], lv.v[j]);
    }
    // fields fa (slot 0) and fb (slot 1) of one pair: one access each for values and residuals
    template <int N>
    static __device__ __forceinline__ void ld2v(const StepArgs<P>& a, int fa, int fb, unsigned o, R (&ra)[N], R (&rb)[N])
    {
        const f16* h = a.f[fa].hi;
        const Pack<f16, 2 * N> hv = ld_pack<2 * N>(h, 2 * o);
        const Pack<int8_t, 2 * N> lv = ld_pack<2 * N>((const int8_t*)h + a.loff, 2 * o);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            ra[j] = c16_decode(hv.v[2 * j], lv.v[2 * j]);
            rb[j] = c16_decode(hv.v[2 * j + 1], lv.v[2 * j + 1]);
        }
        (void)fb;
    }
    static __device__ __forceinline__ void st(const StepArgs<P>& a, int fid, unsigned o, R v)
    {
        f16* h = a.f[fid].hi;
        f16 hh;
        int8_t l;
        c16_encode_rn(v, hh, l, a.zero);
        if (paired(fid)) {
            h[2 * o] = hh;
            ((int8_t*)h + a.loff - slot(fid))[2 * o] = l;
        } else {
            h[o] = hh;
            ((int8_t*)h + a.loff1)[o] = l;
        }
    }
    template <int N>
    static __device__ __forceinline__ void stv(const StepArgs<P>& a, int fid, unsigned o, const R (&r)[N])
    {
        if (paired(fid)) {  // cell by cell: the partner's values lie in between
#pragma unroll
            for (int j = 0; j < N; ++j) st(a, fid, o + j, r[j]);
            return;
        }
        f16* h = a.f[fid].hi;
        Pack<f16, N> hv;
        Pack<int8_t, N> lv;
#pragma unroll
        for (int j = 0; j < N; ++j) c16_encode_rn(r[j], hv.v[j], lv.v[j], a.zero);
        st_pack<N>(h, o, hv);
        st_pack<N>((int8_t*)h + a.loff1, o, lv);
    }
    template <int N>
    static __device__ __forceinline__ void st2v(const StepArgs<P>& a, int fa, int fb, unsigned o, const R (&ra)[N],
                                                const R (&rb)[N])
    {
        f16* h = a.f[fa].hi;
        Pack<f16, 2 * N> hv;
        Pack<int8_t, 2 * N> lv;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            c16_encode_rn(ra[j], hv.v[2 * j], lv.v[2 * j], a.zero);
            c16_encode_rn(rb[j], hv.v[2 * j + 1], lv.v[2 * j + 1], a.zero);
        }
        st_pack<2 * N>(h, 2 * o, hv);
        st_pack<2 * N>((int8_t*)h + a.loff, 2 * o, lv);
        (void)fb;
    }
};

// comp16f coefficient record r (rec_base order, K = rec_width(r) coefficients)
// of cells o .. o+N-1: one access for the N x K FP16 values, one for the
// residual bytes; c[q][j] = coefficient q of cell o + j.
template <int K, int N, class P, class R>
__device__ __forceinline__ void ld_rec(const StepArgs<P>& a, int r, unsigned o, R (&c)[K][N])
{
    const Pack<f16, K * N> h = ld_pack<K * N>(a.rh[r] + K * (size_t)o, 0u);
    const Pack<int8_t, K * N> l = ld_pack<K * N>(a.rl[r] + K * (size_t)o, 0u);
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int q = 0; q < K; ++q) c[q][j] = c16_decode(h.v[K * j + q], l.v[K * j + q]) * a.rs[r][q];
}
// coefficient q of the K-wide record r at one cell
template <int K, class P> __device__ __forceinline__ float ld_rec1(const StepArgs<P>& a, int r, int q, unsigned o)
{
    const size_t k = K * (size_t)o + q;
    return c16_decode(a.rh[r][k], a.rl[r][k]) * a.rs[r][q];
}
// coefficient q of the K-wide record r at cells o .. o+N-1 (the whole records)
template <int K, int N, class P, class R>
__device__ __forceinline__ void ld_rec_slot(const StepArgs<P>& a, int r, int q, unsigned o, R (&c)[N])
{
    R t[K][N];
    ld_rec<K, N>(a, r, o, t);
#pragma unroll
    for (int j = 0; j < N; ++j) c[j] = t[q][j];
}

// Two fields at the same cells: one paired access where the policy stores the
// two as a pair (fa in slot 0, fb in slot 1), two accesses otherwise.
template <class P> __device__ __forceinline__ constexpr bool is_pair(int fa, int fb)
{
    return std::is_same<P, PolC16F>::value && c16f_pair(fa) >= 0 && c16f_pair(fa) == c16f_pair(fb) &&
           c16f_slot(fa) == 0 && c16f_slot(fb) == 1;
}
template <class P, int N>
__device__ __forceinline__ void ld2v(const StepArgs<P>& a, int fa, int fb, unsigned o, typename P::R (&ra)[N],
                                     typename P::R (&rb)[N])
{
    if constexpr (std::is_same<P, PolC16F>::value) {
        if (is_pair<P>(fa, fb)) {
            FA<P>::template ld2v<N>(a, fa, fb, o, ra, rb);
            return;
        }
    }
    FA<P>::template ldv<N>(a, fa, o, ra);
    FA<P>::template ldv<N>(a, fb, o, rb);
}
template <class P, int N>
__device__ __forceinline__ void st2v(const StepArgs<P>& a, int fa, int fb, unsigned o, const typename P::R (&ra)[N],
                                     const typename P::R (&rb)[N])
{
    if constexpr (std::is_same<P, PolC16F>::value) {
        if (is_pair<P>(fa, fb)) {
            FA<P>::template st2v<N>(a, fa, fb, o, ra, rb);
            return;
        }
    }
    FA<P>::template stv<N>(a, fa, o, ra);
    FA<P>::template stv<N>(a, fb, o, rb);
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

// values in a CPML memory record: the four terms' variables, or comp16's
// (psi, prev) pairs
template <class P> __host__ __device__ constexpr int mem_width() { return P::pair_cpml ? 8 : 4; }

// The four CPML terms of one damped axis.  All four memory variables are loaded
// before any is stored: a store could alias the next load as far as the
// compiler knows, so term-by-term code would make four memory round trips.
// With P::pair_cpml the reference Half form instead (Half/main.cu:201-240):
// psi = psi*e + (-c)*(e*prev + der) in float without contraction, psi and
// prev = der stored as FP16, der += psi; the record is {psi x4, prev x4}.
// Every form keeps the axis's variables of cell k as one record at
// phi[0] + width*k, read and written with one access each.
template <class P>
__device__ __forceinline__ void cpml4(typename P::Mem* const* phi, unsigned k, const typename P::R (&c)[4],
                                      const typename P::R (&e)[4], typename P::R* const (&der)[4], unsigned z)
{
    using R = typename P::R;
    using M = typename P::Mem;
    if constexpr (P::pair_cpml) {
#pragma clang fp contract(off)
        M* rec = phi[0] + 8 * (size_t)k;
        const Pack<M, 8> t = ld_pack<8>(rec, 0u);
        Pack<M, 8> u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const R d = *der[q];
            const R nw = (R)t.v[q] * e[q] + (-c[q]) * (e[q] * (R)t.v[q + 4] + d);
            u.v[q] = (M)round32(nw, z);
            u.v[q + 4] = (M)round32(d, z);
            *der[q] = d + nw;
        }
        st_pack<8>(rec, 0u, u);
    } else if constexpr (P::scaled) {
        // the single-variable form (file header), rounded as P::stm rounds
        M* rec = phi[0] + 4 * (size_t)k;
        const Pack<M, 4> t = ld_pack<4>(rec, 0u);
        Pack<M, 4> u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const R d = *der[q];
            const R psi = (R)t.v[q] - c[q] * d;
            u.v[q] = (M)round32(e[q] * (psi - c[q] * d), z);
            *der[q] = d + psi;
        }
        st_pack<4>(rec, 0u, u);
    } else {
        // the four variables of (cell, axis) k as one record at phi[0] + 4k
        using M = typename P::Mem;
        M* rec = phi[0] + 4 * (size_t)k;
        const Pack<M, 4> t = ld_pack<4>(rec, 0u);
        Pack<M, 4> u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const R d = *der[q];
            const R psi = (R)t.v[q] - c[q] * d;
            u.v[q] = (M)(e[q] * (psi - c[q] * d));
            *der[q] = d + psi;
        }
        st_pack<4>(rec, 0u, u);
    }
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

// Staggered difference divided by the grid step as the reference programs do
// (Float/main.cu:115-128, Half/main.cu:186-199).  kFormFast keeps the
// undivided difference (1/h is folded into its launch constants); BF16, which
// has no reference counterpart, multiplies by 1/h.
template <class P> struct Deriv {
    typename P::R h, ih;
    __device__ __forceinline__ explicit Deriv(typename P::R h_) : h(h_), ih((typename P::R)1 / h_) {}
    __device__ __forceinline__ typename P::R operator()(typename P::R d) const
    {
        if constexpr (P::form == kFormFast) return d;
        else if constexpr (P::scaled && !P::pair_cpml) return d * ih;
        else return d / h;
    }
};

// hardware reciprocal (v_rcp_f32, within 1 ulp); rcp(0) = inf, rcp(inf) = 0
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// --------------------------------------------------------------------------
// Row segments.  Each lane owns N consecutive cells of one x row (N = 1, 2, 4),
// loaded and stored as one wide access, so a wave moves 64*N cells of a row.
// The x-1 / x+1 neighbours of a lane's cells are its own registers except at
// the segment ends, which come from the adjacent lane (a cross-lane shuffle);
// only lanes 0 and 63 read the neighbouring tile's cells from memory.
// --------------------------------------------------------------------------
struct Lanes {
    int lane, ix0;
    bool lv;                 // the lane's cells lie inside the padded row
    bool edge_r, edge_l;     // lane 63 / lane 0 with the neighbouring cell inside the row
    unsigned sh;             // lanes past the row end read the row's last N cells instead
};

// Cross-lane moves within the wave (DPP): lane i receives lane i+1's value
// (wave_shl1) or lane i-1's (wave_shr1); the lane that has no source lane
// receives 0 and is overwritten by the caller.
#ifndef PW_X_DPP
#define PW_X_DPP 1
#endif
template <int CTRL> __device__ __forceinline__ float lane_move(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL> __device__ __forceinline__ double lane_move(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;

// r[j] = value at x+1 of element j of c; on = the lane's offset of that value
// for N = 1, of its last element's right neighbour otherwise; edge(q) loads
// offset q.  One cell per lane: the next lane's centre value, and for lane 63
// one load at its own neighbour offset, which is wave-uniform, so the load
// touches one cache line instead of a misaligned 64-lane row.  N cells:
// registers, the adjacent lane, and for lane 63 one load from the next tile.
template <class R, int N, class EdgeLoad, class RowLoad>
__device__ __forceinline__ void seg_right(const Lanes& L, const R (&c)[N], R (&r)[N], unsigned on, EdgeLoad edge,
                                          RowLoad row)
{
    if constexpr (N == 1) {
        if constexpr (PW_X_DPP) {
            const R s = lane_move<kDppWaveShl1>(c[0]);
            const R e = edge(__builtin_amdgcn_readlane(on, 63));
            r[0] = L.lane == 63 ? e : s;
        } else {
            row(r);
        }
    } else {
        // lane 63's neighbour comes from one wave-uniform load (or is 0 past the row end)
        const bool er = __builtin_amdgcn_readlane((int)L.edge_r, 63) != 0;
        const R e = er ? edge(__builtin_amdgcn_readlane(on, 63)) : (R)0;
        const R s = lane_move<kDppWaveShl1>(c[0]);
#pragma unroll
        for (int j = 0; j < N - 1; ++j) r[j] = c[j + 1];
        r[N - 1] = L.lane == 63 ? e : s;
    }
}
// r[j] = value at x-1 of element j; on = the lane's offset of its first
// element's left neighbour (lane 0 loads it)
template <class R, int N, class EdgeLoad, class RowLoad>
__device__ __forceinline__ void seg_left(const Lanes& L, const R (&c)[N], R (&r)[N], unsigned on, EdgeLoad edge,
                                         RowLoad row)
{
    if constexpr (N == 1) {
        if constexpr (PW_X_DPP) {
            const R s = lane_move<kDppWaveShr1>(c[0]);
            const R e = edge(__builtin_amdgcn_readlane(on, 0));
            r[0] = L.lane == 0 ? e : s;
        } else {
            row(r);
        }
    } else {
        const bool el = __builtin_amdgcn_readlane((int)L.edge_l, 0) != 0;
        const R e = el ? edge(__builtin_amdgcn_readlane(on, 0)) : (R)0;
        const R s = lane_move<kDppWaveShr1>(c[N - 1]);
        r[0] = L.lane == 0 ? e : s;
#pragma unroll
        for (int j = 1; j < N; ++j) r[j] = c[j - 1];
    }
}

// 4th order takes its x neighbours by cross-lane moves at FP32 only, where
// that measured 0.8 % faster than loads at 512^3.  The other precisions load
// them: with moves, FP64 (two moves per value) measured 3 % slower, and the
// FP16- and BF16-stored fields, whose neighbours must first be decoded in every
// lane, 12 % (comp16) and 3 % (BF16) slower.
template <class P>
constexpr bool kX4Dpp = PW_X_DPP && (std::is_same<P, PolF32>::value || std::is_same<P, PolC16F>::value);

// 4th order, one cell per lane: a field's values one and two cells along x
// from the lane's centre value c (o = the lane's offset).  The cell one out is
// the adjacent lane's centre value, the cell two out that lane's value one out
// (two chained moves); only the edge lane (63 going right, 0 going left) loads,
// at wave-uniform offsets, so each load touches one cache line.  Every lane of
// the wave must be active.
template <class R, class EdgeLoad>
__device__ __forceinline__ R lane_right1(const Lanes& L, R c, unsigned o, EdgeLoad edge)
{
    const R e = edge((unsigned)__builtin_amdgcn_readlane((int)o, 63) + 1u);
    const R s = lane_move<kDppWaveShl1>(c);
    return L.lane == 63 ? e : s;
}
template <class R, class EdgeLoad>
__device__ __forceinline__ R lane_left1(const Lanes& L, R c, unsigned o, EdgeLoad edge)
{
    const R e = edge((unsigned)__builtin_amdgcn_readlane((int)o, 0) - 1u);
    const R s = lane_move<kDppWaveShr1>(c);
    return L.lane == 0 ? e : s;
}
template <class R, class EdgeLoad>
__device__ __forceinline__ void lane_right2(const Lanes& L, R c, unsigned o, EdgeLoad edge, R& r1, R& r2)
{
    const unsigned oe = (unsigned)__builtin_amdgcn_readlane((int)o, 63);
    const R e1 = edge(oe + 1u), e2 = edge(oe + 2u);
    const R s1 = lane_move<kDppWaveShl1>(c);
    r1 = L.lane == 63 ? e1 : s1;
    const R s2 = lane_move<kDppWaveShl1>(r1);
    r2 = L.lane == 63 ? e2 : s2;
}
template <class R, class EdgeLoad>
__device__ __forceinline__ void lane_left2(const Lanes& L, R c, unsigned o, EdgeLoad edge, R& l1, R& l2)
{
    const unsigned oe = (unsigned)__builtin_amdgcn_readlane((int)o, 0);
    const R e1 = edge(oe - 1u), e2 = edge(oe - 2u);
    const R s1 = lane_move<kDppWaveShr1>(c);
    l1 = L.lane == 0 ? e1 : s1;
    const R s2 = lane_move<kDppWaveShr1>(l1);
    l2 = L.lane == 0 ? e2 : s2;
}

// Lane geometry of a row segment starting at cell xb of a row of `pitch` cells.
template <int N> __device__ __forceinline__ Lanes make_lanes(int xb, int pitch)
{
    Lanes L;
    L.lane = threadIdx.x;
    L.ix0 = xb + L.lane * N;
    // one cell per lane: tiles are 64 wide and the pitch a multiple of 64
    L.lv = N == 1 || L.ix0 < pitch;
    L.edge_r = L.ix0 + N < pitch;
    L.edge_l = L.ix0 >= 1;
    L.sh = L.lv ? 0u : (unsigned)(L.ix0 - (pitch - N));
    return L;
}

// --------------------------------------------------------------------------
// Velocity kernel: Float/main.cu:64-192 (physical), Half/main.cu:156-268 (scaled)
// --------------------------------------------------------------------------
template <class P, int N, int ORD>
__device__ __forceinline__ void velocity_tile(const StepArgs<P>& a, const Tile tl)
{
    using R = typename P::R;
    static_assert(ORD == 2 || (ORD == 4 && N == 1), "4th order runs one cell per lane");
    constexpr int HALO = ORD / 2;  // the outermost HALO layers are never updated
    // wave threadIdx.y owns one row segment of the tile (wx segments side by side)
    const int xb = (tl.bx * a.wx + (int)threadIdx.y % a.wx) * 64 * N;
    const int iy = tl.by * (a.ty / a.wx) + (int)threadIdx.y / a.wx;
    if (iy < HALO || iy > a.ny - 1 - HALO || xb > a.nx - 1 - HALO) return;  // the whole wave (one row)
    const int z0 = a.zb + tl.bz * a.zc;
    const int z1 = min(z0 + a.zc, a.ze);
    if (z0 >= z1) return;
    const Lanes L = make_lanes<N>(xb, a.pitch);
    const int ix0 = L.ix0;
    bool upd[N], all = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        upd[j] = L.lv && ix0 + j >= HALO && ix0 + j <= a.nx - 1 - HALO;
        all = all && upd[j];
    }
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    unsigned o = (unsigned)ix0 + (unsigned)iy * W + (unsigned)z0 * PL;
    const Deriv<P> D(a.h);
    const R dt = a.dt;

    // lanes past the row end (wide segments only) read the row's last cells; their
    // values are never used
    auto LD = [&](int fid, unsigned off, R (&r)[N]) { FA<P>::template ldv<N>(a, fid, off - L.sh, r); };
    auto LD2 = [&](int fa, int fb, unsigned off, R (&ra)[N], R (&rb)[N]) { ld2v<P, N>(a, fa, fb, off - L.sh, ra, rb); };
    auto LDC = [&](int cid, unsigned off, R (&r)[N]) { P::template ldcv<N>(a.c[cid], off - L.sh, r); };
    auto ST = [&](int fid, unsigned off, const R (&r)[N]) {
        if (all) FA<P>::template stv<N>(a, fid, off, r);
        else {
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (upd[j]) FA<P>::st(a, fid, off + j, r[j]);
        }
    };
    auto ST2 = [&](int fa, int fb, unsigned off, const R (&ra)[N], const R (&rb)[N]) {
        if (all) st2v<P, N>(a, fa, fb, off, ra, rb);
        else {
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (upd[j]) {
                    FA<P>::st(a, fa, off + j, ra[j]);
                    FA<P>::st(a, fb, off + j, rb[j]);
                }
        }
    };

    const int ys = slab_index(iy, a.ny, a.lo[1], a.hi[1]);
    const int XW = a.lo[0] + a.hi[0], YW = a.lo[1] + a.hi[1];
    Tab4<R> ty = {0, 1, 0, 1};
    if (ys >= 0) ty = a.tab[1][iy];
    int xs[N];
#pragma unroll
    for (int j = 0; j < N; ++j) xs[j] = upd[j] ? slab_index(ix0 + j, a.nx, a.lo[0], a.hi[0]) : -1;

    // z carries.  Order 2: txz, tyz at z-1; tzz, p at z.  Order 4 also txz, tyz at
    // z-2 and z; tzz, p at z-1 and z+1.
    R txz_m[N], tyz_m[N], tzz_c[N], p_c[N];
    R txz_mm[N], tyz_mm[N], txz_k[N], tyz_k[N], tzz_m[N], p_m[N], tzz_k[N], p_k[N];
    // comp16f, order 2: its pairs (tyy, p) and (txx, tzz) are read a plane ahead,
    // so txx and tyy are carried too
    constexpr bool C16F = std::is_same<P, PolC16F>::value;
    constexpr bool PAIRS = C16F && ORD == 2;
    R txx_p[N], tyy_p[N];
    if constexpr (PAIRS) {
        R txy_m[N];
        LD(PW_TXZ, o - PL, txz_m);
        LD2(PW_TXY, PW_TYZ, o - PL, txy_m, tyz_m);
        LD2(PW_TYY, PW_P, o, tyy_p, p_c);
        LD2(PW_TXX, PW_TZZ, o, txx_p, tzz_c);
    } else {
        LD2(PW_TXZ, PW_TYZ, o - PL, txz_m, tyz_m);
        LD2(PW_TZZ, PW_P, o, tzz_c, p_c);
    }
    if constexpr (ORD == 4) {
        LD(PW_TXZ, o - 2 * PL, txz_mm);
        LD(PW_TYZ, o - 2 * PL, tyz_mm);
        LD(PW_TXZ, o, txz_k);
        LD(PW_TYZ, o, tyz_k);
        LD(PW_TZZ, o - PL, tzz_m);
        LD(PW_P, o - PL, p_m);
        LD(PW_TZZ, o + PL, tzz_k);
        LD(PW_P, o + PL, p_k);
    }
    // physical form: cell-centred rho, C1, C2 of this plane, carried up the column
    // (comp16f: its whole coefficient record, rhof too)
    constexpr bool REC = std::is_same<P, PolC16F>::value;
    R rho_c[N], c1_c[N], c2_c[N], rf_c[N];
    if constexpr (REC) {
        R t[4][N];
        ld_rec<4, N>(a, 0, o - L.sh, t);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            rho_c[j] = t[RV_RHO][j]; c1_c[j] = t[RV_C1][j]; c2_c[j] = t[RV_C2][j]; rf_c[j] = t[RV_RHOF][j];
        }
    } else if constexpr (!P::scaled) {
        LDC(PC_RHO, o, rho_c);
        LDC(PC_C1, o, c1_c);
        LDC(PC_C2, o, c2_c);
    }
    const R C4A = (R)1.125, C4B = (R)(-1.0 / 24.0);

    for (int iz = z0; iz < z1; ++iz, o += PL) {
        // undivided staggered differences, Float/main.cu:115-128 (order 2) or the
        // 4th-order pair 9/8 (f1 - f0) - 1/24 (f2 - f_-1)
        R dx1[N], dx2[N], dx3[N], ds1[N], dy1[N], dy2[N], dy3[N], ds2[N], dz1[N], dz2[N], dz3[N], ds3[N];
        R txz_c[N], tyz_c[N], tzz_n[N], p_n[N];
        if constexpr (ORD == 2) {
            R txx_c[N], txx_r[N], tyy_c[N], tyy_q[N], txy_c[N], txy_l[N], txy_h[N], txz_l[N], tyz_h[N], p_r[N],
                p_q[N];
            if constexpr (PAIRS) {
                R txx_n[N], tyy_n[N];
                LD(PW_TXZ, o, txz_c);
                LD2(PW_TXY, PW_TYZ, o, txy_c, tyz_c);
                LD2(PW_TYY, PW_P, o + PL, tyy_n, p_n);
                LD2(PW_TXX, PW_TZZ, o + PL, txx_n, tzz_n);
                LD2(PW_TYY, PW_P, o + W, tyy_q, p_q);
                LD2(PW_TXY, PW_TYZ, o - W, txy_h, tyz_h);
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    txx_c[j] = txx_p[j]; txx_p[j] = txx_n[j];
                    tyy_c[j] = tyy_p[j]; tyy_p[j] = tyy_n[j];
                }
                seg_right(L, txx_c, txx_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_TXX, q); }, [&](R (&r)[N]) { LD(PW_TXX, o + 1, r); });
                seg_left(L, txy_c, txy_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_TXY, q); }, [&](R (&r)[N]) { LD(PW_TXY, o - 1, r); });
                seg_left(L, txz_c, txz_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_TXZ, q); }, [&](R (&r)[N]) { LD(PW_TXZ, o - 1, r); });
                seg_right(L, p_c, p_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_P, q); }, [&](R (&r)[N]) { LD(PW_P, o + 1, r); });
            } else {
                LD2(PW_TXZ, PW_TYZ, o, txz_c, tyz_c);
                LD2(PW_TZZ, PW_P, o + PL, tzz_n, p_n);
                LD2(PW_TXX, PW_TYY, o, txx_c, tyy_c);
                seg_right(L, txx_c, txx_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_TXX, q); }, [&](R (&r)[N]) { LD(PW_TXX, o + 1, r); });
                LD(PW_TYY, o + W, tyy_q);
                LD(PW_TXY, o, txy_c);
                seg_left(L, txy_c, txy_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_TXY, q); }, [&](R (&r)[N]) { LD(PW_TXY, o - 1, r); });
                LD(PW_TXY, o - W, txy_h);
                seg_left(L, txz_c, txz_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_TXZ, q); }, [&](R (&r)[N]) { LD(PW_TXZ, o - 1, r); });
                LD(PW_TYZ, o - W, tyz_h);
                seg_right(L, p_c, p_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_P, q); }, [&](R (&r)[N]) { LD(PW_P, o + 1, r); });
                LD(PW_P, o + W, p_q);
            }
#pragma unroll
            for (int j = 0; j < N; ++j) {
                dx1[j] = txx_r[j] - txx_c[j]; dx2[j] = txy_c[j] - txy_h[j]; dx3[j] = txz_c[j] - txz_m[j];
                ds1[j] = p_r[j] - p_c[j];
                dy1[j] = tyy_q[j] - tyy_c[j]; dy2[j] = txy_c[j] - txy_l[j]; dy3[j] = tyz_c[j] - tyz_m[j];
                ds2[j] = p_q[j] - p_c[j];
                dz1[j] = tzz_n[j] - tzz_c[j]; dz2[j] = txz_c[j] - txz_l[j]; dz3[j] = tyz_c[j] - tyz_h[j];
                ds3[j] = p_n[j] - p_c[j];
            }
        } else {
            // N == 1: neighbours two cells out along x and y from L1/L2, along z from the carries
            auto F = [&](int fid, long d) { return FA<P>::ld(a, fid, (unsigned)((long)o + d)); };
            const long X = 1, Y = (long)W;
            R txz_p, tyz_p, tzz_nn, p_nn;
            txz_p = F(PW_TXZ, (long)PL);
            tyz_p = F(PW_TYZ, (long)PL);
            tzz_nn = F(PW_TZZ, 2 * (long)PL);
            p_nn = F(PW_P, 2 * (long)PL);
            const R txx0 = F(PW_TXX, 0), txy0 = F(PW_TXY, 0), tyy0 = F(PW_TYY, 0);
            // x neighbours: txx and p at x+1, x+2, x-1; txy and txz at x-1, x-2, x+1
            R txx_r1, txx_r2, txx_l1, p_r1, p_r2, p_l1, txy_l1, txy_l2, txy_r1, txz_l1, txz_l2, txz_r1;
            if constexpr (kX4Dpp<P>) {
                auto E = [&](int fid) { return [&a, fid](unsigned q) { return FA<P>::ld(a, fid, q); }; };
                lane_right2(L, txx0, o, E(PW_TXX), txx_r1, txx_r2);
                txx_l1 = lane_left1(L, txx0, o, E(PW_TXX));
                lane_right2(L, p_c[0], o, E(PW_P), p_r1, p_r2);
                p_l1 = lane_left1(L, p_c[0], o, E(PW_P));
                lane_left2(L, txy0, o, E(PW_TXY), txy_l1, txy_l2);
                txy_r1 = lane_right1(L, txy0, o, E(PW_TXY));
                lane_left2(L, txz_k[0], o, E(PW_TXZ), txz_l1, txz_l2);
                txz_r1 = lane_right1(L, txz_k[0], o, E(PW_TXZ));
            } else {
                txx_r1 = F(PW_TXX, X); txx_r2 = F(PW_TXX, 2 * X); txx_l1 = F(PW_TXX, -X);
                p_r1 = F(PW_P, X); p_r2 = F(PW_P, 2 * X); p_l1 = F(PW_P, -X);
                txy_l1 = F(PW_TXY, -X); txy_l2 = F(PW_TXY, -2 * X); txy_r1 = F(PW_TXY, X);
                txz_l1 = F(PW_TXZ, -X); txz_l2 = F(PW_TXZ, -2 * X); txz_r1 = F(PW_TXZ, X);
            }
            dx1[0] = C4A * (txx_r1 - txx0) + C4B * (txx_r2 - txx_l1);
            dx3[0] = C4A * (txz_k[0] - txz_m[0]) + C4B * (txz_p - txz_mm[0]);
            ds1[0] = C4A * (p_r1 - p_c[0]) + C4B * (p_r2 - p_l1);
            dy2[0] = C4A * (txy0 - txy_l1) + C4B * (txy_r1 - txy_l2);
            dy3[0] = C4A * (tyz_k[0] - tyz_m[0]) + C4B * (tyz_p - tyz_mm[0]);
            dz1[0] = C4A * (tzz_k[0] - tzz_c[0]) + C4B * (tzz_nn - tzz_m[0]);
            dz2[0] = C4A * (txz_k[0] - txz_l1) + C4B * (txz_r1 - txz_l2);
            ds3[0] = C4A * (p_k[0] - p_c[0]) + C4B * (p_nn - p_m[0]);
            if constexpr (std::is_same<P, PolC16F>::value) {
                // the y neighbours in comp16f's pairs: (txy, tyz) and (tyy, p) at
                // the same rows, one paired access each
                auto F2 = [&](int fa, int fb, long d, R& ra, R& rb) {
                    R x[1], y[1];
                    ld2v<P, 1>(a, fa, fb, (unsigned)((long)o + d), x, y);
                    ra = x[0];
                    rb = y[0];
                };
                R txy_ym, txy_yp, txy_ym2, tyz_ym, tyz_yp, tyz_ym2, tyy_yp, tyy_yp2, tyy_ym, p_yp, p_yp2, p_ym;
                F2(PW_TXY, PW_TYZ, -Y, txy_ym, tyz_ym);
                F2(PW_TXY, PW_TYZ, Y, txy_yp, tyz_yp);
                F2(PW_TXY, PW_TYZ, -2 * Y, txy_ym2, tyz_ym2);
                F2(PW_TYY, PW_P, Y, tyy_yp, p_yp);
                F2(PW_TYY, PW_P, 2 * Y, tyy_yp2, p_yp2);
                F2(PW_TYY, PW_P, -Y, tyy_ym, p_ym);
                dx2[0] = C4A * (txy0 - txy_ym) + C4B * (txy_yp - txy_ym2);
                dy1[0] = C4A * (tyy_yp - tyy0) + C4B * (tyy_yp2 - tyy_ym);
                ds2[0] = C4A * (p_yp - p_c[0]) + C4B * (p_yp2 - p_ym);
                dz3[0] = C4A * (tyz_k[0] - tyz_ym) + C4B * (tyz_yp - tyz_ym2);
            } else {
                dx2[0] = C4A * (txy0 - F(PW_TXY, -Y)) + C4B * (F(PW_TXY, Y) - F(PW_TXY, -2 * Y));
                dy1[0] = C4A * (F(PW_TYY, Y) - tyy0) + C4B * (F(PW_TYY, 2 * Y) - F(PW_TYY, -Y));
                ds2[0] = C4A * (F(PW_P, Y) - p_c[0]) + C4B * (F(PW_P, 2 * Y) - F(PW_P, -Y));
                dz3[0] = C4A * (tyz_k[0] - F(PW_TYZ, -Y)) + C4B * (F(PW_TYZ, Y) - F(PW_TYZ, -2 * Y));
            }
            // shift the z carries to the next plane
            txz_mm[0] = txz_m[0]; txz_m[0] = txz_k[0]; txz_k[0] = txz_p;
            tyz_mm[0] = tyz_m[0]; tyz_m[0] = tyz_k[0]; tyz_k[0] = tyz_p;
            tzz_m[0] = tzz_c[0]; tzz_c[0] = tzz_k[0]; tzz_k[0] = tzz_nn;
            p_m[0] = p_c[0]; p_c[0] = p_k[0]; p_k[0] = p_nn;
        }
        // The old velocities are loaded before the CPML memory-variable updates
        // below: those stores could alias them as far as the compiler knows, so
        // loads placed after the stores would wait for a second memory round trip
        // per plane.  (Hoisting the coefficients as well costs a wave per SIMD in
        // registers and measured slower.)
        R wx0[N], wy0[N], wz0[N], ux0[N], uy0[N], uz0[N];
        if constexpr (C16F) {
            LD2(PW_VWX, PW_VWZ, o, wx0, wz0);
            LD2(PW_VUY, PW_VWY, o, uy0, wy0);
            LD2(PW_VUX, PW_VUZ, o, ux0, uz0);
        } else {
            LD2(PW_VWX, PW_VWY, o, wx0, wy0);
            LD2(PW_VUZ, PW_VWZ, o, uz0, wz0);
            LD2(PW_VUX, PW_VUY, o, ux0, uy0);
        }
        const int zs = slab_index(iz, a.nz, a.lo[2], a.hi[2]);
        Tab4<R> tz = {0, 1, 0, 1};
        if (zs >= 0) tz = a.tab[2][iz];

        R sx[N], sy[N], sz[N], s1[N], s2[N], s3[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            // Float/main.cu:115-128
            R x1 = D(dx1[j]), x2 = D(dx2[j]), x3 = D(dx3[j]);
            R y1 = D(dy1[j]), y2 = D(dy2[j]), y3 = D(dy3[j]);
            R z1_ = D(dz1[j]), z2 = D(dz2[j]), z3 = D(dz3[j]);
            R q1 = D(ds1[j]), q2 = D(ds2[j]), q3 = D(ds3[j]);
            const unsigned ix = (unsigned)(ix0 + j);
            // terms 0, 1 on the half-point profile, 2, 3 on the full one
            if (xs[j] >= 0) {
                const Tab4<R> tx = a.tab[0][ix];
                const unsigned k = ((unsigned)iz * a.ny + iy) * XW + xs[j];
                cpml4<P>(a.m[0], k, {tx.ch, tx.ch, tx.cf, tx.cf}, {tx.eh, tx.eh, tx.ef, tx.ef}, {&x1, &q1, &y2, &z2},
                         a.zero);
            }
            if (ys >= 0 && upd[j]) {
                const unsigned k = ((unsigned)iz * YW + ys) * W + ix;
                cpml4<P>(a.m[1], k, {ty.ch, ty.ch, ty.cf, ty.cf}, {ty.eh, ty.eh, ty.ef, ty.ef}, {&y1, &q2, &x2, &z3},
                         a.zero);
            }
            if (zs >= 0 && upd[j]) {
                const unsigned k = ((unsigned)zs * a.ny + iy) * W + ix;
                cpml4<P>(a.m[2], k, {tz.ch, tz.ch, tz.cf, tz.cf}, {tz.eh, tz.eh, tz.ef, tz.ef}, {&z1_, &q3, &x3, &y3},
                         a.zero);
            }
            sx[j] = x1 + x2 + x3;
            sy[j] = y1 + y2 + y3;
            sz[j] = z1_ + z2 + z3;
            s1[j] = q1;
            s2[j] = q2;
            s3[j] = q3;
        }

        R wx1[N], wy1[N], wz1[N], ux1[N], uy1[N], uz1[N];
        if constexpr (!P::scaled) {
            R rho_n[N], rho_x[N], rho_y[N], c1_n[N], c1_x[N], c1_y[N], c2_n[N], c2_x[N], c2_y[N], rf[N];
            if constexpr (REC) {
                // the next plane's records (carried), the x+1 neighbours across
                // lanes, the y+1 row's records; rhof is this plane's
                R tn[4][N], ty[4][N];
                ld_rec<4, N>(a, 0, o + PL - L.sh, tn);
                seg_right(L, rho_c, rho_x, o + N, [&](unsigned q) { return ld_rec1<4>(a, 0, RV_RHO, q); }, [&](R (&r)[N]) { ld_rec_slot<4, N>(a, 0, RV_RHO, o + 1 - L.sh, r); });
                seg_right(L, c1_c, c1_x, o + N, [&](unsigned q) { return ld_rec1<4>(a, 0, RV_C1, q); }, [&](R (&r)[N]) { ld_rec_slot<4, N>(a, 0, RV_C1, o + 1 - L.sh, r); });
                seg_right(L, c2_c, c2_x, o + N, [&](unsigned q) { return ld_rec1<4>(a, 0, RV_C2, q); }, [&](R (&r)[N]) { ld_rec_slot<4, N>(a, 0, RV_C2, o + 1 - L.sh, r); });
                ld_rec<4, N>(a, 0, o + W - L.sh, ty);
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    rho_n[j] = tn[RV_RHO][j]; c1_n[j] = tn[RV_C1][j]; c2_n[j] = tn[RV_C2][j];
                    rho_y[j] = ty[RV_RHO][j]; c1_y[j] = ty[RV_C1][j]; c2_y[j] = ty[RV_C2][j];
                    rf[j] = rf_c[j];
                    rf_c[j] = tn[RV_RHOF][j];
                }
            } else {
            LDC(PC_RHO, o + PL, rho_n);
            seg_right(L, rho_c, rho_x, o + N, [&](unsigned q) { return P::ldc(a.c[PC_RHO], q); }, [&](R (&r)[N]) { LDC(PC_RHO, o + 1, r); });
            LDC(PC_RHO, o + W, rho_y);
            LDC(PC_C1, o + PL, c1_n);
            seg_right(L, c1_c, c1_x, o + N, [&](unsigned q) { return P::ldc(a.c[PC_C1], q); }, [&](R (&r)[N]) { LDC(PC_C1, o + 1, r); });
            LDC(PC_C1, o + W, c1_y);
            LDC(PC_C2, o + PL, c2_n);
            seg_right(L, c2_c, c2_x, o + N, [&](unsigned q) { return P::ldc(a.c[PC_C2], q); }, [&](R (&r)[N]) { LDC(PC_C2, o + 1, r); });
            LDC(PC_C2, o + W, c2_y);
            if constexpr (std::is_same<P, PolF32>::value || std::is_same<P, PolF64>::value) {
                if (a.c[PC_RHOF]) {
                    LDC(PC_RHOF, o, rf);
                } else {
#pragma unroll
                    for (int j = 0; j < N; ++j) rf[j] = a.rhof;
                }
            } else {
                LDC(PC_RHOF, o, rf);
            }
            }
            if constexpr (P::form == kFormFast) {
                // Float/main.cu:106-113, 160-189 in stored units, with the
                // staggered means folded: ir = 1/(rho + rho') (rho^-1 = 2 ir),
                // q2 = 2Q = C2 + C2' - 4 rhof^2 ir, h2 = (C1 + C1') dt/2, so
                // 2 CC1 = q2 + h2 and 2 CC2 = q2 - h2
                const R dt2 = a.kf[KF_DT2], kws = a.kf[KF_WS], kwp = a.kf[KF_WP], ku = a.kf[KF_U],
                        kuw = a.kf[KF_UW];
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const R rf4 = (R)4 * rf[j] * rf[j];
                    auto axis = [&](R r1, R c11, R c21, R S, R dp, R w0, R u0, R& w1, R& u1) {
                        const R ir = fast_rcp(rho_c[j] + r1);
                        const R sc1 = c1_c[j] + c11;
                        const R q2 = fmaf(-rf4, ir, c2_c[j] + c21);
                        const R h2 = sc1 * dt2;
                        const R af = rf[j] * ir;
                        const R num = fmaf(q2 - h2, w0, -fmaf(kws * af, S, kwp * dp));
                        w1 = sc1 == (R)0 ? (R)0 : num * fast_rcp(q2 + h2);
                        u1 = fmaf(ku * ir, S, u0) - kuw * af * (w1 - w0);
                    };
                    axis(rho_x[j], c1_x[j], c2_x[j], sx[j], s1[j], wx0[j], ux0[j], wx1[j], ux1[j]);
                    axis(rho_y[j], c1_y[j], c2_y[j], sy[j], s2[j], wy0[j], uy0[j], wy1[j], uy1[j]);
                    axis(rho_n[j], c1_n[j], c2_n[j], sz[j], s3[j], wz0[j], uz0[j], wz1[j], uz1[j]);
                    rho_c[j] = rho_n[j];
                    c1_c[j] = c1_n[j];
                    c2_c[j] = c2_n[j];
                }
            } else {
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const R rix = inv_mean(rho_c[j], rho_x[j]), riy = inv_mean(rho_c[j], rho_y[j]),
                        riz = inv_mean(rho_c[j], rho_n[j]);
                const R c1x = half_sum(c1_c[j], c1_x[j]), c1y = half_sum(c1_c[j], c1_y[j]),
                        c1z = half_sum(c1_c[j], c1_n[j]);
                const R c2x = half_sum(c2_c[j], c2_x[j]), c2y = half_sum(c2_c[j], c2_y[j]),
                        c2z = half_sum(c2_c[j], c2_n[j]);
                R cc1x, cc2x, cc1y, cc2y, cc1z, cc2z;
                cn_pair(c1x, c2x, rf[j], rix, dt, cc1x, cc2x);
                cn_pair(c1y, c2y, rf[j], riy, dt, cc1y, cc2y);
                cn_pair(c1z, c2z, rf[j], riz, dt, cc1z, cc2z);
                // Float/main.cu:160-189
                wx1[j] = c1x == (R)0 ? (R)0 : (wx0[j] * cc2x - dt * (rf[j] * rix * sx[j] + s1[j])) / cc1x;
                wy1[j] = c1y == (R)0 ? (R)0 : (wy0[j] * cc2y - dt * (rf[j] * riy * sy[j] + s2[j])) / cc1y;
                wz1[j] = c1z == (R)0 ? (R)0 : (wz0[j] * cc2z - dt * (rf[j] * riz * sz[j] + s3[j])) / cc1z;
                ux1[j] = ux0[j] + dt * rix * sx[j] - rf[j] * rix * (wx1[j] - wx0[j]);
                uy1[j] = uy0[j] + dt * riy * sy[j] - rf[j] * riy * (wy1[j] - wy0[j]);
                uz1[j] = uz0[j] + dt * riz * sz[j] - rf[j] * riz * (wz1[j] - wz0[j]);
                rho_c[j] = rho_n[j];
                c1_c[j] = c1_n[j];
                c2_c[j] = c2_n[j];
            }
            }
        } else {
            // Half/main.cu:243-262: vw = Aw*vw - Bt*S - Bp*dp (0 where Aw = Bp = Bt = 0);
            // vu += rho~^-1*S - rhof*rho^-1*(vw_new - vw_old), in the reference's
            // order of float operations, without contraction
#pragma clang fp contract(off)
            // {Aw, Bp, Bt, rho~^-1} per axis and {rhof rho^-1 x, y, z}: one
            // record each (comp16) or four arrays (BF16)
            auto ldq = [&](int k0, int k1, int k2, int k3, R (&v)[4][N]) {
                if constexpr (P::srec) {
                    ld_srec<N>(a.c[k0], o - L.sh, v);
                } else {
                    LDC(k0, o, v[0]); LDC(k1, o, v[1]); LDC(k2, o, v[2]);
                    if (k3 >= 0) LDC(k3, o, v[3]);
                }
            };
            R fr3[4][N], ca[4][N];
            R (&aw)[N] = ca[0];
            R (&bp)[N] = ca[1];
            R (&bt)[N] = ca[2];
            R (&ri)[N] = ca[3];
            ldq(SC_FRX, SC_FRY, SC_FRZ, -1, fr3);
            ldq(SC_AWX, SC_BPX, SC_BTX, SC_RITX, ca);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                wx1[j] = (aw[j] == 0.0f && bp[j] == 0.0f && bt[j] == 0.0f) ? 0.0f : aw[j] * wx0[j] - bt[j] * sx[j] - bp[j] * s1[j];
                ux1[j] = ux0[j] + ri[j] * sx[j] - fr3[0][j] * (wx1[j] - wx0[j]);
            }
            ldq(SC_AWY, SC_BPY, SC_BTY, SC_RITY, ca);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                wy1[j] = (aw[j] == 0.0f && bp[j] == 0.0f && bt[j] == 0.0f) ? 0.0f : aw[j] * wy0[j] - bt[j] * sy[j] - bp[j] * s2[j];
                uy1[j] = uy0[j] + ri[j] * sy[j] - fr3[1][j] * (wy1[j] - wy0[j]);
            }
            ldq(SC_AWZ, SC_BPZ, SC_BTZ, SC_RITZ, ca);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                wz1[j] = (aw[j] == 0.0f && bp[j] == 0.0f && bt[j] == 0.0f) ? 0.0f : aw[j] * wz0[j] - bt[j] * sz[j] - bp[j] * s3[j];
                uz1[j] = uz0[j] + ri[j] * sz[j] - fr3[2][j] * (wz1[j] - wz0[j]);
            }
        }
        if constexpr (C16F) {
            ST2(PW_VWX, PW_VWZ, o, wx1, wz1);
            ST2(PW_VUY, PW_VWY, o, uy1, wy1);
            ST2(PW_VUX, PW_VUZ, o, ux1, uz1);
        } else {
            ST2(PW_VWX, PW_VWY, o, wx1, wy1);
            ST2(PW_VUZ, PW_VWZ, o, uz1, wz1);
            ST2(PW_VUX, PW_VUY, o, ux1, uy1);
        }

        if constexpr (ORD == 2) {
#pragma unroll
            for (int j = 0; j < N; ++j) {
                txz_m[j] = txz_c[j];
                tyz_m[j] = tyz_c[j];
                tzz_c[j] = tzz_n[j];
                p_c[j] = p_n[j];
            }
        }
    }
}

// --------------------------------------------------------------------------
// Stress kernel: Float/main.cu:194-295 (physical), Half/main.cu:270-371 (scaled)
// --------------------------------------------------------------------------
template <class P, int N, int ORD>
__device__ __forceinline__ void stress_tile(const StepArgs<P>& a, const Tile tl)
{
    using R = typename P::R;
    static_assert(ORD == 2 || (ORD == 4 && N == 1), "4th order runs one cell per lane");
    constexpr int HALO = ORD / 2;
    // wave threadIdx.y owns one row segment of the tile (wx segments side by side)
    const int xb = (tl.bx * a.wx + (int)threadIdx.y % a.wx) * 64 * N;
    const int iy = tl.by * (a.ty / a.wx) + (int)threadIdx.y / a.wx;
    if (iy < HALO || iy > a.ny - 1 - HALO || xb > a.nx - 1 - HALO) return;  // the whole wave (one row)
    const int z0 = a.zb + tl.bz * a.zc;
    const int z1 = min(z0 + a.zc, a.ze);
    if (z0 >= z1) return;
    const Lanes L = make_lanes<N>(xb, a.pitch);
    const int ix0 = L.ix0;
    bool upd[N], all = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        upd[j] = L.lv && ix0 + j >= HALO && ix0 + j <= a.nx - 1 - HALO;
        all = all && upd[j];
    }
    const unsigned PL = a.plane, W = (unsigned)a.pitch;
    unsigned o = (unsigned)ix0 + (unsigned)iy * W + (unsigned)z0 * PL;
    const Deriv<P> D(a.h);
    const R dt = a.dt;

    // lanes past the row end (wide segments only) read the row's last cells; their
    // values are never used
    auto LD = [&](int fid, unsigned off, R (&r)[N]) { FA<P>::template ldv<N>(a, fid, off - L.sh, r); };
    auto LD2 = [&](int fa, int fb, unsigned off, R (&ra)[N], R (&rb)[N]) { ld2v<P, N>(a, fa, fb, off - L.sh, ra, rb); };
    auto LDC = [&](int cid, unsigned off, R (&r)[N]) { P::template ldcv<N>(a.c[cid], off - L.sh, r); };
    auto ST = [&](int fid, unsigned off, const R (&r)[N]) {
        if (all) FA<P>::template stv<N>(a, fid, off, r);
        else {
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (upd[j]) FA<P>::st(a, fid, off + j, r[j]);
        }
    };
    auto ST2 = [&](int fa, int fb, unsigned off, const R (&ra)[N], const R (&rb)[N]) {
        if (all) st2v<P, N>(a, fa, fb, off, ra, rb);
        else {
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (upd[j]) {
                    FA<P>::st(a, fa, off + j, ra[j]);
                    FA<P>::st(a, fb, off + j, rb[j]);
                }
        }
    };

    const int ys = slab_index(iy, a.ny, a.lo[1], a.hi[1]);
    const int XW = a.lo[0] + a.hi[0], YW = a.lo[1] + a.hi[1];
    Tab4<R> ty = {0, 1, 0, 1};
    if (ys >= 0) ty = a.tab[1][iy];
    int xs[N];
#pragma unroll
    for (int j = 0; j < N; ++j) xs[j] = upd[j] ? slab_index(ix0 + j, a.nx, a.lo[0], a.hi[0]) : -1;

    // z carries.  Order 2: vuz, vwz at z-1; vux, vuy at z.  Order 4 also vuz, vwz
    // at z-2 and z; vux, vuy at z-1 and z+1.
    R vuz_m[N], vwz_m[N], vux_c[N], vuy_c[N];
    R vuz_mm[N], vwz_mm[N], vuz_k[N], vwz_k[N], vux_m[N], vuy_m[N], vux_k[N], vuy_k[N];
    // comp16f, order 2: its pairs (vux, vuz) and (vuy, vwy) are read a plane
    // ahead, so vuz and vwy are carried too
    constexpr bool C16F = std::is_same<P, PolC16F>::value;
    constexpr bool PAIRS = C16F && ORD == 2;
    R vuz_p[N], vwy_p[N];
    if constexpr (PAIRS) {
        R vux_m1[N], vwx_m1[N];
        LD2(PW_VUX, PW_VUZ, o - PL, vux_m1, vuz_m);
        LD2(PW_VWX, PW_VWZ, o - PL, vwx_m1, vwz_m);
        LD2(PW_VUX, PW_VUZ, o, vux_c, vuz_p);
        LD2(PW_VUY, PW_VWY, o, vuy_c, vwy_p);
    } else {
        LD2(PW_VUZ, PW_VWZ, o - PL, vuz_m, vwz_m);
        LD2(PW_VUX, PW_VUY, o, vux_c, vuy_c);
    }
    if constexpr (ORD == 4) {
        LD(PW_VUZ, o - 2 * PL, vuz_mm);
        LD(PW_VWZ, o - 2 * PL, vwz_mm);
        LD(PW_VUZ, o, vuz_k);
        LD(PW_VWZ, o, vwz_k);
        LD(PW_VUX, o - PL, vux_m);
        LD(PW_VUY, o - PL, vuy_m);
        LD(PW_VUX, o + PL, vux_k);
        LD(PW_VUY, o + PL, vuy_k);
    }
    // physical form: cell-centred mu at (x, y), (x+1, y), (x, y+1) of this plane
    // (kFormFast: their reciprocals, and the cell's own mu in mv_c)
    constexpr bool FAST = P::form == kFormFast;
    auto MU = [](R v) -> R {
        if constexpr (FAST) return fast_rcp(v);
        else return v;
    };
    auto LDMU = [&](unsigned off, R (&r)[N]) {
        LDC(PC_MU, off, r);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = MU(r[j]);
    };
    auto EMU = [&](unsigned q) { return MU(P::ldc(a.c[PC_MU], q)); };
    R mu_c[N], mu_x[N], mu_y[N], mv_c[N];
    // comp16f: this plane's H - 2mu comes with mu in its record, carried
    constexpr bool REC = std::is_same<P, PolC16F>::value;
    R lam_c[N];
    auto EMUR = [&](unsigned q) { return MU(ld_rec1<2>(a, 2, RS_MU, q)); };
    auto LDMUR = [&](unsigned off, R (&r)[N]) {
        ld_rec_slot<2, N>(a, 2, RS_MU, off - L.sh, r);
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = MU(r[j]);
    };
    if constexpr (REC) {
        R t[2][N];
        ld_rec<2, N>(a, 2, o - L.sh, t);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            lam_c[j] = t[RS_H2U][j];
            mv_c[j] = t[RS_MU][j];
            mu_c[j] = MU(mv_c[j]);
        }
        seg_right(L, mu_c, mu_x, o + N, EMUR, [&](R (&r)[N]) { LDMUR(o + 1, r); });
        LDMUR(o + W, mu_y);
    } else if constexpr (!P::scaled) {
        LDC(PC_MU, o, mv_c);
#pragma unroll
        for (int j = 0; j < N; ++j) mu_c[j] = MU(mv_c[j]);
        seg_right(L, mu_c, mu_x, o + N, EMU, [&](R (&r)[N]) { LDMU(o + 1, r); });
        LDMU(o + W, mu_y);
    }
    const R C4A = (R)1.125, C4B = (R)(-1.0 / 24.0);

    for (int iz = z0; iz < z1; ++iz, o += PL) {
        // undivided staggered differences, Float/main.cu:234-249 (order 2) or the
        // 4th-order pair
        R dxx[N], dyy[N], dzz[N], ewx[N], ewy[N], ewz[N], dxy[N], dyx[N], dxz[N], dzx[N], dyz[N], dzy[N];
        R vux_n[N], vuy_n[N], vuz_c[N], vwz_c[N];
        if constexpr (ORD == 2) {
            R vux_l[N], vux_q[N], vuy_h[N], vuy_r[N], vuz_r[N], vuz_q[N], vwx_c[N], vwx_l[N], vwy_c[N], vwy_h[N];
            if constexpr (PAIRS) {
                R vuz_n[N], vwy_n[N];
                LD2(PW_VUX, PW_VUZ, o + PL, vux_n, vuz_n);
                LD2(PW_VUY, PW_VWY, o + PL, vuy_n, vwy_n);
                LD2(PW_VWX, PW_VWZ, o, vwx_c, vwz_c);
                LD2(PW_VUX, PW_VUZ, o + W, vux_q, vuz_q);
                LD2(PW_VUY, PW_VWY, o - W, vuy_h, vwy_h);
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    vuz_c[j] = vuz_p[j]; vuz_p[j] = vuz_n[j];
                    vwy_c[j] = vwy_p[j]; vwy_p[j] = vwy_n[j];
                }
                seg_left(L, vux_c, vux_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_VUX, q); }, [&](R (&r)[N]) { LD(PW_VUX, o - 1, r); });
                seg_right(L, vuy_c, vuy_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_VUY, q); }, [&](R (&r)[N]) { LD(PW_VUY, o + 1, r); });
                seg_right(L, vuz_c, vuz_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_VUZ, q); }, [&](R (&r)[N]) { LD(PW_VUZ, o + 1, r); });
                seg_left(L, vwx_c, vwx_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_VWX, q); }, [&](R (&r)[N]) { LD(PW_VWX, o - 1, r); });
            } else {
                LD2(PW_VUX, PW_VUY, o + PL, vux_n, vuy_n);
                LD2(PW_VUZ, PW_VWZ, o, vuz_c, vwz_c);
                seg_left(L, vux_c, vux_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_VUX, q); }, [&](R (&r)[N]) { LD(PW_VUX, o - 1, r); });
                LD(PW_VUX, o + W, vux_q);
                LD(PW_VUY, o - W, vuy_h);
                seg_right(L, vuy_c, vuy_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_VUY, q); }, [&](R (&r)[N]) { LD(PW_VUY, o + 1, r); });
                seg_right(L, vuz_c, vuz_r, o + N, [&](unsigned q) { return FA<P>::ld(a, PW_VUZ, q); }, [&](R (&r)[N]) { LD(PW_VUZ, o + 1, r); });
                LD(PW_VUZ, o + W, vuz_q);
                LD2(PW_VWX, PW_VWY, o, vwx_c, vwy_c);
                seg_left(L, vwx_c, vwx_l, o - 1, [&](unsigned q) { return FA<P>::ld(a, PW_VWX, q); }, [&](R (&r)[N]) { LD(PW_VWX, o - 1, r); });
                LD(PW_VWY, o - W, vwy_h);
            }
#pragma unroll
            for (int j = 0; j < N; ++j) {
                dxx[j] = vux_c[j] - vux_l[j]; dyy[j] = vuy_c[j] - vuy_h[j]; dzz[j] = vuz_c[j] - vuz_m[j];
                ewx[j] = vwx_c[j] - vwx_l[j]; ewy[j] = vwy_c[j] - vwy_h[j]; ewz[j] = vwz_c[j] - vwz_m[j];
                dxy[j] = vux_q[j] - vux_c[j]; dyx[j] = vuy_r[j] - vuy_c[j];
                dxz[j] = vux_n[j] - vux_c[j]; dzx[j] = vuz_r[j] - vuz_c[j];
                dyz[j] = vuy_n[j] - vuy_c[j]; dzy[j] = vuz_q[j] - vuz_c[j];
            }
        } else {
            auto F = [&](int fid, long d) { return FA<P>::ld(a, fid, (unsigned)((long)o + d)); };
            const long X = 1, Y = (long)W;
            const R vuz_p = F(PW_VUZ, (long)PL), vwz_p = F(PW_VWZ, (long)PL);
            const R vux_nn = F(PW_VUX, 2 * (long)PL), vuy_nn = F(PW_VUY, 2 * (long)PL);
            const R vwx0 = F(PW_VWX, 0), vwy0 = F(PW_VWY, 0);
            // x neighbours: vux and vwx at x-1, x-2, x+1; vuy and vuz at x+1, x+2, x-1
            R vux_l1, vux_l2, vux_r1, vwx_l1, vwx_l2, vwx_r1, vuy_r1, vuy_r2, vuy_l1, vuz_r1, vuz_r2, vuz_l1;
            if constexpr (kX4Dpp<P>) {
                auto E = [&](int fid) { return [&a, fid](unsigned q) { return FA<P>::ld(a, fid, q); }; };
                lane_left2(L, vux_c[0], o, E(PW_VUX), vux_l1, vux_l2);
                vux_r1 = lane_right1(L, vux_c[0], o, E(PW_VUX));
                lane_left2(L, vwx0, o, E(PW_VWX), vwx_l1, vwx_l2);
                vwx_r1 = lane_right1(L, vwx0, o, E(PW_VWX));
                lane_right2(L, vuy_c[0], o, E(PW_VUY), vuy_r1, vuy_r2);
                vuy_l1 = lane_left1(L, vuy_c[0], o, E(PW_VUY));
                lane_right2(L, vuz_k[0], o, E(PW_VUZ), vuz_r1, vuz_r2);
                vuz_l1 = lane_left1(L, vuz_k[0], o, E(PW_VUZ));
            } else {
                vux_l1 = F(PW_VUX, -X); vux_l2 = F(PW_VUX, -2 * X); vux_r1 = F(PW_VUX, X);
                vwx_l1 = F(PW_VWX, -X); vwx_l2 = F(PW_VWX, -2 * X); vwx_r1 = F(PW_VWX, X);
                vuy_r1 = F(PW_VUY, X); vuy_r2 = F(PW_VUY, 2 * X); vuy_l1 = F(PW_VUY, -X);
                vuz_r1 = F(PW_VUZ, X); vuz_r2 = F(PW_VUZ, 2 * X); vuz_l1 = F(PW_VUZ, -X);
            }
            dxx[0] = C4A * (vux_c[0] - vux_l1) + C4B * (vux_r1 - vux_l2);
            dzz[0] = C4A * (vuz_k[0] - vuz_m[0]) + C4B * (vuz_p - vuz_mm[0]);
            ewx[0] = C4A * (vwx0 - vwx_l1) + C4B * (vwx_r1 - vwx_l2);
            ewz[0] = C4A * (vwz_k[0] - vwz_m[0]) + C4B * (vwz_p - vwz_mm[0]);
            dyx[0] = C4A * (vuy_r1 - vuy_c[0]) + C4B * (vuy_r2 - vuy_l1);
            dxz[0] = C4A * (vux_k[0] - vux_c[0]) + C4B * (vux_nn - vux_m[0]);
            dzx[0] = C4A * (vuz_r1 - vuz_k[0]) + C4B * (vuz_r2 - vuz_l1);
            dyz[0] = C4A * (vuy_k[0] - vuy_c[0]) + C4B * (vuy_nn - vuy_m[0]);
            if constexpr (std::is_same<P, PolC16F>::value) {
                // the y neighbours in comp16f's pairs: (vuy, vwy) and (vux, vuz) at
                // the same rows, one paired access each
                auto F2 = [&](int fa, int fb, long d, R& ra, R& rb) {
                    R x[1], y[1];
                    ld2v<P, 1>(a, fa, fb, (unsigned)((long)o + d), x, y);
                    ra = x[0];
                    rb = y[0];
                };
                R vuy_ym, vuy_yp, vuy_ym2, vwy_ym, vwy_yp, vwy_ym2, vux_yp, vux_yp2, vux_ym, vuz_yp, vuz_yp2, vuz_ym;
                F2(PW_VUY, PW_VWY, -Y, vuy_ym, vwy_ym);
                F2(PW_VUY, PW_VWY, Y, vuy_yp, vwy_yp);
                F2(PW_VUY, PW_VWY, -2 * Y, vuy_ym2, vwy_ym2);
                F2(PW_VUX, PW_VUZ, Y, vux_yp, vuz_yp);
                F2(PW_VUX, PW_VUZ, 2 * Y, vux_yp2, vuz_yp2);
                F2(PW_VUX, PW_VUZ, -Y, vux_ym, vuz_ym);
                dyy[0] = C4A * (vuy_c[0] - vuy_ym) + C4B * (vuy_yp - vuy_ym2);
                ewy[0] = C4A * (vwy0 - vwy_ym) + C4B * (vwy_yp - vwy_ym2);
                dxy[0] = C4A * (vux_yp - vux_c[0]) + C4B * (vux_yp2 - vux_ym);
                dzy[0] = C4A * (vuz_yp - vuz_k[0]) + C4B * (vuz_yp2 - vuz_ym);
            } else {
                dyy[0] = C4A * (vuy_c[0] - F(PW_VUY, -Y)) + C4B * (F(PW_VUY, Y) - F(PW_VUY, -2 * Y));
                ewy[0] = C4A * (vwy0 - F(PW_VWY, -Y)) + C4B * (F(PW_VWY, Y) - F(PW_VWY, -2 * Y));
                dxy[0] = C4A * (F(PW_VUX, Y) - vux_c[0]) + C4B * (F(PW_VUX, 2 * Y) - F(PW_VUX, -Y));
                dzy[0] = C4A * (F(PW_VUZ, Y) - vuz_k[0]) + C4B * (F(PW_VUZ, 2 * Y) - F(PW_VUZ, -Y));
            }
            vuz_mm[0] = vuz_m[0]; vuz_m[0] = vuz_k[0]; vuz_k[0] = vuz_p;
            vwz_mm[0] = vwz_m[0]; vwz_m[0] = vwz_k[0]; vwz_k[0] = vwz_p;
            vux_m[0] = vux_c[0]; vux_c[0] = vux_k[0]; vux_k[0] = vux_nn;
            vuy_m[0] = vuy_c[0]; vuy_c[0] = vuy_k[0]; vuy_k[0] = vuy_nn;
        }
        const int zs = slab_index(iz, a.nz, a.lo[2], a.hi[2]);
        Tab4<R> tz = {0, 1, 0, 1};
        if (zs >= 0) tz = a.tab[2][iz];

        R p0[N], xx0[N], yy0[N], zz0[N], xy0[N], yz0[N], xz0[N];
        if constexpr (C16F) {
            LD2(PW_TYY, PW_P, o, yy0, p0);
            LD2(PW_TXY, PW_TYZ, o, xy0, yz0);
            LD2(PW_TXX, PW_TZZ, o, xx0, zz0);
            LD(PW_TXZ, o, xz0);
        } else {
            LD2(PW_TZZ, PW_P, o, zz0, p0);
            LD2(PW_TXX, PW_TYY, o, xx0, yy0);
            LD(PW_TXY, o, xy0);
            LD2(PW_TXZ, PW_TYZ, o, xz0, yz0);
        }
        R p1[N], xx1[N], yy1[N], zz1[N], xy1[N], yz1[N], xz1[N];
        R mu_n[N], mu_xn[N], mu_yn[N], mu_xy[N], mv_n[N];
        R C_[N], M_[N], HH_[N], H2u_[N], mxy[N], myz[N], mxz[N];
        if constexpr (REC) {
            R tn[2][N], tc[2][N];
            ld_rec<2, N>(a, 2, o + PL - L.sh, tn);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                mv_n[j] = tn[RS_MU][j];
                mu_n[j] = MU(mv_n[j]);
            }
            seg_right(L, mu_n, mu_xn, o + PL + N, EMUR, [&](R (&r)[N]) { LDMUR(o + PL + 1, r); });
            LDMUR(o + PL + W, mu_yn);
            seg_right(L, mu_y, mu_xy, o + W + N, EMUR, [&](R (&r)[N]) { LDMUR(o + W + 1, r); });
            ld_rec<2, N>(a, 1, o - L.sh, tc);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                C_[j] = tc[RS_C][j]; M_[j] = tc[RS_M][j]; H2u_[j] = lam_c[j];
                lam_c[j] = tn[RS_H2U][j];
            }
        } else if constexpr (!P::scaled) {
            LDC(PC_MU, o + PL, mv_n);
#pragma unroll
            for (int j = 0; j < N; ++j) mu_n[j] = MU(mv_n[j]);
            seg_right(L, mu_n, mu_xn, o + PL + N, EMU, [&](R (&r)[N]) { LDMU(o + PL + 1, r); });
            LDMU(o + PL + W, mu_yn);
            seg_right(L, mu_y, mu_xy, o + W + N, EMU, [&](R (&r)[N]) { LDMU(o + W + 1, r); });
            LDC(PC_C, o, C_); LDC(PC_M, o, M_); LDC(PC_H2U, o, H2u_);
            // kFormFast forms H as (H - 2 mu) + 2 mu, and so does FP32 when the
            // medium's H is that sum to rounding (c[PC_HH] null): one
            // coefficient array less to stream.  FP64 reads H, to stay
            // bit-identical to the FP64 oracle.
            if constexpr (std::is_same<P, PolF32>::value) {
                if (a.c[PC_HH]) {
                    LDC(PC_HH, o, HH_);
                } else {
#pragma unroll
                    for (int j = 0; j < N; ++j) HH_[j] = H2u_[j] + (R)2 * mu_c[j];
                }
            } else if constexpr (!FAST) {
                LDC(PC_HH, o, HH_);
            }
        } else {
            // {H~, H~ - 2mu~, C~, M~} and {mu~xy, mu~xz, mu~yz}: one record
            // each (comp16) or an array each (BF16)
            R t4[4][N], m4[4][N];
            if constexpr (P::srec) {
                ld_srec<N>(a.c[SC_HH], o - L.sh, t4);
                ld_srec<N>(a.c[SC_MUXY], o - L.sh, m4);
            } else {
                LDC(SC_HH, o, t4[0]); LDC(SC_H2U, o, t4[1]); LDC(SC_C, o, t4[2]); LDC(SC_M, o, t4[3]);
                LDC(SC_MUXY, o, m4[0]); LDC(SC_MUXZ, o, m4[1]); LDC(SC_MUYZ, o, m4[2]);
            }
#pragma unroll
            for (int j = 0; j < N; ++j) {
                HH_[j] = t4[0][j]; H2u_[j] = t4[1][j]; C_[j] = t4[2][j]; M_[j] = t4[3][j];
                mxy[j] = m4[0][j]; mxz[j] = m4[1][j]; myz[j] = m4[2][j];
            }
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            // Float/main.cu:234-249
            R uxx = D(dxx[j]), uyy = D(dyy[j]), uzz = D(dzz[j]);
            R wx = D(ewx[j]), wy = D(ewy[j]), wz = D(ewz[j]);
            R uxy = D(dxy[j]), uyx = D(dyx[j]);
            R uxz = D(dxz[j]), uzx = D(dzx[j]);
            R uyz = D(dyz[j]), uzy = D(dzy[j]);
            const unsigned ix = (unsigned)(ix0 + j);
            // terms 4, 5 on the full profile, 6, 7 on the half-point one
            if (xs[j] >= 0) {
                const Tab4<R> tx = a.tab[0][ix];
                const unsigned k = ((unsigned)iz * a.ny + iy) * XW + xs[j];
                cpml4<P>(a.m[0] + 4, k, {tx.cf, tx.cf, tx.ch, tx.ch}, {tx.ef, tx.ef, tx.eh, tx.eh},
                         {&uxx, &wx, &uyx, &uzx}, a.zero);
            }
            if (ys >= 0 && upd[j]) {
                const unsigned k = ((unsigned)iz * YW + ys) * W + ix;
                cpml4<P>(a.m[1] + 4, k, {ty.cf, ty.cf, ty.ch, ty.ch}, {ty.ef, ty.ef, ty.eh, ty.eh},
                         {&uyy, &wy, &uxy, &uzy}, a.zero);
            }
            if (zs >= 0 && upd[j]) {
                const unsigned k = ((unsigned)zs * a.ny + iy) * W + ix;
                cpml4<P>(a.m[2] + 4, k, {tz.cf, tz.cf, tz.ch, tz.ch}, {tz.ef, tz.ef, tz.eh, tz.eh},
                         {&uzz, &wz, &uxz, &uyz}, a.zero);
            }
            const R du = uxx + uyy + uzz, dw = wx + wy + wz;
            if constexpr (FAST) {
                // Float/main.cu:287-293 in stored units: the shear moduli are
                // 4 / (sum of the four reciprocals) (0 when any mu is 0: its
                // reciprocal is inf), H uxx + (H - 2mu)(uyy + uzz) is
                // (H - 2mu) du + 2mu uxx, and the relative-velocity divergence is
                // brought to velocity units by KF_W
                const R g = a.kf[KF_G], g2 = a.kf[KF_G2], g4 = a.kf[KF_G4];
                const R dwu = dw * a.kf[KF_W];
                const R C = C_[j], M = M_[j], lam = H2u_[j];
                const R com = g * fmaf(lam, du, C * dwu);
                const R gm = g2 * mv_c[j];
                p1[j] = fmaf(-g, fmaf(M, dwu, C * du), p0[j]);
                xx1[j] = fmaf(gm, uxx, xx0[j] + com);
                yy1[j] = fmaf(gm, uyy, yy0[j] + com);
                zz1[j] = fmaf(gm, uzz, zz0[j] + com);
                xy1[j] = fmaf(g4 * fast_rcp(mu_c[j] + mu_x[j] + mu_y[j] + mu_xy[j]), uxy + uyx, xy0[j]);
                yz1[j] = fmaf(g4 * fast_rcp(mu_c[j] + mu_y[j] + mu_n[j] + mu_yn[j]), uyz + uzy, yz0[j]);
                xz1[j] = fmaf(g4 * fast_rcp(mu_c[j] + mu_x[j] + mu_n[j] + mu_xn[j]), uxz + uzx, xz0[j]);
                mu_c[j] = mu_n[j];
                mu_x[j] = mu_xn[j];
                mu_y[j] = mu_yn[j];
                mv_c[j] = mv_n[j];
            } else if constexpr (!P::scaled) {
                // Float/main.cu:480-491
                const R muxz = harm4(mu_c[j], mu_x[j], mu_n[j], mu_xn[j]);
                const R muyz = harm4(mu_c[j], mu_y[j], mu_n[j], mu_yn[j]);
                const R muxy = harm4(mu_c[j], mu_x[j], mu_y[j], mu_xy[j]);
                const R C = C_[j], M = M_[j], HH = HH_[j], H2u = H2u_[j];
                // Float/main.cu:287-293
                p1[j] = p0[j] - dt * (C * du + M * dw);
                xx1[j] = xx0[j] + dt * (H2u * (uyy + uzz) + HH * uxx + C * dw);
                yy1[j] = yy0[j] + dt * (H2u * (uxx + uzz) + HH * uyy + C * dw);
                zz1[j] = zz0[j] + dt * (H2u * (uxx + uyy) + HH * uzz + C * dw);
                xy1[j] = xy0[j] + muxy * dt * (uxy + uyx);
                yz1[j] = yz0[j] + muyz * dt * (uyz + uzy);
                xz1[j] = xz0[j] + muxz * dt * (uxz + uzx);
                mu_c[j] = mu_n[j];
                mu_x[j] = mu_xn[j];
                mu_y[j] = mu_yn[j];
            } else {
                // Half/main.cu:362-370, without contraction
#pragma clang fp contract(off)
                const R C = C_[j], M = M_[j], HH = HH_[j], H2u = H2u_[j];
                p1[j] = p0[j] - C * du - M * dw;
                xx1[j] = xx0[j] + H2u * (uyy + uzz) + HH * uxx + C * dw;
                yy1[j] = yy0[j] + H2u * (uxx + uzz) + HH * uyy + C * dw;
                zz1[j] = zz0[j] + H2u * (uxx + uyy) + HH * uzz + C * dw;
                xy1[j] = xy0[j] + mxy[j] * (uxy + uyx);
                yz1[j] = yz0[j] + myz[j] * (uyz + uzy);
                xz1[j] = xz0[j] + mxz[j] * (uxz + uzx);
            }
        }
        if constexpr (C16F) {
            ST2(PW_TYY, PW_P, o, yy1, p1);
            ST2(PW_TXY, PW_TYZ, o, xy1, yz1);
            ST2(PW_TXX, PW_TZZ, o, xx1, zz1);
            ST(PW_TXZ, o, xz1);
        } else {
            ST2(PW_TZZ, PW_P, o, zz1, p1);
            ST2(PW_TXX, PW_TYY, o, xx1, yy1);
            ST(PW_TXY, o, xy1);
            ST2(PW_TXZ, PW_TYZ, o, xz1, yz1);
        }

        if constexpr (ORD == 2) {
#pragma unroll
            for (int j = 0; j < N; ++j) {
                vuz_m[j] = vuz_c[j];
                vwz_m[j] = vwz_c[j];
                vux_c[j] = vux_n[j];
                vuy_c[j] = vuy_n[j];
            }
        }
    }
}

// (An occupancy hint of 5 waves per SIMD for comp16f, whose kernels take 102 and
// 124 VGPRs, spills 12 and 33 of them to scratch and measured 8 % and 3x
// slower: the register allocation is left to the compiler.)
template <class P, int N, int ORD> __global__ __launch_bounds__(512) void k_velocity(const StepArgs<P> a)
{
    velocity_tile<P, N, ORD>(a, tile_of_block(a.gx, a.gy, a.nblocks, a.xcd_map));
}
template <class P, int N, int ORD> __global__ __launch_bounds__(512) void k_stress(const StepArgs<P> a)
{
    stress_tile<P, N, ORD>(a, tile_of_block(a.gx, a.gy, a.nblocks, a.xcd_map));
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
        const unsigned long long n = (ax == 0 ? nmem0 : (ax == 1 ? nmem1 : nmem2)) * mem_width<P>();
        typename P::Mem* m = a.m[ax][t];
        if (!m) return;
        for (; i < n; i += stride) P::stm(m, (unsigned)i, P::ldm(m, (unsigned)i) * ratio, a.zero);
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

// Tables nobody writes during a launch, read through the constant address
// space: the compiler then loads them into scalar registers.  Read as ordinary
// memory after the stores of a tile loop they came through vector loads, and
// the descriptor's constants took ~50 VGPRs of the tile code's.
// (The host pass, which only parses the kernels, sees a plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using cptr = const __attribute__((address_space(4))) T*;
#else
template <class T> using cptr = const T*;
#endif

template <class T> __device__ __forceinline__ int find_shot_c(cptr<T> start, int n, unsigned t)
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
    const cptr<ShotDesc<P>> D = (cptr<ShotDesc<P>>)c.descs;
    if (!c.xcd_map) {
        const unsigned t = blockIdx.x;
        s = find_shot_c((cptr<unsigned>)c.bstart, c.n, t);
        local = t - ((cptr<unsigned>)c.bstart)[s];
        return true;
    }
    const unsigned x = blockIdx.x % 8u, j = blockIdx.x / 8u;
    const cptr<unsigned> xs = (cptr<unsigned>)(c.xstart + x * (unsigned)(c.n + 1));
    if (j >= xs[c.n]) return false;
    s = find_shot_c(xs, c.n, j);
    const unsigned nb = (unsigned)D[s].a.nblocks, q = nb / 8u, r = nb % 8u;
    local = x * q + (x < r ? x : r) + (j - xs[s]);
    return true;
}

template <class P, int N> __global__ __launch_bounds__(512) void k_velocity_batch(const BatchArgs c, int k)
{
    int s;
    unsigned t;
    if (!batch_block<P>(c, s, t)) return;
    const cptr<ShotDesc<P>> d = (cptr<ShotDesc<P>>)c.descs + s;
    if (((cptr<int>)c.it)[s] + k >= d->nt) return;
    const StepArgs<P> a = d->a;
    velocity_tile<P, N, 2>(a, tile_of_index(t, a.gx, a.gy));
}

template <class P, int N> __global__ __launch_bounds__(512) void k_stress_batch(const BatchArgs c, int k)
{
    int s;
    unsigned t;
    if (!batch_block<P>(c, s, t)) return;
    const cptr<ShotDesc<P>> d = (cptr<ShotDesc<P>>)c.descs + s;
    if (((cptr<int>)c.it)[s] + k >= d->nt) return;
    const StepArgs<P> a = d->a;
    stress_tile<P, N, 2>(a, tile_of_index(t, a.gx, a.gy));
}

// Source channel ch (0..3: txx, tyy, tzz, p) of one shot at step it; the
// arithmetic of k_source
// (Dsc: ShotDesc<P>, or the same read through the constant address space)
template <class P, class Dsc> __device__ __forceinline__ void source_channel(const Dsc& d, int it, int ch)
{
    using R = typename P::R;
    if (it >= d.nt || !d.has_src || d.w[ch] == (R)0) return;
    const int fid[4] = {PW_TXX, PW_TYY, PW_TZZ, PW_P};
    const typename P::Field f = d.a.f[fid[ch]];
    const R amp = d.amp[it];
    P::st(f, d.src_off, P::ld(f, d.src_off) + d.w[ch] * amp);
}

// Receiver q of one shot at step it; the arithmetic of k_gather
template <class P, class Dsc> __device__ __forceinline__ void gather_one(const Dsc& d, int it, unsigned q)
{
    if (it >= d.nt) return;
    const int fid = d.rfield[q];
    const float v = (float)ld_any<P>(d.a, fid, d.roff[q]);
    const float s_stress = d.oscale[2 * it], s_vel = d.oscale[2 * it + 1];
    d.seis[(size_t)q * d.nt + it] = (fid <= PW_VWZ ? s_vel : s_stress) * v;
}

// one thread per (shot, source channel)
template <class P> __global__ void k_source_batch(const BatchArgs c, int k)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = g >> 2, ch = g & 3;
    if (s >= c.n) return;
    source_channel<P>(((const ShotDesc<P>*)c.descs)[s], c.it[s] + k, ch);
}

// Multiplies array arr (0..12 wavefields, then the memory variables) of one
// shot by ratio over the indices i0, i0 + stride, ...  (k_rescale's arithmetic)
// (A: StepArgs<P>, or the same read through the constant address space)
template <class P, class A>
__device__ __forceinline__ void rescale_array(const A& a, int arr, unsigned long long i0, unsigned long long stride,
                                              unsigned long long ncell, const unsigned long long nmem[3], float ratio)
{
    if (arr < PW_NFIELDS) {
        const typename P::Field f = a.f[arr];
        for (unsigned long long i = i0; i < ncell; i += stride) P::st(f, (unsigned)i, P::ld(f, (unsigned)i) * ratio);
    } else {
        const int q = arr - PW_NFIELDS;
        const int ax = q / kMemPerAxis, t = q % kMemPerAxis;
        typename P::Mem* m = a.m[ax][t];
        if (!m) return;
        const unsigned long long n = nmem[ax] * mem_width<P>();  // the records' values
        for (unsigned long long i = i0; i < n; i += stride) P::stm(m, (unsigned)i, P::ldm(m, (unsigned)i) * ratio, a.zero);
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

// one thread per receiver of the batch
template <class P> __global__ void k_gather_batch(const BatchArgs c, int k)
{
    const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= c.nrec) return;
    const int s = find_shot(c.rstart, c.n, r);
    gather_one<P>(((const ShotDesc<P>*)c.descs)[s], c.it[s] + k, r - c.rstart[s]);
}

__global__ void k_advance(int* it, const int* __restrict__ nt, int n, int K)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) it[s] = min(it[s] + K, nt[s]);
}

// --------------------------------------------------------------------------
// Persistent round: ONE cooperative launch advances every shot of a batch by K
// steps.  The grid is as many blocks as fit on the device at once; between the
// phases of a step the whole grid meets at a grid barrier (cooperative groups),
// which orders the phases as the separate launches of batch_step do:
//   [receivers of step k-1, then the source of step k]  (per shot, one block)
//   velocity tiles of step k  |  stress tiles of step k
// A step at which some shot rescales (comp16, BF16) takes the receivers, the
// rescale and the source as three phases instead.  Each block walks the tiles
// of its XCD's share in turn (tiles j0, j0 + blocks per XCD, ...), so the
// blocks of one XCD sweep a contiguous window of its tiles as the batched
// launches do.  Every block passes the same barriers: the loop bounds and the
// rescale test depend only on kernel arguments and on tables nobody writes
// during the launch.  The per-cell code is that of the other step kernels, so
// results are bit-identical to them.
// --------------------------------------------------------------------------
template <class P> __device__ __forceinline__ bool batch_rescales(const BatchArgs& c, int k)
{
    if constexpr (!P::scaled) {
        return false;
    } else {
        const cptr<ShotDesc<P>> D = (cptr<ShotDesc<P>>)c.descs;
        const cptr<int> its = (cptr<int>)c.it;
        for (int s = 0; s < c.n; ++s) {
            const int it = its[s] + k;
            if (it < D[s].nt && D[s].ratio && D[s].ratio[it] != 0.0f) return true;
        }
        return false;
    }
}

// One step tile of one shot (VEL: velocity, else stress).
template <class P, int N, bool VEL> __device__ __forceinline__ void round_tile(cptr<ShotDesc<P>> d, unsigned local)
{
    const StepArgs<P> a = d->a;
    const Tile tl = tile_of_index(local, a.gx, a.gy);
    if constexpr (VEL) velocity_tile<P, N, 2>(a, tl);
    else stress_tile<P, N, 2>(a, tl);
}

// The step tiles of one phase that this block owns.  One call site for the
// tile body: two inlined copies (one per map) measured 124 VGPRs instead of 78.
template <class P, int N, bool VEL> __device__ __forceinline__ void round_tiles(const BatchArgs& c, int k)
{
    const cptr<ShotDesc<P>> D = (cptr<ShotDesc<P>>)c.descs;
    const cptr<int> its = (cptr<int>)c.it;
    // without the XCD map block b takes tiles b, b + grid, ... of the batch's
    // tile prefix; with it the grid is a multiple of 8 blocks, XCD x runs
    // blocks x, x + 8, ... and they take turns over XCD x's share
    const unsigned x = c.xcd_map ? blockIdx.x % 8u : 0u;
    const cptr<unsigned> xs = (cptr<unsigned>)(c.xcd_map ? c.xstart + x * (unsigned)(c.n + 1) : c.bstart);
    const unsigned first = c.xcd_map ? blockIdx.x / 8u : blockIdx.x;
    const unsigned step = c.xcd_map ? gridDim.x / 8u : gridDim.x;
    for (unsigned j = first; j < xs[c.n]; j += step) {
        const int s = find_shot_c(xs, c.n, j);
        if (its[s] + k >= D[s].nt) continue;
        unsigned base = 0;
        if (c.xcd_map) {
            const unsigned nb = (unsigned)D[s].a.nblocks, q = nb / 8u, r = nb % 8u;
            base = x * q + (x < r ? x : r);
        }
        round_tile<P, N, VEL>(D + s, base + (j - xs[s]));
    }
}

template <class P, int N> __global__ __launch_bounds__(512) void k_round(const BatchArgs c, int K)
{
    namespace cg = cooperative_groups;
    cg::grid_group grid = cg::this_grid();
    const cptr<ShotDesc<P>> D = (cptr<ShotDesc<P>>)c.descs;
    const cptr<int> its = (cptr<int>)c.it;
    const cptr<unsigned> rstart = (cptr<unsigned>)c.rstart;
    const unsigned nthr = blockDim.x * blockDim.y, tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int k = 0; k < K; ++k) {
        if (batch_rescales<P>(c, k)) {
            // receivers of step k-1, every shot's rescale, then the sources
            if (k > 0)
                for (unsigned r = blockIdx.x * nthr + tid; r < c.nrec; r += gridDim.x * nthr) {
                    const int s = find_shot_c(rstart, c.n, r);
                    gather_one<P>(D[s], its[s] + k - 1, r - rstart[s]);
                }
            grid.sync();
            const unsigned long long stride = (unsigned long long)gridDim.x * nthr;
            const unsigned long long i0 = (unsigned long long)blockIdx.x * nthr + tid;
            for (int s = 0; s < c.n; ++s) {
                const int it = its[s] + k;
                if (it >= D[s].nt || !D[s].ratio) continue;
                const float ratio = D[s].ratio[it];
                if (ratio == 0.0f) continue;
                const unsigned long long nmem[3] = {D[s].nmem[0], D[s].nmem[1], D[s].nmem[2]};
                for (int arr = 0; arr < PW_NFIELDS + 3 * kMemPerAxis; ++arr)
                    rescale_array<P>(D[s].a, arr, i0, stride, D[s].ncell, nmem, ratio);
            }
            grid.sync();
            for (int s = blockIdx.x; s < c.n; s += gridDim.x)
                if (tid < 4) source_channel<P>(D[s], its[s] + k, (int)tid);
        } else {
            // one block per shot: its receivers of step k-1 are read before its
            // source of step k is added (a receiver may sit at the source cell)
            for (int s = blockIdx.x; s < c.n; s += gridDim.x) {
                if (k > 0)
                    for (unsigned q = tid; q < (unsigned)D[s].nrec; q += nthr) gather_one<P>(D[s], its[s] + k - 1, q);
                __syncthreads();
                if (tid < 4) source_channel<P>(D[s], its[s] + k, (int)tid);
            }
        }
        grid.sync();
        round_tiles<P, N, true>(c, k);
        grid.sync();
        round_tiles<P, N, false>(c, k);
        grid.sync();
    }
    for (int s = blockIdx.x; s < c.n; s += gridDim.x)
        for (unsigned q = tid; q < (unsigned)D[s].nrec; q += nthr) gather_one<P>(D[s], its[s] + K - 1, q);
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

// Calls f(std::integral_constant<int, N>) for N = vx (1, 2 or 4).
template <class F> static hipError_t dispatch_vx(int vx, F&& f)
{
    switch (vx) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 4: return f(std::integral_constant<int, 4>{});
    default: return hipErrorInvalidValue;
    }
}

template <class F> static hipError_t dispatch(int prec, F&& f)
{
    switch (prec) {
    case PW_FP32: return f(PolF32{});
    case PW_FP64: return f(PolF64{});
    case PW_COMP16: return f(PolC16{});
    case PW_BF16: return f(PolB16{});
    case PW_COMP16F: return f(PolC16F{});
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_velocity(const DevBuffers& b, hipStream_t st, int zb, int ze)
{
    if (ze <= zb) return hipSuccess;
    DevBuffers bv = b;
    bv.vx = b.vx_vel;  // the velocity kernel's own row-segment width and tile shape
    bv.wx = b.wx_vel;
    bv.ty = b.ty_vel;
    bv.zc = b.zc_vel;
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        const StepArgs<P> a = range_args<P>(bv, zb, ze);
        if (b.order == 4) {
            k_velocity<P, 1, 4><<<a.nblocks, dim3(64, a.ty), 0, st>>>(a);
            return hipGetLastError();
        }
        return dispatch_vx(a.vx, [&](auto n) {
            k_velocity<P, decltype(n)::value, 2><<<a.nblocks, dim3(64, a.ty), 0, st>>>(a);
            return hipGetLastError();
        });
    });
}

hipError_t launch_stress(const DevBuffers& b, hipStream_t st, int zb, int ze)
{
    if (ze <= zb) return hipSuccess;
    return dispatch(b.prec, [&](auto pol) {
        using P = decltype(pol);
        const StepArgs<P> a = range_args<P>(b, zb, ze);
        if (b.order == 4) {
            k_stress<P, 1, 4><<<a.nblocks, dim3(64, a.ty), 0, st>>>(a);
            return hipGetLastError();
        }
        return dispatch_vx(a.vx, [&](auto n) {
            k_stress<P, decltype(n)::value, 2><<<a.nblocks, dim3(64, a.ty), 0, st>>>(a);
            return hipGetLastError();
        });
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

// The codec alone, element by element (grid-stride), with the device
// functions the step kernels inline.
__global__ void k_codec_decode(const unsigned short* __restrict__ hi, const int8_t* __restrict__ lo,
                               float* __restrict__ out, unsigned long long n)
{
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = c16_decode(__builtin_bit_cast(f16, hi[i]), lo[i]);
}
__global__ void k_codec_encode(const float* __restrict__ in, unsigned short* __restrict__ hi, int8_t* __restrict__ lo,
                               unsigned long long n, int rne)
{
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        f16 h;
        int8_t l;
        if (rne) c16_encode_rn(in[i], h, l, 0u);
        else c16_encode(in[i], h, l, 0u);
        hi[i] = __builtin_bit_cast(unsigned short, h);
        lo[i] = l;
    }
}
static unsigned codec_blocks(unsigned long long n)
{
    const unsigned long long b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : b);
}
hipError_t launch_codec_decode(const unsigned short* hi, const int8_t* lo, float* out, unsigned long long n,
                               hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_codec_decode, dim3(codec_blocks(n)), dim3(256), 0, st, hi, lo, out, n);
    return hipGetLastError();
}
hipError_t launch_codec_encode(const float* in, unsigned short* hi, int8_t* lo, unsigned long long n, bool rne,
                               hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_codec_encode, dim3(codec_blocks(n)), dim3(256), 0, st, in, hi, lo, n, rne ? 1 : 0);
    return hipGetLastError();
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
            if (b.b->prec != prec) return hipErrorInvalidValue;
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
        g.ty = d[0].a.ty;
        g.vx = d[0].a.vx;
        for (int s2 = 1; s2 < n; ++s2)
            if (d[(size_t)s2].a.ty != g.ty || d[(size_t)s2].a.vx != g.vx) return hipErrorInvalidValue;
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
        hipError_t e = dispatch_vx(g.vx, [&](auto n) {
            k_velocity_batch<P, decltype(n)::value><<<g.nblocks, dim3(64, g.ty), 0, st>>>(c, k);
            k_stress_batch<P, decltype(n)::value><<<g.nblocks, dim3(64, g.ty), 0, st>>>(c, k);
            return hipGetLastError();
        });
        if (e != hipSuccess) return e;
        if (g.nrec > 0) k_gather_batch<P><<<(g.nrec + 255) / 256, 256, 0, st>>>(c, k);
        return hipGetLastError();
    });
}

// Blocks of the persistent round kernel: as many as fit on the device at once
// (a cooperative launch needs every block resident), a multiple of 8, and no
// more than the batch has tiles.
template <class K> static hipError_t round_blocks(K kern, const BatchGroup& g, unsigned& blocks)
{
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * g.ty, 0);
    if (e != hipSuccess) return e;
    unsigned b = (unsigned)per_cu * (unsigned)cus;
    const unsigned want = (g.nblocks + 7u) & ~7u;
    if (b > want) b = want;
    b &= ~7u;
    if (b == 0) return hipErrorCooperativeLaunchTooLarge;
    blocks = b;
    return hipSuccess;
}

hipError_t batch_round(const BatchGroup& g, hipStream_t st, int K)
{
    if (g.n <= 0 || K <= 0) return hipSuccess;
    BatchArgs c = batch_args(g);
    return dispatch(g.prec, [&](auto pol) -> hipError_t {
        using P = decltype(pol);
        return dispatch_vx(g.vx, [&](auto n) -> hipError_t {
            auto kern = k_round<P, decltype(n)::value>;
            unsigned blocks = 0;
            hipError_t e = round_blocks(kern, g, blocks);
            if (e != hipSuccess) return e;
            int Kc = K;
            void* args[] = {(void*)&c, (void*)&Kc};
            e = hipLaunchCooperativeKernel((const void*)kern, dim3(blocks), dim3(64, g.ty), args, 0, st);
            if (e != hipSuccess) return e;
            return hipGetLastError();
        });
    });
}

hipError_t batch_advance(const BatchGroup& g, hipStream_t st, int K)
{
    if (g.n <= 0) return hipSuccess;
    k_advance<<<(unsigned)((g.n + 255) / 256), 256, 0, st>>>(g.d_it, g.d_nt, g.n, K);
    return hipGetLastError();
}

}  // namespace pw
