// pw_device.h -- interface between the host API (pw_api.cpp) and the HIP
// kernels (pw_kernels.hip, pw_medium_gpu.hip).  Plain pointers only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "pw_internal.h"

namespace pw {

// v as a rounded FP32 value: its bits pass through an integer XOR with z, which
// is always 0 but arrives at run time (a kernel argument), so the compiler
// cannot see through it.  Without it the compiler may fold a following
// FP32 -> FP16 conversion into the arithmetic that produced v (v_fma_mixlo_f16:
// one rounding of the exact result) where the reference rounds to float first
// and then to half (__float2half of a float).  Whether it folds depends on the
// surrounding code, so identical formulas could otherwise round differently in
// different kernels.
__device__ __forceinline__ float round32(float v, unsigned z) { return __uint_as_float(__float_as_uint(v) ^ z); }

// --------------------------------------------------------------------------
// Compensated FP16 codec (Half/main.cu:59-100): a value is an FP16 "hi" plus an
// int8 residual in units of 2^(e-19), |hi| = f * 2^e with f in [1/2, 1), and
// 2^-32 below 2^-14.  From the FP16 bit pattern: biased FP16 exponent e5 (0 for
// zero/subnormal) gives 2^(max(e5, 1) - 33), i.e. float bits
// (max(e5,1)+94) << 23 with the 127 bias: e5 = 1 -> 2^-32, the subnormal value.
// --------------------------------------------------------------------------
// (The exponent field is used in place, e5 << 10 shifted by 13 to e5 << 23: an
// and, a max and a shift-add per value, where extracting e5 first compiled to
// five instructions.)
__device__ __forceinline__ float c16_unit_bits(unsigned short hb)
{
    const unsigned m = (unsigned)hb & 0x7c00u;
    return __uint_as_float(((m > 0x0400u ? m : 0x0400u) << 13) + (94u << 23));
}
__device__ __forceinline__ float c16_inv_unit_bits(unsigned short hb)
{
    const unsigned m = (unsigned)hb & 0x7c00u;
    return __uint_as_float((160u << 23) - ((m > 0x0400u ? m : 0x0400u) << 13));
}
__device__ __forceinline__ float c16_decode(_Float16 hi, int8_t lo)
{
    return (float)hi + (float)lo * c16_unit_bits(__builtin_bit_cast(unsigned short, hi));
}
__device__ __forceinline__ void c16_encode(float v_in, _Float16& hi, int8_t& lo, unsigned z)
{
    const float v = round32(v_in, z);
    const _Float16 h = (_Float16)v;
    const float hf = (float)h;
    // exact: the unit is a power of two
    const float scaled = (v - hf) * c16_inv_unit_bits(__builtin_bit_cast(unsigned short, h));
    int q = (int)(scaled + (scaled >= 0.0f ? 0.5f : -0.5f));
    q = q > 127 ? 127 : (q < -127 ? -127 : q);
    hi = h;
    lo = (int8_t)q;
}

// comp16f's encode: the same format, with the residual rounded to nearest
// (ties to even) instead of half away from zero.  The inverse unit comes from
// the FP32 exponent field E of the FP16 value, 2^(145 - max(E, 113)); the
// clamped residual plus 1.5 * 2^23 is an FP32 whose low byte holds the rounded
// residual in two's complement.
__device__ __forceinline__ void c16_encode_rn(float v_in, _Float16& hi, int8_t& lo, unsigned z)
{
    const float v = round32(v_in, z);
    const _Float16 h = (_Float16)v;
    const float hf = (float)h;
    const unsigned e = __float_as_uint(hf) & 0x7f800000u;
    const unsigned eb = e > (113u << 23) ? e : (113u << 23);
    const float iu = __uint_as_float((272u << 23) - eb);
    const float q = __builtin_amdgcn_fmed3f((v - hf) * iu, -127.0f, 127.0f);
    hi = h;
    lo = (int8_t)(__float_as_uint(q + 12582912.0f) & 0xffu);
}


// Coefficient slots.  "physical" form (FP32 / FP64 storage): the nine
// cell-centred medium fields on the padded grid (the reference's *_ext arrays
// after extmodel); the kernels form the staggered averages and the
// Crank-Nicolson pair on the fly with the reference's arithmetic
// (Float/main.cu:106-113, 473-513).  "scaled" form (comp16 / bf16): the
// reference's Half FP16 arrays (Half/main.cu:845-1045).
enum PhysCoef { PC_RHO = 0, PC_RHOF, PC_C1, PC_C2, PC_C, PC_M, PC_HH, PC_H2U, PC_MU, PC_N };
enum ScaledCoef {
    SC_RITX = 0, SC_RITY, SC_RITZ, SC_FRX, SC_FRY, SC_FRZ, SC_AWX, SC_AWY, SC_AWZ,
    SC_BPX, SC_BPY, SC_BPZ, SC_BTX, SC_BTY, SC_BTZ, SC_HH, SC_H2U, SC_C, SC_M,
    SC_MUXY, SC_MUXZ, SC_MUYZ, SC_N
};
constexpr int kMaxCoef = 22;

inline bool is_scaled(int prec) { return prec == PW_COMP16 || prec == PW_BF16; }
// the compensated FP16 state: FP16 "hi" arrays plus int8 "lo" residual arrays
inline bool has_lo(int prec) { return prec == PW_COMP16 || prec == PW_COMP16F; }
// CPML memory variables as records, one per damped cell and axis and kernel
// (mem[ax][0] the velocity kernel's, mem[ax][4] the stress kernel's): the four
// variables of the axis's terms, or for comp16 the reference Half program's
// (psi, previous derivative) pairs as {psi x4, prev x4}.
inline int mem_rec_width(int prec) { return prec == PW_COMP16 ? 8 : 4; }
inline int ncoef(int prec) { return is_scaled(prec) ? (int)SC_N : (int)PC_N; }
// comp16 keeps the 22 FP16 coefficients of the scaled form in records (below)
inline bool has_srec(int prec) { return prec == PW_COMP16; }
// the coefficient arrays of a precision (comp16f and comp16 hold records
// instead, below)
inline bool coef_used(int prec, int k) { return k < ncoef(prec) && prec != PW_COMP16F && !has_srec(prec); }
// comp16 keeps the 22 FP16 coefficients as records of four per cell (8 B, one
// access each; 4 % faster than 22 arrays, although the two unused slots add
// 4 B per cell-step; BF16, whose narrower state already moves its bytes at a
// higher rate, keeps the arrays): per axis d, record d = {Aw, Bp, Bt, rho~^-1};
// record 3 = {rhof rho^-1 along x, y, z, 0}; record 4 = {H~, H~ - 2mu~, C~, M~};
// record 5 = {mu~xy, mu~xz, mu~yz, 0}.  coef[k] points at coefficient k's slot
// of its record (element stride 4).
constexpr int kScRecords = 6;
__host__ __device__ constexpr int sc_rec(int k)
{
    return k <= SC_RITZ ? k - SC_RITX : k <= SC_FRZ ? 3 : k <= SC_AWZ ? k - SC_AWX : k <= SC_BPZ ? k - SC_BPX
         : k <= SC_BTZ ? k - SC_BTX : k <= SC_M ? 4 : 5;
}
__host__ __device__ constexpr int sc_slot(int k)
{
    return k <= SC_RITZ ? 3 : k <= SC_FRZ ? k - SC_FRX : k <= SC_AWZ ? 0 : k <= SC_BPZ ? 1 : k <= SC_BTZ ? 2
         : k <= SC_M ? k - SC_HH : k - SC_MUXY;
}
static_assert(sc_rec(SC_AWY) == 1 && sc_slot(SC_AWY) == 0 && sc_rec(SC_RITZ) == 2 && sc_slot(SC_RITZ) == 3 &&
                  sc_rec(SC_FRY) == 3 && sc_slot(SC_FRY) == 1 && sc_rec(SC_BTX) == 0 && sc_slot(SC_BTX) == 2 &&
                  sc_rec(SC_M) == 4 && sc_slot(SC_M) == 3 && sc_rec(SC_MUYZ) == 5 && sc_slot(SC_MUYZ) == 2,
              "scaled coefficient records");
// comp16f keeps its coefficients as compensated records instead: per cell the
// FP16 values of a record's coefficients side by side, and their residual
// bytes side by side in a second array.  Record 0 holds the velocity kernel's
// {rho, C1, C2, rhof}, records 1 and 2 the stress kernel's {C, M} and
// {H - 2 mu, mu} (mu is also read on the neighbouring row, where a 2-wide
// record touches half the cache lines of a 4-wide one).  Each coefficient is
// the cell-centred value times a power of two that brings its largest
// magnitude into [2^14, 2^15) (rec_scale holds the inverse).
constexpr int kRecords = 3;
enum { RV_RHO = 0, RV_C1, RV_C2, RV_RHOF };
enum { RS_C = 0, RS_M };
enum { RS_H2U = 0, RS_MU };
__host__ __device__ constexpr int rec_width(int r) { return r == 0 ? 4 : 2; }
__host__ __device__ constexpr int rec_base(int r, int q)
{
    return r == 0 ? (q == 0 ? PW_RHO : q == 1 ? PW_C1 : q == 2 ? PW_C2 : PW_RHOF)
                  : r == 1 ? (q == 0 ? PW_CC : PW_MM) : (q == 0 ? PW_H2U : PW_MU);
}
inline bool has_records(int prec) { return prec == PW_COMP16F; }
// comp16f stores fields in pairs: a pair's FP16 values alternate in one array
// (slot 0, slot 1 of each cell) and so do its residual bytes, so one 4-byte and
// one 2-byte access per lane move both fields.  The pairs are the fields the
// step kernels read together on a neighbouring row, where a pair costs the
// cache lines of one field and a half: the velocity kernel reads (tyy, p) on
// row y+1 and (txy, tyz) on row y-1, the stress kernel (vux, vuz) on y+1 and
// (vuy, vwy) on y-1.  The rest pair up as (txx, tzz) and (vwx, vwz); txz is
// stored alone.
__host__ __device__ constexpr int c16f_pair(int k)
{
    return k == PW_VUX || k == PW_VUZ ? 0 : k == PW_VUY || k == PW_VWY ? 1 : k == PW_VWX || k == PW_VWZ ? 2
         : k == PW_TYY || k == PW_P ? 3 : k == PW_TXY || k == PW_TYZ ? 4 : k == PW_TXX || k == PW_TZZ ? 5 : -1;
}
__host__ __device__ constexpr int c16f_slot(int k)
{
    return k == PW_VUZ || k == PW_VWY || k == PW_VWZ || k == PW_P || k == PW_TYZ || k == PW_TZZ ? 1 : 0;
}
// the field comp16f stores alone
constexpr int kC16FSingle = PW_TXZ;

// Bytes of one stored wavefield value of field k / coefficient / memory variable.
inline int state_bytes(int prec) { return prec == PW_FP64 ? 8 : prec == PW_FP32 ? 4 : has_lo(prec) ? 3 : 2; }
// (comp16f: per coefficient of a record, 2 B of FP16 value + 1 B of residual)
inline int coef_bytes(int prec) { return prec == PW_FP64 ? 8 : prec == PW_FP32 ? 4 : prec == PW_COMP16F ? 3 : 2; }
// memory-variable slots in use per axis (the rest are null)
inline int mem_slots(int prec) { return prec == PW_COMP16 ? 16 : 8; }
inline int mem_bytes(int prec) { return prec == PW_FP64 ? 8 : (prec == PW_FP32 || prec == PW_COMP16F) ? 4 : 2; }

// Device pointers of one solver.  For comp16, f[] holds the FP16 "hi" arrays
// and lo[] the int8 residuals; other precisions leave lo[] null.
struct DevBuffers {
    int prec;
    Geom g;
    void* f[PW_NFIELDS];
    int8_t* lo[PW_NFIELDS];
    void* coef[kMaxCoef];
    void* rec_hi[kRecords];        // comp16f coefficient records (see rec_base), null otherwise
    void* rec_lo[kRecords];
    float rec_scale[kRecords][4];  // physical coefficient = decoded record value * rec_scale
    void* mem[3][kMemPerAxis];
    void* tab[3];    // per axis: n entries of {c_full, e_full, c_half, e_half} (float or double)
    double dt, h;
    int zc;          // z cells per block in the stress kernel
    int zc_vel;      // ... and in the velocity kernel
    int ty;          // waves per block in the stress kernel (block 64 x ty)
    int ty_vel;      // ... and in the velocity kernel
    int wx;          // waves side by side along x in a stress-kernel tile (1 or 2)
    int wx_vel;      // ... and in a velocity-kernel tile
    int vx;          // cells per lane along x in the stress kernel (1, 2, 4)
    int vx_vel;      // ... and in the velocity kernel
    int order;       // spatial order of the staggered differences: 2 or 4
    float sc_stress, sc_vel, sc_vw;  // comp16f: stored value = physical * scale (powers of two)
    int xcd_map;     // 1: XCD-aware block -> tile mapping in the step kernels
    int derive_h;    // FP32: the stress kernel forms H as (H - 2 mu) + 2 mu (consistent medium)
    double rhof_uniform;  // FP32, FP64, > 0: the medium's one rhof, which the velocity kernel takes as a scalar
};

// Step kernels over the planes [zb, ze) (the updated planes are 1 .. nz-2).
hipError_t launch_velocity(const DevBuffers& b, hipStream_t st, int zb, int ze);
hipError_t launch_stress(const DevBuffers& b, hipStream_t st, int zb, int ze);
hipError_t launch_source(const DevBuffers& b, hipStream_t st, unsigned off, const double w[4], double amp);
hipError_t launch_gather(const DevBuffers& b, hipStream_t st, const int* rfield, const unsigned* roff, int nrec,
                         float* seis, int nt, int it, float s_stress, float s_vel);
hipError_t launch_snapshot(const DevBuffers& b, hipStream_t st, int field, int y, float* dst, float scale);
hipError_t launch_decode(const DevBuffers& b, hipStream_t st, int field, float* dst, float scale);
hipError_t launch_rescale(const DevBuffers& b, hipStream_t st, float ratio);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t st);
hipError_t launch_codec_decode(const unsigned short* hi, const int8_t* lo, float* out, unsigned long long n,
                               hipStream_t st);
hipError_t launch_codec_encode(const float* in, unsigned short* hi, int8_t* lo, unsigned long long n, bool rne,
                               hipStream_t st);

// ---- multi-shot step (shot server) ----
// One member of a batch: the solver buffers of one shot plus its per-step
// tables, all device pointers.  amp is R-typed (float, double for FP64).
struct BatchShot {
    const DevBuffers* b;
    const void* amp;        // [nt] source amplitude
    const float* oscale;    // [nt][2] receiver scale: stress, velocity
    const float* ratio;     // [nt] dynamic rescale ratio, 0 = none (null for FP32/FP64)
    const int* rfield;
    const unsigned* roff;
    float* seis;            // [nrec][nt]
    int nrec, nt;
    unsigned src_off;
    int has_src;
    double w[4];
};
// Device tables of one batch: the shots of one precision, advanced together.
// The step of shot s during a launch is it[s] + k, k = step within the round;
// shots at their last step are skipped, so every launch can be captured once
// and replayed.
struct BatchGroup {
    int prec = -1;
    int n = 0;
    void* d_descs = nullptr;     // per-shot descriptors (policy-typed)
    unsigned* d_bstart = nullptr;  // [n+1] step-kernel block prefix
    unsigned* d_xstart = nullptr;  // [8][n+1] per-XCD block prefix (XCD-aware map)
    unsigned* d_rstart = nullptr;  // [n+1] receiver prefix
    int* d_it = nullptr;         // [n] step of each shot at the start of the round
    int* d_nt = nullptr;         // [n] steps of each shot
    unsigned nblocks = 0, nrec = 0;
    int xcd_map = 1;
    int ty = 4;
    int vx = 1;
    size_t bytes = 0;
};
// (Re)builds g for the given shots (all of precision prec) at their current steps.
hipError_t batch_build(BatchGroup& g, int prec, const BatchShot* shots, int n, const int* it, hipStream_t st);
void batch_free(BatchGroup& g);
// One step (rescale, source, velocity, stress, receivers) of every shot of g at
// round offset k, and the end-of-round advance it[s] += K.
hipError_t batch_step(const BatchGroup& g, hipStream_t st, int k);
hipError_t batch_advance(const BatchGroup& g, hipStream_t st, int K);
// K steps of every shot of g in one persistent cooperative launch (grid barriers
// between the phases of a step); results bit-identical to K batch_step calls.
// The caller still calls batch_advance.
hipError_t batch_round(const BatchGroup& g, hipStream_t st, int K);

// ---- medium precompute (pw_medium_gpu.hip), all FP64 on the device ----
struct MediumArgs {
    const double* base[PW_NBASE];  // interior fields, x fastest
    int nxi, nyi, nzi, np;
    long long row_stride, plane_stride;  // elements; 0 = dense nxi, nxi*nyi (a crop of a larger volume otherwise)
    int nx, ny, nz;                // total grid
    double dt;
};
// res[0] = Kmax, res[1] = Rmax, res[2] = max fast-P velocity squared (Half/main.cu:745-791),
// res[3] = max rhof / CC1 (the size of the relative fluid velocity against the solid's),
// res[4 + f] = max |base field f| (f < PW_NBASE),
// res[kSlotHMismatch] = max |H - (H - 2 mu) - 2 mu| / |H| (0 when the medium's H
// and H - 2 mu come from one formula, as in the reference's model builders),
// res[kSlotRhofSpread] = max |rhof - rhof(first cell)| (1 if that rhof is not
// positive): 0 when one fluid fills the medium.
// d_res (kReduceSlots doubles) is cleared by a fill kernel first.
constexpr int kSlotHMismatch = 4 + PW_NBASE;
constexpr int kSlotRhofSpread = kSlotHMismatch + 1;
constexpr int kReduceSlots = kSlotRhofSpread + 1;
constexpr double kHMismatchTol = 1e-6;  // 8 FP32 ulps
hipError_t medium_reduce(const MediumArgs& m, double* d_res, hipStream_t st);
// The scratch of the allocation-free medium installation (pw_set_medium_device):
// a file-scope device array of kScratchDoubles doubles.  scratch_for_device
// returns its address on device dev (taken once per device with that device
// current, hipGetSymbolAddress); scratch_mutex serialises its users in this
// process.
constexpr int kScratchDoubles = 64;
hipError_t scratch_for_device(int dev, double** out);
std::mutex& scratch_mutex();
// Writes the coefficient arrays of b (physical or scaled form), or comp16f's
// records with each coefficient stored times 1 / b.rec_scale.
hipError_t medium_coefficients(const MediumArgs& m, const DevBuffers& b, double scale_m, double inv_scale_m,
                               hipStream_t st);

// ---- layered synthetic medium generator (pw_medium_gpu.hip) ----
hipError_t generate_layered(const pw_layered& spec, int nxi, int nyi, int nzi, int np, double* d_base /*9*n*/,
                            double* d_work /*4*n*/, double* d_partial /* >= 2*4096 */, hipStream_t st);

}  // namespace pw
