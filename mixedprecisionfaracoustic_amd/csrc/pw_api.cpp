// pw_api.cpp -- the C ABI (include/porowave.h): handle lifetime, medium
// installation, source/receiver/snapshot set-up and the step loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pw_device.h"
#include "pw_solver.h"

namespace pw {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

pw_status fail(pw_status st, const std::string& msg)
{
    set_error(msg);
    return st;
}

}  // namespace pw

using namespace pw;

// comp16f stores stresses times 2^(half_es + kC16FStressExp) (see install_medium)
static constexpr int kC16FStressExp = 6;


namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Faces in pw_config.cpml_open get the interior profile (d = 0, e = 1) on
// their half of the axis: no damping and, from slab_widths, no memory variables.
template <class T> void open_faces(const pw_config& c, int ax, int n, T* d, T* d2, T* e, T* e2)
{
    const bool lo = (c.cpml_open >> (2 * ax)) & 1, hi = (c.cpml_open >> (2 * ax + 1)) & 1;
    for (int i = 0; i < n; ++i)
        if (i < n / 2 ? lo : hi) {
            d[i] = d2[i] = (T)0;
            e[i] = e2[i] = (T)1;
        }
}

// CPML tables in the form the kernels consume: per index {c_full, e_full,
// c_half, e_half} with c = dt*d/2.
void build_tables(const pw_config& c, int prec, int ax, int n, std::vector<double>& out4, int* epml_out,
                  float* isp_out, float* pml_max_out)
{
    out4.assign((size_t)n * 4, 0.0);
    if (prec == PW_FP64) {
        std::vector<double> d(n), d2(n), e(n), e2(n);
        cpml_profile_f64(n, c.np, c.h, c.dt, c.cpml_r, c.cpml_v0, d.data(), d2.data(), e.data(), e2.data());
        open_faces(c, ax, n, d.data(), d2.data(), e.data(), e2.data());
        for (int i = 0; i < n; ++i) {
            out4[4 * i + 0] = c.dt * d[i] * 0.5;
            out4[4 * i + 1] = e[i];
            out4[4 * i + 2] = c.dt * d2[i] * 0.5;
            out4[4 * i + 3] = e2[i];
        }
        return;
    }
    std::vector<float> d(n), d2(n), e(n), e2(n);
    const float DT = (float)c.dt;
    cpml_profile_reference_f32(n, c.np, (float)c.h, DT, (float)c.cpml_r, (float)c.cpml_v0, d.data(), d2.data(),
                               e.data(), e2.data());
    open_faces(c, ax, n, d.data(), d2.data(), e.data(), e2.data());
    if (prec == PW_FP32 || prec == PW_COMP16F) {
        for (int i = 0; i < n; ++i) {
            out4[4 * i + 0] = (float)(DT * d[i]) * 0.5f;
            out4[4 * i + 1] = e[i];
            out4[4 * i + 2] = (float)(DT * d2[i]) * 0.5f;
            out4[4 * i + 3] = e2[i];
        }
        return;
    }
    // Scaled forms: the FP16 tables of Half/main.cu:1031-1042 decoded as
    // pml_dt_d / pml_e do (Half/main.cu:124-132).  epml comes from all three axes
    // (pml_max is passed in by the caller after a first pass).
    const float scale_pml = std::ldexp(1.0f, *epml_out), isp = std::ldexp(1.0f, -*epml_out);
    float pmax = 0.0f;
    for (int i = 0; i < n; ++i) {
        pmax = std::max(pmax, std::max(std::fabs(DT * d[i]), std::fabs(DT * d2[i])));
        pmax = std::max(pmax, std::max(std::fabs(1.0f - e[i]), std::fabs(1.0f - e2[i])));
        const _Float16 hd = (_Float16)(scale_pml * DT * d[i]), hd2 = (_Float16)(scale_pml * DT * d2[i]);
        const _Float16 he = (_Float16)(scale_pml * (1.0f - e[i])), he2 = (_Float16)(scale_pml * (1.0f - e2[i]));
        out4[4 * i + 0] = (isp * (float)hd) * 0.5f;
        out4[4 * i + 1] = 1.0f - isp * (float)he;
        out4[4 * i + 2] = (isp * (float)hd2) * 0.5f;
        out4[4 * i + 3] = 1.0f - isp * (float)he2;
    }
    *isp_out = isp;
    *pml_max_out = pmax;
}

// The tables of all three axes for CPML velocity v0; the scaled forms take
// epml from the three together (Half/main.cu:833-839).
void plan_tables(const pw_config& c, double v0, std::vector<double> (&tab)[3], int* epml, float* isp)
{
    pw_config cv = c;
    cv.cpml_v0 = v0;
    const int prec = c.precision;
    const int n3[3] = {c.nx, c.ny, c.nz};
    *epml = 0;
    *isp = 1.0f;
    if (is_scaled(prec)) {
        float pmax = 0.0f, pm;
        for (int ax = 0; ax < 3; ++ax) {
            build_tables(cv, prec, ax, n3[ax], tab[ax], epml, isp, &pm);
            pmax = std::max(pmax, pm);
        }
        if (pmax > 0.0f) {
            *epml = (int)std::floor(std::log2(2048.0f / pmax));
            *epml = std::min(20, std::max(-20, *epml));
        }
    }
    for (int ax = 0; ax < 3; ++ax) {
        float pm;
        build_tables(cv, prec, ax, n3[ax], tab[ax], epml, isp, &pm);
    }
}

// CPML velocity of the tables: cfg.cpml_v0, or with cpml_v0 <= 0 the medium's
// fastest P velocity (known once a medium is installed)
double table_v0(const pw_config& c, double vp_max) { return c.cpml_v0 > 0 ? c.cpml_v0 : vp_max; }

void slab_widths(const std::vector<double>& t, int n, int* lo, int* hi)
{
    *lo = 0;
    *hi = 0;
    for (int i = 0; i < n / 2; ++i)
        if (t[4 * i] != 0.0 || t[4 * i + 2] != 0.0) *lo = i + 1;
    for (int i = n - 1; i >= n / 2; --i)
        if (t[4 * i] != 0.0 || t[4 * i + 2] != 0.0) *hi = n - i;
}

hipError_t zero_state(pw_solver* s) { return launch_zero(s->pool, s->state_bytes_total, s->stream); }

// Uploads the CPML tables in the kernels' real type.
hipError_t upload_tables(pw_solver* s, const std::vector<double> (&tab)[3])
{
    hipError_t e = hipSuccess;
    for (int ax = 0; ax < 3 && e == hipSuccess; ++ax) {
        if (s->prec == PW_FP64) {
            e = hipMemcpyAsync(s->b.tab[ax], tab[ax].data(), tab[ax].size() * 8, hipMemcpyHostToDevice, s->stream);
        } else {
            std::vector<float> f(tab[ax].begin(), tab[ax].end());
            e = hipMemcpyAsync(s->b.tab[ax], f.data(), f.size() * 4, hipMemcpyHostToDevice, s->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);  // the host vectors go out of scope
    }
    return e;
}

pw_status check_handle(pw_solver* s, const char* fn)
{
    if (!s) return fail(PW_ERR_INVALID, std::string(fn) + ": null handle");
    return PW_OK;
}

}  // namespace

extern "C" void pw_config_default(pw_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->nx = c->ny = c->nz = 128 + 2 * 32;
    c->np = 32;
    c->h = 0.01;
    c->dt = 0.9e-6;
    c->nt = 3500;
    c->precision = PW_FP32;
    c->cpml_r = 1e-6;
    c->cpml_v0 = 4000.0;
    c->half_es = 0;
    c->half_dyn_end = 16.0;
    c->half_dyn_interval = 100;
    c->device = 0;
    c->stream = nullptr;
    c->check_cfl = 1;
    c->space_order = 2;
}

extern "C" const char* pw_last_error(void) { return g_last_error.c_str(); }

namespace {

pw_status check_config(const pw_config& c)
{
    if (c.precision < PW_FP32 || c.precision > PW_COMP16F) return fail(PW_ERR_INVALID, "pw_create: unknown precision");
    if (c.np < 1 || c.nx < 2 * c.np + 3 || c.ny < 2 * c.np + 3 || c.nz < 2 * c.np + 3)
        return fail(PW_ERR_INVALID, "pw_create: grid must be at least 2*np+3 cells per axis");
    if (!(c.h > 0) || !(c.dt > 0) || c.nt < 1) return fail(PW_ERR_INVALID, "pw_create: h, dt and nt must be positive");
    if (!(c.cpml_r > 0 && c.cpml_r < 1) || !std::isfinite(c.cpml_v0))
        return fail(PW_ERR_INVALID, "pw_create: cpml_r must be in (0,1) and cpml_v0 finite");
    if (c.space_order != 0 && c.space_order != 2 && c.space_order != 4)
        return fail(PW_ERR_INVALID, "pw_create: space_order must be 2 or 4");
    if (c.cpml_open < 0 || c.cpml_open > 63)
        return fail(PW_ERR_INVALID, "pw_create: cpml_open must be a mask of the six PW_FACE_* bits");
    if (c.space_order == 4 && (c.nx < 2 * c.np + 5 || c.ny < 2 * c.np + 5 || c.nz < 2 * c.np + 5))
        return fail(PW_ERR_INVALID, "pw_create: 4th order needs at least 2*np+5 cells per axis");
    if (c.tile_cells != 0 && c.tile_cells != 1 && c.tile_cells != 2 && c.tile_cells != 4)
        return fail(PW_ERR_INVALID, "pw_create: tile_cells must be 0, 1, 2 or 4");
    if (c.space_order == 4 && c.tile_cells > 1)
        return fail(PW_ERR_UNSUPPORTED, "pw_create: 4th order runs one cell per lane (tile_cells 0 or 1)");
    if (c.tile_waves_x != 0 && c.tile_waves_x != 1 && c.tile_waves_x != 2)
        return fail(PW_ERR_INVALID, "pw_create: tile_waves_x must be 0, 1 or 2");
    if (c.tile_rows != 0 && c.tile_rows != 1 && c.tile_rows != 2 && c.tile_rows != 4 && c.tile_rows != 8)
        return fail(PW_ERR_INVALID, "pw_create: tile_rows must be 0, 1, 2, 4 or 8");
    if (c.tile_waves_x == 2 && c.tile_rows == 1)
        return fail(PW_ERR_INVALID, "pw_create: two waves along x need tile_rows of 2 or more");
    if (c.tile_planes < 0) return fail(PW_ERR_INVALID, "pw_create: tile_planes must be >= 0");
    const unsigned long long pitch = (unsigned long long)(c.nx + 63) / 64 * 64;
    if (pitch * c.ny * c.nz >= (1ull << 32) - (1ull << 20))
        return fail(PW_ERR_UNSUPPORTED, "pw_create: grid exceeds 2^32 padded cells (32-bit offsets)");
    // a compensated field's residual bytes are addressed by a 32-bit offset from
    // its FP16 values: 2 bytes per cell of values, then 1 per cell
    if (has_lo(c.precision) && 3ull * pitch * c.ny * c.nz >= (1ull << 32) - (1ull << 20))
        return fail(PW_ERR_UNSUPPORTED, "pw_create: comp16 grid exceeds 2^32 / 3 padded cells (32-bit offsets)");
    return PW_OK;
}

// Geometry, CPML tables and the byte layout of the pooled allocation:
// [state | memory variables] (zeroed together) then [coefficients | tables].
struct Layout {
    Geom g{};
    std::vector<double> tab[3];
    int epml = 0;
    float isp = 1.0f;
    std::vector<size_t> off;              // every array after the wavefields, in allocation order
    size_t field_off[PW_NFIELDS] = {};    // a field's values (FP16 values for the compensated forms)
    size_t field_lo[PW_NFIELDS] = {};     // its residual bytes (compensated forms)
    size_t zero_bytes = 0, total = 0;
    size_t pad = 0;   // bytes between consecutive arrays
};

void plan_layout(const pw_config& c, Layout& L)
{
    const int prec = c.precision;
    Geom& g = L.g;
    const int pitch = (c.nx + 63) / 64 * 64;
    g.nx = c.nx; g.ny = c.ny; g.nz = c.nz;
    g.pitch = pitch;
    g.plane = (long long)pitch * c.ny;
    g.cells = (long long)pitch * c.ny * c.nz;
    const size_t cells = (size_t)g.cells;
    // CPML tables.  Where a profile is nonzero does not depend on v0, so the
    // slabs (and the layout) are known before the medium: with cpml_v0 <= 0
    // the tables are planned here with a stand-in velocity and planned again
    // from the medium's vp_max when it is installed.
    const int n3[3] = {c.nx, c.ny, c.nz};
    plan_tables(c, table_v0(c, 4000.0), L.tab, &L.epml, &L.isp);
    for (int ax = 0; ax < 3; ++ax) slab_widths(L.tab[ax], n3[ax], &g.lo[ax], &g.hi[ax]);
    const long long W = pitch;
    g.slab_cells[0] = (long long)c.nz * c.ny * (g.lo[0] + g.hi[0]);
    g.slab_cells[1] = (long long)c.nz * (g.lo[1] + g.hi[1]) * W;
    g.slab_cells[2] = (long long)(g.lo[2] + g.hi[2]) * c.ny * W;

    const size_t sb = (size_t)state_bytes(prec), cb = (size_t)coef_bytes(prec), mb = (size_t)mem_bytes(prec);
    const size_t rb = prec == PW_FP64 ? 8 : 4;
    L.off.clear();
    L.total = 0;
    // 8 KiB between consecutive arrays: the same cell of every array then no
    // longer lies a power-of-two distance apart (256 MiB per FP16 field at
    // 512^3).  Measured on one GPU against no padding: 0.4-1.5 % faster at
    // 512^3 and 1024^3, and on a rank's share, in FP32 and comp16.
    L.pad = 8192;
    auto place = [&](size_t bytes) {
        const size_t o = L.total;
        L.total += align256(bytes) + L.pad;
        return o;
    };
    auto take = [&](size_t bytes) { L.off.push_back(place(bytes)); };
    // compensated fields: FP16 values, then the residual bytes at the same
    // offset (align256(2 * cells)) for every field; comp16f's pairs hold two
    // fields' values and residuals interleaved, the residuals align256(4 * cells)
    // after the values
    for (int k = 0; k < PW_NFIELDS; ++k) {
        if (prec == PW_COMP16F && c16f_pair(k) >= 0) {
            if (c16f_slot(k) != 0) continue;  // placed with its slot-0 partner
            const size_t o = place(align256(cells * 4) + L.pad + 2 * cells);
            for (int q = 0; q < PW_NFIELDS; ++q)
                if (c16f_pair(q) == c16f_pair(k)) {
                    L.field_off[q] = o + 2 * c16f_slot(q);
                    L.field_lo[q] = o + align256(cells * 4) + L.pad + c16f_slot(q);
                }
        } else if (has_lo(prec)) {
            const size_t o = place(align256(cells * 2) + L.pad + cells);
            L.field_off[k] = o;
            L.field_lo[k] = o + align256(cells * 2) + L.pad;
        } else {
            L.field_off[k] = place(cells * sb);
        }
    }
    // memory variables: per axis one array of records for each kernel (the
    // velocity kernel's, then the stress kernel's), a record per damped cell
    for (int ax = 0; ax < 3; ++ax) {
        take((size_t)g.slab_cells[ax] * mem_rec_width(prec) * mb);
        take((size_t)g.slab_cells[ax] * mem_rec_width(prec) * mb);
    }
    L.zero_bytes = L.total;
    for (int k = 0; k < kMaxCoef; ++k)
        if (coef_used(prec, k)) take(cells * cb);
    if (has_srec(prec))
        for (int r = 0; r < kScRecords; ++r) take(cells * 4 * sizeof(uint16_t));
    if (has_records(prec))
        for (int r = 0; r < kRecords; ++r) {
            take(cells * 2 * rec_width(r));  // the FP16 values
            take(cells * rec_width(r));      // the residual bytes
        }
    for (int ax = 0; ax < 3; ++ax) take((size_t)n3[ax] * 4 * rb);
}

}  // namespace

size_t pw::solver_pool_bytes(const pw_config& c)
{
    Layout L;
    plan_layout(c, L);
    return L.total;
}

extern "C" pw_status pw_create(const pw_config* cfg, pw_solver** out) { return solver_create(cfg, out); }

pw_status pw::solver_create(const pw_config* cfg, pw_solver** out)
{
    if (!cfg || !out) return fail(PW_ERR_INVALID, "pw_create: null argument");
    *out = nullptr;
    const pw_config& c = *cfg;
    pw_status cst = check_config(c);
    if (cst != PW_OK) return cst;
    const int order = c.space_order == 4 ? 4 : 2;

    int ndev = 0;
    PW_HIP_TRY(hipGetDeviceCount(&ndev));
    if (c.device < 0 || c.device >= ndev) return fail(PW_ERR_INVALID, "pw_create: no such device");
    PW_HIP_TRY(hipSetDevice(c.device));

    pw_solver* s = new pw_solver();
    s->cfg = c;
    s->prec = c.precision;
    Layout L;
    plan_layout(c, L);
    s->b.g = L.g;
    const Geom& g = s->b.g;
    s->epml = L.epml;
    s->inv_scale_pml = L.isp;
    const std::vector<size_t>& off = L.off;
    const size_t total = L.total;

    hipError_t e = hipMalloc(&s->pool, total);
    if (e != hipSuccess) {
        delete s;
        return fail(e == hipErrorOutOfMemory ? PW_ERR_OUT_OF_MEMORY : PW_ERR_HIP,
                    "pw_create: device allocation of " + std::to_string(total) + " bytes failed: " +
                        hipGetErrorString(e));
    }
    s->pool_bytes = total;
    s->state_bytes_total = L.zero_bytes;
    char* base = (char*)s->pool;
    size_t i = 0;
    for (int k = 0; k < PW_NFIELDS; ++k) s->b.f[k] = base + L.field_off[k];
    for (int k = 0; k < PW_NFIELDS; ++k) s->b.lo[k] = has_lo(s->prec) ? (int8_t*)(base + L.field_lo[k]) : nullptr;
    for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k < kMemPerAxis; ++k) {
            const bool used = k == 0 || k == 4;  // the two record arrays
            s->b.mem[ax][k] = used ? base + off[i++] : nullptr;
        }
    for (int k = 0; k < kMaxCoef; ++k) s->b.coef[k] = coef_used(s->prec, k) ? base + off[i++] : nullptr;
    if (has_srec(s->prec)) {
        char* rec[kScRecords];
        for (int r = 0; r < kScRecords; ++r) rec[r] = base + off[i++];
        for (int k = 0; k < SC_N; ++k) s->b.coef[k] = rec[sc_rec(k)] + sc_slot(k) * sizeof(uint16_t);
    }
    for (int r = 0; r < kRecords; ++r) {
        s->b.rec_hi[r] = has_records(s->prec) ? base + off[i++] : nullptr;
        s->b.rec_lo[r] = has_records(s->prec) ? base + off[i++] : nullptr;
        for (int q = 0; q < 4; ++q) s->b.rec_scale[r][q] = 1.0f;
    }
    for (int ax = 0; ax < 3; ++ax) s->b.tab[ax] = base + off[i++];
    s->b.prec = s->prec;
    s->b.dt = c.dt;
    s->b.h = c.h;
    // planes per block along z: 8 measured best at 512^3 in every precision
    // but comp16f, which runs one-wave blocks (64 x 1 tiles) marching 16
    // planes from nz >= 512: 1.6 % faster at 512^3, 1.2-1.3 % at 192^3 and
    // 256^3 (with 4 and 8 planes) than 64 x 4 tiles
    const bool c16f = s->prec == PW_COMP16F;
    const int zc_def = c.nz >= 512 && c16f ? 16 : (c.nz >= 256 ? 8 : 4);
    s->b.zc = c.tile_planes > 0 ? c.tile_planes : zc_def;
    s->b.ty = c.tile_rows > 0 ? c.tile_rows : (c16f ? 1 : 4);
    // FP32's 4th-order velocity kernel from nz >= 512: one-wave blocks
    // marching 16 planes (3 % faster at 512^3; at order 2 no different, and
    // 2 % slower at 192^3); every other velocity kernel takes the stress
    // kernel's rows and planes
    const bool v1 = s->prec == PW_FP32 && order == 4 && c.nz >= 512;
    s->b.zc_vel = c.tile_planes > 0 ? c.tile_planes : (v1 ? 16 : s->b.zc);
    s->b.ty_vel = c.tile_rows > 0 ? c.tile_rows : (v1 ? 1 : s->b.ty);
    // waves side by side along x per tile: FP32's two kernels and BF16's
    // velocity kernel run fastest with two (128-cell tiles; FP32's velocity
    // kernel 1.6 % faster at 512^3), every other kernel with one.  Only where
    // 128-cell tiles leave no more of a row idle than 64-cell ones: a row of
    // 192 cells would leave half its second 128-cell tile idle (measured 10 %
    // slower there, FP32 stress and BF16 velocity)
    const bool wide = (c.nx + 127) / 128 * 128 == (c.nx + 63) / 64 * 64;
    const bool wx_set = c.tile_waves_x > 0;
    s->b.wx = wx_set ? c.tile_waves_x : (s->prec == PW_FP32 && wide ? 2 : 1);
    s->b.wx_vel = wx_set ? c.tile_waves_x : ((s->prec == PW_FP32 || s->prec == PW_BF16) && wide ? 2 : 1);
    // a two-wave row needs two waves in the block
    if (s->b.ty == 1) s->b.wx = 1;
    if (s->b.ty_vel == 1) s->b.wx_vel = 1;
    // cells per lane along x: wider row segments per wave.  comp16's velocity
    // kernel runs fastest with 2 cells per lane (its FP16 and int8 rows then
    // load whole cache lines) where 128-cell segments fill the row as well as
    // 64-cell ones (at 192 and 320 cells one lane per cell is 22 % and 13 %
    // faster); every other kernel with one
    const bool vx_set = c.tile_cells > 0;
    s->b.vx = vx_set ? c.tile_cells : 1;
    s->b.vx_vel = vx_set ? c.tile_cells : (s->prec == PW_COMP16 && wide ? 2 : 1);
    // XCD-aware block order (measured: without it the step kernels read 15-19 %
    // more from HBM)
    s->b.xcd_map = 1;
    s->b.order = order;
    s->b.sc_stress = s->b.sc_vel = s->b.sc_vw = 1.0f;

    if (c.stream) {
        s->stream = (hipStream_t)c.stream;
    } else {
        e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            (void)hipFree(s->pool);
            delete s;
            return fail(PW_ERR_HIP, std::string("pw_create: stream: ") + hipGetErrorString(e));
        }
        s->own_stream = true;
    }
    e = upload_tables(s, L.tab);
    if (e != hipSuccess) {
        pw_destroy(s);
        return fail(PW_ERR_HIP, std::string("pw_create: table upload: ") + hipGetErrorString(e));
    }
    e = zero_state(s);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        pw_destroy(s);
        return fail(PW_ERR_HIP, std::string("pw_create: zero state: ") + hipGetErrorString(e));
    }
    // default source: the reference's cell and wavelet (Float/main.cu:310-313, 321)
    {
        double w4[4] = {1, 1, 1, 0};
        const int sz = std::min(20 + c.np, c.nz - 2);
        pw_set_source(s, c.nx / 2, c.ny / 2, sz, 3000.0, w4);
    }
    *out = s;
    return PW_OK;
}

namespace {

// Releases the volume-snapshot buffers, events and copy stream's work (after
// both streams have drained).
void free_volumes(pw_solver* s)
{
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
    for (float* p : s->vol_host) (void)hipHostFree(p);
    for (hipEvent_t e : s->vol_done) (void)hipEventDestroy(e);
    for (int j = 0; j < 2; ++j) {
        (void)hipFree(s->d_vstage[j]);
        if (s->ev_decoded[j]) (void)hipEventDestroy(s->ev_decoded[j]);
        if (s->ev_copied[j]) (void)hipEventDestroy(s->ev_copied[j]);
        s->d_vstage[j] = nullptr;
        s->ev_decoded[j] = s->ev_copied[j] = nullptr;
    }
    s->vol_host.clear();
    s->vol_done.clear();
    s->vol_taken.clear();
    s->vol_step.clear();
    s->vol_field.clear();
    s->nvol = 0;
    s->vol_next = 0;
}

}  // namespace

extern "C" void pw_destroy(pw_solver* s)
{
    if (!s) return;
    (void)hipSetDevice(s->cfg.device);
    free_volumes(s);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    (void)hipFree(s->d_rfield);
    (void)hipFree(s->d_roff);
    (void)hipFree(s->d_seis);
    (void)hipFree(s->d_snaps);
    (void)hipFree(s->pool);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

pw_status pw::install_medium(pw_solver* s, const double* const* d_base, int nxi, int nyi, int nzi,
                             long long row_stride, long long plane_stride, double* d_scratch)
{
    const pw_config& c = s->cfg;
    MediumArgs m{};
    for (int k = 0; k < PW_NBASE; ++k) m.base[k] = d_base[k];
    m.row_stride = row_stride;
    m.plane_stride = plane_stride;
    m.nxi = nxi; m.nyi = nyi; m.nzi = nzi; m.np = c.np;
    m.nx = c.nx; m.ny = c.ny; m.nz = c.nz;
    m.dt = c.dt;
    double* d_res = d_scratch;
    if (!d_res) PW_HIP_TRY(hipMalloc(&d_res, kReduceSlots * sizeof(double)));
    double res[kReduceSlots] = {};
    hipError_t e = medium_reduce(m, d_res, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(res, d_res, kReduceSlots * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (!d_scratch) (void)hipFree(d_res);
    if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_set_medium: reduction: ") + hipGetErrorString(e));
    if (!(res[0] > 0) || !std::isfinite(res[0]) || !std::isfinite(res[1]) || !std::isfinite(res[2]))
        return fail(PW_ERR_INVALID, "pw_set_medium: medium has no positive finite modulus");
    // H and H - 2 mu built by one formula agree to rounding (a few FP32 ulps
    // when the three arrays were each rounded to float); FP32 then forms H in
    // the stress kernel instead of streaming it, comp16f always does
    const bool h_consistent = res[kSlotHMismatch] <= kHMismatchTol;
    if (s->prec == PW_COMP16F && !h_consistent)
        return fail(PW_ERR_UNSUPPORTED, "pw_set_medium: comp16f forms H as (H - 2 mu) + 2 mu, but this medium's H "
                                        "differs from it by " + std::to_string(res[kSlotHMismatch]) + " of H");
    s->b.derive_h = s->prec == PW_FP32 && h_consistent ? 1 : 0;
    // one fluid throughout (as in the reference's model, Float/main.cu:1632):
    // the FP32 and FP64 velocity kernels take rhof as a scalar, the same value
    // the array holds
    const bool physical = s->prec == PW_FP32 || s->prec == PW_FP64;
    s->b.rhof_uniform = physical && res[kSlotRhofSpread] == 0.0 ? res[4 + PW_RHOF] : 0.0;
    s->vp_max = std::sqrt(res[2]);
    // von Neumann limit of the staggered leapfrog: vp dt / h * sqrt(3) * sum|c_k| <= 1,
    // sum|c_k| = 1 (2nd order) or 9/8 + 1/24 = 7/6 (4th order)
    s->courant = s->vp_max * c.dt / c.h * std::sqrt(3.0) * (s->b.order == 4 ? 7.0 / 6.0 : 1.0);
    if (c.check_cfl && !(s->courant <= 1.0))
        return fail(PW_ERR_UNSTABLE, "pw_set_medium: Courant number " + std::to_string(s->courant) +
                                         " > 1 (vp_max " + std::to_string(s->vp_max) + " m/s)");
    if (c.cpml_v0 <= 0) {
        // CPML damping from the medium's fastest P velocity: d0 = 3 vp_max ln(1/R) / (2 np h)
        std::vector<double> tab[3];
        plan_tables(c, s->vp_max, tab, &s->epml, &s->inv_scale_pml);
        hipError_t et = upload_tables(s, tab);
        if (et != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_set_medium: CPML tables: ") + hipGetErrorString(et));
    }
    // Half exponent, Half/main.cu:793-807 (float arithmetic as there)
    double sm = 1.0, ism = 1.0;
    if (is_scaled(s->prec) || s->prec == PW_COMP16F) {
        int em = -23;
        const float Kmax = (float)res[0], Rmax = (float)res[1], DT = (float)c.dt;
        if (Kmax > 0.0f && Rmax > 0.0f) {
            const float er = 1.0f / std::sqrt(DT * DT * Kmax * Rmax);
            em = (int)std::lrint(-std::log2(er * DT * Kmax));
        }
        s->em = em;
        sm = std::ldexp(1.0, em);
        ism = std::ldexp(1.0, -em);
        s->source_scale = std::ldexp(1.0f, c.half_es);
        s->stress_base = std::ldexp(1.0f, -c.half_es);
        s->vel_base = std::ldexp(1.0f, em - c.half_es);
        if (s->prec == PW_COMP16F) {
            // Storage scales of the three field classes (coefficients stay
            // unscaled).  The compensated format holds 19 bits relative to a
            // value only above 2^-14 (the residual unit is fixed at 2^-32 below),
            // so stresses are stored times 2^(es + kC16FStressExp): for a
            // unit-amplitude source the waves then stay above 2^-14 far from the
            // source, and the near-source peak (stress of order 10) stays 2^6
            // below FP16's largest value.  On the 512^3 two-layer workload over
            // 2000 steps this exponent cut the seismograms' distance from FP64
            // 3-6x against es = 0.  Solid velocities take the Half exponent em.
            const int es = c.half_es + kC16FStressExp;
            s->b.sc_stress = std::ldexp(1.0f, es);
            s->b.sc_vel = std::ldexp(1.0f, es - em);
            // The relative fluid velocity is up to rhof/CC1 times the solid velocity
            // (res[3], largest in the most permeable rock) and far smaller in tight
            // rock.  Its own scale lifts it so that the largest vw is stored about
            // as large as the solid velocity: the tight-rock vw stays in FP16's
            // normal range and the largest stays below the FP16 maximum.
            int kw = 0;
            if (res[3] > 0.0 && std::isfinite(res[3])) kw = std::max(0, std::min(40, (int)std::lrint(-std::log2(res[3]))));
            s->b.sc_vw = std::ldexp(s->b.sc_vel, kw);
            sm = ism = 1.0;
            // coefficient records: each coefficient times 2^(15 - ex), its largest
            // magnitude m in [2^(ex-1), 2^ex) then stored in [2^14, 2^15)
            for (int r = 0; r < kRecords; ++r)
                for (int q = 0; q < rec_width(r); ++q) {
                    const double mx = res[4 + rec_base(r, q)];
                    int ex = 15;
                    if (mx > 0.0 && std::isfinite(mx)) std::frexp(mx, &ex);
                    s->b.rec_scale[r][q] = std::ldexp(1.0f, ex - 15);
                }
        }
    }
    e = medium_coefficients(m, s->b, sm, ism, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_set_medium: coefficients: ") + hipGetErrorString(e));
    s->has_medium = true;
    return pw_reset(s);
}

namespace {

// Field pointers of a dense interior block of PW_NBASE fields.
struct DenseBase {
    const double* p[PW_NBASE];
    DenseBase(const double* d, size_t ni)
    {
        for (int k = 0; k < PW_NBASE; ++k) p[k] = d + k * ni;
    }
};

}  // namespace

extern "C" pw_status pw_set_medium(pw_solver* s, const double* const* base, int nxi, int nyi, int nzi)
{
    pw_status st = check_handle(s, "pw_set_medium");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (!base) return fail(PW_ERR_INVALID, "pw_set_medium: null base");
    if (nxi != c.nx - 2 * c.np || nyi != c.ny - 2 * c.np || nzi != c.nz - 2 * c.np)
        return fail(PW_ERR_INVALID, "pw_set_medium: interior size must be (nx-2np, ny-2np, nz-2np)");
    for (int k = 0; k < PW_NBASE; ++k)
        if (!base[k]) return fail(PW_ERR_INVALID, "pw_set_medium: null base field");
    PW_HIP_TRY(hipSetDevice(c.device));
    const size_t ni = (size_t)nxi * nyi * nzi;
    double* d_base = nullptr;
    PW_HIP_TRY(hipMalloc(&d_base, PW_NBASE * ni * sizeof(double)));
    hipError_t e = hipSuccess;
    for (int k = 0; k < PW_NBASE && e == hipSuccess; ++k)
        e = hipMemcpyAsync(d_base + k * ni, base[k], ni * sizeof(double), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        (void)hipFree(d_base);
        return fail(PW_ERR_HIP, std::string("pw_set_medium: upload: ") + hipGetErrorString(e));
    }
    st = install_medium(s, DenseBase(d_base, ni).p, nxi, nyi, nzi, 0, 0);
    (void)hipFree(d_base);
    if (st == PW_OK) s->layered = false;
    return st;
}

extern "C" pw_status pw_set_medium_device(pw_solver* s, const double* const* d_base, int nxi, int nyi, int nzi)
{
    pw_status st = check_handle(s, "pw_set_medium_device");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (!d_base) return fail(PW_ERR_INVALID, "pw_set_medium_device: null base");
    if (nxi != c.nx - 2 * c.np || nyi != c.ny - 2 * c.np || nzi != c.nz - 2 * c.np)
        return fail(PW_ERR_INVALID, "pw_set_medium_device: interior size must be (nx-2np, ny-2np, nz-2np)");
    for (int k = 0; k < PW_NBASE; ++k)
        if (!d_base[k]) return fail(PW_ERR_INVALID, "pw_set_medium_device: null base field");
    if (!(c.cpml_v0 > 0))
        return fail(PW_ERR_UNSUPPORTED,
                    "pw_set_medium_device: cpml_v0 <= 0 builds the CPML tables from the medium on the host; "
                    "give the handle a CPML velocity");
    PW_HIP_TRY(hipSetDevice(c.device));
    // the scratch array is shared by every call on this device
    std::lock_guard<std::mutex> lk(scratch_mutex());
    double* d_scratch = nullptr;
    PW_HIP_TRY(scratch_for_device(c.device, &d_scratch));
    st = install_medium(s, d_base, nxi, nyi, nzi, 0, 0, d_scratch);
    if (st == PW_OK) s->layered = false;
    return st;
}

namespace {

pw_status generate_into(pw_solver* s, const pw_layered* spec, double** d_base_out)
{
    const pw_config& c = s->cfg;
    const int nxi = c.nx - 2 * c.np, nyi = c.ny - 2 * c.np, nzi = c.nz - 2 * c.np;
    const size_t ni = (size_t)nxi * nyi * nzi;
    double *d_base = nullptr, *d_work = nullptr, *d_part = nullptr;
    PW_HIP_TRY(hipMalloc(&d_base, PW_NBASE * ni * sizeof(double)));
    hipError_t e = hipMalloc(&d_work, (spec->perturb != 0.0 ? 4 : 3) * ni * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_part, 2 * 4096 * sizeof(double));
    if (e == hipSuccess) e = generate_layered(*spec, nxi, nyi, nzi, c.np, d_base, d_work, d_part, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d_work);
    (void)hipFree(d_part);
    if (e != hipSuccess) {
        (void)hipFree(d_base);
        return fail(e == hipErrorOutOfMemory ? PW_ERR_OUT_OF_MEMORY : PW_ERR_HIP,
                    std::string("pw_set_medium_layered: generator: ") + hipGetErrorString(e));
    }
    *d_base_out = d_base;
    return PW_OK;
}

pw_status check_layered(const pw_solver* s, const pw_layered* L)
{
    if (!L) return fail(PW_ERR_INVALID, "pw_set_medium_layered: null spec");
    if (L->nlayers < 1 || L->nlayers > PW_MAX_LAYERS) return fail(PW_ERR_INVALID, "pw_set_medium_layered: nlayers");
    if (L->z_top[0] != 0) return fail(PW_ERR_INVALID, "pw_set_medium_layered: z_top[0] must be 0");
    for (int k = 1; k < L->nlayers; ++k)
        if (L->z_top[k] <= L->z_top[k - 1]) return fail(PW_ERR_INVALID, "pw_set_medium_layered: z_top not ascending");
    for (int k = 0; k < L->nlayers; ++k) {
        double out[PW_NBASE];
        if (pw_biot_from_rock(&L->rock[k], out) != PW_OK) return PW_ERR_INVALID;
    }
    if (L->perturb < 0 || L->perturb * (L->clamp > 0 ? L->clamp : 3.0) >= 1.0)
        return fail(PW_ERR_INVALID, "pw_set_medium_layered: perturb*clamp must be in [0, 1)");
    (void)s;
    return PW_OK;
}

}  // namespace

extern "C" pw_status pw_set_medium_layered(pw_solver* s, const pw_layered* spec)
{
    pw_status st = check_handle(s, "pw_set_medium_layered");
    if (st != PW_OK) return st;
    if ((st = check_layered(s, spec)) != PW_OK) return st;
    PW_HIP_TRY(hipSetDevice(s->cfg.device));
    double* d_base = nullptr;
    if ((st = generate_into(s, spec, &d_base)) != PW_OK) return st;
    const pw_config& c = s->cfg;
    const int nxi = c.nx - 2 * c.np, nyi = c.ny - 2 * c.np, nzi = c.nz - 2 * c.np;
    st = install_medium(s, DenseBase(d_base, (size_t)nxi * nyi * nzi).p, nxi, nyi, nzi, 0, 0);
    (void)hipFree(d_base);
    if (st == PW_OK) {
        s->layered = true;
        s->spec = *spec;
    }
    return st;
}

extern "C" pw_status pw_get_base_fields(pw_solver* s, double* out)
{
    pw_status st = check_handle(s, "pw_get_base_fields");
    if (st != PW_OK) return st;
    if (!out) return fail(PW_ERR_INVALID, "pw_get_base_fields: null output");
    if (!s->layered) return fail(PW_ERR_STATE, "pw_get_base_fields: only media from pw_set_medium_layered are kept");
    PW_HIP_TRY(hipSetDevice(s->cfg.device));
    double* d_base = nullptr;
    if ((st = generate_into(s, &s->spec, &d_base)) != PW_OK) return st;
    const pw_config& c = s->cfg;
    const size_t ni = (size_t)(c.nx - 2 * c.np) * (c.ny - 2 * c.np) * (c.nz - 2 * c.np);
    hipError_t e = hipMemcpy(out, d_base, PW_NBASE * ni * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d_base);
    if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_get_base_fields: ") + hipGetErrorString(e));
    return PW_OK;
}

extern "C" pw_status pw_set_source(pw_solver* s, int sx, int sy, int sz, double f0, const double* w4)
{
    pw_status st = check_handle(s, "pw_set_source");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (sx < 1 || sy < 1 || sz < 1 || sx > c.nx - 2 || sy > c.ny - 2 || sz > c.nz - 2)
        return fail(PW_ERR_INVALID, "pw_set_source: source cell must be inside the updated grid");
    if (!(f0 > 0)) return fail(PW_ERR_INVALID, "pw_set_source: f0 must be positive");
    s->src_off = (unsigned)((long long)sx + (long long)sy * s->b.g.pitch + (long long)sz * s->b.g.plane);
    s->f0 = f0;
    const double ref[4] = {1, 1, 1, 0};
    for (int k = 0; k < 4; ++k) s->w[k] = w4 ? w4[k] : ref[k];
    s->has_source = true;
    return PW_OK;
}

extern "C" pw_status pw_set_receivers(pw_solver* s, int n, const int* field, const long long* index)
{
    pw_status st = check_handle(s, "pw_set_receivers");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (n < 0 || (n > 0 && (!field || !index))) return fail(PW_ERR_INVALID, "pw_set_receivers: bad arguments");
    const long long ncell = (long long)c.nx * c.ny * c.nz;
    std::vector<unsigned> off(n);
    for (int r = 0; r < n; ++r) {
        if (field[r] < 0 || field[r] >= PW_NFIELDS || index[r] < 0 || index[r] >= ncell)
            return fail(PW_ERR_INVALID, "pw_set_receivers: receiver " + std::to_string(r) + " out of range");
        const long long ix = index[r] % c.nx, iy = (index[r] / c.nx) % c.ny, iz = index[r] / ((long long)c.nx * c.ny);
        off[r] = (unsigned)(ix + iy * s->b.g.pitch + iz * s->b.g.plane);
    }
    PW_HIP_TRY(hipSetDevice(c.device));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    (void)hipFree(s->d_rfield); (void)hipFree(s->d_roff); (void)hipFree(s->d_seis);
    s->d_rfield = nullptr; s->d_roff = nullptr; s->d_seis = nullptr;
    s->nrec = 0;
    s->ref_receivers = false;
    if (n == 0) return PW_OK;
    PW_HIP_TRY(hipMalloc(&s->d_rfield, n * sizeof(int)));
    PW_HIP_TRY(hipMalloc(&s->d_roff, n * sizeof(unsigned)));
    const size_t seis_bytes = align256((size_t)n * c.nt * sizeof(float));
    PW_HIP_TRY(hipMalloc(&s->d_seis, seis_bytes));
    PW_HIP_TRY(hipMemcpy(s->d_rfield, field, n * sizeof(int), hipMemcpyHostToDevice));
    PW_HIP_TRY(hipMemcpy(s->d_roff, off.data(), n * sizeof(unsigned), hipMemcpyHostToDevice));
    PW_HIP_TRY(launch_zero(s->d_seis, seis_bytes, s->stream));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    s->nrec = n;
    return PW_OK;
}

extern "C" pw_status pw_set_receivers_reference(pw_solver* s, int sx, int sy)
{
    pw_status st = check_handle(s, "pw_set_receivers_reference");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (sx < 0 || sy < 0 || sx >= c.nx || sy >= c.ny) return fail(PW_ERR_INVALID, "pw_set_receivers_reference: bad line");
    const int ch[5] = {PW_TXX, PW_TZZ, PW_VUX, PW_VWX, PW_P};  // sisx, sisz, sisvu, sisvw, sisp
    std::vector<int> f;
    std::vector<long long> idx;
    for (int k = 0; k < 5; ++k)
        for (int iz = 0; iz < c.nz; ++iz) {
            f.push_back(ch[k]);
            idx.push_back((long long)sx + (long long)sy * c.nx + (long long)iz * c.nx * c.ny);
        }
    st = pw_set_receivers(s, (int)f.size(), f.data(), idx.data());
    if (st == PW_OK) s->ref_receivers = true;
    return st;
}

extern "C" pw_status pw_set_snapshots(pw_solver* s, int n, const int* step, const int* field, int plane_y)
{
    pw_status st = check_handle(s, "pw_set_snapshots");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (n < 0 || (n > 0 && (!step || !field)) || plane_y < 0 || plane_y >= c.ny)
        return fail(PW_ERR_INVALID, "pw_set_snapshots: bad arguments");
    for (int k = 0; k < n; ++k)
        if (field[k] < 0 || field[k] >= PW_NFIELDS || step[k] < 0 || step[k] >= c.nt)
            return fail(PW_ERR_INVALID, "pw_set_snapshots: snapshot " + std::to_string(k) + " out of range");
    PW_HIP_TRY(hipSetDevice(c.device));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    (void)hipFree(s->d_snaps);
    s->d_snaps = nullptr;
    s->nsnap = 0;
    s->snap_step.assign(step, step + n);
    s->snap_field.assign(field, field + n);
    s->snap_y = plane_y;
    if (n == 0) return PW_OK;
    const size_t bytes = align256((size_t)n * c.nz * c.nx * sizeof(float));
    PW_HIP_TRY(hipMalloc(&s->d_snaps, bytes));
    PW_HIP_TRY(launch_zero(s->d_snaps, bytes, s->stream));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    s->nsnap = n;
    return PW_OK;
}

extern "C" pw_status pw_reset(pw_solver* s)
{
    pw_status st = check_handle(s, "pw_reset");
    if (st != PW_OK) return st;
    PW_HIP_TRY(hipSetDevice(s->cfg.device));
    PW_HIP_TRY(zero_state(s));
    if (s->d_seis) PW_HIP_TRY(launch_zero(s->d_seis, align256((size_t)s->nrec * s->cfg.nt * 4), s->stream));
    if (s->d_snaps)
        PW_HIP_TRY(launch_zero(s->d_snaps, align256((size_t)s->nsnap * s->cfg.nz * s->cfg.nx * 4), s->stream));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->copy_stream) PW_HIP_TRY(hipStreamSynchronize(s->copy_stream));
    std::fill(s->vol_taken.begin(), s->vol_taken.end(), 0);
    s->step = 0;
    s->dyn_scale = 1.0f;
    return PW_OK;
}

namespace {

// dynamic_scale_for_step, Half/main.cu:116-122
float dynamic_scale_for_step(int it, int nt, float end)
{
    if (nt <= 1) return end;
    const float alpha = (float)it / (float)(nt - 1);
    return 1.0f + (end - 1.0f) * alpha;
}

}  // namespace

StepPlan pw::plan_step(const pw_solver* s, int it, float* dyn)
{
    const pw_config& c = s->cfg;
    const bool scaled = is_scaled(s->prec);
    StepPlan p{0.0f, 0.0, 1.0f, 1.0f};
    if (scaled && (it == 0 || (c.half_dyn_interval > 0 && it % c.half_dyn_interval == 0))) {
        const float dyn_end = c.half_dyn_end > 0 ? (float)c.half_dyn_end : 16.0f;
        const float target = dynamic_scale_for_step(it, c.nt, dyn_end);
        const float ratio = target / *dyn;
        if (std::fabs(ratio - 1.0f) > 1.0e-6f) {
            p.ratio = ratio;
            *dyn = target;
        }
    }
    if (s->has_source) {
        if (s->prec == PW_FP64) p.amp = ricker_f64(it, c.dt, s->f0);
        else if (scaled) p.amp = *dyn * s->source_scale * ricker_f32(it, (float)c.dt, (float)s->f0);
        else p.amp = ricker_f32(it, (float)c.dt, (float)s->f0);
    }
    if (scaled) {
        p.s_stress = s->stress_base / *dyn;
        p.s_vel = s->vel_base / *dyn;
    }
    return p;
}

extern "C" pw_status pw_run(pw_solver* s, int nsteps)
{
    pw_status st = check_handle(s, "pw_run");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (!s->has_medium) return fail(PW_ERR_STATE, "pw_run: no medium installed");
    // written so that no sum can overflow: s->step <= c.nt always holds
    if (nsteps < 0 || nsteps > c.nt - s->step)
        return fail(PW_ERR_INVALID, "pw_run: steps beyond the configured nt");
    PW_HIP_TRY(hipSetDevice(c.device));
    for (int k = 0; k < nsteps; ++k) {
        const int it = s->step;
        const StepPlan plan = plan_step(s, it, &s->dyn_scale);
        if (plan.ratio != 0.0f) PW_HIP_TRY(launch_rescale(s->b, s->stream, plan.ratio));
        if (s->has_source) PW_HIP_TRY(launch_source(s->b, s->stream, s->src_off, s->w, plan.amp));
        const int nz = c.nz;
        const int halo = s->b.order == 4 ? 2 : 1;  // planes [halo, nz - halo) are updated
        // one step kernel launch, bracketed by events when profiling
        auto step_kernel = [&](int kind, int zb, int ze) -> pw_status {
            if (ze <= zb) return PW_OK;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (s->profiling) {
                for (hipEvent_t* e : {&e0, &e1}) {
                    if (s->ev_pool.empty()) {
                        hipEvent_t ne = nullptr;
                        PW_HIP_TRY(hipEventCreate(&ne));
                        s->ev_pool.push_back(ne);
                    }
                    *e = s->ev_pool.back();
                    s->ev_pool.pop_back();
                    s->ev_pending.push_back(*e);
                }
                s->ev_kind.push_back((char)kind);
                PW_HIP_TRY(hipEventRecord(e0, s->stream));
            }
            PW_HIP_TRY(kind == 0 ? launch_velocity(s->b, s->stream, zb, ze) : launch_stress(s->b, s->stream, zb, ze));
            if (e1) PW_HIP_TRY(hipEventRecord(e1, s->stream));
            return PW_OK;
        };
        pw_status ks = PW_OK;
        if ((ks = step_kernel(0, halo, nz - halo)) != PW_OK) return ks;
        if ((ks = step_kernel(1, halo, nz - halo)) != PW_OK) return ks;
        if (s->profiling) s->prof_steps++;
        const float ss = plan.s_stress, sv = plan.s_vel;
        if (s->nrec > 0)
            PW_HIP_TRY(launch_gather(s->b, s->stream, s->d_rfield, s->d_roff, s->nrec, s->d_seis, c.nt, it, ss, sv));
        for (int q = 0; q < s->nsnap; ++q)
            if (s->snap_step[q] == it) {
                const float sc = s->snap_field[q] <= PW_VWZ ? sv : ss;
                PW_HIP_TRY(launch_snapshot(s->b, s->stream, s->snap_field[q], s->snap_y,
                                           s->d_snaps + (size_t)q * c.nz * c.nx, sc));
            }
        for (int q = 0; q < s->nvol; ++q)
            if (s->vol_step[q] == it) {
                const float sc = s->vol_field[q] <= PW_VWZ ? sv : ss;
                const size_t vbytes = (size_t)c.nx * c.ny * c.nz * sizeof(float);
                const int j = s->vol_next;
                s->vol_next ^= 1;
                // the staging volume's previous copy must be done before it is overwritten
                PW_HIP_TRY(hipStreamWaitEvent(s->stream, s->ev_copied[j], 0));
                PW_HIP_TRY(launch_decode(s->b, s->stream, s->vol_field[q], s->d_vstage[j], sc));
                PW_HIP_TRY(hipEventRecord(s->ev_decoded[j], s->stream));
                PW_HIP_TRY(hipStreamWaitEvent(s->copy_stream, s->ev_decoded[j], 0));
                PW_HIP_TRY(hipMemcpyAsync(s->vol_host[q], s->d_vstage[j], vbytes, hipMemcpyDeviceToHost,
                                          s->copy_stream));
                PW_HIP_TRY(hipEventRecord(s->ev_copied[j], s->copy_stream));
                PW_HIP_TRY(hipEventRecord(s->vol_done[q], s->copy_stream));
                s->vol_taken[q] = 1;
            }
        s->step++;
    }
    return PW_OK;
}

extern "C" pw_status pw_synchronize(pw_solver* s)
{
    pw_status st = check_handle(s, "pw_synchronize");
    if (st != PW_OK) return st;
    PW_HIP_TRY(hipSetDevice(s->cfg.device));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    return PW_OK;
}

extern "C" int pw_current_step(const pw_solver* s) { return s ? s->step : -1; }

extern "C" pw_status pw_get_seismogram(pw_solver* s, float* host)
{
    pw_status st = check_handle(s, "pw_get_seismogram");
    if (st != PW_OK) return st;
    if (!host) return fail(PW_ERR_INVALID, "pw_get_seismogram: null output");
    if (s->nrec == 0) return PW_OK;
    PW_HIP_TRY(hipSetDevice(s->cfg.device));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    PW_HIP_TRY(hipMemcpy(host, s->d_seis, (size_t)s->nrec * s->cfg.nt * sizeof(float), hipMemcpyDeviceToHost));
    return PW_OK;
}

extern "C" pw_status pw_get_snapshots(pw_solver* s, float* host)
{
    pw_status st = check_handle(s, "pw_get_snapshots");
    if (st != PW_OK) return st;
    if (!host) return fail(PW_ERR_INVALID, "pw_get_snapshots: null output");
    if (s->nsnap == 0) return PW_OK;
    PW_HIP_TRY(hipSetDevice(s->cfg.device));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    PW_HIP_TRY(hipMemcpy(host, s->d_snaps, (size_t)s->nsnap * s->cfg.nz * s->cfg.nx * sizeof(float),
                         hipMemcpyDeviceToHost));
    return PW_OK;
}

extern "C" pw_status pw_set_volume_snapshots(pw_solver* s, int n, const int* step, const int* field)
{
    pw_status st = check_handle(s, "pw_set_volume_snapshots");
    if (st != PW_OK) return st;
    const pw_config& c = s->cfg;
    if (n < 0 || (n > 0 && (!step || !field))) return fail(PW_ERR_INVALID, "pw_set_volume_snapshots: bad arguments");
    for (int k = 0; k < n; ++k)
        if (step[k] < 0 || step[k] >= c.nt || field[k] < 0 || field[k] >= PW_NFIELDS)
            return fail(PW_ERR_INVALID, "pw_set_volume_snapshots: step or field out of range");
    PW_HIP_TRY(hipSetDevice(c.device));
    free_volumes(s);
    if (n == 0) return PW_OK;
    const size_t vbytes = (size_t)c.nx * c.ny * c.nz * sizeof(float);
    hipError_t e = hipSuccess;
    if (!s->copy_stream) e = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking);
    for (int k = 0; k < n && e == hipSuccess; ++k) {
        float* p = nullptr;
        e = hipHostMalloc((void**)&p, vbytes, hipHostMallocDefault);
        if (e != hipSuccess) break;
        s->vol_host.push_back(p);
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) s->vol_done.push_back(ev);
    }
    for (int j = 0; j < 2 && e == hipSuccess; ++j) {
        e = hipMalloc((void**)&s->d_vstage[j], vbytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_decoded[j], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_copied[j], hipEventDisableTiming);
        // recorded once so that the first wait on it passes
        if (e == hipSuccess) e = hipEventRecord(s->ev_copied[j], s->copy_stream);
    }
    if (e != hipSuccess) {
        free_volumes(s);
        (void)hipGetLastError();
        return fail(e == hipErrorOutOfMemory ? PW_ERR_OUT_OF_MEMORY : PW_ERR_HIP,
                    std::string("pw_set_volume_snapshots: ") + hipGetErrorString(e));
    }
    s->vol_step.assign(step, step + n);
    s->vol_field.assign(field, field + n);
    s->vol_taken.assign(n, 0);
    s->nvol = n;
    return PW_OK;
}

extern "C" pw_status pw_get_volume_snapshot(pw_solver* s, int k, float* host)
{
    pw_status st = check_handle(s, "pw_get_volume_snapshot");
    if (st != PW_OK) return st;
    if (k < 0 || k >= s->nvol || !host) return fail(PW_ERR_INVALID, "pw_get_volume_snapshot: bad arguments");
    if (!s->vol_taken[k]) return fail(PW_ERR_STATE, "pw_get_volume_snapshot: its step has not run yet");
    const pw_config& c = s->cfg;
    PW_HIP_TRY(hipSetDevice(c.device));
    PW_HIP_TRY(hipEventSynchronize(s->vol_done[k]));
    std::memcpy(host, s->vol_host[k], (size_t)c.nx * c.ny * c.nz * sizeof(float));
    return PW_OK;
}

extern "C" pw_status pw_get_field(pw_solver* s, int field, float* host)
{
    pw_status st = check_handle(s, "pw_get_field");
    if (st != PW_OK) return st;
    if (!host || field < 0 || field >= PW_NFIELDS) return fail(PW_ERR_INVALID, "pw_get_field: bad arguments");
    const pw_config& c = s->cfg;
    PW_HIP_TRY(hipSetDevice(c.device));
    const size_t bytes = (size_t)c.nx * c.ny * c.nz * sizeof(float);
    float* d = nullptr;
    PW_HIP_TRY(hipMalloc(&d, bytes));
    const bool scaled = is_scaled(s->prec);
    const float sc = !scaled ? 1.0f : (field <= PW_VWZ ? s->vel_base : s->stress_base) / s->dyn_scale;
    hipError_t e = launch_decode(s->b, s->stream, field, d, sc);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(host, d, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_get_field: ") + hipGetErrorString(e));
    return PW_OK;
}

extern "C" pw_status pw_get_info(pw_solver* s, pw_info* out)
{
    pw_status st = check_handle(s, "pw_get_info");
    if (st != PW_OK) return st;
    if (!out) return fail(PW_ERR_INVALID, "pw_get_info: null output");
    const Geom& g = s->b.g;
    out->precision = s->prec;
    out->em = s->em;
    out->epml = s->epml;
    // Compulsory traffic of the two-kernel step: 39 wavefield touches, the
    // coefficient arrays once, and 16 memory-variable touches per damped
    // (cell, axis) (8 variables read + written).
    const double n3[3] = {(double)g.nx, (double)g.ny, (double)g.nz};
    double f = 0.0;
    for (int ax = 0; ax < 3; ++ax) f += (g.lo[ax] + g.hi[ax]) / n3[ax];
    // velocity: reads tau x6 + p, read-modify-writes vu x3 + vw x3 (19 touches);
    // stress: reads vu x3 + vw x3, read-modify-writes tau x6 + p (20 touches).
    // coefficients: scaled form 15 (velocity) + 7 (stress) FP16 values;
    // physical form 4 (rho, rhof, C1, C2) + 5 (C, M, H, H-2mu, mu) cell-centred
    // arrays (comp16f, and FP32 on a medium whose H is (H-2mu) + 2mu, form H
    // from H-2mu and mu; FP32 takes a uniform rhof as a scalar; comp16f: two
    // records of 4 compensated values)
    const bool form_h = s->prec == PW_COMP16F || s->b.derive_h;
    // (comp16: 15 and 7 coefficients in records of 4, so 16 and 8 FP16 values
    // move)
    const int cv = has_srec(s->prec) ? 16 : is_scaled(s->prec) ? 15 : (s->b.rhof_uniform > 0 ? 3 : 4),
              ct = has_srec(s->prec) ? 8 : is_scaled(s->prec) ? 7 : (form_h ? 4 : 5);
    // memory variables: each slot of a kernel read and written per damped axis
    const double mt = mem_slots(s->prec) * f * mem_bytes(s->prec);
    out->bytes_velocity = 19.0 * state_bytes(s->prec) + cv * coef_bytes(s->prec) + mt;
    out->bytes_stress = 20.0 * state_bytes(s->prec) + ct * coef_bytes(s->prec) + mt;
    out->bytes_per_cell_step = out->bytes_velocity + out->bytes_stress;
    out->vp_max = s->vp_max;
    out->courant = s->courant;
    out->device_bytes = (long long)s->pool_bytes;
    for (int ax = 0; ax < 3; ++ax) {
        out->slab_lo[ax] = g.lo[ax];
        out->slab_hi[ax] = g.hi[ax];
    }
    return PW_OK;
}

extern "C" pw_status pw_codec_decode(const unsigned short* hi, const signed char* lo, float* out, long long n,
                                     void* stream)
{
    if (n < 0) return fail(PW_ERR_INVALID, "pw_codec_decode: n < 0");
    if (n > 0 && (!hi || !lo || !out)) return fail(PW_ERR_INVALID, "pw_codec_decode: null buffer");
    hipStream_t st = (hipStream_t)stream;
    PW_HIP_TRY(launch_codec_decode(hi, (const int8_t*)lo, out, (unsigned long long)n, st));
    PW_HIP_TRY(hipStreamSynchronize(st));
    return PW_OK;
}

extern "C" pw_status pw_codec_encode(const float* in, unsigned short* hi, signed char* lo, long long n, int precision,
                                     void* stream)
{
    if (n < 0) return fail(PW_ERR_INVALID, "pw_codec_encode: n < 0");
    if (precision != PW_COMP16 && precision != PW_COMP16F)
        return fail(PW_ERR_INVALID, "pw_codec_encode: precision must be PW_COMP16 or PW_COMP16F");
    if (n > 0 && (!in || !hi || !lo)) return fail(PW_ERR_INVALID, "pw_codec_encode: null buffer");
    hipStream_t st = (hipStream_t)stream;
    PW_HIP_TRY(launch_codec_encode(in, hi, (int8_t*)lo, (unsigned long long)n, precision == PW_COMP16F, st));
    PW_HIP_TRY(hipStreamSynchronize(st));
    return PW_OK;
}

extern "C" pw_status pw_set_profiling(pw_solver* s, int enable)
{
    pw_status st = check_handle(s, "pw_set_profiling");
    if (st != PW_OK) return st;
    s->profiling = enable != 0;
    return PW_OK;
}

extern "C" pw_status pw_get_kernel_times(pw_solver* s, double* ms_velocity, double* ms_stress, long long* launches)
{
    pw_status st = check_handle(s, "pw_get_kernel_times");
    if (st != PW_OK) return st;
    PW_HIP_TRY(hipSetDevice(s->cfg.device));
    PW_HIP_TRY(hipStreamSynchronize(s->stream));
    double tv = 0.0, tt = 0.0;
    for (size_t k = 0; k < s->ev_kind.size(); ++k) {
        float a = 0.0f;
        PW_HIP_TRY(hipEventElapsedTime(&a, s->ev_pending[2 * k], s->ev_pending[2 * k + 1]));
        (s->ev_kind[k] == 1 ? tt : tv) += a;
    }
    for (hipEvent_t e : s->ev_pending) s->ev_pool.push_back(e);
    s->ev_pending.clear();
    s->ev_kind.clear();
    if (ms_velocity) *ms_velocity = tv;
    if (ms_stress) *ms_stress = tt;
    if (launches) *launches = s->prof_steps;
    s->prof_steps = 0;
    return PW_OK;
}
