// pw_output.cpp -- output files: the binary seismogram / snapshot format and the
// reference program's output set (Float/main.cu:1187-1222) written from a handle.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pw_device.h"
#include "pw_solver.h"

using namespace pw;

namespace {

const char kMagic[4] = {'P', 'W', 'B', '1'};

pw_status write_file(const std::string& path, const void* head, size_t hbytes, const float* data, size_t n)
{
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return fail(PW_ERR_IO, "cannot open " + path);
    bool ok = std::fwrite(head, 1, hbytes, fp) == hbytes;
    if (ok && n) ok = std::fwrite(data, sizeof(float), n, fp) == n;
    if (std::fclose(fp) != 0) ok = false;
    return ok ? PW_OK : fail(PW_ERR_IO, "write failed: " + path);
}

}  // namespace

extern "C" pw_status pw_write_bin(const char* path, const pw_bin_header* hdr, const float* data)
{
    if (!path || !hdr || hdr->nr < 0 || hdr->nc < 0 || (!data && (size_t)hdr->nr * hdr->nc > 0))
        return fail(PW_ERR_INVALID, "pw_write_bin: bad arguments");
    pw_bin_header h = *hdr;
    std::memcpy(h.magic, kMagic, 4);
    h.version = 1;
    h.label[sizeof(h.label) - 1] = 0;
    return write_file(path, &h, sizeof(h), data, (size_t)h.nr * h.nc);
}

extern "C" pw_status pw_read_bin(const char* path, pw_bin_header* hdr, float* data)
{
    if (!path || !hdr) return fail(PW_ERR_INVALID, "pw_read_bin: bad arguments");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(PW_ERR_IO, std::string("pw_read_bin: cannot open ") + path);
    pw_bin_header h;
    bool ok = std::fread(&h, 1, sizeof(h), fp) == sizeof(h) && std::memcmp(h.magic, kMagic, 4) == 0 &&
              h.version == 1 && h.nr >= 0 && h.nc >= 0;
    if (ok && data) {
        const size_t n = (size_t)h.nr * h.nc;
        ok = std::fread(data, sizeof(float), n, fp) == n;
    }
    std::fclose(fp);
    if (!ok) return fail(PW_ERR_IO, std::string("pw_read_bin: not a readable porowave binary file: ") + path);
    *hdr = h;
    return PW_OK;
}

namespace {

const char* const kFieldName[PW_NFIELDS] = {"vux", "vuy", "vuz", "vwx", "vwy", "vwz", "txx",
                                            "tyy", "tzz", "txy", "txz", "tyz", "p"};

pw_status write_one(const std::string& base, int format, int kind, int field, int step, const pw_solver* s,
                    const float* data, int nr, int nc, const char* label)
{
    if (format == PW_FORMAT_DAT) return pw_write_dat((base + ".dat").c_str(), data, nr, nc);
    pw_bin_header h;
    std::memset(&h, 0, sizeof(h));
    h.kind = kind;
    h.nr = nr;
    h.nc = nc;
    h.field = field;
    h.step = step;
    h.dt = s->cfg.dt;
    h.h = s->cfg.h;
    h.scale = 1.0;
    std::snprintf(h.label, sizeof(h.label), "%s", label);
    return pw_write_bin((base + ".bin").c_str(), &h, data);
}

}  // namespace

extern "C" pw_status pw_write_outputs(pw_solver* s, const char* dir, int format)
{
    if (!s || !dir) return fail(PW_ERR_INVALID, "pw_write_outputs: null argument");
    if (format != PW_FORMAT_DAT && format != PW_FORMAT_BIN) return fail(PW_ERR_INVALID, "pw_write_outputs: format");
    const pw_config& c = s->cfg;
    const std::string d = std::string(dir) + "/";
    pw_status st;
    if (s->nrec > 0) {
        std::vector<float> seis((size_t)s->nrec * c.nt);
        if ((st = pw_get_seismogram(s, seis.data())) != PW_OK) return st;
        if (s->ref_receivers) {
            // the reference's five files, one row per depth (Float/main.cu:1213-1222)
            const char* names[5] = {"sisx", "sisz", "sisvu", "sisvw", "sisp"};
            const int fields[5] = {PW_TXX, PW_TZZ, PW_VUX, PW_VWX, PW_P};
            for (int k = 0; k < 5; ++k)
                if ((st = write_one(d + names[k], format, PW_BIN_SEISMOGRAM, fields[k], -1, s,
                                    seis.data() + (size_t)k * c.nz * c.nt, c.nz, c.nt, names[k])) != PW_OK)
                    return st;
        } else if ((st = write_one(d + "seismogram", format, PW_BIN_SEISMOGRAM, -1, -1, s, seis.data(), s->nrec,
                                   c.nt, "seismogram")) != PW_OK) {
            return st;
        }
    }
    if (s->nsnap > 0) {
        std::vector<float> snaps((size_t)s->nsnap * c.nz * c.nx);
        if ((st = pw_get_snapshots(s, snaps.data())) != PW_OK) return st;
        for (int k = 0; k < s->nsnap; ++k) {
            // txx500, vux500, ... (Float/main.cu:1187-1212)
            const std::string name = std::string(kFieldName[s->snap_field[k]]) + std::to_string(s->snap_step[k]);
            if ((st = write_one(d + name, format, PW_BIN_SNAPSHOT, s->snap_field[k], s->snap_step[k], s,
                                snaps.data() + (size_t)k * c.nz * c.nx, c.nz, c.nx, name.c_str())) != PW_OK)
                return st;
        }
    }
    return PW_OK;
}
