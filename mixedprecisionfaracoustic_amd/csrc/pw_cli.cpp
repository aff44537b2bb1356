// porowave -- command-line replacement for the reference programs' main()
// (Float/main.cu:300-1477 and, with --precision comp16, Half/main.cu's main()).
//
// With no options it runs the reference shot: 192^3 total grid (128^3 interior
// + 32-cell CPML), H = 1 cm, DT = 0.9 us, 3500 steps, 3 kHz Ricker into txx,
// tyy, tzz at (96, 96, 52), the borehole medium of create_model
// (Float/main.cu:1590-1655), and writes the reference's output files into the
// output directory (Float/main.cu:1187-1222):
//   sisx.dat sisz.dat sisvu.dat sisvw.dat sisp.dat      [nz][nt] receiver lines
//   txx500.dat ... txx3000.dat, txx.dat (after the last step),
//   vux500.dat vuz500.dat vwx500.dat vwz500.dat          y = ny/2 slices [nz][nx]
// in the reference's text format, or as binary files (--format bin).
// Everything runs through the C ABI of libporowave.so.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "porowave.h"

namespace {

void usage(FILE* f)
{
    std::fprintf(f,
                 "usage: porowave [options]\n"
                 "  Runs the reference shot (Float/main.cu; Half/main.cu with --precision comp16)\n"
                 "  and writes its output files.\n"
                 "  --precision P  fp32 (default), comp16, bf16, fp64 or comp16f\n"
                 "  --out DIR      output directory (default: .)\n"
                 "  --format F     dat (default: the reference's text files) or bin\n"
                 "  --nt N         time steps (default 3500)\n"
                 "  --n N          total grid cells per axis, CPML included (default 192)\n"
                 "  --np N         CPML thickness in cells (default 32)\n"
                 "  --h H          grid step in m (default 0.01)\n"
                 "  --dt DT        time step in s (default 0.9e-6)\n"
                 "  --f0 F         Ricker peak frequency in Hz (default 3000)\n"
                 "  --order K      2 (default, the reference) or 4\n"
                 "  --device D     HIP device (default 0)\n"
                 "  --quiet        no summary line\n"
                 "  --help\n");
}

int precision_of(const char* s)
{
    const char* names[] = {"fp32", "comp16", "bf16", "fp64", "comp16f"};
    const int ids[] = {PW_FP32, PW_COMP16, PW_BF16, PW_FP64, PW_COMP16F};
    for (int k = 0; k < 5; ++k)
        if (std::strcmp(s, names[k]) == 0) return ids[k];
    return -1;
}

bool parse_int(const char* s, int* v)
{
    char* end = nullptr;
    const long x = std::strtol(s, &end, 10);
    if (!end || *end || end == s || x < -2147483647L || x > 2147483647L) return false;
    *v = (int)x;
    return true;
}

bool parse_double(const char* s, double* v)
{
    char* end = nullptr;
    *v = std::strtod(s, &end);
    return end && !*end && end != s;
}

int die(const char* what)
{
    std::fprintf(stderr, "porowave: %s: %s\n", what, pw_last_error());
    return 1;
}

pw_status write_one(const std::string& dir, const char* name, bool bin, int kind, int field, int step,
                    const pw_config& c, const float* data, int nr, int nc)
{
    const std::string path = dir + "/" + name + (bin ? ".bin" : ".dat");
    if (!bin) return pw_write_dat(path.c_str(), data, nr, nc);
    pw_bin_header h;
    std::memset(&h, 0, sizeof(h));
    h.kind = kind;
    h.nr = nr;
    h.nc = nc;
    h.field = field;
    h.step = step;
    h.dt = c.dt;
    h.h = c.h;
    h.scale = 1.0;
    std::snprintf(h.label, sizeof(h.label), "%s", name);
    return pw_write_bin(path.c_str(), &h, data);
}

}  // namespace

int main(int argc, char** argv)
{
    pw_config cfg;
    pw_config_default(&cfg);  // Float/main.cu:20-23, 305-322
    std::string out = ".", prec_name = "fp32";
    bool bin = false, quiet = false;
    double f0 = 3000.0;
    int n = cfg.nx;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help" || a == "-h") {
            usage(stdout);
            return 0;
        }
        if (a == "--quiet") {
            quiet = true;
            continue;
        }
        if (i + 1 >= argc) {
            std::fprintf(stderr, "porowave: %s needs a value\n", a.c_str());
            usage(stderr);
            return 2;
        }
        const char* v = argv[++i];
        bool ok = true;
        if (a == "--precision") ok = (cfg.precision = precision_of(v)) >= 0, prec_name = v;
        else if (a == "--out") out = v;
        else if (a == "--format") ok = (bin = std::strcmp(v, "bin") == 0) || std::strcmp(v, "dat") == 0;
        else if (a == "--nt") ok = parse_int(v, &cfg.nt) && cfg.nt >= 1;
        else if (a == "--n") ok = parse_int(v, &n);
        else if (a == "--np") ok = parse_int(v, &cfg.np);
        else if (a == "--h") ok = parse_double(v, &cfg.h);
        else if (a == "--dt") ok = parse_double(v, &cfg.dt);
        else if (a == "--f0") ok = parse_double(v, &f0) && f0 > 0;
        else if (a == "--order") ok = parse_int(v, &cfg.space_order);
        else if (a == "--device") ok = parse_int(v, &cfg.device);
        else {
            std::fprintf(stderr, "porowave: unknown option %s\n", a.c_str());
            usage(stderr);
            return 2;
        }
        if (!ok) {
            std::fprintf(stderr, "porowave: bad value for %s: %s\n", a.c_str(), v);
            return 2;
        }
    }
    cfg.nx = cfg.ny = cfg.nz = n;

    pw_solver* s = nullptr;
    if (pw_create(&cfg, &s) != PW_OK) return die("pw_create");
    const int ni = cfg.nx - 2 * cfg.np;
    const size_t nc3 = (size_t)ni * ni * ni;
    std::vector<double> base((size_t)PW_NBASE * nc3);
    if (pw_medium_borehole(ni, ni, ni, base.data()) != PW_OK) return die("pw_medium_borehole");  // create_model
    const double* f[PW_NBASE];
    for (int k = 0; k < PW_NBASE; ++k) f[k] = base.data() + (size_t)k * nc3;
    if (pw_set_medium(s, f, ni, ni, ni) != PW_OK) return die("pw_set_medium");  // extmodel .. H2D
    base.clear();
    base.shrink_to_fit();

    // source and receivers of Float/main.cu:310-313, 321, 1085-1093
    const int sx = cfg.nx / 2, sy = cfg.ny / 2, sz = 20 + cfg.np;
    const double w[4] = {1.0, 1.0, 1.0, 0.0};
    if (pw_set_source(s, sx, sy, sz, f0, w) != PW_OK) return die("pw_set_source");
    if (pw_set_receivers_reference(s, sx, sy) != PW_OK) return die("pw_set_receivers_reference");

    // snapshots of Float/main.cu:1094-1163: txx every 500 steps, txx after the
    // last step, and the four velocity slices at step 500
    std::vector<int> steps, fields;
    std::vector<std::string> names;
    for (int it = 500; it < cfg.nt; it += 500) {
        steps.push_back(it);
        fields.push_back(PW_TXX);
        names.push_back("txx" + std::to_string(it));
    }
    steps.push_back(cfg.nt - 1);
    fields.push_back(PW_TXX);
    names.push_back("txx");
    if (500 < cfg.nt) {
        const int vf[4] = {PW_VUX, PW_VUZ, PW_VWX, PW_VWZ};
        const char* vn[4] = {"vux500", "vuz500", "vwx500", "vwz500"};
        for (int k = 0; k < 4; ++k) {
            steps.push_back(500);
            fields.push_back(vf[k]);
            names.push_back(vn[k]);
        }
    }
    if (pw_set_snapshots(s, (int)steps.size(), steps.data(), fields.data(), cfg.ny / 2) != PW_OK)
        return die("pw_set_snapshots");

    const auto t0 = std::chrono::steady_clock::now();
    if (pw_run(s, cfg.nt) != PW_OK || pw_synchronize(s) != PW_OK) return die("pw_run");
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    std::vector<float> seis((size_t)5 * cfg.nz * cfg.nt), snaps(steps.size() * (size_t)cfg.nz * cfg.nx);
    if (pw_get_seismogram(s, seis.data()) != PW_OK) return die("pw_get_seismogram");
    if (pw_get_snapshots(s, snaps.data()) != PW_OK) return die("pw_get_snapshots");
    const char* sis[5] = {"sisx", "sisz", "sisvu", "sisvw", "sisp"};
    const int sis_field[5] = {PW_TXX, PW_TZZ, PW_VUX, PW_VWX, PW_P};
    for (int c = 0; c < 5; ++c)  // wfile (Float/main.cu:1213-1222)
        if (write_one(out, sis[c], bin, PW_BIN_SEISMOGRAM, sis_field[c], -1, cfg,
                      seis.data() + (size_t)c * cfg.nz * cfg.nt, cfg.nz, cfg.nt) != PW_OK)
            return die("writing seismograms");
    for (size_t k = 0; k < steps.size(); ++k)  // wfile3d (Float/main.cu:1187-1211)
        if (write_one(out, names[k].c_str(), bin, PW_BIN_SNAPSHOT, fields[k], steps[k], cfg,
                      snaps.data() + k * (size_t)cfg.nz * cfg.nx, cfg.nz, cfg.nx) != PW_OK)
            return die("writing snapshots");
    pw_destroy(s);
    if (!quiet)
        std::printf("porowave: %d^3, %d steps, %s: %.3f s, %.3f Gcell-steps/s; outputs in %s\n", n, cfg.nt,
                    prec_name.c_str(), sec, (double)n * n * n * cfg.nt / sec / 1e9, out.c_str());
    return 0;
}
