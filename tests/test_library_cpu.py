"""CPU-side checks of the product library: it loads, exports every entry point
include/porowave.h declares, and its host-only functions (no GPU needed) agree
with the reference's formulas."""
import os
import re

import numpy as np
import pytest

import mixedprecisionfaracoustic_amd as pw
from oracle import pw_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "porowave.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pw_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = pw.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) == set(pw._SIGS), set(syms) ^ set(pw._SIGS)


def test_borehole_medium_matches_create_model_bitwise():
    ref = O.borehole_model(24, 20, 28)
    got = pw.medium_borehole(28, 20, 24)
    assert got.dtype == np.float64
    assert np.array_equal(got.astype(np.float32), ref)
    assert np.array_equal(got, ref.astype(np.float64))


def test_biot_from_rock_matches_oracle():
    r = pw.rock(Ks=35e9, Kb=15e9, mu=14e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.1, kappa=1e-14,
                eta=1e-3, tortuosity=2.0, pore_m=8.0)
    got = pw.biot_from_rock(r)
    ref = O.biot_from_moduli(35e9, 15e9, 14e9, 2.25e9, 2650.0, 1000.0, 0.1, 1e-14, 1e-3, 2.0, 8.0)
    assert np.allclose(got, ref, rtol=1e-15, atol=0)
    bad = pw.rock(Ks=35e9, Kb=15e9, mu=14e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.0, kappa=1e-14,
                  eta=1e-3)
    with pytest.raises(pw.PorowaveError) as ei:
        pw.biot_from_rock(bad)
    assert ei.value.status == 1 and "non-physical" in str(ei.value)


def test_write_dat_reproduces_reference_text_format(tmp_path):
    gold = np.load(os.path.join(ROOT, "tests", "golden", "float_snapshots.npz"))["vwz500"]
    p = tmp_path / "vwz500.dat"
    pw.write_dat(str(p), gold)
    text = p.read_text()
    lines = text.split("\n")
    # wfile3d: every row is 192 "%e " tokens then "\n", then an empty line
    assert len(lines) == 2 * 192 + 1 and all(lines[2 * i + 1] == "" for i in range(192))
    row0 = lines[0]
    assert row0.endswith(" ") and row0.count(" ") == 192
    assert lines[0].split() == ["%e" % float(v) for v in gold[0]]
    back = np.array([[float(t) for t in ln.split()] for ln in lines if ln.strip()], np.float32)
    assert np.array_equal(back, gold)


def test_write_dat_error_path(tmp_path):
    with pytest.raises(pw.PorowaveError) as ei:
        pw.write_dat(str(tmp_path / "no" / "such" / "dir.dat"), np.zeros((2, 2), np.float32))
    assert ei.value.status == 7


def test_create_without_gpu_reports_error():
    # In a container without a GPU pw_create must fail cleanly (never fall back).
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(pw.PorowaveError):
        pw.Solver(20, 20, 20, 4, 0.01, 0.9e-6, 10)


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirror of every struct in include/porowave.h has the C layout
    (size and field offsets, as gcc lays the header out)."""
    import ctypes as C
    import subprocess
    structs = {"pw_config": pw.Config, "pw_info": pw.Info, "pw_rock": pw.Rock, "pw_layered": pw.Layered,
               "pw_server_config": pw.ServerConfig, "pw_shot": pw.ShotSpec, "pw_shot_status": pw.ShotStatus,
               "pw_server_stats": pw.ServerStats, "pw_bin_header": pw.BinHeader}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "porowave.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, v = ln.split()
            got[(a, b)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_server_needs_a_device():
    """Without a GPU the server refuses to start (no CPU fallback); server calls
    on no server are refused."""
    with pytest.raises(pw.PorowaveError):
        pw.Server()
    assert pw.lib().pw_server_cancel(None, 1) == 1  # PW_ERR_INVALID


def test_config_validation_errors():
    """Bad configurations are refused with their reason before any device call."""
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Solver(40, 40, 40, 6, 5.0, 5e-4, 10, space_order=3)
    assert ei.value.status == 1 and "space_order" in str(ei.value)
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Solver(16, 40, 40, 6, 5.0, 5e-4, 10, space_order=4)  # 4th order needs 2*np+5 cells
    assert ei.value.status == 1
    # 1152 x 1100 x 1200 padded cells (1.52e9) fit 32-bit offsets for FP32, but a
    # compensated field's residual bytes, addressed 2 B per cell past its FP16
    # values, would not
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Solver(1100, 1100, 1200, 20, 5.0, 5e-4, 10, precision="comp16")
    assert ei.value.status == 6 and "2^32 / 3" in str(ei.value)
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Solver(1700, 1700, 1600, 20, 5.0, 5e-4, 10, precision="fp32")
    assert ei.value.status == 6 and "2^32 padded cells" in str(ei.value)
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Solver(40, 40, 40, 6, 5.0, 5e-4, 10, tiling={"cells": 3})
    assert ei.value.status == 1 and "tile_cells" in str(ei.value)
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Solver(40, 40, 40, 6, 5.0, 5e-4, 10, space_order=4, tiling={"cells": 2})
    assert ei.value.status == 6 and "one cell per lane" in str(ei.value)  # PW_ERR_UNSUPPORTED
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Solver(40, 40, 40, 6, 5.0, 5e-4, 10, tiling={"rows": 1, "waves_x": 2})
    assert ei.value.status == 1 and "tile_rows" in str(ei.value)
    with pytest.raises(ValueError):
        pw.Solver(40, 40, 40, 6, 5.0, 5e-4, 10, tiling={"zc": 4})
    # the persistent round kernel is a form of the batched rounds
    with pytest.raises(pw.PorowaveError) as ei:
        pw.Server(persistent=True, batched=False)
    assert ei.value.status == 1 and "persistent" in str(ei.value)
    for bad in (-1.0, float("nan"), float("inf")):
        with pytest.raises(pw.PorowaveError) as ei:
            pw.Server(max_wait_s=bad)
        assert ei.value.status == 1 and "max_wait_s" in str(ei.value)


def test_binary_files_round_trip(tmp_path):
    """pw_write_bin / pw_read_bin: an 80-byte header then the values, bit-exact."""
    import ctypes as C
    assert C.sizeof(pw.BinHeader) == 80
    rng = np.random.default_rng(3)
    a = rng.standard_normal((7, 33)).astype(np.float32)
    a[0, 0] = np.float32(1e-40)  # a subnormal survives
    p = tmp_path / "sisx.bin"
    pw.write_bin(str(p), a, kind=0, field=6, dt=0.5e-3, h=5.0, label="sisx")
    raw = p.read_bytes()
    assert len(raw) == 80 + a.size * 4 and raw[:4] == b"PWB1"
    assert np.array_equal(np.frombuffer(raw[80:], "<f4").reshape(a.shape), a)
    head, b = pw.read_bin(str(p))
    assert np.array_equal(a, b)
    assert (head["kind"], head["nr"], head["nc"], head["field"], head["step"]) == (0, 7, 33, 6, -1)
    assert head["dt"] == 0.5e-3 and head["h"] == 5.0 and head["label"] == "sisx" and head["version"] == 1
    q = tmp_path / "junk.bin"
    q.write_bytes(b"not a porowave file" * 10)
    with pytest.raises(pw.PorowaveError) as ei:
        pw.read_bin(str(q))
    assert ei.value.status == 7  # PW_ERR_IO


def test_cli_usage_and_argument_errors():
    """The porowave command line (the reference main()'s replacement) parses its
    options before touching a device: --help succeeds, bad values exit 2, and
    without a GPU pw_create's error comes back as exit 1 with its text."""
    import subprocess
    cli = os.path.join(ROOT, "mixedprecisionfaracoustic_amd", "porowave")
    if not os.path.exists(cli):
        pytest.skip("porowave not built")
    r = subprocess.run([cli, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--precision" in r.stdout
    for bad in (["--precision", "fp8"], ["--format", "csv"], ["--nt", "0"], ["--n", "x"], ["--bogus", "1"], ["--nt"]):
        r = subprocess.run([cli] + bad, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (bad, r.stderr)
    import torch
    if not torch.cuda.is_available():
        r = subprocess.run([cli, "--nt", "2", "--quiet"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 1 and "pw_create" in r.stderr


def test_codec_entry_points_refuse_bad_arguments():
    """pw_codec_decode / pw_codec_encode check their arguments before any device call."""
    for call in (lambda: pw.codec_decode(0, 0, 0, -1), lambda: pw.codec_decode(0, 0, 0, 8),
                 lambda: pw.codec_encode(0, 0, 0, -1, "comp16"), lambda: pw.codec_encode(0, 0, 0, 8, "comp16f"),
                 lambda: pw.codec_encode(0, 0, 0, 8, "bf16")):
        with pytest.raises(pw.PorowaveError) as ei:
            call()
        assert ei.value.status == 1


def test_integration_stub_config_matches_the_library():
    """The ctypes stub shown in INTEGRATION.md declares pw_config with every
    field, in order (pw_config_default writes all of them)."""
    import ctypes as C
    src = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    i = src.index("class Config(C.Structure):")
    j = src.index("# every field of pw_config", i)
    ns = {"C": C}
    exec(src[i:j], ns)
    stub = ns["Config"]
    assert C.sizeof(stub) == C.sizeof(pw.Config)
    assert [f for f, _ in stub._fields_] == [f for f, _ in pw.Config._fields_]
