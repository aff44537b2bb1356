"""The porowave command line, end to end through its output files: the
reference shot (192^3 borehole, 3 kHz) run by the program that replaces
Float/main.cu:main(), its text files read back and compared with the snapshot
files the reference committed (tests/golden/float_snapshots.npz) and with the
same run through the Python binding.  A comp16 run and a 4th-order run write
binary files that equal the same runs through the binding."""
import os
import subprocess

import numpy as np
import pytest

import mixedprecisionfaracoustic_amd as pw
from cases import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "mixedprecisionfaracoustic_amd", "porowave")
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "float_snapshots.npz"))
SIS = ("sisx", "sisz", "sisvu", "sisvw", "sisp")


def run_cli(args, tmp_path):
    assert os.path.exists(CLI), "porowave is not built"
    print("running:", " ".join([CLI] + args), flush=True)
    r = subprocess.run([CLI, "--out", str(tmp_path)] + args, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    return r


def test_cli_reference_shot_files_match_reference(tmp_path):
    nt = 501
    run_cli(["--nt", str(nt)], tmp_path)
    names = sorted(os.listdir(tmp_path))
    assert names == sorted([s + ".dat" for s in SIS] +
                           ["txx500.dat", "txx.dat", "vux500.dat", "vuz500.dat", "vwx500.dat", "vwz500.dat"])
    for n in ("txx500", "vux500", "vuz500", "vwx500", "vwz500"):
        got = np.loadtxt(tmp_path / f"{n}.dat", dtype=np.float64)
        assert got.shape == (192, 192)
        e = rel_l2(got, GOLD[n])
        print(f"{n}.dat vs the reference's file: rel-L2 {e:.3e}")
        assert e <= 1e-5
    # txx.dat is the slice after the last step (here step 500)
    assert np.array_equal(np.loadtxt(tmp_path / "txx.dat"), np.loadtxt(tmp_path / "txx500.dat"))
    # the seismogram files hold the Python binding's run of the same shot, to text precision
    s = pw.reference_solver("fp32", nt=nt)
    s.run(nt)
    seis = s.reference_seismograms()
    s.close()
    for c in SIS:
        got = np.loadtxt(tmp_path / f"{c}.dat", dtype=np.float64)
        assert got.shape == (192, nt)
        ref = seis[c].astype(np.float64)
        assert np.all(np.abs(got - ref) <= 5.01e-7 * np.abs(ref)), c


def test_cli_comp16_binary_outputs(tmp_path):
    nt = 200
    run_cli(["--precision", "comp16", "--nt", str(nt), "--format", "bin"], tmp_path)
    s = pw.reference_solver("comp16", nt=nt)
    s.run(nt)
    seis = s.reference_seismograms()
    s.close()
    for c in SIS:
        h, data = pw.read_bin(str(tmp_path / f"{c}.bin"))
        assert (h["nr"], h["nc"]) == (192, nt) and h["kind"] == 0
        assert np.array_equal(data, seis[c]), c
    h, data = pw.read_bin(str(tmp_path / "txx.bin"))
    assert h["step"] == nt - 1 and h["kind"] == 1 and np.isfinite(data).all() and np.abs(data).max() > 0


def test_cli_fourth_order_matches_the_binding(tmp_path):
    """--order 4: the program's seismogram files equal the same 4th-order run
    through the Python binding, bit for bit."""
    nt = 150
    run_cli(["--order", "4", "--nt", str(nt), "--format", "bin"], tmp_path)
    s = pw.reference_solver("fp32", nt=nt, space_order=4)
    s.run(nt)
    seis = s.reference_seismograms()
    s.close()
    for c in SIS:
        h, data = pw.read_bin(str(tmp_path / f"{c}.bin"))
        assert (h["nr"], h["nc"]) == (192, nt), c
        assert np.array_equal(data, seis[c]), c
    assert np.abs(seis["sisp"]).max() > 0
