"""bench.py's host-side contract, without a GPU: the byte accounting behind
`roofline` and `reference_scheme`, the serving trace generator, the accuracy
figures against FP64, and the one-GPU guard."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("prec,total", [("fp32", 254.0), ("comp16", 183.5), ("bf16", 137.0), ("fp64", 508.0)])
def test_reference_scheme_bytes_at_512(prec, total):
    """The reference two-kernel scheme at 512^3 with a 20-cell CPML: 39 state
    touches, 17 medium arrays (Half: 22) and 32 CPML touches per damped
    (cell, axis), f = 3 * 40 / 512 = 0.234375.  Totals worked by hand:
    FP32 156 + 68 + 30 = 254; comp16 117 + 44 + 22.5 = 183.5;
    BF16 78 + 44 + 15 = 137; FP64 312 + 136 + 60 = 508."""
    v, t = bench.reference_scheme_bytes(prec, 512, 20)
    assert v + t == pytest.approx(total, abs=1e-9)
    # the velocity kernel carries the 3 extra medium arrays (and the Half scheme's 8 more)
    assert v - t == pytest.approx({"fp32": 8.0, "comp16": 13.0, "bf16": 14.0, "fp64": 16.0}[prec], abs=1e-9)


def test_reference_scheme_bytes_no_cpml():
    """With no damped cells the figure is the interior one: 39 s + M c."""
    v, t = bench.reference_scheme_bytes("fp32", 512, 0)
    assert (v, t) == (19 * 4 + 10 * 4, 20 * 4 + 7 * 4)


def test_serving_trace_is_seeded_and_in_range():
    """The shot trace: seeded, arrivals from 0 in order, grids from the four
    sizes, crops inside the 512^3 model's interior, sources and receivers
    inside each shot's own interior."""
    a = bench.serving_trace(64, 4.0, 0, 512, 20)
    assert a == bench.serving_trace(64, 4.0, 0, 512, 20)
    assert a != bench.serving_trace(64, 4.0, 1, 512, 20)
    assert len(a) == 64 and a[0]["t"] == 0.0
    assert all(x["t"] <= y["t"] for x, y in zip(a, a[1:]))
    for s in a:
        n, ni = s["n"], s["n"] - 40
        assert n in (96, 128, 192, 256)
        assert all(0 <= c and c + ni <= 512 - 40 for c in s["crop"])
        assert all(20 <= c < n - 20 for c in s["src"])
        assert 500 <= s["nt"] <= 1500 and s["prec"] in ("fp32", "comp16f")
        assert len(s["idx"]) == len(s["fields"]) == 128
        assert all(0 <= i < n ** 3 for i in s["idx"])


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_accuracy_against_fp64_per_channel():
    """Relative L2 per receiver channel and of the last snapshot; a channel
    whose FP64 trace is all zero (no arrival yet) reports None."""
    import numpy as np
    rng = np.random.default_rng(1)
    ref_seis = rng.standard_normal((6, 40)).astype(np.float32)
    ref_seis[4:] = 0.0
    ref_snap = rng.standard_normal((2, 5, 7)).astype(np.float32)
    seis = ref_seis.copy()
    seis[0:2] *= np.float32(1.5)   # txx: error 0.5
    seis[2:4] += ref_seis[2:4]      # p: error 1
    snap = ref_snap.copy()
    snap[-1] *= np.float32(0.75)    # last snapshot: error 0.25
    acc = bench.accuracy((seis, snap), (ref_seis, ref_snap), ["txx", "txx", "p", "p", "vwx", "vwx"])
    assert acc["seis_rel_l2"]["txx"] == pytest.approx(0.5, rel=1e-6)
    assert acc["seis_rel_l2"]["p"] == pytest.approx(1.0, rel=1e-6)
    assert acc["seis_rel_l2"]["vwx"] is None
    assert acc["snapshot_txx_rel_l2"] == pytest.approx(0.25, rel=1e-6)
    assert acc["fp64_seis_peak"] == pytest.approx(float(np.abs(ref_seis).max()), rel=1e-3)
