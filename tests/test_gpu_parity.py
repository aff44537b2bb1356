"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the
same seeded inputs.

Tolerances (relative L2 over each channel / field):
  FP32    vs oracle Float path (Float/main.cu)   <= 1e-5
  FP64    vs oracle FP64 run                      <= 1e-10
  comp16  vs oracle Half path (Half/main.cu)      <= 1e-3
  comp16  vs oracle FP64 run                      <= 1e-3
  BF16    vs oracle FP64 run                      <= 3e-2 (reported; 8-bit mantissa storage)
"""
import numpy as np
import pytest

import cases
from cases import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vux", "vuz", "vwx", "vwz", "txx", "tzz", "txz", "p")


def per_channel(seis, recs):
    out = {}
    for f in sorted(set(f for f, _ in recs)):
        idx = [k for k, (g, _) in enumerate(recs) if g == f]
        out[f] = seis[idx]
    return out


def check(gpu, ref, recs, tol, label):
    worst = 0.0
    g, r = per_channel(gpu["seis"], recs), per_channel(ref["seis"], recs)
    for f in g:
        e = rel_l2(g[f], r[f])
        worst = max(worst, e)
        assert e <= tol, f"{label}: seismogram {f} rel-L2 {e:.3e} > {tol}"
    if ref.get("final") is not None:
        for f in gpu["fields"]:
            e = rel_l2(gpu["fields"][f], ref["final"][cases.O.FIELD_ID[f]])
            worst = max(worst, e)
            assert e <= tol, f"{label}: field {f} rel-L2 {e:.3e} > {tol}"
    print(f"{label}: worst rel-L2 {worst:.3e}")
    return worst


@pytest.fixture(scope="module")
def borehole():
    grid, base, nt = cases.borehole_case()
    recs = cases.receivers(grid)
    ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    return grid, base, nt, recs, ref32


def test_fp32_borehole_matches_oracle(borehole):
    grid, base, nt, recs, ref32 = borehole
    gpu, _ = cases.run_gpu(grid, nt, "fp32", recs, base=base, fields=FIELDS)
    check(gpu, ref32, recs, 1e-5, "fp32 borehole")
    # the fluid column keeps vw exactly zero (Float/main.cu:160-163)
    vwx = gpu["fields"]["vwx"]
    nx, ny = grid["nx"], grid["ny"]
    assert np.all(vwx[:, ny // 2, nx // 2] == 0.0)


def test_fp32_chunked_run_is_identical(borehole):
    grid, base, nt, recs, _ = borehole
    a, _ = cases.run_gpu(grid, nt, "fp32", recs, base=base, fields=("txx",))
    b, _ = cases.run_gpu(grid, nt, "fp32", recs, base=base, fields=("txx",), chunks=[1, 63, 100, 36])
    assert np.array_equal(a["seis"], b["seis"])
    assert np.array_equal(a["fields"]["txx"], b["fields"]["txx"])


def test_fp64_borehole_matches_oracle(borehole):
    grid, base, nt, recs, _ = borehole
    ref64 = cases.run_oracle("f64", grid, nt, recs, base=base, final=True)
    gpu, _ = cases.run_gpu(grid, nt, "fp64", recs, base=base, fields=FIELDS)
    check(gpu, ref64, recs, 1e-10, "fp64 borehole")


def test_comp16_borehole_matches_half_oracle(borehole):
    grid, base, nt, recs, _ = borehole
    refh = cases.run_oracle("half", grid, nt, recs, base=base, final=True)
    gpu, _ = cases.run_gpu(grid, nt, "comp16", recs, base=base, fields=FIELDS)
    check(gpu, refh, recs, 1e-3, "comp16 borehole vs Half")


@pytest.fixture(scope="module")
def layered():
    grid, spec, nt = cases.layered_case()
    recs = cases.receivers(grid)
    gpu32, base = cases.run_gpu(grid, nt, "fp32", recs, spec=spec, fields=FIELDS)
    return grid, spec, nt, recs, gpu32, base


def test_fp32_layered_matches_oracle(layered):
    grid, spec, nt, recs, gpu32, base = layered
    ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    check(gpu32, ref32, recs, 1e-5, "fp32 layered")


def test_precisions_vs_fp64_layered(layered):
    grid, spec, nt, recs, gpu32, base = layered
    ref64 = cases.run_oracle("f64", grid, nt, recs, base=base, final=True)
    g64, _ = cases.run_gpu(grid, nt, "fp64", recs, spec=spec, fields=FIELDS)
    check(g64, ref64, recs, 1e-10, "fp64 layered")
    check(gpu32, ref64, recs, 1e-4, "fp32 layered vs fp64")
    g16, _ = cases.run_gpu(grid, nt, "comp16", recs, spec=spec, fields=FIELDS)
    check(g16, ref64, recs, 1e-3, "comp16 layered vs fp64")
    refh = cases.run_oracle("half", grid, nt, recs, base=base, final=True)
    check(g16, refh, recs, 1e-3, "comp16 layered vs Half")
    gb, _ = cases.run_gpu(grid, nt, "bf16", recs, spec=spec, fields=FIELDS)
    check(gb, ref64, recs, 3e-2, "bf16 layered vs fp64")


def test_layered_generator_is_deterministic(layered):
    grid, spec, nt, recs, gpu32, base = layered
    _, base2 = cases.run_gpu(grid, 1, "fp32", recs[:1], spec=spec)
    assert np.array_equal(base, base2)
    # two layers, each perturbed by at most perturb * clamp
    nzi = grid["nz"] - 2 * grid["np"]
    mu = base[8]
    top = mu[: nzi // 2 - 1]
    assert np.all(np.abs(top / 9e9 - 1) <= 0.05 * 3 + 1e-12)
    assert 0.01 < np.std(top / 9e9 - 1) < 0.08


def test_snapshots_match_fields(borehole):
    grid, base, nt, recs, _ = borehole
    snaps = [(nt - 1, "txx"), (nt - 1, "vuz"), (50, "p")]
    gpu, _ = cases.run_gpu(grid, nt, "fp32", recs, base=base, snaps=snaps, fields=("txx", "vuz"))
    y = grid["ny"] // 2
    assert np.array_equal(gpu["snaps"][0], gpu["fields"]["txx"][:, y, :])
    assert np.array_equal(gpu["snaps"][1], gpu["fields"]["vuz"][:, y, :])
    ref = cases.run_oracle("f32", grid, nt, recs, base=base, snaps=snaps)
    for k in range(3):
        assert rel_l2(gpu["snaps"][k], ref["snaps"][k]) <= 1e-5
