"""GPU parity with open faces (pw_config.cpml_open): faces without an absorbing
layer, as on one rank's z-slab share of a decomposed grid.  The GPU against
the oracle with the same faces open (the oracle's open faces are pinned in
tests/test_oracle_open_faces.py), same tolerances as the damped grid:
FP32 <= 1e-5 (floor rule off the reference channels), FP64 <= 1e-10, comp16
against the Half path <= 1e-3 (floor rule)."""
import numpy as np
import pytest

import cases
from cases import rel_l2
from test_gpu_parity import FIELDS, check, check_exact, fp32_floor

pytestmark = pytest.mark.gpu

MASKS = {"z slab (interior rank)": 16 | 32, "low z face only": 16, "every high face": 2 | 8 | 32}


@pytest.fixture(scope="module")
def layered_open():
    grid, spec, nt = cases.layered_case()
    recs = cases.receivers(grid)
    out = {}
    for label, m in MASKS.items():
        g = dict(grid, open=m)
        gpu32, base = cases.run_gpu(g, nt, "fp32", recs, spec=spec, fields=FIELDS)
        out[label] = (g, gpu32, base)
    return nt, recs, spec, out


@pytest.mark.parametrize("label", list(MASKS))
def test_fp32_open_faces_match_oracle(layered_open, label):
    nt, recs, spec, out = layered_open
    g, gpu32, base = out[label]
    ref32 = cases.run_oracle("f32", g, nt, recs, base=base, final=True)
    floor, _ = fp32_floor(g, nt, recs, base, ref32)
    check(gpu32, ref32, recs, 1e-5, f"fp32 open {label}", floor=floor)


def test_open_faces_change_the_solution(layered_open):
    """The z-slab run is not the damped run relabelled: the waves the z layers
    would absorb come back (by 150 steps the first reflections have just
    reached the receivers: about 1e-3 of the seismogram, against ~1e-6 of FP32
    rounding between schedules)."""
    nt, recs, spec, out = layered_open
    g, gpu32, base = out["z slab (interior rank)"]
    damped, _ = cases.run_gpu(dict(g, open=0), nt, "fp32", recs, spec=spec)
    assert rel_l2(gpu32["seis"], damped["seis"]) > 1e-4
    info = gpu32["info"]
    assert info["slab_lo"][2] == 0 and info["slab_hi"][2] == 0
    assert info["slab_lo"][0] > 0 and info["slab_hi"][1] > 0


def test_fp64_z_slab_matches_oracle(layered_open):
    nt, recs, spec, out = layered_open
    g, _, base = out["z slab (interior rank)"]
    ref64 = cases.run_oracle("f64", g, nt, recs, base=base, final=True)
    g64, _ = cases.run_gpu(g, nt, "fp64", recs, spec=spec, fields=FIELDS)
    check(g64, ref64, recs, 1e-10, "fp64 open z slab")


def test_comp16_z_slab_matches_half_oracle(layered_open):
    nt, recs, spec, out = layered_open
    g, _, base = out["z slab (interior rank)"]
    refh = cases.run_oracle("half", g, nt, recs, base=base, final=True)
    _, ref64 = fp32_floor(g, nt, recs, base, refh, "Half")
    g16, _ = cases.run_gpu(g, nt, "comp16", recs, spec=spec, fields=FIELDS)
    check_exact(g16, refh, recs, "comp16 open z slab vs Half", ref64)
    assert np.isfinite(g16["seis"]).all()
