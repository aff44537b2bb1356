"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the
same seeded inputs.

Tolerances (relative L2 over each channel / field):
  FP32    vs oracle Float path (Float/main.cu)   <= 1e-5 on the reference's
          seismogram channels (txx, tzz, vux, vwx, p on the source line).  Other
          channels and full fields: <= max(1e-5, 2 * floor), floor = the oracle
          Float path's own distance to the FP64 oracle on that channel (the
          relative fluid velocity vw in stiff media is a difference of nearly
          equal terms, so FP32 itself carries ~1e-4 there).
  FP64    vs oracle FP64 run                      <= 1e-10 (outputs are FP32)
  comp16  vs oracle Half path (Half/main.cu)      bit-identical: every seismogram
          and every full field (check_exact).  The GPU path keeps the
          reference's (psi, previous derivative) FP16 pairs in the CPML, divides
          the differences by h and evaluates the Half update in the reference's
          order of float operations without contraction, as the oracle does; the
          Half path's own distance to FP64 (1e-6 to 1e-4 on these cases) is
          printed beside it, so that equality is a check the Half arithmetic
          must pass, not a tolerance it sits inside
  BF16    vs oracle FP64 run                      <= 5e-2 (8-bit mantissa storage)
  comp16 / BF16 accuracy against FP64 is reported (test_precision_report).
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


REFERENCE_CHANNELS = {"seis:txx", "seis:tzz", "seis:vux", "seis:vwx", "seis:p"}


def errors(gpu, ref, recs):
    errs = {}
    g, r = per_channel(gpu["seis"], recs), per_channel(ref["seis"], recs)
    for f in g:
        errs["seis:" + f] = rel_l2(g[f], r[f])
    if ref.get("final") is not None:
        for f in gpu["fields"]:
            errs["field:" + f] = rel_l2(gpu["fields"][f], ref["final"][cases.O.FIELD_ID[f]])
    return errs


def check(gpu, ref, recs, tol, label, floor=None):
    """tol: scalar, or per-key via floor (dict of the Float path's own error)."""
    errs = errors(gpu, ref, recs)
    print(f"{label}: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    bad = {}
    for k, v in errs.items():
        t = tol
        if floor is not None and k not in REFERENCE_CHANNELS:
            t = max(tol, 2.0 * floor.get(k, 0.0))
        if not v <= t:
            bad[k] = (v, t)
    assert not bad, f"{label}: rel-L2 above tolerance: {bad}"
    return max(errs.values())


def check_exact(gpu, ref, recs, label, ref64=None):
    """Seismograms and full fields bit-identical to the oracle run."""
    assert np.array_equal(gpu["seis"], ref["seis"]), f"{label}: seismograms differ, rel-L2 {rel_l2(gpu['seis'], ref['seis']):.2e}"
    for f in gpu["fields"]:
        b = ref["final"][cases.O.FIELD_ID[f]]
        assert np.array_equal(gpu["fields"][f], b), f"{label}: field {f} differs, rel-L2 {rel_l2(gpu['fields'][f], b):.2e}"
    if ref64 is not None:
        print(f"{label}: bit-identical; the Half path vs FP64 on the seismograms: {rel_l2(ref['seis'], ref64['seis']):.2e}")


def fp32_floor(grid, nt, recs, base, ref32, label="Float"):
    """The oracle path's own distance to the FP64 oracle, per channel."""
    ref64 = cases.run_oracle("f64", grid, nt, recs, base=base, final=True)
    fake = {"seis": ref32["seis"], "fields": {f: ref32["final"][cases.O.FIELD_ID[f]] for f in FIELDS}}
    fl = errors(fake, ref64, recs)
    print(f"oracle {label} vs FP64 floor: " + " ".join(f"{k}={v:.2e}" for k, v in fl.items()))
    return fl, ref64


@pytest.fixture(scope="module")
def borehole():
    grid, base, nt = cases.borehole_case()
    recs = cases.receivers(grid)
    ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    return grid, base, nt, recs, ref32


def test_fp32_borehole_matches_oracle(borehole):
    grid, base, nt, recs, ref32 = borehole
    gpu, _ = cases.run_gpu(grid, nt, "fp32", recs, base=base, fields=FIELDS)
    floor, _ = fp32_floor(grid, nt, recs, base, ref32)
    check(gpu, ref32, recs, 1e-5, "fp32 borehole", floor=floor)
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
    _, ref64 = fp32_floor(grid, nt, recs, base, refh, "Half")
    gpu, _ = cases.run_gpu(grid, nt, "comp16", recs, base=base, fields=FIELDS)
    check_exact(gpu, refh, recs, "comp16 borehole vs Half", ref64)


@pytest.fixture(scope="module")
def layered():
    grid, spec, nt = cases.layered_case()
    recs = cases.receivers(grid)
    gpu32, base = cases.run_gpu(grid, nt, "fp32", recs, spec=spec, fields=FIELDS)
    return grid, spec, nt, recs, gpu32, base


def test_fp32_layered_matches_oracle(layered):
    grid, spec, nt, recs, gpu32, base = layered
    ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    floor, ref64 = fp32_floor(grid, nt, recs, base, ref32)
    check(gpu32, ref32, recs, 1e-5, "fp32 layered", floor=floor)
    # and the GPU FP32 run is itself within FP32's floor of the FP64 oracle
    check(gpu32, ref64, recs, 1e-5, "fp32 layered vs fp64", floor=floor)


def test_fp64_layered_matches_oracle(layered):
    grid, spec, nt, recs, gpu32, base = layered
    ref64 = cases.run_oracle("f64", grid, nt, recs, base=base, final=True)
    g64, _ = cases.run_gpu(grid, nt, "fp64", recs, spec=spec, fields=FIELDS)
    check(g64, ref64, recs, 1e-10, "fp64 layered")


def test_comp16_layered_matches_half_oracle(layered):
    grid, spec, nt, recs, gpu32, base = layered
    refh = cases.run_oracle("half", grid, nt, recs, base=base, final=True)
    _, ref64 = fp32_floor(grid, nt, recs, base, refh, "Half")
    g16, _ = cases.run_gpu(grid, nt, "comp16", recs, spec=spec, fields=FIELDS)
    check_exact(g16, refh, recs, "comp16 layered vs Half", ref64)


def test_precision_report(layered):
    """Accuracy of the reduced-precision storages against the FP64 oracle
    (reported; BF16 bounded)."""
    grid, spec, nt, recs, gpu32, base = layered
    ref64 = cases.run_oracle("f64", grid, nt, recs, base=base, final=True)
    refh = cases.run_oracle("half", grid, nt, recs, base=base, final=True)
    fake = {"seis": refh["seis"], "fields": {f: refh["final"][cases.O.FIELD_ID[f]] for f in FIELDS}}
    print("oracle Half vs FP64: " + " ".join(f"{k}={v:.2e}" for k, v in errors(fake, ref64, recs).items()))
    for prec in ("comp16", "comp16f", "bf16"):
        g, _ = cases.run_gpu(grid, nt, prec, recs, spec=spec, fields=FIELDS)
        e = errors(g, ref64, recs)
        print(f"{prec} vs FP64: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
        if prec == "bf16":
            assert all(v <= 5e-2 for k, v in e.items() if k in REFERENCE_CHANNELS), e


@pytest.mark.parametrize("case", ["borehole", "layered"])
def test_comp16f_matches_fp64_within_1e3(case, borehole, layered):
    """comp16f: the compensated state (FP16 + int8 residual, stresses stored
    times 2^6, velocities on the Half exponent, the relative fluid velocity on
    its own scale) under this build's FP32 form of the Float update (undivided
    differences, hardware reciprocals, the cell-centred coefficients held as
    compensated values, FP32 memory variables).  It has
    no oracle of its own, so it is held to the FP64 oracle: every seismogram
    channel <= 1e-3 (or 2x the Float path's own distance to FP64 where larger).
    Full fields: <= max(1e-3, 64 x that distance): the relative fluid velocity vw
    is a near-cancellation of pressure gradient and inertia, so it carries the
    stored stresses' relative precision amplified, and the 19-bit compensated
    state holds 2^5 less of it than FP32 (FP32 itself is at ~1e-4 on vw here).
    A wrong coefficient, sign or neighbour in the kernels gives O(1) errors."""
    if case == "borehole":
        grid, base, nt, recs, ref32 = borehole
        spec = None
    else:
        grid, spec, nt, recs, gpu32, base = layered
        ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    floor, ref64 = fp32_floor(grid, nt, recs, base, ref32)
    g, _ = cases.run_gpu(grid, nt, "comp16f", recs, base=base if spec is None else None, spec=spec, fields=FIELDS)
    e = errors(g, ref64, recs)
    print(f"comp16f {case} vs fp64: " + " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    bad = {}
    for k, v in e.items():
        t = max(1e-3, (2.0 if k.startswith("seis:") else 64.0) * floor.get(k, 0.0))
        if not v <= t:
            bad[k] = (v, t)
    assert not bad, bad


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


def test_host_medium_equals_generated_medium(layered):
    """pw_set_medium with the base fields in host memory (what a caller holding
    its own model does) gives the same run, bit for bit, as generating the same
    medium on the device with pw_set_medium_layered."""
    grid, spec, nt, recs, gpu32, base = layered
    for prec in ("fp32", "comp16"):
        g_dev = gpu32 if prec == "fp32" else cases.run_gpu(grid, nt, prec, recs, spec=spec, fields=FIELDS)[0]
        g_host, _ = cases.run_gpu(grid, nt, prec, recs, base=base, fields=FIELDS)
        assert np.array_equal(g_dev["seis"], g_host["seis"]), prec
        for f in FIELDS:
            assert np.array_equal(g_dev["fields"][f], g_host["fields"][f]), (prec, f)


def test_two_handles_run_concurrently(layered):
    """Two handles driven at once from two host threads, each on the stream
    the library created for it, give the same bits as each run alone (one
    handle is not thread-safe; distinct handles may run concurrently)."""
    import threading
    grid, spec, nt, recs, gpu32, base = layered
    solo16, _ = cases.run_gpu(grid, nt, "comp16", recs, spec=spec, fields=FIELDS)
    handles = {}
    for prec in ("fp32", "comp16"):
        h = cases.make_solver(grid, nt, prec, spec=spec)
        h.set_receivers([f for f, _ in recs], [i for _, i in recs])
        handles[prec] = h
    errs = []

    def drive(h):
        try:
            done = 0
            while done < nt:
                k = min(7, nt - done)
                h.run(k)
                done += k
            h.synchronize()
        except Exception as e:  # reported below, on the test's thread
            errs.append(e)

    threads = [threading.Thread(target=drive, args=(h,)) for h in handles.values()]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errs, errs
    for prec, solo in (("fp32", gpu32), ("comp16", solo16)):
        h = handles[prec]
        assert np.array_equal(h.seismogram(), solo["seis"]), prec
        for f in FIELDS:
            assert np.array_equal(h.field(f), solo["fields"][f]), (prec, f)
        h.close()


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


SCHEDULES = (
    {},                               # each precision's default tiles
    {"cells": 2},                     # two cells per lane (128-cell row segments)
    {"cells": 4, "rows": 8},          # four cells per lane, 8-row tiles
    {"waves_x": 2},                   # two waves side by side along x (128-cell tiles)
    {"waves_x": 1, "cells": 1},       # every kernel one wave per 64-cell row
    {"planes": 3, "rows": 2},         # 3-plane z chunks (ragged at the top), 64 x 2 tiles
    {"rows": 4},                      # 64 x 4 tiles (comp16f's default is 64 x 1)
    {"rows": 1, "planes": 16},        # one-wave blocks, 16-plane chunks
    {"rows": 2, "waves_x": 2},        # 128 x 1 tiles
)


def test_step_schedules_are_bit_identical(layered):
    """Every tiling of the two step kernels applies the same per-cell
    arithmetic: results are bit-identical to the default, in every precision."""
    grid, spec, nt, recs, gpu32, base = layered
    bad = []
    for prec in ("fp32", "comp16", "comp16f", "bf16", "fp64"):
        outs = []
        for tiling in SCHEDULES:
            t = dict(tiling)
            if prec == "fp64" and t.get("cells") == 4:
                t["cells"] = 2
            g, _ = cases.run_gpu(dict(grid, tiling=t), nt, prec, recs, spec=spec, fields=("vux", "vwz", "txz", "p"))
            outs.append(g)
        for env, g in zip(SCHEDULES[1:], outs[1:]):
            same = np.array_equal(g["seis"], outs[0]["seis"]) and all(
                np.array_equal(g["fields"][f], outs[0]["fields"][f]) for f in ("vux", "vwz", "txz", "p"))
            if same:
                continue
            err = rel_l2(g["seis"], outs[0]["seis"])
            if (prec, tuple(sorted(env.items()))) in KNOWN_DIFFERENT and err <= 1e-5:
                print(f"known difference {prec} {env}: {err:.2e}")
                continue
            bad.append((prec, env, err))
    assert not bad, f"schedules differing from the default step: {bad}"


# (schedule, precision) pairs allowed to differ from the default in the last
# bits; empty: every schedule is bit-identical in every precision.
KNOWN_DIFFERENT = set()
