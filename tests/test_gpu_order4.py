"""The 4th-order staggered operator (space_order=4) on the GPU.

Parity: the GPU against the oracle's 4th-order update (the same differences,
9/8 and -1/24, whose stencil the oracle tests pin) on the two-layer case:
FP32 within max(1e-5, 2 x the oracle FP32 path's distance to its FP64 path),
FP64 within 1e-10.  The reduced-precision storages are checked against the GPU
FP64 run.  Physics: on a coarse grid the 4th-order result is closer than the
2nd-order one to a fine-grid solution of the same problem."""
import numpy as np
import pytest

import cases
import mixedprecisionfaracoustic_amd as pw
from cases import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vux", "vuz", "vwx", "vwz", "txx", "tzz", "txz", "p")


def run_gpu4(grid, nt, prec, recs, base, fields=FIELDS):
    s = pw.Solver(grid["nx"], grid["ny"], grid["nz"], grid["np"], grid["h"], grid["dt"], nt, precision=prec,
                  cpml_r=grid["rc"], cpml_v0=grid["v0"], space_order=4, tiling=grid.get("tiling"))
    s.set_medium(base)
    s.set_source(*grid["src"], grid["f0"], grid["src_w"])
    s.set_receivers([f for f, _ in recs], [i for _, i in recs])
    s.run(nt)
    out = {"seis": s.seismogram(), "fields": {f: s.field(f) for f in fields}}
    s.close()
    return out


def errs(g, ref, recs, final=True):
    out = {}
    for f in sorted(set(f for f, _ in recs)):
        k = [i for i, (q, _) in enumerate(recs) if q == f]
        out["seis:" + f] = rel_l2(g["seis"][k], ref["seis"][k])
    if final:
        for f in g["fields"]:
            out["field:" + f] = rel_l2(g["fields"][f], ref["final"][cases.O.FIELD_ID[f]])
    return out


@pytest.fixture(scope="module")
def case4():
    grid, spec, nt = cases.layered_case()
    grid = dict(grid, order=4)
    recs = cases.receivers(grid)
    s = pw.Solver(grid["nx"], grid["ny"], grid["nz"], grid["np"], grid["h"], grid["dt"], 2, cpml_r=grid["rc"])
    s.set_medium_layered(spec)
    base = s.base_fields()
    s.close()
    ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    ref64 = cases.run_oracle("f64", grid, nt, recs, base=base, final=True)
    return grid, base, nt, recs, ref32, ref64


def test_order4_fp32_matches_oracle(case4):
    grid, base, nt, recs, ref32, ref64 = case4
    g = run_gpu4(grid, nt, "fp32", recs, base)
    fake = {"seis": ref32["seis"], "fields": {f: ref32["final"][cases.O.FIELD_ID[f]] for f in FIELDS}}
    floor = errs(fake, ref64, recs)
    e = errs(g, ref32, recs)
    print("order4 fp32 vs oracle:", " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    print("order4 oracle FP32 vs FP64:", " ".join(f"{k}={v:.2e}" for k, v in floor.items()))
    bad = {k: v for k, v in e.items() if not v <= max(1e-5, 2 * floor[k])}
    assert not bad, bad
    # the 4th-order run is a different (not a re-labelled 2nd-order) solution
    g2, _ = cases.run_gpu(dict(grid, order=2), nt, "fp32", recs, base=base)
    assert rel_l2(g2["seis"], g["seis"]) > 1e-3


def test_order4_fp64_matches_oracle(case4):
    grid, base, nt, recs, ref32, ref64 = case4
    g = run_gpu4(grid, nt, "fp64", recs, base)
    e = errs(g, ref64, recs)
    print("order4 fp64 vs oracle:", " ".join(f"{k}={v:.2e}" for k, v in e.items()))
    assert all(v <= 1e-10 for v in e.values()), e


def test_order4_reduced_precisions(case4):
    grid, base, nt, recs, ref32, ref64 = case4
    g64 = run_gpu4(grid, nt, "fp64", recs, base)
    ref = {"seis": g64["seis"]}
    for prec, tol in (("comp16", 5e-3), ("comp16f", 1e-3), ("bf16", 5e-2)):
        g = run_gpu4(grid, nt, prec, recs, base)
        e = errs(g, ref, recs, final=False)
        print(f"order4 {prec} vs FP64:", " ".join(f"{k}={v:.2e}" for k, v in e.items()))
        for c in ("seis:txx", "seis:tzz", "seis:vux", "seis:p"):
            assert e[c] <= tol, (prec, c, e[c])


def test_order4_tilings_are_bit_identical(case4):
    """The 4th-order kernels tile like the 2nd-order ones (BF16's velocity kernel
    puts two waves side by side by default): one or two waves per tile row, and
    64 x 2 or 64 x 8 tiles, give the same bits."""
    grid, base, nt, recs, _, _ = case4
    for prec in ("bf16", "fp32", "comp16", "comp16f"):
        outs = []
        for t in ({}, {"waves_x": 1}, {"waves_x": 2}, {"rows": 8}, {"rows": 2, "waves_x": 2}):
            outs.append((t, run_gpu4(dict(grid, tiling=t), nt, prec, recs, base)))
        for env, g in outs[1:]:
            assert np.array_equal(g["seis"], outs[0][1]["seis"]), (prec, env)
            for f in FIELDS:
                assert np.array_equal(g["fields"][f], outs[0][1]["fields"][f]), (prec, env, f)


def test_order4_converges_faster():
    """Homogeneous sandstone, the same physical source (moment scaled by the
    cell volume) and receiver (120 m from the source along x) on grids of 10 m
    and 5 m, against a 2.5 m 4th-order solution: at 10 m the 4th-order trace is
    closer to it than the 2nd-order one, and so it is at 5 m."""
    b = pw.biot_from_rock(cases.sandstone())
    dt, nt, f0 = 2.5e-4, 700, 30.0

    def trace(h, order):
        ni = int(round(320.0 / h))
        npml = int(round(40.0 / h))
        n = ni + 2 * npml
        base = np.broadcast_to(b[:, None, None, None], (9, ni, ni, ni)).copy()
        s = pw.Solver(n, n, n, npml, h, dt, nt, precision="fp32", cpml_r=1e-4, cpml_v0=4000.0, space_order=order)
        s.set_medium(base)
        c = n // 2
        w = (10.0 / h) ** 3
        s.set_source(c, c, c, f0, (w, w, w, -w))
        rx = c + int(round(120.0 / h))
        s.set_receivers(["txx", "vux"], [rx + c * n + c * n * n] * 2)
        s.run(nt)
        out = s.seismogram()
        s.close()
        return out

    fine = trace(2.5, 4)
    e = {(h, o): rel_l2(trace(h, o)[0], fine[0]) for h in (10.0, 5.0) for o in (2, 4)}
    print("convergence (h, order) -> rel-L2 vs 2.5 m 4th order:", {k: f"{v:.2e}" for k, v in e.items()})
    assert e[(10.0, 4)] < e[(10.0, 2)]
    assert e[(5.0, 4)] < e[(5.0, 2)]
