"""Edge cases on the GPU, against the oracle where there is a result to compare.

* Ragged and minimal grids: the smallest grid the CPML allows (2 np + 3 per
  axis), and odd sizes just past a tile edge (67 or 130 cells in x: 3 or 2 live
  lanes in the last 64-lane tile; 9 or 66 rows in y against 4-row tiles; z
  chunks that do not divide the updated planes), on a heterogeneous medium.
  FP32 against the Float oracle (same rule as the other parity tests), FP64
  against the FP64 oracle (1e-10), comp16 against the Half path (1e-3, floor
  rule).
* Every tiling (1, 2 or 4 cells per lane, 1 or 2 waves per row, 2- to 8-row
  tiles) bit-identical to the default on grids whose x extent ends inside the
  last 64-, 128- and 256-cell tile, with the wave crossing those tile edges.
* Mirror symmetry about the middle plane of every axis: homogeneous medium,
  odd extents, explosive source in the middle cell; the x tile edges, y block
  edges and z chunk edges are not symmetric about it, so a wrong neighbour or
  carry at any of them breaks the symmetry.
* A source in the CPML corner cell (1, 1, 1).
* Receivers on the edge planes, which are never updated: their traces stay 0.
* No receivers; a zero-step run; running past nt is an error that leaves the
  handle's state unchanged.
"""
import numpy as np
import pytest

import cases
import mixedprecisionfaracoustic_amd as pw
from cases import rel_l2
from test_gpu_parity import FIELDS, check, check_exact, fp32_floor

pytestmark = pytest.mark.gpu

SHAPES = [
    ((7, 7, 7), 2),      # minimum: 2 np + 3 on every axis, a 3^3 interior
    ((67, 9, 8), 2),     # 3 live lanes in the second x tile, ragged y rows
    ((130, 7, 23), 2),   # 2 live lanes in the third x tile, ragged z chunks
    ((9, 66, 11), 3),    # one partial x tile, 17 y tiles with the last half full
]


HH, H2U, MU = (pw.BASE.index(k) for k in ("HH", "H2u", "mu"))


def hetero_case(shape, np_, nt=60, src=None, seed=0, consistent_h=False):
    """Sandstone with every base field perturbed by up to 5 % on its own, so a
    coefficient array read in place of another gives an O(1) error.  H is then
    not (H - 2 mu) + 2 mu, and the FP32 stress kernel streams it; with
    consistent_h H is set to that sum (as the reference's model builder makes
    it), and FP32 and comp16f form H from the other two."""
    nx, ny, nz = shape
    grid = {"nx": nx, "ny": ny, "nz": nz, "np": np_, "h": 5.0, "dt": 0.5e-3, "rc": 1e-4, "v0": 4000.0,
            "f0": 30.0, "src": src or (nx // 2, ny // 2, nz // 2), "src_w": (1.0, 1.0, 1.0, -1.0)}
    b = pw.biot_from_rock(cases.sandstone())
    rng = np.random.default_rng(seed)
    shape_i = (nz - 2 * np_, ny - 2 * np_, nx - 2 * np_)
    base = b[:, None, None, None] * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, (9,) + shape_i))
    if consistent_h:
        base[HH] = base[H2U] + 2.0 * base[MU]
    return grid, base, nt


def line_receivers(grid):
    """Every cell of the x line through the source, edge cells included, for
    the five reference channels."""
    nx, ny = grid["nx"], grid["ny"]
    sx, sy, sz = grid["src"]
    return [(f, x + sy * nx + sz * nx * ny) for f in ("txx", "tzz", "vux", "vwx", "p") for x in range(nx)]


@pytest.mark.parametrize("consistent_h", [False, True], ids=["H-streamed", "H-formed"])
@pytest.mark.parametrize("shape,np_", SHAPES)
def test_ragged_grids_match_oracle(shape, np_, consistent_h):
    grid, base, nt = hetero_case(shape, np_, consistent_h=consistent_h)
    recs = line_receivers(grid)
    g32, _ = cases.run_gpu(grid, nt, "fp32", recs, base=base, fields=FIELDS)
    ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    floor, ref64 = fp32_floor(grid, nt, recs, base, ref32)
    check(g32, ref32, recs, 1e-5, f"fp32 {shape}", floor=floor)
    g64, _ = cases.run_gpu(grid, nt, "fp64", recs, base=base, fields=FIELDS)
    check(g64, ref64, recs, 1e-10, f"fp64 {shape}")
    refh = cases.run_oracle("half", grid, nt, recs, base=base, final=True)
    _, ref64h = fp32_floor(grid, nt, recs, base, refh, "Half")
    g16, _ = cases.run_gpu(grid, nt, "comp16", recs, base=base, fields=FIELDS)
    check_exact(g16, refh, recs, f"comp16 {shape} vs Half", ref64h)
    # edge cells are never updated: their traces are exactly zero
    nx = grid["nx"]
    for k, (f, i) in enumerate(recs):
        if i % nx in (0, nx - 1):
            assert not g32["seis"][k].any(), (f, i)


TILINGS = (
    {"waves_x": 1, "cells": 1},  # one wave per 64-cell row in every kernel
    {"cells": 2},                # 128-cell row segments, 2 cells per lane
    {"cells": 4, "rows": 8},     # 256-cell row segments, 8-row tiles
    {"waves_x": 2},              # two waves per 128-cell row
    {"waves_x": 2, "rows": 8},
    {"rows": 2},
)


@pytest.mark.parametrize("shape,np_,src", [((300, 11, 19), 3, (270, 5, 9)), ((130, 9, 13), 2, (100, 4, 6))])
def test_ragged_tilings_are_bit_identical(shape, np_, src):
    """Row tails: 300 cells end 44 cells into the fifth 64-cell tile, the third
    128-cell one and the second 256-cell one; 130 cells leave 2 cells (both
    edge or CPML) in the last tile of every width. The source sits near the high
    x end so that the wave crosses those tile edges within the run."""
    grid, base, _ = hetero_case(shape, np_, src=src)
    _, base_c, _ = hetero_case(shape, np_, src=src, consistent_h=True)
    base_cu = base_c.copy()  # one fluid: FP32 forms H and takes rhof as a scalar
    base_cu[pw.BASE.index("rhof")] = 1000.0
    nt = 80
    recs = line_receivers(grid)
    bad = []
    for prec, b in (("fp32", base), ("fp32", base_cu), ("comp16", base), ("comp16f", base_c), ("bf16", base),
                    ("fp64", base)):
        outs = []
        for tiling in ({},) + TILINGS:
            t = dict(tiling)
            if prec == "fp64" and t.get("cells") == 4:
                t["cells"] = 2
            g, _ = cases.run_gpu(dict(grid, tiling=t), nt, prec, recs, base=b, fields=FIELDS)
            outs.append(g)
        ref = outs[0]
        assert np.abs(ref["seis"]).max() > 0, prec
        # the wave has reached the tile edges at 256 (or 128) cells
        edge = 256 if grid["nx"] == 300 else 128
        assert np.abs(ref["fields"]["vux"][:, :, edge - 2:edge + 2]).max() > 0, prec
        for env, g in zip(TILINGS, outs[1:]):
            same = np.array_equal(g["seis"], ref["seis"]) and all(
                np.array_equal(g["fields"][f], ref["fields"][f]) for f in FIELDS)
            if not same:
                bad.append((prec, env, rel_l2(g["seis"], ref["seis"])))
    assert not bad, f"tilings differing from the default: {bad}"


@pytest.mark.parametrize("prec,order", [("fp32", 2), ("fp32", 4), ("fp64", 4), ("comp16", 2), ("comp16", 4),
                                        ("comp16f", 2), ("comp16f", 4), ("bf16", 2)])
def test_mirror_symmetry(prec, order):
    """Homogeneous sandstone on a 131 x 99 x 99 grid, explosive source in the
    middle cell (65, 49, 49): every field is even or odd about the middle plane
    of each axis (cases.mirror_deviation).  Negation and round-to-nearest are
    symmetric, so the same update at mirrored cells gives mirrored bits.  The
    tile edges along x (64, 128: lanes 63 / 0 of a wave), the 4-row block edges
    along y and the 4-plane chunk edges along z sit at different places relative
    to the source than their mirror images, so a wrong neighbour or carry at any
    of them is an O(1) error.  In 120 steps the P wave moves about 40 cells, short
    of the CPML 45 or more cells out, whose profiles (the reference's) are not
    mirror-exact.  The bound is 1e-6 of the field's peak rather than 0: the
    update spreads values one to two cells per step, so values near 1e-25 of
    the peak do reach the CPML."""
    nx, ny, nz, np_ = 131, 99, 99, 4
    grid = {"nx": nx, "ny": ny, "nz": nz, "np": np_, "h": 5.0, "dt": 0.5e-3, "rc": 1e-4, "v0": 4000.0,
            "f0": 30.0, "src": (nx // 2, ny // 2, nz // 2), "src_w": (1.0, 1.0, 1.0, -1.0), "order": order}
    b = pw.biot_from_rock(cases.sandstone())
    ni = (nz - 2 * np_, ny - 2 * np_, nx - 2 * np_)
    base = np.broadcast_to(b[:, None, None, None], (9,) + ni).copy()
    names = tuple(cases.STAGGER)
    g, _ = cases.run_gpu(grid, 120, prec, [], base=base, fields=names)
    for f in names:
        dev, exact, peak = cases.mirror_deviation(g["fields"][f], f)
        print(f"{prec} order {order} {f}: max mirror deviation {dev:.1e} of the peak, {exact:.4f} of cells bit-exact")
        assert peak > 0, f
        assert dev <= 1e-6, (prec, order, f, dev)


def test_source_in_cpml_corner_matches_oracle():
    grid, base, nt = hetero_case((20, 18, 22), 4, nt=80, src=(1, 1, 1))
    recs = [(f, x + 1 * 20 + 1 * 20 * 18) for f in ("txx", "vux", "p") for x in range(20)]
    recs += [(f, 10 + 9 * 20 + 11 * 20 * 18) for f in ("txx", "vux", "p")]
    g32, _ = cases.run_gpu(grid, nt, "fp32", recs, base=base, fields=FIELDS)
    ref32 = cases.run_oracle("f32", grid, nt, recs, base=base, final=True)
    floor, _ = fp32_floor(grid, nt, recs, base, ref32)
    check(g32, ref32, recs, 1e-5, "fp32 corner source", floor=floor)


def test_api_misuse_is_refused_and_the_handle_stays_usable():
    """Calls out of order or with bad arguments fail with their status code and
    a message, before touching the device state; the handle then runs normally."""
    grid, base, nt = hetero_case((20, 18, 22), 3, nt=12)
    s = pw.Solver(20, 18, 22, 3, 5.0, 0.5e-3, nt, precision="fp32", cpml_r=1e-4, cpml_v0=4000.0)
    with pytest.raises(pw.PorowaveError) as ei:
        s.run(1)                                         # before any medium: PW_ERR_STATE
    assert ei.value.status == 4 and "no medium" in str(ei.value)
    s.set_medium(base)
    for call, status in [
        (lambda: s.base_fields(), 4),                    # only generated media are kept
        (lambda: s.set_source(0, 9, 11, 30.0), 1),       # x = 0 is never updated
        (lambda: s.set_source(10, 9, 11, -1.0), 1),      # f0 must be positive
        (lambda: s.set_receivers(["txx"], [20 * 18 * 22]), 1),  # one past the last cell
        (lambda: s.field(99), 1),                        # no such field
    ]:
        with pytest.raises(pw.PorowaveError) as ei:
            call()
        assert ei.value.status == status and str(ei.value), (status, str(ei.value))
    with pytest.raises(ValueError):
        s.set_medium(base[:, :-1])                       # wrong interior shape (checked in Python)
    assert s.step == 0
    s.set_source(10, 9, 11, 30.0, (1.0, 1.0, 1.0, -1.0))
    s.set_receivers(["txx", "p"], [10 + 9 * 20 + 11 * 20 * 18, 11 + 9 * 20 + 11 * 20 * 18])
    s.run(nt)
    seis = s.seismogram()
    assert seis.shape == (2, nt) and np.isfinite(seis).all() and np.abs(seis).max() > 0
    s.close()


def test_medium_properties_select_the_coefficient_streams():
    """pw_set_medium compares H with (H - 2 mu) + 2 mu: FP32 streams H only when
    they differ (4 B per cell more in the stress kernel), and comp16f, which
    always forms H, refuses such a medium and then takes a consistent one.  A
    medium with one rhof gives FP32's velocity kernel a scalar instead of the
    rhof array (4 B per cell less); a rhof that differs in one cell does not."""
    shape, np_ = (20, 18, 22), 3
    _, base_i, nt = hetero_case(shape, np_, nt=12)
    _, base_c, _ = hetero_case(shape, np_, nt=12, consistent_h=True)
    s = pw.Solver(*shape, np_, 5.0, 0.5e-3, nt, precision="fp32", cpml_r=1e-4, cpml_v0=4000.0)
    s.set_medium(base_i)
    streamed, vel_arr = s.info()["bytes_stress"], s.info()["bytes_velocity"]
    s.set_medium(base_c)
    formed = s.info()["bytes_stress"]
    assert s.info()["bytes_velocity"] == vel_arr
    # H rounded to float on its own stays within the tolerance (8 FP32 ulps)
    b32 = base_c.astype(np.float32).astype(np.float64)
    s.set_medium(b32)
    assert s.info()["bytes_stress"] == formed
    rhof = pw.BASE.index("rhof")
    b_u = base_c.copy()
    b_u[rhof] = 1000.0
    s.set_medium(b_u)
    assert vel_arr - s.info()["bytes_velocity"] == 4.0
    b_u[rhof][5, 6, 7] = 1000.5
    s.set_medium(b_u)
    assert s.info()["bytes_velocity"] == vel_arr
    s.close()
    assert streamed - formed == 4.0, (streamed, formed)
    # FP64 takes one rhof as a scalar too (8 B per cell less) and always reads H
    d = pw.Solver(*shape, np_, 5.0, 0.5e-3, nt, precision="fp64", cpml_r=1e-4, cpml_v0=4000.0)
    d.set_medium(base_c)
    v64, t64 = d.info()["bytes_velocity"], d.info()["bytes_stress"]
    b_u[rhof] = 1000.0
    d.set_medium(b_u)
    assert v64 - d.info()["bytes_velocity"] == 8.0 and d.info()["bytes_stress"] == t64
    d.close()
    c = pw.Solver(*shape, np_, 5.0, 0.5e-3, nt, precision="comp16f", cpml_r=1e-4, cpml_v0=4000.0)
    with pytest.raises(pw.PorowaveError) as ei:
        c.set_medium(base_i)
    assert ei.value.status == 6 and "H - 2 mu" in str(ei.value)
    c.set_medium(base_c)
    c.set_source(10, 9, 11, 30.0, (1.0, 1.0, 1.0, -1.0))
    c.set_receivers(["p"], [11 + 9 * 20 + 11 * 20 * 18])
    c.run(nt)
    assert np.isfinite(c.seismogram()).all() and np.abs(c.seismogram()).max() > 0
    c.close()


def test_no_receivers_zero_steps_and_overrun():
    grid, base, nt = hetero_case((40, 36, 44), 6, nt=50)
    s = cases.make_solver(grid, nt, "fp32", base=base)
    s.run(0)
    assert s.step == 0
    assert not any(s.field(f).any() for f in FIELDS)
    s.run(nt)
    ref = cases.run_oracle("f32", grid, nt, [], base=base, final=True)
    for f in ("txx", "vux", "p"):
        e = rel_l2(s.field(f), ref["final"][cases.O.FIELD_ID[f]])
        print(f"no receivers, field {f}: rel-L2 {e:.2e}")
        assert e <= 1e-5
    before = s.field("txx").copy()
    with pytest.raises(pw.PorowaveError):
        s.run(1)
    assert s.step == nt
    assert np.array_equal(s.field("txx"), before)
    assert s.seismogram().size == 0
    s.close()
