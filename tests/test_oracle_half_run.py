"""The oracle's Half run (or_run_half, Half/main.cu), pinned end to end without
the reference's Half outputs (it committed none).  Its parts are pinned
elsewhere: the codec exhaustively (test_oracle_codec.py) and the exponents and
scaled coefficients against method.md (test_oracle_closed_forms.py).  Here the
whole run is tied to the FP64 oracle, which the reference's Float snapshot
files pin, through properties the mathematics fixes:

* convergence: the Half scheme is the same discretisation with its state and
  coefficients rounded to the compensated format, so it must stay within that
  format's error of the FP64 run (a dropped term or a wrong exponent is O(1));
* invariance: stored values are power-of-two scalings of physical ones, and
  the outputs are unscaled, so the internal scaling must not show in them
  beyond rounding: the 1 -> 16 dynamic rescale switched off, the rescale
  applied every step instead of every 100, and a wavefield exponent es = 4
  all give the same seismograms to the format's error (forgetting to rescale
  the memory variables, the source or an output breaks this);
* linearity: a source twice as strong gives outputs twice as large, exactly
  wherever the values stay in the normal FP16 range; below it the residual unit
  is fixed (2^-32) and rounding no longer scales, which leaves ~5e-6 here (the
  same bound, 1e-5, as the GPU comp16 run at full size).
"""
import numpy as np
import pytest

from oracle import pw_oracle as O

SANDSTONE = (35e9, 7e9, 9e9, 2.25e9, 2650.0, 1000.0, 0.2, 1e-12, 1e-3, 2.0, 8.0)
NT = 160


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def case():
    nx, ny, nz, np_ = 26, 22, 30, 4
    grid = {"nx": nx, "ny": ny, "nz": nz, "np": np_, "h": 5.0, "dt": 0.5e-3, "rc": 1e-4, "v0": 4000.0,
            "f0": 30.0, "src": (nx // 2, ny // 2, nz // 2), "src_w": (1.0, 1.0, 1.0, -1.0)}
    rng = np.random.default_rng(7)
    b = np.array(O.biot_from_moduli(*SANDSTONE))
    shape = (nz - 2 * np_, ny - 2 * np_, nx - 2 * np_)
    base = b[:, None, None, None] * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, (9,) + shape))
    sx, sy, sz = grid["src"]
    recs = [(f, sx + sy * nx + iz * nx * ny) for f in ("txx", "tzz", "vux", "p") for iz in range(np_, nz - np_)]
    ref = O.run("f64", grid, base, NT, receivers=recs)
    return grid, base, recs, ref


def channels(out, recs):
    return {f: out["seis"][[k for k, (g, _) in enumerate(recs) if g == f]] for f in ("txx", "tzz", "vux", "p")}


def errs(out, ref, recs):
    a, b = channels(out, recs), channels(ref, recs)
    return {f: rel_l2(a[f], b[f]) for f in a}


def test_half_run_converges_to_the_fp64_oracle(case):
    grid, base, recs, ref = case
    half = O.run("half", grid, base, NT, receivers=recs)
    e = errs(half, ref, recs)
    print("Half vs FP64:", e)
    assert all(v <= 1e-3 for v in e.values()), e
    # and it is a reduced-precision run, not the FP32 one relabelled
    f32 = O.run("f32", grid, base, NT, receivers=recs)
    assert max(errs(f32, ref, recs).values()) < min(e.values())


@pytest.mark.parametrize("half", [{"dyn_end": 1.0}, {"dyn_interval": 1}, {"es": 4}],
                         ids=["no-dynamic-rescale", "rescale-every-step", "es=4"])
def test_half_outputs_do_not_depend_on_internal_scaling(case, half):
    grid, base, recs, ref = case
    default = O.run("half", grid, base, NT, receivers=recs)
    other = O.run("half", grid, base, NT, receivers=recs, half=half)
    d = errs(other, default, recs)
    e = errs(other, ref, recs)
    print(half, "vs default Half:", d, "vs FP64:", e)
    assert all(v <= 1e-4 for v in d.values()), d
    assert all(v <= 1e-3 for v in e.values()), e


def test_half_run_is_linear_in_the_source(case):
    grid, base, recs, _ = case
    one = O.run("half", grid, base, NT, receivers=recs)
    g2 = dict(grid, src_w=tuple(2.0 * w for w in grid["src_w"]))
    two = O.run("half", g2, base, NT, receivers=recs)
    a, b = np.asarray(one["seis"], np.float64), np.asarray(two["seis"], np.float64)
    assert np.abs(a).max() > 0
    print("Half linearity: rel-L2", rel_l2(b, 2.0 * a), "bit-exact fraction", float(np.mean(b == 2.0 * a)))
    assert rel_l2(b, 2.0 * a) <= 1e-5
