"""The GPU codec of the compensated formats alone (pw_codec_decode,
pw_codec_encode), value by value:

* decode: every one of the 65,536 x 256 (hi, lo) pairs, bit for bit against the
  oracle's load_cr_value (Half/main.cu:69-73);
* encode, comp16 rounding: a dense sweep (every finite FP16 value, offsets of
  half a residual unit from -130 to +130 units around each, so every tie and
  every clamp, both signs) plus random floats over 2^-40..65504, bit for bit
  against the oracle's store_ch (Half/main.cu:85-100);
* encode, comp16f rounding: the same sweep against the definition of its
  residual, clamp(rint((v - hi) / u(hi)), -127, 127) with rint rounding ties
  to even, evaluated in FP64 (exact here: v - hi and the power-of-two division
  are exact), and the same FP16 value as the oracle's conversion."""
import numpy as np
import pytest
import torch

import mixedprecisionfaracoustic_amd as pw
from oracle import pw_oracle as O

pytestmark = pytest.mark.gpu


def unit(hf):
    """The residual unit of the oracle's or_residual_unit over an array: 2^-32
    below FP16's normal range, else the FP16 ULP of |hf| over 256."""
    a = np.abs(np.asarray(hf, np.float64))
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    return np.where(a < 2.0 ** -14, 2.0 ** -32, 2.0 ** (e - 18))


def dense_sweep():
    h = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float64)
    u = unit(h)
    k = np.arange(-260, 261, dtype=np.float64) * 0.5
    v = (h[:, None] + k[None, :] * u[:, None]).ravel()
    rng = np.random.default_rng(11)
    extra = np.exp2(rng.uniform(-40, np.log2(65504.0), 1 << 20))
    v = np.concatenate([v, extra])
    v = v[np.abs(v) < 65504.0]
    v = np.concatenate([v, -v]).astype(np.float32)
    return v


def gpu_encode(v, precision):
    x = torch.from_numpy(v).cuda()
    hi = torch.empty(v.size, dtype=torch.int16, device="cuda")
    lo = torch.empty(v.size, dtype=torch.int8, device="cuda")
    pw.codec_encode(x.data_ptr(), hi.data_ptr(), lo.data_ptr(), v.size, precision)
    return hi.cpu().numpy().view(np.uint16), lo.cpu().numpy()


def test_decode_every_pair_matches_the_oracle():
    hi = np.repeat(np.arange(65536, dtype=np.uint32).astype(np.uint16), 256)
    lo = np.tile(np.arange(-128, 128, dtype=np.int16).astype(np.int8), 65536)
    print(f"decode: {hi.size} pairs", flush=True)
    th = torch.from_numpy(hi.view(np.int16)).cuda()
    tl = torch.from_numpy(lo).cuda()
    out = torch.empty(hi.size, dtype=torch.float32, device="cuda")
    pw.codec_decode(th.data_ptr(), tl.data_ptr(), out.data_ptr(), hi.size)
    got = out.cpu().numpy()
    ref = O.load_ch(hi, lo)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))
    # the residual moves the value by lo units of ULP/256 (a unit 2x off would fail here)
    fin = np.isfinite(ref) & (lo != 0)
    hf = hi.view(np.float16).astype(np.float64)
    step = (got[fin].astype(np.float64) - hf[fin]) / lo[fin]
    assert np.array_equal(step, unit(hf[fin]))


def test_encode_comp16_matches_the_oracle_store():
    v = dense_sweep()
    print(f"encode comp16: {v.size} values", flush=True)
    hi, lo = gpu_encode(v, "comp16")
    rh, rl = O.store_ch(v)
    assert np.array_equal(hi, rh)
    assert np.array_equal(lo, rl)
    # ties (an odd number of half units from an FP16 value) round away from zero
    hf = rh.view(np.float16).astype(np.float64)
    s = (v.astype(np.float64) - hf) / unit(hf)
    tie = (np.abs(s - np.trunc(s)) == 0.5) & (np.abs(s) < 127)
    assert tie.sum() > 1000
    assert np.array_equal(lo[tie].astype(np.float64), np.trunc(s[tie]) + np.sign(s[tie]))


def test_encode_comp16f_rounds_residuals_to_nearest_even():
    v = dense_sweep()
    print(f"encode comp16f: {v.size} values", flush=True)
    hi, lo = gpu_encode(v, "comp16f")
    rh, _ = O.store_ch(v)
    assert np.array_equal(hi, rh)
    hf = rh.view(np.float16).astype(np.float64)
    s = (v.astype(np.float64) - hf) / unit(hf)
    ref = np.clip(np.rint(s), -127, 127).astype(np.int8)
    assert np.array_equal(lo, ref)
    tie = (np.abs(s - np.trunc(s)) == 0.5) & (np.abs(s) < 127)
    assert tie.sum() > 1000 and np.all(lo[tie] % 2 == 0)
    # decoding gives back v within half a unit wherever the residual is not clamped
    dec = O.load_ch(hi, lo).astype(np.float64)
    ok = np.abs(s) <= 127
    assert np.all(np.abs(dec[ok] - v[ok]) <= 0.5 * unit(hf[ok]))


def test_empty_and_refused_calls():
    z = torch.empty(0, device="cuda")
    pw.codec_decode(0, 0, 0, 0)
    pw.codec_encode(z.data_ptr(), 0, 0, 0, "comp16f")
    with pytest.raises(pw.PorowaveError) as ei:
        pw.codec_encode(0, 0, 0, 4, "fp32")
    assert ei.value.status == 1
