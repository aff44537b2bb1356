"""Pins the oracle's compensated-FP16 codec (Half/main.cu:49-100).

binary16 conversions are checked exhaustively against numpy's IEEE float16
(round to nearest, ties to even), and the residual codec against its defining
error bound.
"""
import numpy as np

from oracle import pw_oracle as O


def f2h(x):
    return O.lib().or_f32_to_f16(float(x))


def test_f16_decode_all_codes():
    L = O.lib()
    codes = np.arange(65536, dtype=np.uint16)
    ref = codes.view(np.float16).astype(np.float32)
    got = np.array([L.or_f16_to_f32(int(c)) for c in codes], np.float32)
    finite = np.isfinite(ref)
    assert np.array_equal(got[finite], ref[finite])
    assert np.array_equal(np.isnan(got), np.isnan(ref))


def test_f16_encode_matches_numpy_rne():
    rng = np.random.default_rng(0)
    # every float16 value, every midpoint between neighbours (ties), and random floats
    h = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float64)
    mids = (h[:-1] + h[1:]) / 2
    vals = np.concatenate([h, mids, np.nextafter(mids, 0), np.nextafter(mids, 1e9),
                           rng.uniform(-70000, 70000, 20000), rng.standard_normal(20000) * 1e-6,
                           np.array([65504, 65519.99, 65520, 1e9, 2 ** -25, 2 ** -24, 3 * 2 ** -26])])
    vals = vals.astype(np.float32)
    vals = np.concatenate([vals, -vals])
    ref = vals.astype(np.float16).view(np.uint16)
    got = np.array([f2h(v) for v in vals], np.uint16)
    assert np.array_equal(got, ref)


def test_residual_unit_is_half_ulp_over_256():
    L = O.lib()
    # |hi| in [2^k, 2^(k+1)) has FP16 ULP 2^(k-10); the residual unit is ULP/256
    for k in range(-14, 16):
        for m in (1.0, 1.5, 1.999):
            assert L.or_residual_unit(m * 2.0 ** k) == 2.0 ** (k - 10) / 256
    # below the FP16 normal range the unit is 2^-32 (subnormal ULP 2^-24 / 256)
    assert L.or_residual_unit(2.0 ** -15) == 2.0 ** -32
    assert L.or_residual_unit(0.0) == 2.0 ** -32


def test_compensated_roundtrip_bound():
    L = O.lib()
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-8, 4, 20000),
                           np.array([0.0, 1.0, -1.0, 6.1e-5, 6.2e-5, 65000.0])]).astype(np.float32)
    hi = (O.C.c_ushort * 1)()
    lo = (O.C.c_byte * 1)()
    for v in vals:
        L.or_store_ch(float(v), hi, lo)
        dec = L.or_load_ch(hi[0], lo[0])
        hf = L.or_f16_to_f32(hi[0])
        unit = L.or_residual_unit(hf)
        assert hi[0] == f2h(v)
        assert -127 <= lo[0] <= 127
        if abs(lo[0]) < 127:
            assert abs(dec - float(v)) <= 0.5 * unit * (1 + 1e-6)
        # the pair is ~8 bits finer than FP16 alone
        assert abs(dec - float(v)) <= abs(hf - float(v)) + 1e-30


def test_array_forms_match_the_scalar_codec():
    """O.load_ch / O.store_ch (the array forms the GPU codec test compares
    against) give the scalar functions' results element by element."""
    L = O.lib()
    rng = np.random.default_rng(5)
    hi = rng.integers(0, 0x7c00, 3000).astype(np.uint16)
    hi[::2] |= 0x8000
    lo = rng.integers(-128, 128, 3000).astype(np.int8)
    got = O.load_ch(hi, lo)
    ref = np.array([L.or_load_ch(int(h), int(l)) for h, l in zip(hi, lo)], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    v = (rng.standard_normal(3000) * 10.0 ** rng.uniform(-9, 4, 3000)).astype(np.float32)
    h2, l2 = O.store_ch(v)
    hb = (O.C.c_ushort * 1)()
    lb = (O.C.c_byte * 1)()
    for k in range(v.size):
        L.or_store_ch(float(v[k]), hb, lb)
        assert (h2[k], l2[k]) == (hb[0], lb[0])
