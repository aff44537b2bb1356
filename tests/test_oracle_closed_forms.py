"""Pins the oracle's set-up functions against closed forms the reference documents.

Known answers: method.md:919-1022 (Biot moduli of the reference formation),
method.md:1073-1169 / 1187-1307 / 1482-1522 (Half exponent and FP16 coefficients),
the CPML d0 of Float/main.cu:600 and the Ricker wavelet of Float/main.cu:1059.
"""
import math

import numpy as np
import pytest

from oracle import pw_oracle as O


def rel(a, b):
    return abs(a - b) / abs(b)


def test_create_model_formation_matches_method_md():
    base = O.borehole_model(32, 32, 32)
    # a formation cell far from the borehole axis (ix, iy) = (0, 0)
    rho, rhof, C1, C2, HH, H2u, C, M, mu = (float(base[k, 5, 0, 0]) for k in range(9))
    assert rel(rho, 2485.0) < 1e-7
    assert rel(mu, 1.738665e10) < 1e-7
    assert rel(M, 2.25e10) < 1e-6          # M = Kf*Ks/(phi Ks + (alpha-phi) Kf) = 10 Kf
    assert rel(C, 2.25e9) < 1e-6           # C = alpha M
    assert rel(HH, 3.8385e10) < 1e-6       # H = alpha^2 M + Kb + 4/3 mu
    assert rel(H2u, 3.6117e9) < 1e-4       # H - 2 mu (5 significant digits in method.md)
    assert rel(C1, 5.0e8) < 1e-7           # eta / kappa
    assert rel(C2, 3.75e4) < 1e-7          # (1 + 2/m) tau rhof / phi
    assert rhof == 1000.0
    # Q = C2 - rhof^2 / rho (method.md:1011-1018)
    assert rel(C2 - rhof * rhof / rho, 3.7097585513e4) < 1e-8


def test_create_model_borehole_fluid_cells():
    nz = ny = nx = 32
    base = O.borehole_model(nz, ny, nx)
    for iy in range(ny):
        for ix in range(nx):
            inside = (ix - nx // 2) ** 2 + (iy - ny // 2) ** 2 <= 100
            c1 = base[2, 3, iy, ix]
            mu = base[8, 3, iy, ix]
            assert (c1 == 0.0) == inside and (mu == 0.0) == inside
    kf = np.float32(1000.0) * np.float32(1500.0) * np.float32(1500.0)
    cell = base[:, 0, ny // 2, nx // 2]
    assert list(cell) == [1000.0, 1000.0, 0.0, 1000.0, kf, kf, kf, kf, 0.0]


def test_biot_from_moduli_sandstone():
    # water-saturated sandstone: alpha = 1 - Kb/Ks = 0.8, M = 9.431e9, H = 2.504e10
    out = O.biot_from_moduli(35e9, 7e9, 9e9, 2.25e9, 2650.0, 1000.0, 0.2, 1e-12, 1e-3, 2.0, 8.0)
    rho, rhof, C1, C2, HH, H2u, C, M, mu = out
    alpha = 1 - 7e9 / 35e9
    assert rel(M, 2.25e9 * 35e9 / (0.2 * 35e9 + (alpha - 0.2) * 2.25e9)) < 1e-14
    assert rel(M, 9.4311377e9) < 1e-7
    assert rel(HH, alpha ** 2 * M + 7e9 + 4.0 / 3.0 * 9e9) < 1e-14
    assert rel(rho, 2320.0) < 1e-14 and rel(C1, 1e9) < 1e-14 and rel(C2, 12500.0) < 1e-14
    assert rel(C, alpha * M) < 1e-14 and rel(H2u, HH - 18e9) < 1e-14


def test_pml_profile_closed_form():
    n, np_, h, dt, rc, v0 = 192, 32, 0.01, 0.9e-6, 1e-6, 4000.0
    d, d2, e, e2 = O.pml_profile_f32(n, np_, h, dt, rc, v0)
    d0 = 3.0 * v0 * math.log(1.0 / rc) / (2.0 * np_ * h)
    assert rel(d0, 259040.82) < 1e-7
    # quadratic profile: d(0) = d0, d(i) = d0 ((NP - i)/NP)^2 on the left
    assert rel(float(d[0]), d0) < 1e-6
    for i in (1, 7, 20, 31):
        # positions are float: the difference DP - i*H carries the float cancellation
        assert rel(float(d[i]), d0 * ((np_ - i) / np_) ** 2) < 2e-5
        assert rel(float(d2[i]), d0 * ((np_ - i - 0.5) / np_) ** 2) < 2e-5
    # interior undamped, e = 1 exactly
    assert np.all(d[np_:n - 1 - np_] == 0) and np.all(e[np_:n - 1 - np_] == 1.0)
    # right side covers NP+1 cells (index n-1-NP has d = 0 but d2 > 0)
    assert d[n - 1 - np_] == 0 and d2[n - 1 - np_] > 0 and d[n - 1] > 0
    assert rel(float(e[5]), math.exp(-float(d[5]) * dt)) < 1e-6
    # the FP64 profile agrees with the float one
    D, D2, E, E2 = O.pml_profile_f64(n, np_, h, dt, rc, v0)
    assert np.allclose(D, d, rtol=1e-4, atol=0) and np.allclose(E2, e2, rtol=1e-6)


def test_ricker_closed_form():
    # I(0) = -(1 - 2 pi^2 f0^2 t0^2) exp(-pi^2 f0^2 t0^2) with f0 t0 = 1.2
    a = math.pi ** 2 * 1.2 ** 2
    assert rel(O.lib().or_ricker_f64(0, 0.9e-6, 3000.0), -(1 - 2 * a) * math.exp(-a)) < 1e-9
    assert rel(O.lib().or_ricker_f64(0, 0.9e-6, 3000.0), 1.84436e-5) < 1e-4
    vals = np.array([O.lib().or_ricker_f32(i, 0.9e-6, 3000.0) for i in range(1000)])
    assert int(np.argmin(vals)) == 444 and abs(vals.min() + 0.99996) < 1e-4
    # at t = t0 the wavelet is exactly -1 (f0 = 1 Hz, dt = 0.1 s, it = 12)
    assert abs(O.lib().or_ricker_f64(12, 0.1, 1.0) + 1.0) < 1e-12


def test_half_exponents_and_coefficients_match_method_md():
    g = O.reference_grid()
    # a uniform formation block (no borehole) gives method.md's per-cell numbers
    base = O.borehole_model(128, 128, 128)
    em, epml, coef = O.half_precompute(g, base, want_coef=True)
    assert em == -23            # method.md:1109-1124
    assert epml == 13           # floor(log2(2048 / (dt d0 (32.5/32)^2)))
    c = coef[:, 40, 2 + 32, 2 + 32].astype(np.uint16).view(np.float16).astype(np.float64)
    # FP16 rounding bounds each value to 2^-11 relative
    tol = 2.0 ** -11
    expect = {0: 3.0381276459e-3, 3: 0.4024144869, 6: 0.9879429575, 9: 2.0228360646e-4,
              12: 8.1401853706e-5, 15: 4.1182637215e-3, 16: 3.8749337196e-4, 17: 2.4139881134e-4,
              18: 2.4139881134e-3, 19: 1.8653851748e-3}
    for k, v in expect.items():
        assert rel(c[k], v) <= tol, (k, c[k], v)


def test_precompute_staggered_means():
    g = {"nx": 20, "ny": 20, "nz": 20, "np": 4, "h": 0.01, "dt": 0.9e-6}
    rng = np.random.default_rng(1)
    base = np.zeros((9, 12, 12, 12), np.float32)
    for k in range(9):
        base[k] = rng.uniform(1.0, 2.0, (12, 12, 12)).astype(np.float32)
    base[8, 3, 3, 3] = 0.0  # a zero shear modulus zeroes the neighbouring harmonic means
    out = O.precompute_f32(g, base)
    pad = np.pad(base, ((0, 0), (4, 4), (4, 4), (4, 4)), mode="edge")
    iz, iy, ix = 9, 10, 6
    rho, mu = pad[0], pad[8]
    assert rel(float(out[0, iz, iy, ix]), 2.0 / (rho[iz, iy, ix] + rho[iz, iy, ix + 1])) < 1e-6
    assert rel(float(out[2, iz, iy, ix]), 2.0 / (rho[iz, iy, ix] + rho[iz + 1, iy, ix])) < 1e-6
    h = 1.0 / (0.25 * (1 / mu[iz, iy, ix] + 1 / mu[iz, iy, ix + 1] + 1 / mu[iz, iy + 1, ix] + 1 / mu[iz, iy + 1, ix + 1]))
    assert rel(float(out[14, iz, iy, ix]), h) < 1e-6  # muxy
    assert out[14, 7, 7, 7] == 0.0 and out[15, 6, 7, 7] == 0.0  # cells touching the zero
    # outer layer of staggered arrays is never computed (stays zero)
    assert np.all(out[0, :, :, 0] == 0) and np.all(out[14, -1] == 0)


def test_staggered_differences_pin_the_fourth_order_stencil():
    """The oracle's staggered difference (the macro its update expands, for both
    orders): exact for polynomials up to degree 4 at order 4 (only up to degree 2
    at order 2), second/fourth-order convergence on a smooth function, and the
    forward/backward points o +- 1/2.  The call sites (which field, which stride,
    which direction per term) are pinned at order 2 by the reference snapshots."""
    import numpy as np
    x = np.arange(40, dtype=np.float64)
    for deg in (1, 2, 3, 4):
        f = (x - 17.3) ** deg
        for o in (10, 20):
            for fwd, xm in ((True, o + 0.5), (False, o - 0.5)):
                exact = deg * (xm - 17.3) ** (deg - 1)
                d4 = O.stagger_diff(4, fwd, f, o)
                d2 = O.stagger_diff(2, fwd, f, o)
                assert abs(d4 - exact) <= 1e-9 * max(1.0, abs(exact)), (deg, fwd, d4, exact)
                if deg <= 2:
                    assert abs(d2 - exact) <= 1e-9 * max(1.0, abs(exact))
                else:
                    assert abs(d2 - exact) > 1e-3 * abs(exact)  # order 2 is not exact for cubics
    # convergence: error ratio on halving h is 2^order
    errs = {2: [], 4: []}
    for h in (0.2, 0.1, 0.05):
        xs = np.arange(0, 4, h)
        f = np.sin(1.3 * xs)
        o = len(xs) // 2
        exact = 1.3 * np.cos(1.3 * (xs[o] + 0.5 * h))
        for order in (2, 4):
            errs[order].append(abs(O.stagger_diff(order, True, f, o) / h - exact))
    for order in (2, 4):
        r = errs[order][0] / errs[order][1], errs[order][1] / errs[order][2]
        assert all(abs(v - 2 ** order) < 0.1 * 2 ** order for v in r), (order, r)
    # strides: the same stencil along a strided axis
    g = np.zeros(200)
    g[::10] = (np.arange(20) - 7.1) ** 3
    assert abs(O.stagger_diff(4, True, g, 80, 10) - 3 * (8.5 - 7.1) ** 2) < 1e-9
