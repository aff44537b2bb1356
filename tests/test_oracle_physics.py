"""The oracle against textbook physics: a medium made of the borehole's fluid
cells everywhere (create_model's fluid: rho = rhof = 1000, C1 = 0, mu = 0,
H = H - 2mu = C = M = Kf = 2.25 GPa) turns the Biot update into 3-D acoustics.
With C1 = 0 the relative fluid velocity stays exactly zero (Float/main.cu:160-185
sets it to 0), with mu = 0 the shear stresses stay exactly zero, and the
normal stresses obey sigma_t = Kf div(vu), rho vu_t = grad(sigma): waves at
c = sqrt(Kf / rho) = 1500 m/s.  For a monopole (the same sample added to
txx, tyy, tzz) the 3-D field has no near-field term, so along a line from the
source the trace at r2 is the trace at r1 delayed by (r2 - r1) / c and scaled
by r1 / r2.  A wrong wave speed, a dropped gradient term or a mis-staggered
difference fails one of these."""
import numpy as np

from oracle import pw_oracle as O


def test_fluid_medium_is_acoustic_at_sqrt_kf_over_rho():
    np_, ni = 10, 60
    n = ni + 2 * np_
    h, dt, nt, f0 = 1.0, 2.0e-4, 400, 40.0
    fluid = O.borehole_model(32, 32, 32)[:, 16, 16, 16].astype(np.float64)
    assert fluid[2] == 0.0 and fluid[8] == 0.0  # C1 = 0, mu = 0
    c = float(np.sqrt(fluid[4] / fluid[0]))
    assert abs(c - 1500.0) < 1e-3
    base = np.broadcast_to(fluid[:, None, None, None], (9, ni, ni, ni)).copy()
    sx = sy = sz = n // 2
    r1, r2 = 12, 24
    grid = {"nx": n, "ny": n, "nz": n, "np": np_, "h": h, "dt": dt, "rc": 1e-4, "v0": 1500.0, "f0": f0,
            "src": (sx, sy, sz), "src_w": (1.0, 1.0, 1.0, 0.0)}
    recs = [("txx", sx + r + sy * n + sz * n * n) for r in (r1, r2)]
    out = O.run("f64", grid, base, nt, receivers=recs, final=True)
    a, b = out["seis"][0].astype(np.float64), out["seis"][1].astype(np.float64)

    # sub-sample delay from the cross-correlation peak (parabolic fit)
    xc = np.correlate(b, a, mode="full")
    k = int(np.argmax(xc))
    y0, y1, y2 = xc[k - 1], xc[k], xc[k + 1]
    lag = (k - (len(a) - 1)) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    c_est = (r2 - r1) * h / (lag * dt)
    amp = np.abs(b).max() / np.abs(a).max()
    print(f"acoustic limit: c = {c_est:.2f} m/s (expected {c:.2f}), amplitude ratio {amp:.4f} (1/r: {r1 / r2})")
    assert abs(c_est / c - 1.0) < 0.01
    assert abs(amp / (r1 / r2) - 1.0) < 0.05

    fin = out["final"]
    for f in ("vwx", "vwy", "vwz", "txy", "txz", "tyz"):
        assert not fin[O.FIELD_ID[f]].any(), f
    # the three normal stresses stay equal (isotropic source, isotropic fluid, H = H - 2mu):
    # the same sum in a different order, so equal to FP64 rounding
    txx = fin[O.FIELD_ID["txx"]].astype(np.float64)
    for f in ("tyy", "tzz"):
        d = np.linalg.norm(fin[O.FIELD_ID[f]].astype(np.float64) - txx) / np.linalg.norm(txx)
        print(f"txx vs {f}: rel-L2 {d:.2e}")
        assert d < 1e-7, (f, d)


def test_fp64_and_fp32_runs_are_linear_in_the_source():
    """The update is linear in the state, so a source twice as strong doubles
    every output; scaling by 2 is exact in binary floating point except where
    a result is subnormal, so nearly every sample doubles bit for bit."""
    np_, n, nt = 4, 24, 80
    rng = np.random.default_rng(3)
    b = np.array(O.biot_from_moduli(35e9, 7e9, 9e9, 2.25e9, 2650.0, 1000.0, 0.2, 1e-12, 1e-3, 2.0, 8.0))
    ni = n - 2 * np_
    base = b[:, None, None, None] * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, (9, ni, ni, ni)))
    grid = {"nx": n, "ny": n, "nz": n, "np": np_, "h": 5.0, "dt": 0.5e-3, "rc": 1e-4, "v0": 4000.0, "f0": 30.0,
            "src": (n // 2, n // 2, n // 2), "src_w": (1.0, 1.0, 1.0, -1.0)}
    recs = [(f, x + (n // 2) * n + (n // 2) * n * n) for f in ("txx", "vux", "vwx", "p") for x in range(n)]
    for kind in ("f64", "f32"):
        one = O.run(kind, grid, base, nt, receivers=recs)["seis"].astype(np.float64)
        two = O.run(kind, dict(grid, src_w=(2.0, 2.0, 2.0, -2.0)), base, nt, receivers=recs)["seis"].astype(np.float64)
        exact = float(np.mean(two == 2.0 * one))
        e = np.linalg.norm(two - 2.0 * one) / np.linalg.norm(2.0 * one)
        print(f"{kind} linearity: rel-L2 {e:.2e}, bit-exact {exact:.4f}")
        assert e < 1e-12 and exact > 0.99


def test_oracle_is_mirror_symmetric():
    """Homogeneous sandstone on a 99 x 67 x 67 grid, explosive source in the
    middle cell (49, 33, 33): every field is even or odd about the middle plane
    of each axis (cases.mirror_deviation), to 1e-6 of its peak.  In 60 steps
    the P wave moves about 20 cells, short of the CPML 29 or more cells out,
    whose profiles (the reference's) are not mirror-exact; values far below
    the peak do reach it.  Float and FP64 runs at 2nd and 4th order, Half at 2nd.  A
    difference taken on the wrong side, or a sign error in one term, is an O(1)
    asymmetry along its axis."""
    import cases
    sand = O.biot_from_moduli(35e9, 7e9, 9e9, 2.25e9, 2650.0, 1000.0, 0.2, 1e-12, 1e-3, 2.0, 8.0)
    nx, ny, nz, np_ = 99, 67, 67, 4
    ni = (nz - 2 * np_, ny - 2 * np_, nx - 2 * np_)
    base = np.broadcast_to(np.asarray(sand, np.float64)[:, None, None, None], (9,) + ni).copy()
    for order in (2, 4):
        grid = {"nx": nx, "ny": ny, "nz": nz, "np": np_, "h": 5.0, "dt": 0.5e-3, "rc": 1e-4, "v0": 4000.0,
                "f0": 30.0, "src": (nx // 2, ny // 2, nz // 2), "src_w": (1.0, 1.0, 1.0, -1.0), "order": order}
        for kind in ("f32", "f64", "half") if order == 2 else ("f32", "f64"):  # the Half program is 2nd order
            b = base.astype(np.float64 if kind == "f64" else np.float32)
            fin = O.run(kind, grid, b, 60, final=True)["final"]
            for f in cases.STAGGER:
                dev, _, peak = cases.mirror_deviation(fin[O.FIELD_ID[f]], f)
                assert peak > 0, (kind, order, f)
                assert dev <= 1e-6, (kind, order, f, dev)
