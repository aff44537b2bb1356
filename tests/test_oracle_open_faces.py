"""Faces without an absorbing layer (or_args.pml_open / pw_config.cpml_open),
the setting of one rank's z-slab share in a decomposed grid.

What pins it, independently of the open-face code itself:
* all six faces open is the reference's own profile formula at RC = 1, where
  d0 = 3 v0 ln(1/RC) / (2 NP H) = 0: every profile is d = 0, e = 1 through
  Float/main.cu:570-654's arithmetic, so the two runs must be bit-identical
  (FP32, FP64 and the Half path);
* opening a face changes nothing before the wave reaches that face's layer
  (causality), and everything after: with the z faces open the energy that
  the z layers would absorb stays in the grid;
* an open face keeps its outermost plane fixed like every edge plane.
"""
import numpy as np
import pytest

from oracle import pw_oracle as O

SANDSTONE = (35e9, 7e9, 9e9, 2.25e9, 2650.0, 1000.0, 0.2, 1e-12, 1e-3, 2.0, 8.0)


def case(open_mask=0, rc=1e-4, n=(24, 20, 30), np_=5, src=None):
    nx, ny, nz = n
    g = {"nx": nx, "ny": ny, "nz": nz, "np": np_, "h": 5.0, "dt": 0.5e-3, "rc": rc, "v0": 4000.0,
         "f0": 30.0, "src": src or (nx // 2, ny // 2, nz // 2), "src_w": (1.0, 1.0, 1.0, -1.0), "open": open_mask}
    b = np.array(O.biot_from_moduli(*SANDSTONE))
    base = np.broadcast_to(b[:, None, None, None], (9, nz - 2 * np_, ny - 2 * np_, nx - 2 * np_)).copy()
    return g, base


def probe_receivers(g):
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    sx, sy, _ = g["src"]
    return [(f, sx + sy * nx + iz * nx * ny) for f in ("txx", "vuz", "p") for iz in range(nz)]


@pytest.mark.parametrize("kind", ["f32", "f64", "half"])
def test_all_faces_open_is_the_reference_profile_at_rc_1(kind):
    g_open, base = case(open_mask=63)
    g_rc1, _ = case(open_mask=0, rc=1.0)
    nt = 60
    recs = probe_receivers(g_open)
    a = O.run(kind, g_open, base, nt, receivers=recs, final=True)
    b = O.run(kind, g_rc1, base, nt, receivers=recs, final=True)
    assert np.abs(a["final"]).max() > 0
    assert np.array_equal(a["seis"], b["seis"])
    assert np.array_equal(a["final"], b["final"])
    # and it differs from the damped grid once waves reach the layers
    c = O.run(kind, case()[0], base, nt, receivers=recs, final=True)
    assert not np.array_equal(a["final"], c["final"])


def test_open_z_faces_change_nothing_before_the_wave_arrives():
    """Source at the centre of a 30-plane grid with 5-plane z layers: the first
    steps are bit-identical with and without the z layers, later ones are not."""
    g0, base = case()
    gz, _ = case(open_mask=16 | 32)
    recs = probe_receivers(g0)
    early0 = O.run("f32", g0, base, 4, receivers=recs, final=True)
    earlyz = O.run("f32", gz, base, 4, receivers=recs, final=True)
    assert np.abs(early0["final"]).max() > 0
    assert np.array_equal(early0["final"], earlyz["final"])
    late0 = O.run("f32", g0, base, 120, receivers=recs, final=True)
    latez = O.run("f32", gz, base, 120, receivers=recs, final=True)
    assert not np.array_equal(late0["final"], latez["final"])


def test_open_z_faces_keep_the_energy_the_layers_absorb():
    """After the wave has crossed the z layers several times, the grid with
    open (reflecting) z faces holds more wavefield than the damped one, and the
    x/y layers still damp both alike: opening only the x faces leaves more than
    the all-damped grid as well."""
    nt = 300
    out = {}
    for name, m in (("damped", 0), ("z_open", 48), ("x_open", 3)):
        g, base = case(open_mask=m)
        out[name] = O.run("f64", g, base, nt, final=True)["final"]
    norm = {k: float(np.sqrt((v[6:9] ** 2).sum())) for k, v in out.items()}  # normal stresses
    assert norm["z_open"] > 1.5 * norm["damped"], norm
    assert norm["x_open"] > 1.5 * norm["damped"], norm


def test_open_face_planes_stay_fixed():
    """The outermost plane of an open face is never updated (zero from the
    start), as every edge plane of the reference grid."""
    g, base = case(open_mask=16 | 32)
    f = O.run("f32", g, base, 200, final=True)["final"]
    assert np.abs(f).max() > 0
    assert not f[:, 0].any() and not f[:, -1].any()


def test_pml_open_is_validated():
    g, base = case(open_mask=64)
    with pytest.raises(RuntimeError):
        O.run("f32", g, base, 2)
