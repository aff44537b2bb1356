"""Pins the oracle's FP32 path end to end against the reference's own outputs:
the five step-500 snapshot files Float/{txx,vux,vuz,vwx,vwz}500.dat (reference
configuration, 192^3, 501 updates; tests/golden/float_snapshots.npz).  Also the
symmetries the fixtures themselves show.  About 2-3 minutes on 8 cores."""
import os

import numpy as np
import pytest

from oracle import pw_oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "float_snapshots.npz"))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def step500():
    g = O.reference_grid()
    base = O.borehole_model(128, 128, 128)
    snaps = [(500, "txx"), (500, "vux"), (500, "vuz"), (500, "vwx"), (500, "vwz")]
    return O.run("f32", g, base, 501, snapshots=snaps)["snaps"]


@pytest.mark.parametrize("k,name", list(enumerate(["txx500", "vux500", "vuz500", "vwx500", "vwz500"])))
def test_oracle_reproduces_reference_snapshot(step500, k, name):
    e = rel_l2(step500[k], GOLD[name])
    print(f"{name}: oracle rel-L2 vs reference file {e:.3e}")
    assert e <= 1e-5


def test_fixture_structure(step500):
    # txx500 peaks next to the source cell (iz = 52, ix = 96) in both
    assert np.unravel_index(np.argmax(np.abs(GOLD["txx500"])), (192, 192)) == \
        np.unravel_index(np.argmax(np.abs(step500[0])), (192, 192))
    # outer layer never updated, fluid vw zero on the borehole columns
    for arr in (GOLD["txx500"], step500[0]):
        assert np.all(arr[0] == 0) and np.all(arr[-1] == 0) and np.all(arr[:, 0] == 0) and np.all(arr[:, -1] == 0)
    assert np.all(step500[3][:, 87:105] == 0) and np.all(GOLD["vwx500"][:, 87:105] == 0)


@pytest.mark.slow
def test_oracle_late_snapshots():
    g = O.reference_grid()
    base = O.borehole_model(128, 128, 128)
    steps = [1000, 1500, 2000, 2500, 3000, 3499]
    names = ["txx1000", "txx1500", "txx2000", "txx2500", "txx3000", "txx"]
    r = O.run("f32", g, base, 3500, snapshots=[(s, "txx") for s in steps])["snaps"]
    n500 = float(np.linalg.norm(GOLD["txx500"].astype(np.float64)))
    for k, n in enumerate(names):
        assert float(np.linalg.norm(r[k] - GOLD[n]) / n500) <= 1e-5, n
