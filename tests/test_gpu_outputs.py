"""The reference's output set written from a run (pw_write_outputs): the five
sis*.dat seismograms and the field+step snapshot files, in the reference's text
format and in the binary format, equal to what the handle returns."""
import os

import numpy as np
import pytest

import cases
import mixedprecisionfaracoustic_amd as pw

pytestmark = pytest.mark.gpu


def test_write_outputs_reference_set(tmp_path):
    grid, base, _ = cases.borehole_case()
    nt = 60
    s = cases.make_solver(grid, nt, "fp32", base=base)
    sx, sy, _ = grid["src"]
    s.set_receivers_reference(sx, sy)
    s.set_snapshots([20, nt - 1], ["txx", "vuz"], grid["ny"] // 2)
    s.run(nt)
    seis = s.reference_seismograms()
    snaps = s.snapshots()
    for fmt in ("dat", "bin"):
        d = tmp_path / fmt
        d.mkdir()
        s.write_outputs(str(d), fmt)
        names = sorted(os.listdir(d))
        assert names == sorted(f"{n}.{fmt}" for n in ("sisx", "sisz", "sisvu", "sisvw", "sisp", "txx20",
                                                      f"vuz{nt - 1}")), names
        for name, ref in list(seis.items()) + [("txx20", snaps[0]), (f"vuz{nt - 1}", snaps[1])]:
            p = str(d / f"{name}.{fmt}")
            if fmt == "bin":
                head, got = pw.read_bin(p)
                assert np.array_equal(got, ref), name
                assert head["label"] == name and head["dt"] == grid["dt"] and head["h"] == grid["h"]
            else:
                got = np.loadtxt(p, dtype=np.float64).reshape(ref.shape)
                # "%e" keeps 7 significant digits: within half a unit of the 7th
                err = np.abs(got - ref.astype(np.float64))
                lim = 5.01e-7 * np.abs(ref.astype(np.float64))
                bad = np.argwhere(err > lim)
                assert len(bad) == 0, (name, [(tuple(i), float(ref[tuple(i)]), float(got[tuple(i)])) for i in bad[:4]])
    s.close()
