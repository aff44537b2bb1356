"""The allocation-free medium installation (pw_set_medium_device, through the
C ABI) against pw_set_medium, on two calls in a row on one device, the second
on a smaller grid whose medium is softer than the first's.

pw_set_medium_device reduces the medium's maxima into a file-scope device
array that every call on the device shares; each call must clear it, or the
second call would take the first medium's fastest velocity and moduli (its
stability check, and comp16's exponents).  Both calls must give the same
handle state as pw_set_medium with the same values in host memory: the same
vp_max, Courant number and Half exponents, and bit-identical seismograms and
fields after a run.  The base fields are put in device memory by torch before
the call (plumbing, outside the entry point)."""
import numpy as np
import pytest
import torch

import cases
import mixedprecisionfaracoustic_amd as pw

pytestmark = pytest.mark.gpu


def case(n, rock, prec, nt):
    grid = {"nx": n, "ny": n + 2, "nz": n - 2, "np": 6, "h": 5.0, "dt": 0.5e-3, "rc": 1e-4, "v0": 4000.0,
            "f0": 30.0, "src": (n // 2, n // 2 + 1, n // 2 - 3), "src_w": (1.0, 1.0, 1.0, -1.0)}
    spec = pw.layered([0, (n - 2) // 2], [rock, rock], perturb=0.05, corr_cells=3.0, seed=11)
    s = pw.Solver(grid["nx"], grid["ny"], grid["nz"], grid["np"], grid["h"], grid["dt"], 2)
    s.set_medium_layered(spec)
    base = s.base_fields()
    s.close()
    return grid, base, prec, nt


def run(grid, base, prec, nt, device_medium):
    s = pw.Solver(grid["nx"], grid["ny"], grid["nz"], grid["np"], grid["h"], grid["dt"], nt, precision=prec,
                  cpml_r=grid["rc"], cpml_v0=grid["v0"])
    keep = None
    if device_medium:
        keep = torch.from_numpy(np.ascontiguousarray(base)).to("cuda:0")
        torch.cuda.synchronize()
        s.set_medium_device([keep[k].data_ptr() for k in range(9)])
    else:
        s.set_medium(base)
    s.set_source(*grid["src"], grid["f0"], grid["src_w"])
    recs = cases.receivers(grid)
    s.set_receivers([f for f, _ in recs], [i for _, i in recs])
    s.run(nt)
    out = {"seis": s.seismogram(), "info": s.info(), "fields": {f: s.field(f) for f in ("vux", "txz", "p")}}
    s.close()
    del keep
    return out


def test_medium_from_device_memory_matches_host_medium_twice():
    big = case(40, cases.lower_layer(), "comp16", 90)  # stiffer rock: larger maxima
    small = case(30, cases.sandstone(), "comp16", 70)  # a smaller grid, a softer rock
    for grid, base, prec, nt in (big, small):
        ref = run(grid, base, prec, nt, device_medium=False)
        got = run(grid, base, prec, nt, device_medium=True)
        for k in ("vp_max", "courant", "em", "epml"):
            assert got["info"][k] == ref["info"][k], (grid["nx"], k, got["info"][k], ref["info"][k])
        assert np.abs(ref["seis"]).max() > 0
        assert np.array_equal(got["seis"], ref["seis"]), grid["nx"]
        for f in ref["fields"]:
            assert np.array_equal(got["fields"][f], ref["fields"][f]), (grid["nx"], f)
    assert small[0]["nx"] < big[0]["nx"]


def test_medium_from_device_memory_refuses_what_it_cannot_do():
    grid, base, prec, nt = case(30, cases.sandstone(), "fp32", 10)
    keep = torch.from_numpy(np.ascontiguousarray(base)).to("cuda:0")
    ptrs = [keep[k].data_ptr() for k in range(9)]
    s = pw.Solver(grid["nx"], grid["ny"], grid["nz"], grid["np"], grid["h"], grid["dt"], nt, cpml_r=1e-4,
                  cpml_v0=0.0)
    with pytest.raises(pw.PorowaveError) as ei:
        s.set_medium_device(ptrs)  # CPML velocity from the medium: host tables
    assert ei.value.status == 6
    s.close()
    # an interior size other than the handle's is refused on the host, before any
    # kernel runs (the C entry point is called directly: the Python method always
    # passes the handle's own size)
    import ctypes as C
    s = pw.Solver(grid["nx"], grid["ny"], grid["nz"], grid["np"], grid["h"], grid["dt"], nt, cpml_r=1e-4)
    nzi, nyi, nxi = s.interior
    arr = (C.c_void_p * 9)(*ptrs)
    assert pw.lib().pw_set_medium_device(s._h, arr, nxi + 2, nyi, nzi) == 1  # PW_ERR_INVALID
    assert "interior size" in pw.lib().pw_last_error().decode()
    with pytest.raises(pw.PorowaveError) as ei:
        s.run(1)  # nothing installed
    assert ei.value.status == 4
    s.close()
    del keep
