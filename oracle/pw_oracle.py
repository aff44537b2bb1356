"""ctypes binding of the CPU oracle (oracle/libpw_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpw_oracle.so")

FIELDS = ["vux", "vuy", "vuz", "vwx", "vwy", "vwz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "p"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
BASE = ["rho", "rhof", "C1", "C2", "HH", "H2u", "C", "M", "mu"]


class OrArgs(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("np", C.c_int),
        ("h", C.c_double), ("dt", C.c_double), ("rc", C.c_double), ("v0", C.c_double),
        ("nt", C.c_int), ("f0", C.c_double),
        ("sx", C.c_int), ("sy", C.c_int), ("sz", C.c_int),
        ("src_w", C.c_double * 4),
        ("nrec", C.c_int), ("rec_field", C.POINTER(C.c_int)), ("rec_index", C.POINTER(C.c_longlong)),
        ("seis", C.POINTER(C.c_float)),
        ("nsnap", C.c_int), ("snap_step", C.POINTER(C.c_int)), ("snap_field", C.POINTER(C.c_int)),
        ("snap_y", C.c_int), ("snaps", C.POINTER(C.c_float)),
        ("final_state", C.POINTER(C.c_float)),
        ("pml_everywhere", C.c_int), ("nthreads", C.c_int),
        ("half_es", C.c_int), ("half_dyn_end", C.c_double), ("half_dyn_interval", C.c_int),
        ("half_em_out", C.c_int), ("half_epml_out", C.c_int),
        ("order", C.c_int), ("pml_open", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        L.or_run_f32.argtypes = [C.POINTER(OrArgs), fp]
        L.or_run_f64.argtypes = [C.POINTER(OrArgs), dp]
        L.or_run_half.argtypes = [C.POINTER(OrArgs), fp]
        for f in (L.or_run_f32, L.or_run_f64, L.or_run_half):
            f.restype = C.c_int
        L.or_create_model_borehole.argtypes = [C.c_int, C.c_int, C.c_int, fp]
        L.or_create_model_borehole.restype = None
        L.or_biot_from_moduli.argtypes = [C.c_double] * 11 + [dp]
        L.or_biot_from_moduli.restype = None
        L.or_pml_profile_f32.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp, fp, fp, fp]
        L.or_pml_profile_f32.restype = None
        L.or_pml_profile_f64.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
        L.or_pml_profile_f64.restype = None
        L.or_ricker_f32.argtypes = [C.c_int, C.c_float, C.c_float]
        L.or_ricker_f32.restype = C.c_float
        L.or_ricker_f64.argtypes = [C.c_int, C.c_double, C.c_double]
        L.or_ricker_f64.restype = C.c_double
        L.or_time_f32.argtypes = [C.POINTER(OrArgs), fp, C.c_int, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_int)]
        L.or_time_f32.restype = C.c_int
        L.or_precompute_f32.argtypes = [C.POINTER(OrArgs), fp, fp]
        L.or_precompute_f32.restype = C.c_int
        L.or_f32_to_f16.argtypes = [C.c_float]
        L.or_f32_to_f16.restype = C.c_ushort
        L.or_f16_to_f32.argtypes = [C.c_ushort]
        L.or_f16_to_f32.restype = C.c_float
        L.or_residual_unit.argtypes = [C.c_float]
        L.or_residual_unit.restype = C.c_float
        L.or_store_ch.argtypes = [C.c_float, C.POINTER(C.c_ushort), C.POINTER(C.c_byte)]
        L.or_store_ch.restype = None
        L.or_load_ch.argtypes = [C.c_ushort, C.c_byte]
        L.or_load_ch.restype = C.c_float
        L.or_load_ch_n.argtypes = [C.POINTER(C.c_ushort), C.POINTER(C.c_byte), C.c_long, fp]
        L.or_load_ch_n.restype = None
        L.or_store_ch_n.argtypes = [fp, C.c_long, C.POINTER(C.c_ushort), C.POINTER(C.c_byte)]
        L.or_store_ch_n.restype = None
        L.or_half_precompute.argtypes = [C.POINTER(OrArgs), fp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_ushort)]
        L.or_half_precompute.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def load_ch(hi, lo):
    """load_cr_value (Half/main.cu:69-73) over arrays: FP16 bit patterns and int8 residuals to float32."""
    hi = np.ascontiguousarray(hi, np.uint16)
    lo = np.ascontiguousarray(lo, np.int8)
    out = np.empty(hi.shape, np.float32)
    lib().or_load_ch_n(_ptr(hi, C.c_ushort), _ptr(lo, C.c_byte), hi.size, _ptr(out, C.c_float))
    return out


def store_ch(v):
    """store_ch (Half/main.cu:85-100) over an array: float32 values to (FP16 bit patterns, int8 residuals)."""
    v = np.ascontiguousarray(v, np.float32)
    hi = np.empty(v.shape, np.uint16)
    lo = np.empty(v.shape, np.int8)
    lib().or_store_ch_n(_ptr(v, C.c_float), v.size, _ptr(hi, C.c_ushort), _ptr(lo, C.c_byte))
    return hi, lo


def borehole_model(nzi, nyi, nxi):
    """Base fields of create_model (Float/main.cu:1590-1655), float32 [9, nz, ny, nx]."""
    out = np.zeros((9, nzi, nyi, nxi), np.float32)
    lib().or_create_model_borehole(nzi, nyi, nxi, _ptr(out, C.c_float))
    return out


def biot_from_moduli(Ks, Kb, mu, Kf, rhos, rhof, phi, kappa, eta, tortuosity, pore_m=8.0):
    out = np.zeros(9, np.float64)
    lib().or_biot_from_moduli(Ks, Kb, mu, Kf, rhos, rhof, phi, kappa, eta, tortuosity, pore_m,
                              _ptr(out, C.c_double))
    return out


def pml_profile_f32(n, np_, h, dt, rc, v0):
    d, d2, e, e2 = (np.zeros(n, np.float32) for _ in range(4))
    lib().or_pml_profile_f32(n, np_, h, dt, rc, v0, *(_ptr(a, C.c_float) for a in (d, d2, e, e2)))
    return d, d2, e, e2


def pml_profile_f64(n, np_, h, dt, rc, v0):
    d, d2, e, e2 = (np.zeros(n, np.float64) for _ in range(4))
    lib().or_pml_profile_f64(n, np_, h, dt, rc, v0, *(_ptr(a, C.c_double) for a in (d, d2, e, e2)))
    return d, d2, e, e2


def precompute_f32(grid, base):
    args, keep = _args(grid, 0, [], [], None, False)
    n = grid["nx"] * grid["ny"] * grid["nz"]
    out = np.zeros((17, n), np.float32)
    b = np.ascontiguousarray(base, np.float32)
    rc = lib().or_precompute_f32(C.byref(args), _ptr(b, C.c_float), _ptr(out, C.c_float))
    if rc:
        raise RuntimeError(f"or_precompute_f32 failed: {rc}")
    return out.reshape(17, grid["nz"], grid["ny"], grid["nx"])


def half_precompute(grid, base, want_coef=False):
    args, keep = _args(grid, 0, [], [], None, False)
    b = np.ascontiguousarray(base, np.float32)
    em, epml = C.c_int(0), C.c_int(0)
    n = grid["nx"] * grid["ny"] * grid["nz"]
    coef = np.zeros((22, n), np.uint16) if want_coef else None
    rc = lib().or_half_precompute(C.byref(args), _ptr(b, C.c_float), C.byref(em), C.byref(epml),
                                  _ptr(coef, C.c_ushort) if want_coef else None)
    if rc:
        raise RuntimeError(f"or_half_precompute failed: {rc}")
    return em.value, epml.value, (coef.reshape(22, grid["nz"], grid["ny"], grid["nx"]) if want_coef else None)


def _args(grid, nt, receivers, snapshots, snap_y, final, pml_everywhere=False, nthreads=0, half=None):
    a = OrArgs()
    for k in ("nx", "ny", "nz", "np"):
        setattr(a, k, int(grid[k]))
    a.h, a.dt = float(grid["h"]), float(grid["dt"])
    a.rc, a.v0 = float(grid.get("rc", 1e-6)), float(grid.get("v0", 4000.0))
    a.nt = int(nt)
    a.f0 = float(grid.get("f0", 3000.0))
    src = grid.get("src", (grid["nx"] // 2, grid["ny"] // 2, min(20 + grid["np"], grid["nz"] - 2)))
    a.sx, a.sy, a.sz = (int(v) for v in src)
    w = grid.get("src_w", (1.0, 1.0, 1.0, 0.0))
    for i in range(4):
        a.src_w[i] = float(w[i])
    keep = []
    nrec = len(receivers)
    a.nrec = nrec
    seis = np.zeros((max(nrec, 1), max(nt, 1)), np.float32)
    if nrec:
        rf = np.array([FIELD_ID[f] if isinstance(f, str) else int(f) for f, _ in receivers], np.int32)
        ri = np.array([int(i) for _, i in receivers], np.int64)
        keep += [rf, ri]
        a.rec_field, a.rec_index = _ptr(rf, C.c_int), _ptr(ri, C.c_longlong)
        a.seis = _ptr(seis, C.c_float)
    keep.append(seis)
    ns = len(snapshots)
    a.nsnap = ns
    snaps = np.zeros((max(ns, 1), grid["nz"], grid["nx"]), np.float32)
    if ns:
        st = np.array([int(s) for s, _ in snapshots], np.int32)
        sf = np.array([FIELD_ID[f] if isinstance(f, str) else int(f) for _, f in snapshots], np.int32)
        keep += [st, sf]
        a.snap_step, a.snap_field = _ptr(st, C.c_int), _ptr(sf, C.c_int)
        a.snap_y = int(grid["ny"] // 2 if snap_y is None else snap_y)
        a.snaps = _ptr(snaps, C.c_float)
    keep.append(snaps)
    fin = None
    if final:
        fin = np.zeros((13, grid["nz"], grid["ny"], grid["nx"]), np.float32)
        a.final_state = _ptr(fin, C.c_float)
    keep.append({"seis": seis, "snaps": snaps, "final": fin})
    a.pml_everywhere = 1 if pml_everywhere else 0
    a.nthreads = int(nthreads)
    half = half or {}
    a.half_es = int(half.get("es", 0))
    a.half_dyn_end = float(half.get("dyn_end", 16.0))
    a.half_dyn_interval = int(half.get("dyn_interval", 100))
    a.order = int(grid.get("order", 2))
    a.pml_open = int(grid.get("open", 0))
    return a, keep


def run(kind, grid, base, nt, receivers=(), snapshots=(), snap_y=None, final=False,
        pml_everywhere=False, nthreads=0, half=None):
    """Run the oracle.

    kind: "f32" (Float/main.cu as written), "f64" (FP64 oracle) or "half"
    (Half/main.cu).  grid: dict with nx, ny, nz (extended), np, h, dt, rc, v0, f0,
    src (sx, sy, sz), src_w (txx, tyy, tzz, p weights).  base: [9, nzi, nyi, nxi].
    receivers: list of (field, flat_index); snapshots: list of (step, field).
    Returns dict(seis=[nrec, nt], snaps=[nsnap, nz, nx], final=[13, nz, ny, nx] or None,
    em/epml for half).
    """
    args, keep = _args(grid, nt, list(receivers), list(snapshots), snap_y, final,
                       pml_everywhere, nthreads, half)
    L = lib()
    if kind == "f64":
        b = np.ascontiguousarray(base, np.float64)
        rc = L.or_run_f64(C.byref(args), _ptr(b, C.c_double))
    elif kind == "f32":
        b = np.ascontiguousarray(base, np.float32)
        rc = L.or_run_f32(C.byref(args), _ptr(b, C.c_float))
    elif kind == "half":
        b = np.ascontiguousarray(base, np.float32)
        rc = L.or_run_half(C.byref(args), _ptr(b, C.c_float))
    else:
        raise ValueError(kind)
    if rc:
        raise RuntimeError(f"oracle run ({kind}) failed with code {rc}")
    outs = keep[-1]
    seis, snaps, fin = outs["seis"], outs["snaps"], outs["final"]
    out = {"seis": seis[:len(receivers), :nt], "snaps": snaps[:len(snapshots)], "final": fin}
    if kind == "half":
        out["em"], out["epml"] = args.half_em_out, args.half_epml_out
    return out


def time_f32(grid, base, warm, steps, nthreads=0):
    """Seconds for `steps` oracle FP32 steps after `warm` untimed ones, and threads used."""
    args, keep = _args(grid, warm + steps, [], [], None, False, False, nthreads)
    b = np.ascontiguousarray(base, np.float32)
    sec, thr = C.c_double(), C.c_int()
    rc = lib().or_time_f32(C.byref(args), _ptr(b, C.c_float), int(warm), int(steps), C.byref(sec), C.byref(thr))
    if rc:
        raise RuntimeError(f"or_time_f32 failed: {rc}")
    return sec.value, thr.value


def reference_grid():
    """The reference's configuration (Float/main.cu:20-23, 305-322)."""
    n = 128 + 2 * 32
    return {"nx": n, "ny": n, "nz": n, "np": 32, "h": 0.01, "dt": 0.9e-6, "rc": 1e-6, "v0": 4000.0,
            "f0": 3000.0, "src": (n // 2, n // 2, 20 + 32), "src_w": (1.0, 1.0, 1.0, 0.0)}


def stagger_diff(order, forward, f, o, stride=1):
    """The oracle's staggered difference (the update's own macro) of the FP64
    array f at index o: forward (to o + stride/2) or backward, undivided."""
    L = lib()
    L.or_stagger_diff_f64.restype = C.c_double
    L.or_stagger_diff_f64.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.c_long, C.c_long]
    a = np.ascontiguousarray(f, np.float64)
    return L.or_stagger_diff_f64(int(order), 1 if forward else 0, a.ctypes.data_as(C.POINTER(C.c_double)), int(o),
                                 int(stride))

