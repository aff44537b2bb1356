"""porowave: MI355X-native 3-D Biot poroelastic staggered-grid FD solver.

Python mirror of the C ABI in include/porowave.h (ctypes over the in-tree
libporowave.so).  The operator surface follows the reference programs
(Float/main.cu, Half/main.cu): the same medium / source / receiver parameters,
the same seismogram channels (sisx/sisz/sisvu/sisvw/sisp) and snapshot slices,
and a .dat writer in the reference's text format.  There is no CPU fallback:
every call runs the HIP kernels, and importing the library raises if the built
extension is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libporowave.so")

FIELDS = ["vux", "vuy", "vuz", "vwx", "vwy", "vwz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "p"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
BASE = ["rho", "rhof", "C1", "C2", "HH", "H2u", "C", "M", "mu"]
PRECISIONS = {"fp32": 0, "comp16": 1, "bf16": 2, "fp64": 3, "comp16f": 4}
# The reference's seismogram channels (Float/main.cu:1085-1093, 1213-1222).
REFERENCE_CHANNELS = [("sisx", "txx"), ("sisz", "tzz"), ("sisvu", "vux"), ("sisvw", "vwx"), ("sisp", "p")]
STATUS = {0: "PW_OK", 1: "PW_ERR_INVALID", 2: "PW_ERR_HIP", 3: "PW_ERR_OUT_OF_MEMORY", 4: "PW_ERR_STATE",
          5: "PW_ERR_UNSTABLE", 6: "PW_ERR_UNSUPPORTED", 7: "PW_ERR_IO", 8: "PW_ERR_TIMEOUT",
          9: "PW_ERR_CANCELLED", 10: "PW_ERR_NONFINITE"}
SHOT_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}
MAX_LAYERS = 16


class PorowaveError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("np", C.c_int),
                ("h", C.c_double), ("dt", C.c_double), ("nt", C.c_int), ("precision", C.c_int),
                ("cpml_r", C.c_double), ("cpml_v0", C.c_double),
                ("half_es", C.c_int), ("half_dyn_end", C.c_double), ("half_dyn_interval", C.c_int),
                ("device", C.c_int), ("stream", C.c_void_p), ("check_cfl", C.c_int), ("space_order", C.c_int),
                ("cpml_open", C.c_int), ("tile_cells", C.c_int), ("tile_waves_x", C.c_int), ("tile_rows", C.c_int),
                ("tile_planes", C.c_int)]


class Info(C.Structure):
    _fields_ = [("precision", C.c_int), ("em", C.c_int), ("epml", C.c_int),
                ("bytes_per_cell_step", C.c_double), ("bytes_velocity", C.c_double), ("bytes_stress", C.c_double),
                ("vp_max", C.c_double), ("courant", C.c_double),
                ("device_bytes", C.c_longlong), ("slab_lo", C.c_int * 3), ("slab_hi", C.c_int * 3)]


class Rock(C.Structure):
    _fields_ = [("Ks", C.c_double), ("Kb", C.c_double), ("mu", C.c_double), ("Kf", C.c_double),
                ("rhos", C.c_double), ("rhof", C.c_double), ("phi", C.c_double),
                ("kappa", C.c_double), ("eta", C.c_double), ("tortuosity", C.c_double), ("pore_m", C.c_double)]


class Layered(C.Structure):
    _fields_ = [("nlayers", C.c_int), ("z_top", C.c_int * MAX_LAYERS), ("rock", Rock * MAX_LAYERS),
                ("perturb", C.c_double), ("corr_cells", C.c_double), ("clamp", C.c_double),
                ("seed", C.c_ulonglong)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("stream", C.c_void_p), ("hbm_budget", C.c_longlong),
                ("max_active", C.c_int), ("round_steps", C.c_int), ("use_graphs", C.c_int), ("batched", C.c_int),
                ("fill_cells", C.c_longlong), ("persistent", C.c_int), ("max_wait_s", C.c_double)]


class ShotSpec(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("np", C.c_int),
                ("h", C.c_double), ("dt", C.c_double), ("nt", C.c_int), ("precision", C.c_int),
                ("cpml_r", C.c_double), ("cpml_v0", C.c_double), ("crop", C.c_int * 3),
                ("sx", C.c_int), ("sy", C.c_int), ("sz", C.c_int), ("f0", C.c_double), ("w", C.c_double * 4),
                ("nrec", C.c_int), ("rec_field", C.POINTER(C.c_int)), ("rec_index", C.POINTER(C.c_longlong))]


class ShotStatus(C.Structure):
    _fields_ = [("state", C.c_int), ("precision", C.c_int), ("nt", C.c_int), ("nrec", C.c_int),
                ("cells", C.c_longlong), ("device_bytes", C.c_longlong),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("error", C.c_int), ("error_text", C.c_char * 256)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("failed", C.c_longlong),
                ("rounds", C.c_longlong), ("graphs_built", C.c_longlong), ("launches", C.c_longlong),
                ("busy_s", C.c_double), ("setup_s", C.c_double), ("graph_s", C.c_double),
                ("retire_s", C.c_double), ("cell_steps", C.c_double), ("peak_bytes", C.c_longlong),
                ("peak_active", C.c_int)]


class BinHeader(C.Structure):
    _fields_ = [("magic", C.c_char * 4), ("version", C.c_int), ("kind", C.c_int), ("nr", C.c_int), ("nc", C.c_int),
                ("field", C.c_int), ("step", C.c_int), ("reserved", C.c_int), ("dt", C.c_double), ("h", C.c_double),
                ("scale", C.c_double), ("label", C.c_char * 24)]


FORMATS = {"dat": 0, "bin": 1}

# symbol -> (restype, argtypes); the same list the header declares.
_SIGS = {
    "pw_config_default": (None, [C.POINTER(Config)]),
    "pw_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "pw_destroy": (None, [C.c_void_p]),
    "pw_last_error": (C.c_char_p, []),
    "pw_set_medium": (C.c_int, [C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.c_int, C.c_int, C.c_int]),
    "pw_set_medium_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int]),
    "pw_set_source": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "pw_set_receivers": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "pw_set_receivers_reference": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "pw_set_snapshots": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
    "pw_run": (C.c_int, [C.c_void_p, C.c_int]),
    "pw_synchronize": (C.c_int, [C.c_void_p]),
    "pw_current_step": (C.c_int, [C.c_void_p]),
    "pw_reset": (C.c_int, [C.c_void_p]),
    "pw_get_seismogram": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pw_get_snapshots": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pw_get_field": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float)]),
    "pw_set_volume_snapshots": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "pw_get_volume_snapshot": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float)]),
    "pw_write_dat": (C.c_int, [C.c_char_p, C.POINTER(C.c_float), C.c_int, C.c_int]),
    "pw_get_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "pw_medium_borehole": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "pw_biot_from_rock": (C.c_int, [C.POINTER(Rock), C.POINTER(C.c_double)]),
    "pw_set_medium_layered": (C.c_int, [C.c_void_p, C.POINTER(Layered)]),
    "pw_get_base_fields": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pw_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "pw_get_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_longlong)]),
    "pw_codec_decode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p]),
    "pw_codec_encode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p]),
    "pw_write_bin": (C.c_int, [C.c_char_p, C.POINTER(BinHeader), C.POINTER(C.c_float)]),
    "pw_read_bin": (C.c_int, [C.c_char_p, C.POINTER(BinHeader), C.POINTER(C.c_float)]),
    "pw_write_outputs": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "pw_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "pw_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "pw_server_destroy": (None, [C.c_void_p]),
    "pw_server_set_medium": (C.c_int, [C.c_void_p, C.POINTER(C.POINTER(C.c_double)), C.c_int, C.c_int, C.c_int]),
    "pw_server_set_medium_layered": (C.c_int, [C.c_void_p, C.POINTER(Layered), C.c_int, C.c_int, C.c_int,
                                               C.c_int]),
    "pw_server_submit": (C.c_int, [C.c_void_p, C.POINTER(ShotSpec), C.POINTER(C.c_longlong)]),
    "pw_server_wait": (C.c_int, [C.c_void_p, C.c_longlong, C.c_double]),
    "pw_server_status": (C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(ShotStatus)]),
    "pw_server_result": (C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(C.c_float)]),
    "pw_server_release": (C.c_int, [C.c_void_p, C.c_longlong]),
    "pw_server_cancel": (C.c_int, [C.c_void_p, C.c_longlong]),
    "pw_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "pw_server_clock": (C.c_double, [C.c_void_p]),
}

_lib = None


def lib():
    """Load libporowave.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"porowave HIP library not built: {LIB_PATH} "
                              "(run python -m mixedprecisionfaracoustic_amd.build)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


FACES = {"x_lo": 1, "x_hi": 2, "y_lo": 4, "y_hi": 8, "z_lo": 16, "z_hi": 32}


def face_mask(faces):
    """PW_FACE_* bit mask from face names (or an int mask, passed through)."""
    if isinstance(faces, (int, np.integer)):
        return int(faces)
    m = 0
    for f in faces:
        if f not in FACES:
            raise ValueError(f"unknown face {f!r}; faces are {sorted(FACES)}")
        m |= FACES[f]
    return m


def _check(status):
    if status != 0:
        raise PorowaveError(status, lib().pw_last_error().decode(errors="replace"))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def medium_borehole(nxi, nyi, nzi):
    """create_model (Float/main.cu:1590-1655): base fields [9, nzi, nyi, nxi] float64."""
    out = np.zeros((9, nzi, nyi, nxi), np.float64)
    _check(lib().pw_medium_borehole(nxi, nyi, nzi, _dp(out)))
    return out


def rock(**kw):
    r = Rock()
    defaults = {"tortuosity": 1.0, "pore_m": 8.0}
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(r, k, float(v))
    return r


def biot_from_rock(r):
    out = np.zeros(9, np.float64)
    _check(lib().pw_biot_from_rock(C.byref(r), _dp(out)))
    return out


def layered(z_tops, rocks, perturb=0.0, corr_cells=8.0, clamp=3.0, seed=0):
    L = Layered()
    L.nlayers = len(rocks)
    for k, (z, r) in enumerate(zip(z_tops, rocks)):
        L.z_top[k] = int(z)
        L.rock[k] = r
    L.perturb, L.corr_cells, L.clamp, L.seed = float(perturb), float(corr_cells), float(clamp), int(seed)
    return L


def write_dat(path, data):
    """wfile / wfile3d text format (Float/main.cu:1501-1516, 1563-1588)."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim != 2:
        raise ValueError("write_dat expects a 2-D array")
    _check(lib().pw_write_dat(os.fsencode(path), _fp(a), a.shape[0], a.shape[1]))


def write_bin(path, data, kind=0, field=-1, step=-1, dt=0.0, h=0.0, label=""):
    """Binary seismogram (kind 0) or snapshot (kind 1) file: 80-byte header + FP32 rows."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim != 2:
        raise ValueError("write_bin expects a 2-D array")
    hd = BinHeader()
    hd.kind, hd.nr, hd.nc, hd.field, hd.step = int(kind), a.shape[0], a.shape[1], int(field), int(step)
    hd.dt, hd.h, hd.scale = float(dt), float(h), 1.0
    hd.label = label.encode()[:23]
    _check(lib().pw_write_bin(os.fsencode(path), C.byref(hd), _fp(a)))


def read_bin(path):
    """(header dict, [nr, nc] float32 array) of a binary file written by write_bin / write_outputs."""
    hd = BinHeader()
    _check(lib().pw_read_bin(os.fsencode(path), C.byref(hd), None))
    out = np.zeros((hd.nr, hd.nc), np.float32)
    _check(lib().pw_read_bin(os.fsencode(path), C.byref(hd), _fp(out) if out.size else None))
    head = {k: getattr(hd, k) for k, _ in BinHeader._fields_}
    head["magic"] = head["magic"].decode()
    head["label"] = head["label"].decode()
    return head, out


def codec_decode(hi_ptr, lo_ptr, out_ptr, n, stream=None):
    """pw_codec_decode on device addresses (ints): n FP16 bit patterns and int8
    residuals to n floats, with the step kernels' decode."""
    _check(lib().pw_codec_decode(hi_ptr, lo_ptr, out_ptr, int(n), stream))


def codec_encode(in_ptr, hi_ptr, lo_ptr, n, precision, stream=None):
    """pw_codec_encode on device addresses (ints): n floats to FP16 bit patterns
    and int8 residuals, rounded as the precision's ("comp16" or "comp16f") step
    kernels round them."""
    _check(lib().pw_codec_encode(in_ptr, hi_ptr, lo_ptr, int(n), PRECISIONS[precision], stream))


class Solver:
    """One shot on one GPU (a pw_solver handle).

    nx, ny, nz are the total grid including the CPML (the reference's NX_ext).
    precision: "fp32" (Float path), "comp16" (Half path: FP16 + int8 residual),
    "bf16", or "fp64" (the FP64 oracle run).
    cpml_open: faces without an absorbing layer, any of "x_lo", "x_hi", "y_lo",
    "y_hi", "z_lo", "z_hi" (or the PW_FACE_* bit mask); e.g. ("z_lo", "z_hi")
    is one interior rank's z-slab, whose outermost planes are its ghost planes.
    tiling: optional dict of the step kernels' tile shape, keys "cells" (per
    lane along x: 1, 2, 4), "waves_x" (1, 2), "rows" (waves per block: 1, 2, 4,
    8), "planes" (z planes per block); omitted keys take the precision's
    default.  Every tiling gives bit-identical results.
    """

    def __init__(self, nx, ny, nz, np_, h, dt, nt, precision="fp32", cpml_r=1e-6, cpml_v0=4000.0,
                 device=0, check_cfl=True, half_es=0, half_dyn_end=16.0, half_dyn_interval=100, stream=None,
                 space_order=2, cpml_open=0, tiling=None):
        L = lib()
        cfg = Config()
        L.pw_config_default(C.byref(cfg))
        cfg.nx, cfg.ny, cfg.nz, cfg.np = int(nx), int(ny), int(nz), int(np_)
        cfg.h, cfg.dt, cfg.nt = float(h), float(dt), int(nt)
        cfg.precision = PRECISIONS[precision]
        cfg.cpml_r, cfg.cpml_v0 = float(cpml_r), float(cpml_v0)
        cfg.half_es, cfg.half_dyn_end, cfg.half_dyn_interval = int(half_es), float(half_dyn_end), int(half_dyn_interval)
        cfg.device = int(device)
        cfg.stream = stream
        cfg.check_cfl = 1 if check_cfl else 0
        cfg.space_order = int(space_order)
        cfg.cpml_open = face_mask(cpml_open)
        for k, v in (tiling or {}).items():
            if k not in ("cells", "waves_x", "rows", "planes"):
                raise ValueError(f"unknown tiling key {k!r}")
            setattr(cfg, "tile_" + k, int(v))
        h_ = C.c_void_p()
        _check(L.pw_create(C.byref(cfg), C.byref(h_)))
        self._h = h_
        self.cfg = cfg
        self.precision = precision
        self.nrec = 0
        self.nsnap = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().pw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.cfg.nz, self.cfg.ny, self.cfg.nx)

    @property
    def interior(self):
        c = self.cfg
        return (c.nz - 2 * c.np, c.ny - 2 * c.np, c.nx - 2 * c.np)

    # -- set-up -----------------------------------------------------------
    def set_medium(self, base):
        nzi, nyi, nxi = self.interior
        b = np.ascontiguousarray(base, np.float64)
        if b.shape != (9, nzi, nyi, nxi):
            raise ValueError(f"base must be [9, {nzi}, {nyi}, {nxi}], got {b.shape}")
        ptrs = (C.POINTER(C.c_double) * 9)(*[_dp(b[k]) for k in range(9)])
        _check(lib().pw_set_medium(self._h, ptrs, nxi, nyi, nzi))

    def set_medium_device(self, ptrs):
        """pw_set_medium_device: nine device addresses (ints) of the base fields on
        the handle's device; allocates nothing.  The caller guarantees that each
        address holds nzi*nyi*nxi FP64 values (this handle's interior, x fastest):
        device buffers carry no size the library could check."""
        nzi, nyi, nxi = self.interior
        p = [int(x) for x in ptrs]
        if len(p) != 9:
            raise ValueError("set_medium_device takes 9 device addresses")
        arr = (C.c_void_p * 9)(*p)
        _check(lib().pw_set_medium_device(self._h, arr, nxi, nyi, nzi))

    def set_medium_layered(self, spec):
        _check(lib().pw_set_medium_layered(self._h, C.byref(spec)))

    def base_fields(self):
        nzi, nyi, nxi = self.interior
        out = np.zeros((9, nzi, nyi, nxi), np.float64)
        _check(lib().pw_get_base_fields(self._h, _dp(out)))
        return out

    def set_source(self, sx, sy, sz, f0, weights=(1.0, 1.0, 1.0, 0.0)):
        w = np.ascontiguousarray(weights, np.float64)
        _check(lib().pw_set_source(self._h, int(sx), int(sy), int(sz), float(f0), _dp(w)))

    def set_receivers(self, fields, indices):
        f = np.ascontiguousarray([FIELD_ID[x] if isinstance(x, str) else int(x) for x in fields], np.int32)
        i = np.ascontiguousarray(indices, np.int64)
        if f.shape != i.shape:
            raise ValueError("fields and indices must have the same length")
        _check(lib().pw_set_receivers(self._h, len(f), f.ctypes.data_as(C.POINTER(C.c_int)),
                                      i.ctypes.data_as(C.POINTER(C.c_longlong))))
        self.nrec = len(f)

    def set_receivers_reference(self, sx, sy):
        _check(lib().pw_set_receivers_reference(self._h, int(sx), int(sy)))
        self.nrec = 5 * self.cfg.nz

    def set_snapshots(self, steps, fields, plane_y=None):
        s = np.ascontiguousarray(steps, np.int32)
        f = np.ascontiguousarray([FIELD_ID[x] if isinstance(x, str) else int(x) for x in fields], np.int32)
        y = self.cfg.ny // 2 if plane_y is None else int(plane_y)
        _check(lib().pw_set_snapshots(self._h, len(s), s.ctypes.data_as(C.POINTER(C.c_int)),
                                      f.ctypes.data_as(C.POINTER(C.c_int)), y))
        self.nsnap = len(s)

    def set_volume_snapshots(self, steps, fields):
        """Full 3-D snapshots after the given steps, copied to pinned host memory
        on a second stream while the run goes on (volume_snapshot(k) reads one)."""
        s = np.ascontiguousarray(steps, np.int32)
        f = np.ascontiguousarray([FIELD_ID[x] if isinstance(x, str) else int(x) for x in fields], np.int32)
        _check(lib().pw_set_volume_snapshots(self._h, len(s), s.ctypes.data_as(C.POINTER(C.c_int)),
                                             f.ctypes.data_as(C.POINTER(C.c_int))))
        self.nvol = len(s)

    def volume_snapshot(self, k):
        out = np.zeros(self.shape, np.float32)
        _check(lib().pw_get_volume_snapshot(self._h, int(k), _fp(out)))
        return out

    # -- stepping -----------------------------------------------------------
    def run(self, nsteps):
        _check(lib().pw_run(self._h, int(nsteps)))

    def synchronize(self):
        _check(lib().pw_synchronize(self._h))

    def reset(self):
        _check(lib().pw_reset(self._h))

    @property
    def step(self):
        return lib().pw_current_step(self._h)

    # -- outputs ------------------------------------------------------------
    def seismogram(self):
        out = np.zeros((max(self.nrec, 1), self.cfg.nt), np.float32)
        _check(lib().pw_get_seismogram(self._h, _fp(out)))
        return out[:self.nrec]

    def reference_seismograms(self):
        """Split the reference receiver set into {sisx, sisz, sisvu, sisvw, sisp}: [nz, nt] each."""
        s = self.seismogram()
        nz = self.cfg.nz
        return {name: s[k * nz:(k + 1) * nz] for k, (name, _) in enumerate(REFERENCE_CHANNELS)}

    def snapshots(self):
        out = np.zeros((max(self.nsnap, 1), self.cfg.nz, self.cfg.nx), np.float32)
        _check(lib().pw_get_snapshots(self._h, _fp(out)))
        return out[:self.nsnap]

    def field(self, name):
        out = np.zeros(self.shape, np.float32)
        _check(lib().pw_get_field(self._h, FIELD_ID[name] if isinstance(name, str) else int(name), _fp(out)))
        return out

    def write_outputs(self, directory, fmt="dat"):
        """The reference's output files (sisx/sisz/sisvu/sisvw/sisp, txx500, ...) as
        .dat text or .bin binary into directory."""
        _check(lib().pw_write_outputs(self._h, os.fsencode(directory), FORMATS[fmt]))

    def set_profiling(self, enable=True):
        _check(lib().pw_set_profiling(self._h, 1 if enable else 0))

    def kernel_times(self):
        """(ms_velocity_total, ms_stress_total, launches) since the last call."""
        v, t, n = C.c_double(), C.c_double(), C.c_longlong()
        _check(lib().pw_get_kernel_times(self._h, C.byref(v), C.byref(t), C.byref(n)))
        return v.value, t.value, n.value

    def info(self):
        i = Info()
        _check(lib().pw_get_info(self._h, C.byref(i)))
        return {"precision": self.precision, "em": i.em, "epml": i.epml,
                "bytes_per_cell_step": i.bytes_per_cell_step, "bytes_velocity": i.bytes_velocity,
                "bytes_stress": i.bytes_stress, "vp_max": i.vp_max, "courant": i.courant,
                "device_bytes": i.device_bytes, "slab_lo": list(i.slab_lo), "slab_hi": list(i.slab_hi)}


def reference_solver(precision="fp32", nt=3500, **kw):
    """The reference configuration (Float/main.cu:20-23, 305-322, 1590-1655):
    192^3 total (128^3 + 32-cell PML), H = 1 cm, DT = 0.9 us, 3 kHz Ricker at
    (96, 96, 52), borehole medium, receivers on the vertical line through the source."""
    n = 128 + 64
    s = Solver(n, n, n, 32, 0.01, 0.9e-6, nt, precision=precision, cpml_r=1e-6, cpml_v0=4000.0, **kw)
    s.set_medium(medium_borehole(128, 128, 128))
    s.set_source(n // 2, n // 2, 20 + 32, 3000.0, (1.0, 1.0, 1.0, 0.0))
    s.set_receivers_reference(n // 2, n // 2)
    return s


class Server:
    """Shot server (a pw_server handle): admits shots first come, first served (or,
    with max_wait_s > 0, least work first unless a shot has waited longer than
    that) by device-memory budget and slot count, keeps them resident and advances all
    active shots together (one launch per step kernel per precision, rounds of
    steps replayed as hipGraphs; or, with persistent=True, one persistent
    cooperative launch per precision and round).  Shots crop their medium from
    one volume set with set_medium / set_medium_layered."""

    def __init__(self, device=0, max_active=16, round_steps=32, use_graphs=True, batched=True, hbm_budget=0,
                 stream=None, fill_cells=None, persistent=False, max_wait_s=0.0):
        L = lib()
        cfg = ServerConfig()
        L.pw_server_config_default(C.byref(cfg))
        cfg.device, cfg.max_active, cfg.round_steps = int(device), int(max_active), int(round_steps)
        cfg.use_graphs, cfg.batched, cfg.hbm_budget = int(bool(use_graphs)), int(bool(batched)), int(hbm_budget)
        cfg.stream = stream
        if fill_cells is not None:  # None keeps the library's default
            cfg.fill_cells = int(fill_cells)
        cfg.persistent = int(bool(persistent))
        cfg.max_wait_s = float(max_wait_s)
        h = C.c_void_p()
        _check(L.pw_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.cfg = cfg
        self._nrec = {}
        self._nt = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().pw_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_medium(self, base):
        """base: [9, nzi, nyi, nxi] float64 interior fields (the volume shots are cropped from)."""
        b = np.ascontiguousarray(base, np.float64)
        if b.ndim != 4 or b.shape[0] != 9:
            raise ValueError("base must be [9, nzi, nyi, nxi]")
        ptrs = (C.POINTER(C.c_double) * 9)(*[_dp(b[k]) for k in range(9)])
        _check(lib().pw_server_set_medium(self._h, ptrs, b.shape[3], b.shape[2], b.shape[1]))

    def set_medium_layered(self, spec, nxi, nyi, nzi, np_):
        _check(lib().pw_server_set_medium_layered(self._h, C.byref(spec), int(nxi), int(nyi), int(nzi), int(np_)))

    def submit(self, nx, ny, nz, np_, h, dt, nt, precision, crop, source, f0, fields, indices,
               weights=(1.0, 1.0, 1.0, 0.0), cpml_r=1e-4, cpml_v0=4000.0):
        q = ShotSpec()
        q.nx, q.ny, q.nz, q.np = int(nx), int(ny), int(nz), int(np_)
        q.h, q.dt, q.nt = float(h), float(dt), int(nt)
        q.precision = PRECISIONS[precision]
        q.cpml_r, q.cpml_v0 = float(cpml_r), float(cpml_v0)
        for a in range(3):
            q.crop[a] = int(crop[a])
        q.sx, q.sy, q.sz = (int(v) for v in source)
        q.f0 = float(f0)
        for k in range(4):
            q.w[k] = float(weights[k])
        f = np.ascontiguousarray([FIELD_ID[x] if isinstance(x, str) else int(x) for x in fields], np.int32)
        i = np.ascontiguousarray(indices, np.int64)
        if f.shape != i.shape:
            raise ValueError("fields and indices must have the same length")
        q.nrec = len(f)
        q.rec_field = f.ctypes.data_as(C.POINTER(C.c_int))
        q.rec_index = i.ctypes.data_as(C.POINTER(C.c_longlong))
        sid = C.c_longlong()
        _check(lib().pw_server_submit(self._h, C.byref(q), C.byref(sid)))
        self._nrec[sid.value] = len(f)
        self._nt[sid.value] = int(nt)
        return sid.value

    def wait(self, sid, timeout=None):
        _check(lib().pw_server_wait(self._h, int(sid), -1.0 if timeout is None else float(timeout)))

    def status(self, sid):
        st = ShotStatus()
        _check(lib().pw_server_status(self._h, int(sid), C.byref(st)))
        return {"state": SHOT_STATES[st.state], "precision": st.precision, "nt": st.nt, "nrec": st.nrec,
                "cells": st.cells, "device_bytes": st.device_bytes, "t_submit": st.t_submit,
                "t_admit": st.t_admit, "t_done": st.t_done, "error": st.error,
                "error_text": st.error_text.decode(errors="replace")}

    def result(self, sid):
        n = self._nrec[sid]
        out = np.zeros((max(n, 1), self._nt[sid]), np.float32)
        _check(lib().pw_server_result(self._h, int(sid), _fp(out)))
        return out[:n]

    def cancel(self, sid):
        """Cancels a queued or running shot (it then fails with PW_ERR_CANCELLED)."""
        _check(lib().pw_server_cancel(self._h, int(sid)))

    def release(self, sid):
        _check(lib().pw_server_release(self._h, int(sid)))
        self._nrec.pop(sid, None)
        self._nt.pop(sid, None)

    def stats(self):
        st = ServerStats()
        _check(lib().pw_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def clock(self):
        return lib().pw_server_clock(self._h)
