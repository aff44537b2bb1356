"""Shared test cases: the same medium, source and receivers fed to the HIP
library and to the CPU oracle."""
import numpy as np

import mixedprecisionfaracoustic_amd as pw
from oracle import pw_oracle as O


# each field's position along (x, y, z): 1 = staggered to the half-cell point
STAGGER = {"p": (0, 0, 0), "txx": (0, 0, 0), "tyy": (0, 0, 0), "tzz": (0, 0, 0),
           "vux": (1, 0, 0), "vwx": (1, 0, 0), "vuy": (0, 1, 0), "vwy": (0, 1, 0), "vuz": (0, 0, 1), "vwz": (0, 0, 1),
           "txy": (1, 1, 0), "txz": (1, 0, 1), "tyz": (0, 1, 1)}


def mirror_deviation(a, name):
    """For a run whose medium and explosive source are symmetric about the
    middle plane of every axis (odd extents, source in the middle cell): a field
    at whole cells along an axis is even there, F[i] = F[n-1-i], and one
    staggered to i + 1/2 is odd, F[i] = -F[n-2-i].  Returns the largest
    deviation over the three axes relative to the field's peak, the smallest
    fraction of bit-exact cells, and the peak.  a is [z][y][x]."""
    a = np.asarray(a, np.float64)
    peak = float(np.abs(a).max())
    worst, exact = 0.0, 1.0
    for ax, stag in zip((2, 1, 0), STAGGER[name]):  # x, y, z
        if stag:
            b = np.take(a, range(a.shape[ax] - 1), axis=ax)
            m = -np.flip(b, axis=ax)
        else:
            b, m = a, np.flip(a, axis=ax)
        worst = max(worst, float(np.abs(b - m).max()) / peak if peak > 0 else 0.0)
        exact = min(exact, float(np.mean(b == m)))
    return worst, exact, peak


def rel_l2(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(got - ref) / den) if den > 0 else float(np.linalg.norm(got))


def borehole_case(nxi=32, nyi=32, nzi=32, np_=8, nt=200):
    """The reference's borehole model and physical parameters on a reduced grid."""
    nx, ny, nz = nxi + 2 * np_, nyi + 2 * np_, nzi + 2 * np_
    base32 = O.borehole_model(nzi, nyi, nxi)
    grid = {"nx": nx, "ny": ny, "nz": nz, "np": np_, "h": 0.01, "dt": 0.9e-6, "rc": 1e-6, "v0": 4000.0,
            "f0": 3000.0, "src": (nx // 2, ny // 2, min(np_ + 10, nz - 2)), "src_w": (1.0, 1.0, 1.0, 0.0)}
    return grid, base32.astype(np.float64), nt


def sandstone():
    return pw.rock(Ks=35e9, Kb=7e9, mu=9e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.2, kappa=1e-12,
                   eta=1e-3, tortuosity=2.0, pore_m=8.0)


def lower_layer():
    return pw.rock(Ks=35e9, Kb=15e9, mu=14e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.1, kappa=1e-14,
                   eta=1e-3, tortuosity=2.0, pore_m=8.0)


def layered_case(nx=40, ny=36, nz=44, np_=6, nt=150, perturb=0.05, seed=0):
    """Two-layer perturbed sandstone, metre-scale grid (the synthetic workload's physics)."""
    grid = {"nx": nx, "ny": ny, "nz": nz, "np": np_, "h": 5.0, "dt": 0.5e-3, "rc": 1e-4, "v0": 4000.0,
            "f0": 30.0, "src": (nx // 2, ny // 2, nz // 2 - 3), "src_w": (1.0, 1.0, 1.0, -1.0)}
    spec = pw.layered([0, nz // 2], [sandstone(), lower_layer()], perturb=perturb, corr_cells=3.0, seed=seed)
    return grid, spec, nt


def make_solver(grid, nt, precision, base=None, spec=None):
    s = pw.Solver(grid["nx"], grid["ny"], grid["nz"], grid["np"], grid["h"], grid["dt"], nt, precision=precision,
                  cpml_r=grid["rc"], cpml_v0=grid["v0"], space_order=grid.get("order", 2),
                  cpml_open=grid.get("open", 0), tiling=grid.get("tiling"))
    if spec is not None:
        s.set_medium_layered(spec)
    else:
        s.set_medium(base)
    sx, sy, sz = grid["src"]
    s.set_source(sx, sy, sz, grid["f0"], grid["src_w"])
    return s


def receivers(grid, n_line=None):
    """The reference's five channels on the vertical line through the source
    (Float/main.cu:1085-1093), plus a horizontal line off the source's symmetry
    planes (iy = sy + 3) for the shear and transverse channels: on a symmetry
    plane those are pure rounding noise and carry no parity information."""
    nx, ny, nz = grid["nx"], grid["ny"], grid["nz"]
    sx, sy, sz = grid["src"]
    recs = []
    for f in ("txx", "tzz", "vux", "vwx", "p"):
        for iz in range(nz):
            recs.append((f, sx + sy * nx + iz * nx * ny))
    for f in ("vuy", "vuz", "vwz", "tyy", "tyz", "txy", "txz"):
        for ix in range(0, nx, 3):
            recs.append((f, ix + (sy + 3) * nx + (sz + 2) * nx * ny))
    return recs


def run_gpu(grid, nt, precision, recs, snaps=(), base=None, spec=None, chunks=None, fields=()):
    s = make_solver(grid, nt, precision, base=base, spec=spec)
    s.set_receivers([f for f, _ in recs], [i for _, i in recs])
    if snaps:
        s.set_snapshots([st for st, _ in snaps], [f for _, f in snaps], grid["ny"] // 2)
    if chunks:
        for c in chunks:
            s.run(c)
    else:
        s.run(nt)
    out = {"seis": s.seismogram(), "snaps": s.snapshots() if snaps else None,
           "fields": {f: s.field(f) for f in fields}, "info": s.info()}
    base_out = s.base_fields() if spec is not None else None
    s.close()
    return out, base_out


def run_oracle(kind, grid, nt, recs, snaps=(), base=None, final=False):
    b = base.astype(np.float64 if kind == "f64" else np.float32)
    return O.run(kind, grid, b, nt, receivers=recs, snapshots=list(snaps), final=final)
