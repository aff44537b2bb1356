"""Benchmark: Gcell-steps/s of the Biot staggered-grid step on one MI355X.

Workload (default): 512^3 total grid, dx = 5 m, dt = 0.5 ms, 20-cell CPML
(R = 1e-4, d0 from the medium's fastest P velocity), two layers split at z = 256 (upper: water-saturated
sandstone Ks 35 GPa, Kb 7, mu 9, Kf 2.25, rho_s 2650, rho_f 1000, phi 0.2,
kappa 1e-12, eta 1e-3, tortuosity 2; lower: Kb 15, mu 14, phi 0.1, kappa 1e-14),
Kb/mu/phi perturbed by 5 % smoothed Gaussian noise (sigma 8 cells, seed 0),
25 Hz Ricker at the centre into txx/tyy/tzz/p, 256 receivers x 5 channels on a
horizontal line 10 cells below the top CPML, txx snapshots every 500 steps.
Synthetic medium generated on the GPU (no data files).

A "step" is one pass of the hot path: source -> velocity kernel -> stress
kernel -> receiver gather.  Inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0).  --gpus > 1 is refused: the library runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-steps/s & HBM GB/s, FP32 vs comp-FP16 @512^3; shot p50/p99 latency"
# CPML velocity: 0 lets the library take d0 from the medium's fastest P velocity
CPML_V0 = 0.0


def workload(n, np_):
    import mixedprecisionfaracoustic_amd as pw
    upper = pw.rock(Ks=35e9, Kb=7e9, mu=9e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.2, kappa=1e-12,
                    eta=1e-3, tortuosity=2.0, pore_m=8.0)
    lower = pw.rock(Ks=35e9, Kb=15e9, mu=14e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.1, kappa=1e-14,
                    eta=1e-3, tortuosity=2.0, pore_m=8.0)
    return pw.layered([0, n // 2], [upper, lower], perturb=0.05, corr_cells=8.0, clamp=3.0, seed=0)


def large_workload(n, np_, seed=3):
    """The 1024^3 model: 8 horizontal layers, porosity U[0.05, 0.3], permeability
    log-uniform in [1e-15, 1e-12] m^2, dry moduli from a porosity trend (stiffer
    when less porous), layer tops spread with a seeded jitter, each cell's Kb,
    mu and phi perturbed by 5 % (8-cell correlation)."""
    import numpy as np
    import mixedprecisionfaracoustic_amd as pw
    rng = np.random.default_rng(seed)
    tops = [0] + sorted(int(v) for v in rng.choice(np.arange(np_ + 40, n - np_ - 40), 7, replace=False))
    rocks = []
    for _ in range(8):
        phi = float(rng.uniform(0.05, 0.3))
        kb = float(35e9 * (1 - phi / 0.4) ** 2 + 2e9)
        rocks.append(pw.rock(Ks=35e9, Kb=kb, mu=0.8 * kb, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=phi,
                             kappa=float(10 ** rng.uniform(-15, -12)), eta=1e-3, tortuosity=2.0, pore_m=8.0))
    return pw.layered(tops, rocks, perturb=0.05, corr_cells=8.0, clamp=3.0, seed=seed)


def receivers(n, np_):
    import numpy as np
    xs = np.linspace(np_, n - 1 - np_, 256).round().astype(np.int64)
    iz = np_ + 10
    iy = n // 2
    fields, idx = [], []
    for f in ("txx", "tzz", "vux", "vwx", "p"):
        for x in xs:
            fields.append(f)
            idx.append(int(x) + iy * n + iz * n * n)
    return fields, idx


def measured_traffic(precision, kernel, n):
    """HBM bytes per launch of the kernel on the n^3 grid from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/hbm_traffic_<n>.md), or None."""
    import re
    path = os.path.join(HERE, "profiles", f"hbm_traffic_{n}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"traffic (\S+) (\S+) (\d+) (\d+)\s*$", line)
        if m and m.group(1) == precision and m.group(2) == kernel and int(m.group(3)) == n:
            return float(m.group(4))
    return None


def cpu_baseline(n, np_, spec_base, dt, h, steps, bytes_per_cell_step, order=2, v0=4000.0):
    """The oracle (OpenMP C restatement of Float/main.cu) on the same grid and medium,
    slab-CPML mode, timed on a bounded number of steps; its effective GB/s is priced
    at the same bytes per cell-step as the GPU line.  v0: the CPML velocity the
    library took from the medium."""
    import numpy as np
    from oracle import pw_oracle as O
    grid = {"nx": n, "ny": n, "nz": n, "np": np_, "h": h, "dt": dt, "rc": 1e-4, "v0": float(v0), "f0": 25.0,
            "src": (n // 2, n // 2, n // 2), "src_w": (1.0, 1.0, 1.0, -1.0), "order": order}
    base = spec_base.astype(np.float32)
    t0 = time.perf_counter()
    sec, thr = O.time_f32(grid, base, 2, steps)
    total = time.perf_counter() - t0
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cells = float(n) ** 3
    rate = cells * steps / sec / 1e9
    return {"value": round(rate, 5), "unit": "Gcell-steps/s", "cores": thr, "kind": "port", "cpu_model": model,
            "effective_gb_s": round(rate * bytes_per_cell_step, 2),
            "sample": f"oracle FP32 (OpenMP C restatement of Float/main.cu FD_V/FD_T/Source, slab CPML"
                      f"{', 4th-order differences' if order == 4 else ''}) on "
                      f"the same {n}^3 medium: {steps} timed steps after 2 warm-up steps, {sec:.2f} s timed "
                      f"({total:.1f} s with set-up)"}


def serving_trace(nshots, rate, seed, n_vol, np_):
    """Shot requests: Poisson arrivals (mean rate per second), total grid N uniform
    in {96, 128, 192, 256}, source uniform in the non-CPML interior, 500-1500
    steps, FP32 or compensated FP16 (comp16f) with equal odds, medium cropped from the interior of
    the 512^3 two-layer model at a random origin, 64 receivers x (txx, vux) on
    a line along x at iz = np + 10, iy = N / 2."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, nshots))
    t -= t[0]
    vol_i = n_vol - 2 * np_
    out = []
    for k in range(nshots):
        n = int(rng.choice([96, 128, 192, 256]))
        ni = n - 2 * np_
        crop = tuple(int(v) for v in rng.integers(0, vol_i - ni + 1, 3))
        src = tuple(int(v) for v in rng.integers(np_, n - np_, 3))
        nt = int(rng.integers(500, 1501))
        prec = "fp32" if rng.random() < 0.5 else "comp16f"
        xs = np.linspace(np_, n - 1 - np_, 64).round().astype(np.int64)
        iz, iy = np_ + 10, n // 2
        fields = ["txx"] * 64 + ["vux"] * 64
        idx = [int(x) + iy * n + iz * n * n for x in xs] * 2
        out.append(dict(t=float(t[k]), n=n, crop=crop, src=src, nt=nt, prec=prec, fields=fields, idx=idx))
    return out


def run_serving(trace, n_vol, np_, h, dt, spec, force_fp32=False, **server_kw):
    """Replays the trace against one server: each request is submitted at its
    arrival time; returns per-shot timings and the server's statistics.
    server_kw: pw.Server keyword arguments; none gives the server's defaults."""
    import threading
    import numpy as np
    import mixedprecisionfaracoustic_amd as pw
    srv = pw.Server(**server_kw)
    vol_i = n_vol - 2 * np_
    srv.set_medium_layered(spec, vol_i, vol_i, vol_i, np_)
    ids = [None] * len(trace)
    t0 = srv.clock()

    def feeder():
        for k, q in enumerate(trace):
            delay = q["t"] - (srv.clock() - t0)
            if delay > 0:
                time.sleep(delay)
            prec = "fp32" if force_fp32 else q["prec"]
            ids[k] = srv.submit(q["n"], q["n"], q["n"], np_, h, dt, q["nt"], prec, q["crop"], q["src"], 25.0,
                                q["fields"], q["idx"], weights=(1.0, 1.0, 1.0, -1.0), cpml_r=1e-4, cpml_v0=CPML_V0)

    th = threading.Thread(target=feeder)
    th.start()
    done = 0
    for k in range(len(trace)):
        while ids[k] is None:
            time.sleep(0.001)
        srv.wait(ids[k])
        done += 1
        if done % 32 == 0:
            print(f"  {done}/{len(trace)} shots done at {srv.clock() - t0:.1f} s", file=sys.stderr, flush=True)
    th.join()
    st = [srv.status(i) for i in ids]
    seis = [srv.result(i) for i in ids]
    stats = srv.stats()
    srv.close()
    lat = np.array([s["t_done"] - s["t_submit"] for s in st])
    span = max(s["t_done"] for s in st) - min(s["t_submit"] for s in st)
    work = sum(float(q["n"]) ** 3 * q["nt"] for q in trace)
    finite = all(bool(np.isfinite(x).all()) for x in seis)
    return {"shots": len(trace), "makespan_s": round(span, 3), "shots_per_s": round(len(trace) / span, 4),
            "gcell_steps_per_s": round(work / span / 1e9, 3), "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3), "mean_latency_s": round(float(lat.mean()), 3),
            "busy_s": round(stats["busy_s"], 3), "setup_s": round(stats["setup_s"], 3),
            "graph_s": round(stats["graph_s"], 3), "retire_s": round(stats["retire_s"], 3),
            "rounds": stats["rounds"], "graphs_built": stats["graphs_built"],
            "peak_active": stats["peak_active"], "peak_gb": round(stats["peak_bytes"] / 1e9, 2),
            "finite": finite}, seis


def serving_main(args):
    n_vol, np_, h, dt = 512, 20, 5.0, 0.5e-3
    spec = workload(n_vol, np_)
    trace = serving_trace(args.shots, args.rate, args.seed, n_vol, np_)
    print(f"serving: {len(trace)} shots over {trace[-1]['t']:.1f} s of arrivals", file=sys.stderr, flush=True)
    batched, seis_b = run_serving(trace, n_vol, np_, h, dt, spec, max_active=args.slots, round_steps=args.round_steps,
                                  use_graphs=not args.no_graphs, batched=True, fill_cells=args.fill_cells,
                                  persistent=args.persistent, max_wait_s=args.max_wait)
    print("  batched:", json.dumps(batched), file=sys.stderr, flush=True)
    seq, seis_s = run_serving(trace, n_vol, np_, h, dt, spec, force_fp32=True, max_active=1, round_steps=64,
                              use_graphs=False, batched=False)
    print("  sequential fp32:", json.dumps(seq), file=sys.stderr, flush=True)
    # accuracy of the mixed-precision serving run against the FP32 run, per shot
    import numpy as np
    err = []
    for q, a, b in zip(trace, seis_b, seis_s):
        if q["prec"] == "fp32":
            err.append(float(np.max(np.abs(a - b))))
    acc = comp16f_vs_fp32(trace, seis_b, seis_s)
    out = {"metric": "shot-serving throughput and p50/p99 latency vs sequential FP32 FCFS",
           "value": batched["gcell_steps_per_s"], "unit": "Gcell-steps/s", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic (Poisson trace, seeded)",
           "config": {"workload": "shot_serving_trace", "shots": args.shots, "rate_per_s": args.rate,
                      "seed": args.seed, "grids": [96, 128, 192, 256], "steps": [500, 1500],
                      "precisions": "fp32/comp16f 50/50", "slots": args.slots, "round_steps": args.round_steps,
                      "graphs": not args.no_graphs, "fill_cells": args.fill_cells,
                      "persistent": args.persistent, "max_wait_s": args.max_wait},
           "serving": batched, "sequential_fp32": seq,
           "throughput_ratio": round(batched["gcell_steps_per_s"] / seq["gcell_steps_per_s"], 3),
           "p50_ratio": round(batched["p50_latency_s"] / seq["p50_latency_s"], 3),
           "p99_ratio": round(batched["p99_latency_s"] / seq["p99_latency_s"], 3),
           "fp32_shots_identical_to_sequential": all(e == 0.0 for e in err),
           "comp16f_shots_vs_fp32": acc}
    print(json.dumps(out), flush=True)
    return 0


def precision_report(args):
    """The 512^3 two-layer run in every storage precision against the FP64 run on
    the same medium: relative L2 of every seismogram channel and of the step-1000
    txx snapshot, and each precision's Gcell-steps/s."""
    import numpy as np
    import torch
    import mixedprecisionfaracoustic_amd as pw
    n, np_, h, dt = args.n, args.np, 5.0, 0.5e-3
    nt = args.report_steps
    spec = workload(n, np_)
    fields, idx = receivers(n, np_)
    snap_step = min(1000, nt - 1)
    res = {}
    for prec in ("fp64", "fp32", "comp16", "comp16f", "bf16"):
        stream = torch.cuda.Stream()
        s = pw.Solver(n, n, n, np_, h, dt, nt, precision=prec, cpml_r=1e-4, cpml_v0=CPML_V0, stream=stream.cuda_stream,
                      space_order=args.space_order)
        s.set_medium_layered(spec)
        s.set_source(n // 2, n // 2, n // 2, 25.0, (1.0, 1.0, 1.0, -1.0))
        s.set_receivers(fields, idx)
        s.set_snapshots([snap_step], ["txx"], n // 2)
        s.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        done = 0
        while done < nt:  # in chunks, so a long run reports progress
            k = min(250, nt - done)
            s.run(k)
            done += k
            s.synchronize()
            print(f"  {prec}: {done}/{nt} steps", file=sys.stderr, flush=True)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1)
        res[prec] = {"seis": s.seismogram(), "snap": s.snapshots()[0], "gcell": float(n) ** 3 * nt / (ms / 1e3) / 1e9}
        s.close()
    ref = res["fp64"]
    chans = ("txx", "tzz", "vux", "vwx", "p")

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = {"metric": "rel-L2 vs FP64 per seismogram channel and step-1000 txx snapshot; Gcell-steps/s per precision",
           "config": {"workload": f"biot_two_layer_{n}^3" + ("_order4" if args.space_order == 4 else ""), "steps": nt,
                      "receivers": len(idx), "snapshot_step": snap_step, "space_order": args.space_order},
           "data": "synthetic (two-layer perturbed Biot medium generated on the GPU, seed 0)", "precisions": {}}
    nr = len(idx) // len(chans)
    for prec in ("fp32", "comp16", "comp16f", "bf16", "fp64"):
        r = res[prec]
        e = {c: rel(r["seis"][k * nr:(k + 1) * nr], ref["seis"][k * nr:(k + 1) * nr]) for k, c in enumerate(chans)}
        out["precisions"][prec] = {"gcell_steps_per_s": round(r["gcell"], 3),
                                   "ratio_vs_fp32": round(r["gcell"] / res["fp32"]["gcell"], 3),
                                   "seis_rel_l2": {c: float(f"{v:.3e}") for c, v in e.items()},
                                   "snapshot_rel_l2": float(f"{rel(r['snap'], ref['snap']):.3e}")}
    print(json.dumps(out), flush=True)
    return 0


def reference_scheme_bytes(prec, n, np_):
    """Compulsory bytes per cell-step of the reference program's two-kernel
    scheme (Float/main.cu:64-295, Half/main.cu:156-371) with perfect neighbour
    reuse: the velocity kernel reads 7 state fields and updates 6 (19 touches)
    and 10 medium arrays (Half: 15); the stress kernel reads 6 and updates 7 (20
    touches) and 7 medium arrays; each kernel makes 16 CPML touches (psi and
    prev, read and written) per damped (cell, axis), with f = 3 * 2 * np / n
    damped axes per cell.  Returns (velocity, stress) bytes per cell.  The
    kernels here run the same update with fewer compulsory bytes (staggered
    coefficients formed on the fly from 9 cell-centred arrays, single-variable
    CPML), so this is an equivalent-work figure, not their traffic."""
    s, c, m, mv, mt = {"fp32": (4, 4, 4, 10, 7), "fp64": (8, 8, 8, 10, 7), "comp16": (3, 2, 3, 15, 7),
                       "comp16f": (3, 4, 4, 10, 7), "bf16": (2, 2, 2, 15, 7)}[prec]
    f = 3.0 * 2.0 * np_ / n
    return 19 * s + mv * c + 16 * f * m, 20 * s + mt * c + 16 * f * m


def survey_fracs(ref_b, cells, ms_v, ms_t):
    """Each kernel's throughput priced at the reference scheme's bytes of that
    kernel (velocity, stress), as a fraction of the 8 TB/s peak."""
    return {"k_velocity": round(ref_b[0] * cells / (ms_v / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "k_stress": round(ref_b[1] * cells / (ms_t / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "bytes": {"k_velocity": round(ref_b[0], 1), "k_stress": round(ref_b[1], 1)}}


def comp16f_vs_fp32(trace, seis_srv, seis_seq):
    """The comp16f-tagged shots of a serving run against the same shots run in
    FP32 (the sequential baseline): relative L2 of each shot's seismogram."""
    import numpy as np
    e = [float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64)))
         for q, a, b in zip(trace, seis_srv, seis_seq) if q["prec"] == "comp16f" and np.linalg.norm(b) > 0]
    if not e:
        return None
    return {"shots": len(e), "median_rel_l2": float(f"{np.median(e):.3e}"), "max_rel_l2": float(f"{max(e):.3e}")}


def serving_summary(shots, rate, seed):
    """The default line's serving object: the first `shots` requests of the
    shot-serving trace (Poisson arrivals, grids of 96^3-256^3, 500-1500 steps,
    FP32 / comp16f) replayed against the server with its default configuration,
    then as sequential FP32 shots (one at a time, first come first served) on the
    same arrival times."""
    n_vol, np_, h, dt = 512, 20, 5.0, 0.5e-3
    spec = workload(n_vol, np_)
    trace = serving_trace(shots, rate, seed, n_vol, np_)
    srv, seis_srv = run_serving(trace, n_vol, np_, h, dt, spec)
    seq, seis_seq = run_serving(trace, n_vol, np_, h, dt, spec, force_fp32=True, max_active=1, round_steps=64,
                                use_graphs=False, batched=False)
    keys = ("shots_per_s", "gcell_steps_per_s", "p50_latency_s", "p99_latency_s", "mean_latency_s", "makespan_s",
            "peak_active", "finite")
    return {"trace": {"shots": shots, "rate_per_s": rate, "seed": seed, "grids": [96, 128, 192, 256],
                      "steps": [500, 1500], "precisions": "fp32/comp16f 50/50"},
            "server": {k: srv[k] for k in keys}, "sequential_fp32": {k: seq[k] for k in keys},
            "throughput_ratio": round(srv["gcell_steps_per_s"] / seq["gcell_steps_per_s"], 3),
            "p50_ratio": round(srv["p50_latency_s"] / seq["p50_latency_s"], 3),
            "p99_ratio": round(srv["p99_latency_s"] / seq["p99_latency_s"], 3),
            # "at equal accuracy": the comp16f shots against the same shots in FP32
            "comp16f_shots_vs_fp32": comp16f_vs_fp32(trace, seis_srv, seis_seq)}


def rank_share_main(args):
    """One interior rank's share of the 512^3 workload split into z-slabs: the
    step kernels at that shape, n x n x planes owned planes plus one ghost
    plane on each z face.  The z faces are open (no absorbing layer, no z
    memory variables); the ghost planes are held like edge planes, where a
    decomposed run would write its neighbours' planes between the two kernels.
    Throughput counts owned cells only."""
    import numpy as np
    import torch
    import mixedprecisionfaracoustic_amd as pw

    n, planes, np_, h, dt = args.n, args.rank_planes, args.np, 5.0, 0.5e-3
    nz = planes + 2
    nt = args.warmup + args.steps + args.prof_steps
    stream = torch.cuda.Stream()
    solver = pw.Solver(n, n, nz, np_, h, dt, nt, precision=args.precision, cpml_r=1e-4, cpml_v0=CPML_V0,
                       stream=stream.cuda_stream, cpml_open=("z_lo", "z_hi"))
    # the two-layer model with its interface mid-slab (z_top is a total-grid index)
    spec = workload(n, np_)
    spec.z_top[1] = nz // 2
    solver.set_medium_layered(spec)
    solver.set_source(n // 2, n // 2, nz // 2, 25.0, (1.0, 1.0, 1.0, -1.0))
    xs = np.linspace(np_, n - 1 - np_, 256).round().astype(np.int64)
    idx = [int(x) + (n // 2) * n + (nz // 2) * n * n for x in xs] * 2
    solver.set_receivers(["txx"] * 256 + ["vux"] * 256, idx)
    info = solver.info()
    solver.run(args.warmup)
    solver.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    solver.run(args.steps)
    ev1.record(stream)
    solver.synchronize()
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    finite = bool(np.isfinite(solver.seismogram()).all())
    solver.set_profiling(True)
    solver.run(min(args.prof_steps, nt - solver.step))
    ms_v, ms_t, launches = solver.kernel_times()
    solver.close()
    cells_total = float(n) * n * nz
    owned = float(n) * n * planes
    sec = ms_total / 1e3
    avg_v, avg_t = ms_v / launches, ms_t / launches
    if ms_v >= ms_t:
        kname, kbytes, kms = "k_velocity", info["bytes_velocity"] * cells_total, avg_v
    else:
        kname, kbytes, kms = "k_stress", info["bytes_stress"] * cells_total, avg_t
    achieved = kbytes / (kms / 1e3) / 1e9
    # halo traffic a decomposed run would add per rank and step: after each of
    # the two kernels, 2 fields per face leave and 2 arrive on each z face
    state_b = {"fp32": 4, "fp64": 8, "comp16": 3, "comp16f": 3, "bf16": 2}[args.precision]
    halo_bytes = 2 * 2 * 2 * n * n * state_b
    print(json.dumps({
        "metric": "rank share: Gcell-steps/s over owned cells", "value": round(owned * args.steps / sec / 1e9, 4),
        "unit": "Gcell-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_total / args.steps, 4), "higher_is_better": True,
        "config": {"workload": f"biot_two_layer_rank_share_{n}x{n}x{planes}+2_{args.precision}",
                   "grid": [n, n, nz], "owned_planes": planes, "cpml_xy": np_, "cpml_open": ["z_lo", "z_hi"],
                   "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": kbytes,
                     "avg_launch_ms": round(kms, 4)},
        "bytes_per_cell_step": round(info["bytes_per_cell_step"], 2),
        "kernel_ms": {"k_velocity": round(avg_v, 4), "k_stress": round(avg_t, 4)},
        "halo_bytes_per_step": halo_bytes, "finite": finite}), flush=True)
    return 0


def side_run(pw, torch, precision, n, np_, h, dt, f0, spec, steps, warmup, prof_steps, order=2, total_steps=0):
    """The same workload in another storage precision, timed the same way: its
    throughput, step time, kernel times and dominant-kernel roofline; then,
    untimed, on to total_steps for the accuracy comparison."""
    nt = max(warmup + steps + prof_steps, total_steps)
    stream = torch.cuda.Stream()
    s = pw.Solver(n, n, n, np_, h, dt, nt, precision=precision, cpml_r=1e-4, cpml_v0=CPML_V0,
                  stream=stream.cuda_stream, space_order=order)
    s.set_medium_layered(spec)
    s.set_source(n // 2, n // 2, n // 2, f0, (1.0, 1.0, 1.0, -1.0))
    fields, idx = receivers(n, np_)
    s.set_receivers(fields, idx)
    snap_steps = [k for k in range(500, nt, 500)] or [nt - 1]
    s.set_snapshots(snap_steps, ["txx"] * len(snap_steps), n // 2)
    info = s.info()
    s.run(warmup)
    s.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    s.run(steps)
    ev1.record(stream)
    s.synchronize()
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    s.set_profiling(True)
    s.run(min(prof_steps, nt - s.step))
    ms_v, ms_t, launches = s.kernel_times()
    s.set_profiling(False)
    if s.step < nt:
        s.run(nt - s.step)  # untimed
        s.synchronize()
    outputs = (s.seismogram(), s.snapshots(), s.step)
    s.close()
    cells = float(n) ** 3
    avg_v, avg_t = ms_v / launches, ms_t / launches
    kname, kbytes, kms = (("k_velocity", info["bytes_velocity"] * cells, avg_v) if ms_v >= ms_t
                          else ("k_stress", info["bytes_stress"] * cells, avg_t))
    achieved = kbytes / (kms / 1e3) / 1e9
    return {"value": round(cells * steps / (ms_total / 1e3) / 1e9, 4), "ms_per_step": round(ms_total / steps, 4),
            "kernel_ms": {"k_velocity": round(avg_v, 4), "k_stress": round(avg_t, 4)},
            "roofline": {"kernel": kname, "achieved": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4)},
            "bytes_per_cell_step": round(info["bytes_per_cell_step"], 2)}, outputs


def fp64_outputs(pw, n, np_, h, dt, f0, spec, nt, order):
    """The same run in FP64 (untimed): the reference for every line's accuracy."""
    s = pw.Solver(n, n, n, np_, h, dt, nt, precision="fp64", cpml_r=1e-4, cpml_v0=CPML_V0, space_order=order)
    s.set_medium_layered(spec)
    s.set_source(n // 2, n // 2, n // 2, f0, (1.0, 1.0, 1.0, -1.0))
    fields, idx = receivers(n, np_)
    s.set_receivers(fields, idx)
    snap_steps = [k for k in range(500, nt, 500)] or [nt - 1]
    s.set_snapshots(snap_steps, ["txx"] * len(snap_steps), n // 2)
    s.run(nt)
    out = (s.seismogram(), s.snapshots())
    s.close()
    return out


def accuracy(outputs, ref, fields):
    """Relative L2 of each seismogram channel and of the last txx snapshot
    against the FP64 run, with the FP64 seismogram's peak (before the first
    arrival at the receiver line it is rounding noise, and so is the error)."""
    import numpy as np
    seis, snaps = outputs
    rseis, rsnaps = ref

    def rel(a, b):
        nb = float(np.linalg.norm(b.astype(np.float64)))
        return float(np.linalg.norm(a.astype(np.float64) - b.astype(np.float64)) / nb) if nb > 0 else None

    chans = {}
    for ch in sorted(set(fields)):
        k = [i for i, f in enumerate(fields) if f == ch]
        e = rel(seis[k], rseis[k])
        chans[ch] = None if e is None else float(f"{e:.3g}")
    return {"seis_rel_l2": chans, "fp64_seis_peak": float(f"{float(np.abs(rseis).max()):.3g}"),
            "snapshot_txx_rel_l2": float(f"{rel(snaps[-1], rsnaps[-1]):.3g}")}


def hbm_ceiling():
    """Run the standalone bandwidth probe built next to the library (plain copy
    and 2/6/12-read streaming kernels over 256 MiB arrays) as a child process,
    and print its JSON line: the rates a kernel of the simplest shape reaches,
    against which the step kernels' rates are read."""
    import subprocess
    probe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mixedprecisionfaracoustic_amd", "pw_hbm_probe")
    if not os.path.exists(probe):
        raise SystemExit(f"{probe} is not built (python -m mixedprecisionfaracoustic_amd.build)")
    r = subprocess.run([probe], capture_output=True, text=True)
    sys.stderr.write(r.stdout + r.stderr)
    if r.returncode != 0:
        return r.returncode
    line = json.loads(r.stdout.strip().splitlines()[-1])
    print(json.dumps({"metric": "hbm_ceiling", "peak_datasheet_gb_s": 8000.0, **line}), flush=True)
    return 0


def host_job_main(args):
    """The whole job as a caller that holds its medium in host memory runs it:
    create the solver, hand over the nine cell-centred FP64 base volumes
    (pw_set_medium: upload from pageable memory plus the coefficient
    precompute), set source, receivers and snapshots, run every step, and read
    the seismograms and snapshots back. Its rate includes the PCIe transfers
    and set-up; the bench line's value does not (inputs resident in HBM)."""
    import numpy as np
    import torch
    import mixedprecisionfaracoustic_amd as pw
    n, np_, h, dt, f0 = args.n, args.np, 5.0, 0.5e-3, 25.0
    nt = args.job_steps
    spec = workload(n, np_)
    s0 = pw.Solver(n, n, n, np_, h, dt, 2, precision="fp32", cpml_r=1e-4, cpml_v0=CPML_V0)
    s0.set_medium_layered(spec)
    base = s0.base_fields()  # the caller's host buffers (not timed)
    s0.close()
    fields, idx = receivers(n, np_)
    snap_steps = [s for s in range(500, nt, 500)] or [nt - 1]
    cells = float(n) ** 3
    out = {"metric": "host-to-host job rate (PCIe and set-up included)", "unit": "Gcell-steps/s",
           "config": {"workload": f"biot_two_layer_{n}^3", "steps": nt, "receivers": len(idx),
                      "snapshots": len(snap_steps), "medium_bytes": int(base.nbytes)}}
    for prec in args.job_precisions.split(","):
        torch.cuda.synchronize()
        t = [time.perf_counter()]
        s = pw.Solver(n, n, n, np_, h, dt, nt, precision=prec, cpml_r=1e-4, cpml_v0=CPML_V0)
        t.append(time.perf_counter())
        s.set_medium(base)
        t.append(time.perf_counter())
        s.set_source(n // 2, n // 2, n // 2, f0, (1.0, 1.0, 1.0, -1.0))
        s.set_receivers(fields, idx)
        s.set_snapshots(snap_steps, ["txx"] * len(snap_steps), n // 2)
        s.run(nt)
        s.synchronize()
        t.append(time.perf_counter())
        seis = s.seismogram()
        snaps = s.snapshots()
        t.append(time.perf_counter())
        s.close()
        t.append(time.perf_counter())
        job = t[-1] - t[0]
        out[prec] = {"value": round(cells * nt / job / 1e9, 4), "job_s": round(job, 3),
                     "stepping_rate": round(cells * nt / (t[3] - t[2]) / 1e9, 4),
                     "create_s": round(t[1] - t[0], 3), "set_medium_s": round(t[2] - t[1], 3),
                     "run_s": round(t[3] - t[2], 3), "outputs_s": round(t[4] - t[3], 3),
                     "close_s": round(t[5] - t[4], 3),
                     "finite": bool(np.isfinite(seis).all() and np.isfinite(snaps).all())}
        print(f"host job {prec}: {out[prec]}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return 0


def medium_digest_main(args):
    """sha256 of each benchmark medium's nine cell-centred FP64 base fields,
    exactly as pw_get_base_fields returns them ([field][z][y][x], little-endian
    doubles): the record that a generated medium is the one measured.  The
    512^3 two-layer model (also the one the shot-serving trace crops from);
    with --large also the 1024^3 eight-layer model (64 GB of host memory)."""
    import hashlib
    import mixedprecisionfaracoustic_amd as pw
    todo = [("biot_two_layer_512", 512, 20, 5.0, 0.5e-3, workload(512, 20))]
    if args.large:
        todo.append(("biot_eight_layer_1024", 1024, 32, 4.0, 0.4e-3, large_workload(1024, 32)))
    out = {"metric": "medium_digest", "layout": "9 x [nz-2np][ny-2np][nx-2np] float64 little-endian"}
    rec_path = os.path.join(HERE, "tests", "golden", "medium_digests.json")
    recorded = json.load(open(rec_path)) if os.path.exists(rec_path) else {}
    for name, n, np_, h, dt, spec in todo:
        s = pw.Solver(n, n, n, np_, h, dt, 2, precision="fp32", cpml_r=1e-4, cpml_v0=CPML_V0)
        s.set_medium_layered(spec)
        base = s.base_fields()
        s.close()
        out[name] = {"sha256": hashlib.sha256(base).hexdigest(), "bytes": int(base.nbytes)}
        if name in recorded:
            out[name]["matches_recorded"] = out[name]["sha256"] == recorded[name]["sha256"]
        print(f"{name}: {out[name]}", file=sys.stderr, flush=True)
        del base
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--np", type=int, default=20)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "comp16", "comp16f", "bf16", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-comp16", action="store_true",
                    help="skip the compensated-FP16 run reported beside the FP32 line")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--prof-steps", type=int, default=40)
    ap.add_argument("--serving", action="store_true", help="replay the shot-serving trace instead")
    ap.add_argument("--precision-report", action="store_true",
                    help="every storage precision against FP64 on the same workload instead")
    ap.add_argument("--report-steps", type=int, default=2000)
    ap.add_argument("--shots", type=int, default=256)
    ap.add_argument("--serving-shots", type=int, default=48,
                    help="trace prefix replayed for the default line's serving object (0: none)")
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--round-steps", type=int, default=32)
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--max-wait", type=float, default=0.0,
                    help="--serving: admit the least-work shot first unless one has waited this many seconds "
                         "(0: first come first served)")
    ap.add_argument("--persistent", action="store_true",
                    help="--serving: the server runs each round as one persistent cooperative launch")
    ap.add_argument("--fill-cells", type=int, default=None,
                    help="server admission limit in grid cells (default: the library's)")
    ap.add_argument("--volume-every", type=int, default=0,
                    help="also take a full 3-D txx snapshot every K steps (asynchronous copy to the host)")
    ap.add_argument("--rank-share", action="store_true",
                    help="one interior rank's z-slab share of the workload (open z faces) instead")
    ap.add_argument("--rank-planes", type=int, default=64)
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="run the streaming-kernel bandwidth probe (the practical HBM roof) instead")
    ap.add_argument("--large", action="store_true",
                    help="the 1024^3 eight-layer model (NP 32, dx 4 m, dt 0.4 ms, 20 Hz) instead of 512^3; "
                         "no CPU baseline (its FP64 base fields alone take 77 GB of host memory)")
    ap.add_argument("--space-order", type=int, default=2, choices=[2, 4],
                    help="staggered differences of 2nd (the reference's) or 4th order")
    ap.add_argument("--medium-digest", action="store_true",
                    help="print the sha256 of the workload media (with --large, also the 1024^3 one) instead")
    ap.add_argument("--host-job", action="store_true",
                    help="time whole jobs from host buffers (set-up, medium upload, run, read-back) instead")
    ap.add_argument("--job-steps", type=int, default=2000)
    ap.add_argument("--job-precisions", default="fp32,comp16f,comp16")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        return 2
    if args.hbm_ceiling:
        return hbm_ceiling()
    if args.rank_share:
        return rank_share_main(args)
    if args.host_job:
        return host_job_main(args)
    if args.medium_digest:
        return medium_digest_main(args)
    if args.serving:
        return serving_main(args)
    if args.precision_report:
        return precision_report(args)

    import numpy as np
    import torch
    import mixedprecisionfaracoustic_amd as pw

    if args.large:
        n, np_, h, dt, f0 = 1024, 32, 4.0, 0.4e-3, 20.0
        spec = large_workload(n, np_)
        name, medium = "biot_eight_layer", "eight-layer perturbed Biot medium generated on the GPU, seed 3"
    else:
        n, np_, h, dt, f0 = args.n, args.np, 5.0, 0.5e-3, 25.0
        spec = workload(n, np_)
        name, medium = "biot_two_layer", "two-layer perturbed Biot medium generated on the GPU, seed 0"
    # the default line also checks FP32 and the compensated modes against an
    # FP64 run; those runs go on, untimed, to 1000 steps: the first arrival
    # reaches the receiver line (226 cells from the source) near step 780
    acc_steps = 1000 if (args.precision == "fp32" and not args.no_comp16 and not args.large) else 0
    nt = max(args.warmup + args.steps + args.prof_steps, acc_steps)
    stream = torch.cuda.Stream()
    solver = pw.Solver(n, n, n, np_, h, dt, nt, precision=args.precision, cpml_r=1e-4, cpml_v0=CPML_V0,
                       stream=stream.cuda_stream, space_order=args.space_order)
    solver.set_medium_layered(spec)
    solver.set_source(n // 2, n // 2, n // 2, f0, (1.0, 1.0, 1.0, -1.0))
    fields, idx = receivers(n, np_)
    solver.set_receivers(fields, idx)
    snap_steps = [s for s in range(500, nt, 500)] or [nt - 1]
    solver.set_snapshots(snap_steps, ["txx"] * len(snap_steps), n // 2)
    nt_run = args.warmup + args.steps + args.prof_steps
    vol_steps = list(range(args.volume_every - 1, nt_run, args.volume_every)) if args.volume_every > 0 else []
    if vol_steps:
        # full 3-D txx volumes, copied to pinned host memory while the run goes on
        solver.set_volume_snapshots(vol_steps, ["txx"] * len(vol_steps))
    info = solver.info()

    # warm-up (untimed)
    solver.run(args.warmup)
    solver.synchronize()
    torch.cuda.synchronize()

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    solver.run(args.steps)
    ev1.record(stream)
    solver.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    seis = solver.seismogram()

    # per-kernel split: a separate short pass with HIP events around every
    # velocity / stress launch on the solver's stream (kept out of the timed run)
    nprof = min(args.prof_steps, nt - solver.step)
    solver.set_profiling(True)
    solver.run(nprof)
    ms_v, ms_t, launches = solver.kernel_times()
    solver.set_profiling(False)
    if acc_steps and solver.step < nt:
        solver.run(nt - solver.step)  # untimed
        solver.synchronize()
    main_outputs = (solver.seismogram(), solver.snapshots(), solver.step)

    finite = bool(np.isfinite(seis).all())
    cells = float(n) ** 3
    sec = ms_total / 1e3
    gcell = cells * args.steps / sec / 1e9
    step_bytes = info["bytes_per_cell_step"] * cells
    # per step: each kernel's time over the profiled steps
    avg_v, avg_t = ms_v / launches, ms_t / launches
    # dominant kernel: the one with the larger total time
    if ms_v >= ms_t:
        kname, kbytes, kms = "k_velocity", info["bytes_velocity"] * cells, avg_v
    else:
        kname, kbytes, kms = "k_stress", info["bytes_stress"] * cells, avg_t
    achieved = kbytes / (kms / 1e3) / 1e9
    ref_b = reference_scheme_bytes(args.precision, n, np_)
    dtype = {"fp32": "f32", "comp16": "f16+i8 storage, f32 math", "bf16": "bf16 storage, f32 math",
             "comp16f": "f16+i8 state, f32 coefficients and math", "fp64": "f64"}[args.precision]
    out = {
        "metric": METRIC,
        "value": round(gcell, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_total / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": f"synthetic ({medium})",
        "config": {"workload": f"{name}_{n}^3_{args.precision}" + ("_order4" if args.space_order == 4 else ""),
                   "grid": [n, n, n], "cpml": np_, "space_order": args.space_order,
                   "dx_m": h, "dt_s": dt, "f0_hz": f0, "receivers": len(idx), "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     # profiled for the 2nd-order step only
                     "traffic": measured_traffic(args.precision, kname, n) if args.space_order == 2 else None,
                     "bytes_per_launch": kbytes, "avg_launch_ms": round(kms, 4),
                     # each kernel priced at the reference scheme's per-kernel bytes
                     "frac_survey": survey_fracs(ref_b, cells, avg_v, avg_t)},
        "step_hbm_gbs": round(step_bytes * args.steps / sec / 1e9, 1),
        "bytes_per_cell_step": round(info["bytes_per_cell_step"], 2),
        # the same throughput priced at the reference scheme's compulsory bytes
        "reference_scheme": {"bytes_per_cell_step": round(sum(ref_b), 2),
                             "step_gb_s": round(sum(ref_b) * cells * args.steps / sec / 1e9, 1),
                             "frac": round(sum(ref_b) * cells * args.steps / sec / 1e9 / HBM_PEAK_GBS, 4)},
        "kernel_ms": {"k_velocity": round(avg_v, 4), "k_stress": round(avg_t, 4)},
        "wall_s": round(wall, 3),
        "finite": finite,
    }
    if vol_steps:
        timed = [v for v in vol_steps if args.warmup <= v < args.warmup + args.steps]
        out["volume_snapshots"] = {"every": args.volume_every, "in_timed_steps": len(timed),
                                   "bytes_each": int(cells * 4)}
    solver.close()
    if args.precision == "fp32" and not args.no_comp16:
        # the metric's comparison, compensated FP16 on the same workload: comp16f
        # (the compensated state under an FP32 update; the accurate and fast mode)
        # and comp16 (the reference Half format, bit-identical to the Half oracle)
        side = {}
        for prec in ("comp16f", "comp16"):
            obj, outs = side_run(pw, torch, prec, n, np_, h, dt, f0, spec, args.steps, args.warmup, args.prof_steps,
                                 order=args.space_order, total_steps=acc_steps)
            obj["ratio_vs_fp32"] = round(obj["value"] / gcell, 4)
            obj["roofline"]["frac_survey"] = survey_fracs(reference_scheme_bytes(prec, n, np_), cells,
                                                          obj["kernel_ms"]["k_velocity"], obj["kernel_ms"]["k_stress"])
            out[prec] = obj
            side[prec] = outs
        if acc_steps and main_outputs[2] == nt and all(o[2] == nt for o in side.values()):
            # every line's accuracy against the FP64 run of the same steps
            ref = fp64_outputs(pw, n, np_, h, dt, f0, spec, nt, args.space_order)
            out["accuracy_vs_fp64"] = accuracy(main_outputs[:2], ref, fields)
            for prec, outs in side.items():
                out[prec]["accuracy_vs_fp64"] = accuracy(outs[:2], ref, fields)
            del ref
    if (args.precision == "fp32" and not args.large and args.space_order == 2 and args.n == 512
            and args.serving_shots > 0):
        # the metric's serving half, on the default line only
        out["serving"] = serving_summary(args.serving_shots, args.rate, args.seed)
    if not args.no_cpu_baseline and not args.large:
        base = None
        s2 = pw.Solver(n, n, n, np_, h, dt, 2, precision="fp32", cpml_r=1e-4, cpml_v0=CPML_V0,
                       space_order=args.space_order)
        s2.set_medium_layered(spec)
        base = s2.base_fields()
        s2.close()
        out["cpu_baseline"] = cpu_baseline(n, np_, base, dt, h, args.cpu_steps, info["bytes_per_cell_step"],
                                           order=args.space_order, v0=info["vp_max"])
        del base
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
