"""Benchmark: Gcell-steps/s of the Biot staggered-grid step on one MI355X.

Workload (default): 512^3 total grid, dx = 5 m, dt = 0.5 ms, 20-cell CPML
(R = 1e-4, v0 = 4000 m/s), two layers split at z = 256 (upper: water-saturated
sandstone Ks 35 GPa, Kb 7, mu 9, Kf 2.25, rho_s 2650, rho_f 1000, phi 0.2,
kappa 1e-12, eta 1e-3, tortuosity 2; lower: Kb 15, mu 14, phi 0.1, kappa 1e-14),
Kb/mu/phi perturbed by 5 % smoothed Gaussian noise (sigma 8 cells, seed 0),
25 Hz Ricker at the centre into txx/tyy/tzz/p, 256 receivers x 5 channels on a
horizontal line 10 cells below the top CPML, txx snapshots every 500 steps.
Synthetic medium generated on the GPU (no data files).

A "step" is one pass of the hot path: source -> velocity kernel -> stress
kernel -> receiver gather.  Inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0).  --gpus > 1 is refused: this build targets one GPU.
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


def workload(n, np_):
    import mixedprecisionfaracoustic_amd as pw
    upper = pw.rock(Ks=35e9, Kb=7e9, mu=9e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.2, kappa=1e-12,
                    eta=1e-3, tortuosity=2.0, pore_m=8.0)
    lower = pw.rock(Ks=35e9, Kb=15e9, mu=14e9, Kf=2.25e9, rhos=2650.0, rhof=1000.0, phi=0.1, kappa=1e-14,
                    eta=1e-3, tortuosity=2.0, pore_m=8.0)
    return pw.layered([0, n // 2], [upper, lower], perturb=0.05, corr_cells=8.0, clamp=3.0, seed=0)


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


def cpu_baseline(n, np_, spec_base, dt, h, steps):
    """The oracle (OpenMP C restatement of Float/main.cu) on the same grid and medium,
    slab-CPML mode, timed on a bounded number of steps."""
    import numpy as np
    from oracle import pw_oracle as O
    grid = {"nx": n, "ny": n, "nz": n, "np": np_, "h": h, "dt": dt, "rc": 1e-4, "v0": 4000.0, "f0": 25.0,
            "src": (n // 2, n // 2, n // 2), "src_w": (1.0, 1.0, 1.0, -1.0)}
    base = spec_base.astype(np.float32)
    t0 = time.perf_counter()
    sec, thr = O.time_f32(grid, base, 1, steps)
    total = time.perf_counter() - t0
    cells = float(n) ** 3
    return {"value": round(cells * steps / sec / 1e9, 5), "unit": "Gcell-steps/s", "cores": thr, "kind": "port",
            "sample": f"oracle FP32 (OpenMP C restatement of Float/main.cu FD_V/FD_T/Source, slab CPML) on "
                      f"the same {n}^3 medium: {steps} timed steps after 1 warm-up step, {sec:.2f} s timed "
                      f"({total:.1f} s with set-up)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--np", type=int, default=20)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "comp16", "bf16", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--prof-steps", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch
    import mixedprecisionfaracoustic_amd as pw

    n, np_, h, dt = args.n, args.np, 5.0, 0.5e-3
    nt = args.warmup + args.steps + args.prof_steps
    stream = torch.cuda.Stream()
    solver = pw.Solver(n, n, n, np_, h, dt, nt, precision=args.precision, cpml_r=1e-4, cpml_v0=4000.0,
                       stream=stream.cuda_stream)
    spec = workload(n, np_)
    solver.set_medium_layered(spec)
    solver.set_source(n // 2, n // 2, n // 2, 25.0, (1.0, 1.0, 1.0, -1.0))
    fields, idx = receivers(n, np_)
    solver.set_receivers(fields, idx)
    snap_steps = [s for s in range(500, nt, 500)] or [nt - 1]
    solver.set_snapshots(snap_steps, ["txx"] * len(snap_steps), n // 2)
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

    finite = bool(np.isfinite(seis).all())
    cells = float(n) ** 3
    sec = ms_total / 1e3
    gcell = cells * args.steps / sec / 1e9
    step_bytes = info["bytes_per_cell_step"] * cells
    # per step: the velocity (stress) kernel's time summed over its launches of
    # that step (one launch per step, or one per z-slab with PW_SLAB)
    avg_v, avg_t = ms_v / launches, ms_t / launches
    # dominant kernel: the one with the larger total time
    if ms_v >= ms_t:
        kname, kbytes, kms = "k_velocity", info["bytes_velocity"] * cells, avg_v
    else:
        kname, kbytes, kms = "k_stress", info["bytes_stress"] * cells, avg_t
    achieved = kbytes / (kms / 1e3) / 1e9
    dtype = {"fp32": "f32", "comp16": "f16+i8 storage, f32 math", "bf16": "bf16 storage, f32 math",
             "fp64": "f64"}[args.precision]
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
        "data": "synthetic (two-layer perturbed Biot medium generated on the GPU, seed 0)",
        "config": {"workload": f"biot_two_layer_{n}^3_{args.precision}", "grid": [n, n, n], "cpml": np_,
                   "dx_m": h, "dt_s": dt, "f0_hz": 25.0, "receivers": len(idx), "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": kbytes, "avg_launch_ms": round(kms, 4)},
        "step_hbm_gbs": round(step_bytes * args.steps / sec / 1e9, 1),
        "bytes_per_cell_step": round(info["bytes_per_cell_step"], 2),
        "kernel_ms": {"k_velocity": round(avg_v, 4), "k_stress": round(avg_t, 4)},
        "wall_s": round(wall, 3),
        "finite": finite,
    }
    solver.close()
    if not args.no_cpu_baseline:
        base = None
        s2 = pw.Solver(n, n, n, np_, h, dt, 2, precision="fp32", cpml_r=1e-4, cpml_v0=4000.0)
        s2.set_medium_layered(spec)
        base = s2.base_fields()
        s2.close()
        out["cpu_baseline"] = cpu_baseline(n, np_, base, dt, h, args.cpu_steps)
        del base
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
