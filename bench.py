"""Benchmark of the D3Q19 phase-field bubble LBM step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: the 256^3 periodic case with two seeded non-overlapping bubbles
(R in [16, 32], sigma 0.1, tau_f 0.52, full phase relaxation), FP32
populations, synthetic initial state. A step is one full time step of the
lattice (phase recovery pass + fused collision pass). The timed region is K
steps with all state resident in HBM, bracketed by device synchronisation;
HIP events around each launch inside it give the per-kernel durations for
the roofline. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multic-bubble-cuda_amd"))

HBM_PEAK_GBS = 8000.0                 # MI355X HBM3E peak
B_ALG_STEP = 2 * 19 * 2 * 4           # f and g, 19 pops, read + write, FP32: 304 B per lattice update
B_ALG_COLLIDE = 3 * 19 * 4            # k_collide: f read, f write, g write: 228 B per cell
B_ALG_PHASE = 19 * 4                  # k_phase: g read: 76 B per cell
METRIC = "MLUPS at 256³ + % of HBM roofline; serving p50/p99 latency vs plain solves"


def cpu_baseline(case, seconds_target=15.0):
    """OpenMP oracle (the transcription) on this host: a few steps of the same case."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    import mcb
    p = case.params
    o = oracle.Oracle(p.nx, p.ny, p.nz, prec=32, bc_mode=p.bc_mode, weights=p.weight_set,
                      tau_f=float(p.tau_f), tau_g=float(p.tau_g), sigma=float(p.sigma))
    o.init([tuple(float(np.float32(v)) for v in b) for b in mcb.case_bubbles(case)], case.width)
    cells = p.nx * p.ny * p.nz
    t1 = o.time_steps(1)               # warm-up step, also sizes the sample
    n = max(1, min(50, int(seconds_target / max(t1, 1e-3))))
    t = o.time_steps(n)
    threads = oracle.max_threads()
    o.close()
    return {"value": round(cells * n / t / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the same 256^3 case after 1 warm-up step, OpenMP over {threads} threads"}


def traffic_for(workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh)
    return data.get(workload)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only time a plain device copy of BYTES (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(whole cases or z-slabs would shard across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import mcb
    if args.copy_probe:
        ms = mcb.bench_copy(args.copy_probe, 3)
        print(json.dumps({"copy_bytes": args.copy_probe, "ms_per_copy": ms,
                          "gbs": 2 * args.copy_probe / (ms * 1e-3) / 1e9}), flush=True)
        return
    case = mcb.case_preset("cfg2", args.seed)
    p = case.params
    cells = p.nx * p.ny * p.nz
    lat = mcb.Lattice.from_case(case)
    tiling = lat.tiling()

    lat.step(args.warmup)
    mcb.lib().mcb_device_synchronize()
    t0 = time.perf_counter()
    t_phase, t_collide = lat.kernel_times(args.steps)   # the timed steps, events around each launch
    mcb.lib().mcb_device_synchronize()
    wall = time.perf_counter() - t0

    ms_per_step = wall * 1e3 / args.steps
    mlups = cells * args.steps / wall / 1e6
    avg_collide_ms = t_collide / args.steps
    avg_phase_ms = t_phase / args.steps
    achieved = B_ALG_COLLIDE * cells / (avg_collide_ms * 1e-3) / 1e9
    phase_gbs = B_ALG_PHASE * cells / (avg_phase_ms * 1e-3) / 1e9

    lat.compute_macros()
    sum_phi, kinetic, finite = lat.reduce()
    lat.close()

    workload = "cfg2-256-fp32"
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "256^3 periodic D3Q19 f+g, two seeded bubbles R in [16,32] (seed %d), sigma 0.1, "
                        "tau_f 0.52, tau_g 1, FP32 populations" % args.seed,
            "lattice": [p.nx, p.ny, p.nz],
            "storage": "fp32",
            "path": "fused (k_phase + k_collide)",
            "tiling": list(tiling),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_collide",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_for(workload),
            "alg_bytes_per_launch": B_ALG_COLLIDE * cells,
            "avg_launch_ms": round(avg_collide_ms, 5),
        },
        "step_roofline": {
            "b_alg_per_lu": B_ALG_STEP,
            "achieved_gbs": round(mlups * 1e6 * B_ALG_STEP / 1e9, 1),
            "frac": round(mlups * 1e6 * B_ALG_STEP / 1e9 / HBM_PEAK_GBS, 4),
            "k_phase_ms": round(avg_phase_ms, 5),
            "k_phase_gbs": round(phase_gbs, 1),
        },
        "diagnostics": {"sum_phi": sum_phi, "kinetic": kinetic, "finite": finite},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
