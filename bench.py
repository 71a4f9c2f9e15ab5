"""Benchmark of the D3Q19 phase-field bubble LBM step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: the 256^3 periodic case with two seeded non-overlapping bubbles
(R in [16, 32], sigma 0.1, tau_f 0.52, full phase relaxation), FP32
populations, synthetic initial state. A step is one full time step of the
lattice (phase recovery pass + fused collision pass). The timed region is K
steps with all state resident in HBM, bracketed by device synchronisation;
HIP events around each launch inside it give the per-kernel durations for
the roofline. Prints one JSON line.

Beside the plain line (unless --plain-only): the same case with FP16 and
BF16 population storage; the 128^3 tau_g = 0.8 case with g stored as
populations; the reference preset (64x64x256, ghost-layer boundaries) on the
fused and the staged kernels; the uneven 64-case batch (2000 steps each)
advanced together vs one case at a time; and the serving trace (Poisson
arrivals, FP16 continuous batching) vs the plain FP32 path serving one
request at a time.
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
# Stored g populations (tau_g != 1): the BGK phase collision reads the
# pre-collision g (the gathered post-stream populations) besides writing g*.
B_ALG_COLLIDE = 4 * 19 * 4 + 4        # k_collide: f read + write, g read + write, phi read: 308 B per cell
B_ALG_PHASE = 19 * 4 + 4              # k_phase: g read, phi write: 80 B per cell
# g in moment form (tau_g = 1, the benchmark case): g*_l(y) is recomputed from
# phi and u of its source cell and its normal n, itself recomputed from phi,
# instead of being stored.
B_ALG_COLLIDE_GM = 2 * 19 * 4 + 3 * 4 + 4   # f read, f write, u write, phi read: 168 B per cell
B_ALG_PHASE_GM = 4 * 4 + 4                  # phi, u read; phi write: 20 B per cell
# f in moment form as well (the 13 values collisionFluid reads, u among them):
# the collide pass reads a cell's 13 moments and phi, writes its 13 moments.
B_ALG_COLLIDE_FM = 2 * 13 * 4 + 4           # 108 B per cell
METRIC = "MLUPS at 256³ + % of HBM roofline; serving p50/p99 latency vs plain solves"


def host_cpu():
    """CPU model and core counts of this host (for the CPU baseline record)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count()}


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
    out = {"value": round(cells * n / t / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
           "sample": f"{n} steps of the same 256^3 case after 1 warm-up step, OpenMP over {threads} threads"}
    out.update(host_cpu())
    return out


def stored_g(mcb, steps, warmup, seed=0):
    """cfg1: 128^3, seeded off-centre bubble, tau_f 0.55, tau_g 0.8, FP32.
    With tau_g != 1 the phase collision relaxes towards the reference's
    equilibrium (BGK), so g is kept as 19 populations: k_phase sums the 19
    gathered g (80 B per cell), k_collide reads f and the gathered g, writes
    f* and g* (308 B per cell): 388 B moved per lattice update against the
    304 B of the population roofline."""
    case = mcb.case_preset("cfg1", seed)
    p = case.params
    cells = p.nx * p.ny * p.nz
    lat = mcb.Lattice.from_case(case, path=mcb.PATH_FUSED)
    lat.step(warmup)
    t_phase, t_collide = lat.kernel_times(steps)
    ms = lat.step_timed(steps) / steps
    lat.observables()
    s, k, finite = lat.reduce()
    tiling = lat.tiling()
    lat.close()
    mlups = cells / (ms * 1e-3) / 1e6
    ph, co = t_phase / steps, t_collide / steps
    return {"workload": "128^3 periodic D3Q19 f+g, seeded off-centre bubble R in [12,24] (seed %d), sigma 0.1, "
                        "tau_f 0.55, tau_g 0.8, FP32, g as populations" % seed,
            "mlups": round(mlups, 1), "ms_per_step": round(ms, 4), "tiling": list(tiling),
            "roofline_frac_304": round(mlups * 1e6 * B_ALG_STEP / 1e9 / HBM_PEAK_GBS, 4),
            "b_moved_per_lu": B_ALG_COLLIDE + B_ALG_PHASE,
            "moved_frac": round(mlups * 1e6 * (B_ALG_COLLIDE + B_ALG_PHASE) / 1e9 / HBM_PEAK_GBS, 4),
            "k_phase_ms": round(ph, 5), "k_phase_gbs": round(B_ALG_PHASE * cells / (ph * 1e-3) / 1e9, 1),
            "k_collide_ms": round(co, 5), "k_collide_gbs": round(B_ALG_COLLIDE * cells / (co * 1e-3) / 1e9, 1),
            "k_collide_traffic": traffic_for("cfg1-128-fp32-storedg"),
            "sum_phi": s, "finite": finite}


def reference_preset(mcb, steps, warmup):
    """The reference's own case (src/var.cu:3-4, src/init.cu:39, src/main.cu:35):
    64x64x256 with ghost-layer boundaries, tau 0.505, sigma 0.1, one R = 20
    bubble at the centre, FP32 with the literal weights. The fused path and the
    staged kernels (the reference's decomposition) give the same bits."""
    case = mcb.case_preset("reference")
    p = case.params
    cells = p.nx * p.ny * p.nz
    out = {"workload": "the reference preset: 64x64x256, ghost-layer boundaries, tau 0.505, sigma 0.1, "
                       "one R = 20 bubble at the centre, FP32"}
    for name, path in (("fused", mcb.PATH_FUSED), ("staged", mcb.PATH_STAGED)):
        lat = mcb.Lattice.from_case(case, path=path)
        lat.step(warmup)
        ms = lat.step_timed(steps) / steps
        lat.close()
        out[name] = {"mlups": round(cells / (ms * 1e-3) / 1e6, 1), "ms_per_step": round(ms, 4)}
    out["fused_vs_staged"] = round(out["fused"]["mlups"] / out["staged"]["mlups"], 3)
    return out


def f_moments(mcb, case, steps, warmup):
    """The same 256^3 case with f kept as the 13 moments collisionFluid reads
    (rho, u, Pi^neq, F) instead of 19 populations: bit-identical results,
    108 B per cell in the collide pass instead of 168."""
    cells = case.params.nx * case.params.ny * case.params.nz
    lat = mcb.Lattice.from_case(case, path=mcb.PATH_FUSED, f_form=mcb.F_MOMENTS)
    lat.step(warmup)
    t_phase, t_collide = lat.kernel_times(steps)
    ms = lat.step_timed(steps) / steps
    lat.observables()
    s, k, finite = lat.reduce()
    tiling = lat.tiling()
    lat.close()
    mlups = cells / (ms * 1e-3) / 1e6
    ph, co = t_phase / steps, t_collide / steps
    return {"mlups": round(mlups, 1), "ms_per_step": round(ms, 4), "tiling": list(tiling),
            "b_moved_per_lu": B_ALG_COLLIDE_FM + B_ALG_PHASE_GM,
            "k_phase_ms": round(ph, 5), "k_collide_ms": round(co, 5),
            "k_collide_gbs": round(B_ALG_COLLIDE_FM * cells / (co * 1e-3) / 1e9, 1),
            "k_collide_traffic": traffic_for("cfg2-256-fp32-fmoments"),
            "sum_phi": s, "finite": finite}


def storage_form(mcb, case, storage, steps, warmup):
    """MLUPS of the same 256^3 case with 16-bit population storage."""
    cells = case.params.nx * case.params.ny * case.params.nz
    lat = mcb.Lattice.from_case(case, storage=storage, path=mcb.PATH_FUSED)
    lat.step(warmup)
    ms = lat.step_timed(steps) / steps
    lat.observables()
    s, k, finite = lat.reduce()
    tiling = lat.tiling()
    lat.close()
    mlups = cells / (ms * 1e-3) / 1e6
    return {"mlups": round(mlups, 1), "ms_per_step": round(ms, 4), "tiling": list(tiling),
            "b_alg_per_lu": B_ALG_STEP // 2,
            "roofline_frac": round(mlups * 1e6 * (B_ALG_STEP // 2) / 1e9 / HBM_PEAK_GBS, 4),
            "sum_phi": s, "finite": finite}


def batching(mcb, n_cases=64, steps=2000, seed=1000):
    """The uneven batch: n_cases FP32 cases (sides 32..128) advanced together
    (one tile shape for all, so one phase and one collide launch per step,
    graph-replayed) vs the same cases advanced one at a time on single-case
    contexts with each side's autotuned tile."""
    cases = [mcb.case_preset("batch", seed + i) for i in range(n_cases)]
    sides = sorted({c.params.nx for c in cases})
    tuned = {}
    for s in sides:
        tuned[s] = list(mcb.autotune(cases[[c.params.nx for c in cases].index(s)].params)[:3])
    cells = sum(c.params.nx * c.params.ny * c.params.nz for c in cases)
    with mcb.Batch(mcb.STORE_FP32) as b:
        for c in cases:
            b.add_case(c)
        b.step(4)
        ms_batched = b.step_timed(steps)
        nbytes = b.nbytes()
    ms_single = 0.0
    for c in cases:
        lat = mcb.Lattice.from_case(c, path=mcb.PATH_FUSED)
        tx, ty, zc = tuned[c.params.nx]
        lat.set_tiling(tx, ty, zc)
        lat.step(2)
        ms_single += lat.step_timed(steps)
        lat.close()
    agg = cells * steps / (ms_batched * 1e-3) / 1e6
    seq = cells * steps / (ms_single * 1e-3) / 1e6
    return {"cases": n_cases, "cells": cells, "steps": steps, "sides": sides, "autotuned": tuned,
            "mlups_batched": round(agg, 1), "mlups_one_at_a_time": round(seq, 1),
            "speedup": round(agg / seq, 3), "device_bytes": nbytes,
            "workload": "%d 'batch' presets (seeds %d..%d): cube sides in {32,48,64,96,128}, R in [0.15,0.3] side, "
                        "sigma in [0.01,0.1], tau_f in [0.52,0.8], FP32; %d timed steps each"
                        % (n_cases, seed, seed + n_cases - 1, steps)}


def serving(mcb, n=500, step_scale=0.1, seed=0, saturated_rate=1000.0, moderate_rate=20.0):
    """The serving trace replayed through FP16 continuous batching and through
    the plain FP32 path one request at a time (single-case tile shapes
    autotuned per side, so the baseline runs at its best).

    Throughput is compared with every request arriving within the first half
    second (saturated_rate): both systems are then busy the whole time and
    completed lattice updates / makespan is each one's capacity. Time to result
    is stated at that load and at a moderate Poisson rate."""
    for side in range(48, 193, 16):
        mcb.autotune(mcb.make_params(side, side, side, tau_g=1.0, storage=mcb.STORE_FP32))
    out = {"requests": n, "step_scale": step_scale,
           "workload": "%d 'serve' presets: sides 48..192 step 16, R in [0.1,0.3] side, sigma in [0.01,0.1], "
                       "tau_f in [0.52,0.8], tau_g 1, steps log-uniform [500,5000] x %g, Poisson arrivals"
                       % (n, step_scale)}

    def run(rate, storage, seq):
        reqs = mcb.serve_trace(n, rate, step_scale, seed)
        res, st = mcb.serve(reqs, storage=storage, sequential=seq)
        return {"arrival_rate_per_s": rate, "mlups": round(st.mlups, 1), "p50_s": round(st.p50_s, 4),
                "p99_s": round(st.p99_s, 4), "mean_s": round(st.mean_s, 4), "makespan_s": round(st.makespan_s, 3),
                "completed": st.completed, "rejected": st.rejected, "peak_cases": st.peak_cases,
                "rounds": st.rounds}

    out["saturated"] = {"batched_fp16": run(saturated_rate, mcb.STORE_FP16, False),
                        "sequential_fp32": run(saturated_rate, mcb.STORE_FP32, True)}
    sat = out["saturated"]
    out["throughput_vs_sequential"] = round(sat["batched_fp16"]["mlups"] / sat["sequential_fp32"]["mlups"], 3)
    out["moderate"] = {"batched_fp16": run(moderate_rate, mcb.STORE_FP16, False),
                       "sequential_fp32": run(moderate_rate, mcb.STORE_FP32, True)}
    return out


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
    ap.add_argument("--plain-only", action="store_true",
                    help="only the plain FP32 256^3 line (no 16-bit, batching or serving sections)")
    ap.add_argument("--serve-scale", type=float, default=0.1,
                    help="step-count scale of the serving trace (1 = the full 500..5000 steps)")
    ap.add_argument("--serve-rate", type=float, default=20.0,
                    help="moderate serving arrival rate, requests per second (latency section)")
    ap.add_argument("--form16", choices=("fp16", "bf16"), default=None,
                    help="only time the 256^3 case with this 16-bit storage form (profiling)")
    ap.add_argument("--stored-g", action="store_true",
                    help="only time the 128^3 tau_g = 0.8 case (g as populations; profiling)")
    ap.add_argument("--reference-only", action="store_true",
                    help="only time the reference preset on the fused and staged kernels (profiling)")
    ap.add_argument("--f-moments", action="store_true",
                    help="only time the 256^3 case with f in moment form (profiling)")
    ap.add_argument("--batching-only", action="store_true",
                    help="only the 64-case batch section (profiling); --steps sets its step count")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only time a plain device copy of BYTES (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(whole cases or z-slabs would shard across GPUs; not built)", file=sys.stderr)
        sys.exit(2)

    import mcb
    if args.copy_probe:
        ms = mcb.bench_copy(args.copy_probe, 3)
        print(json.dumps({"copy_bytes": args.copy_probe, "ms_per_copy": ms,
                          "gbs": 2 * args.copy_probe / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.form16:
        # Only the 16-bit storage form of the 256^3 case (for profiling its kernels).
        case = mcb.case_preset("cfg2", args.seed)
        st = mcb.STORE_FP16 if args.form16 == "fp16" else mcb.STORE_BF16
        print(json.dumps({args.form16: storage_form(mcb, case, st, args.steps, args.warmup)}), flush=True)
        return
    if args.stored_g:
        print(json.dumps({"stored_g_128": stored_g(mcb, args.steps, args.warmup)}), flush=True)
        return
    if args.reference_only:
        print(json.dumps({"reference_preset": reference_preset(mcb, args.steps, args.warmup)}), flush=True)
        return
    if args.batching_only:
        print(json.dumps({"batching": batching(mcb, steps=args.steps)}), flush=True)
        return
    if args.f_moments:
        print(json.dumps({"f_moments": f_moments(mcb, mcb.case_preset("cfg2", args.seed), args.steps, args.warmup)}),
              flush=True)
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
    moment_form = p.tau_g == 1.0
    b_collide = B_ALG_COLLIDE_GM if moment_form else B_ALG_COLLIDE
    b_phase = B_ALG_PHASE_GM if moment_form else B_ALG_PHASE
    achieved = b_collide * cells / (avg_collide_ms * 1e-3) / 1e9
    phase_gbs = b_phase * cells / (avg_phase_ms * 1e-3) / 1e9

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
                        "tau_f 0.52, tau_g 1, FP32 f populations, g in moment form" % args.seed,
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
            "alg_bytes_per_launch": b_collide * cells,
            "avg_launch_ms": round(avg_collide_ms, 5),
        },
        "step_roofline": {
            "b_alg_per_lu": B_ALG_STEP,
            "achieved_gbs": round(mlups * 1e6 * B_ALG_STEP / 1e9, 1),
            "frac": round(mlups * 1e6 * B_ALG_STEP / 1e9 / HBM_PEAK_GBS, 4),
            "g_storage": "moments (phi, u; n recomputed)" if moment_form else "populations",
            "b_moved_per_lu": b_collide + b_phase,
            "moved_frac": round(mlups * 1e6 * (b_collide + b_phase) / 1e9 / HBM_PEAK_GBS, 4),
            "k_phase_ms": round(avg_phase_ms, 5),
            "k_phase_gbs": round(phase_gbs, 1),
        },
        "diagnostics": {"sum_phi": sum_phi, "kinetic": kinetic, "finite": finite},
    }
    if not args.plain_only:
        for name, st in (("fp16_storage", mcb.STORE_FP16), ("bf16_storage", mcb.STORE_BF16)):
            out[name] = storage_form(mcb, case, st, args.steps, args.warmup)
            out[name]["vs_plain"] = round(out[name]["mlups"] / out["value"], 3)
        out["f_moments"] = f_moments(mcb, case, args.steps, args.warmup)
        out["f_moments"]["vs_plain"] = round(out["f_moments"]["mlups"] / out["value"], 3)
        out["stored_g_128"] = stored_g(mcb, args.steps, args.warmup)
        out["reference_preset"] = reference_preset(mcb, args.steps, args.warmup)
        out["batching"] = batching(mcb)
        out["serving"] = serving(mcb, step_scale=args.serve_scale, moderate_rate=args.serve_rate)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
