"""Benchmark of the Lee-Liu D3Q19 step on one MI355X.

One step = one full time step of the hot path (moments, c stencil, mu
stencil, fused MRT collision with gather streaming, FP64 mass correction) over
the 300^3 random-sphere pack at porosity 0.37 (about 1e7 fluid sites), FP32
distributions with FP64 reductions -- the configuration the headline MLUPS is
quoted on.  MLUPS = fluid sites x steps / wall seconds, as the reference's
driver reports it (src/felbm_gpu_main.cu:785-788).

Prints one JSON line (rank 0).  The CPU baseline beside it is the FP64 OpenMP
oracle (a restatement of the reference step; the reference's own CPU engine
is not available) timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic values per site per launch (FP32 -> x4 bytes): moments reads h and g
# (38) and writes c_raw, p0, j (5); the c pass reads c_raw, p0, j (5) and writes
# c, mu, p* (3); the mu pass reads mu (1) and packs c, p*, lapmu into one record
# per site (+2 read, +3 written); the collision reads g (19), the own record (3)
# and mu (1) and writes h and g (38).  Neighbour values come from cache-resident
# lines and are not counted.
ALG_VALUES = {"moments": 43, "cfield": 8, "mufield": 6, "collide": 61}
B_ALG_PER_LU = 452  # compulsory bytes per lattice update, FP32 (113 values)


def _profile_traffic(workload):
    """HBM bytes per collide launch from the committed rocprof counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("collide_hbm_bytes_per_launch")
    return None


def cpu_baseline(workload, w, steps):
    from oracle.pyoracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle(w["solid"], w["params"], use_mrt=w["switches"]["use_mrt"], fuse_collision=True,
               stream_inplace=False, mass_mode=1 if w["switches"]["correct_op_mass"] else 0)
    o.init_rest(w["c0"], w["p0"])
    o.record_target_mass()
    o.run(1)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    n = o.n
    del o
    return {"value": n * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "%s FP64 OpenMP oracle, %d steps x %d sites, %.1f s" % (workload, steps, n, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="porous_300")
    ap.add_argument("--cpu-steps", type=int, default=int(os.environ.get("FELBM_CPU_STEPS", "40")),
                    help="steps of the FP64 OpenMP oracle on the same lattice (a bounded sample, ~30 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-steps", type=int, default=1000,
                    help="steps of the 48-instance batched sweep (0 skips it)")
    ap.add_argument("--serve", type=int, default=16, help="requests of the served stream (0 skips it)")
    ap.add_argument("--serve-max-sites", type=float, default=6e6,
                    help="lattice sites in flight the server admits up to (0: the HBM budget alone)")
    ap.add_argument("--err-steps", type=int, default=10000,
                    help="steps of the FP32-vs-FP64 field comparison on the same lattice, the 10,000 steps of "
                         "the 300^3 throughput case by default (0 skips it)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep 256 sites per workgroup in every pass instead of the autotuned shapes")
    ap.add_argument("--probe-copy-mb", type=int, default=0,
                    help="only run the device-to-device copy probe of this size (counter calibration)")
    ap.add_argument("--probe-stream-mb", type=int, default=0,
                    help="only run the hand-written streaming probes (read, write, copy) of this size")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; runs across several GPUs are not supported "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from felbm_gpu_amd import MultiPhaseGPU
    from felbm_gpu_amd import workloads as W
    import ctypes as C
    from felbm_gpu_amd import _lib

    def copy_probe(mb, reps):
        gbs = C.c_double()
        _lib.check(_lib.load().felbm_copy_probe(0, int(mb) << 20, int(reps), C.byref(gbs)))
        return gbs.value

    def stream_probe(mb, reps):
        out = {}
        for mode, name in ((0, "read"), (1, "write"), (2, "copy")):
            gbs = C.c_double()
            _lib.check(_lib.load().felbm_stream_probe(0, int(mb) << 20, int(reps), mode, C.byref(gbs)))
            out[name] = gbs.value
        return out

    if args.probe_copy_mb:
        print(json.dumps({"copy_probe_mb": args.probe_copy_mb, "GBs": copy_probe(args.probe_copy_mb, 5)}))
        return
    if args.probe_stream_mb:
        print(json.dumps({"stream_probe_mb": args.probe_stream_mb, "GBs": stream_probe(args.probe_stream_mb, 5)}))
        return

    t0 = time.time()
    w = W.workload(args.workload)
    print("[bench] workload %s built in %.1fs" % (args.workload, time.time() - t0), file=sys.stderr, flush=True)
    eng = MultiPhaseGPU(precision=w["precision"]).init(w["solid"], w["params"], w["switches"])
    eng.init_rest(w["c0"], w["p0"])
    eng.record_target_mass()
    n = eng.n
    print("[bench] sites %d, precision %s" % (n, w["precision"]), file=sys.stderr, flush=True)

    tune = None
    if not args.no_autotune:
        # workgroup shape per pass, timed on this lattice; the state is left as it was
        tune = eng.batch.autotune(10)
        print("[bench] autotuned sites per workgroup %s" % json.dumps(tune), file=sys.stderr, flush=True)
    eng.step(args.warmup)
    eng.batch.synchronize()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.batch.synchronize()
    dt = time.perf_counter() - t0
    mlups = n * args.steps / dt / 1e6
    ms = eng.batch.kernel_times(max(5, min(50, args.steps)))
    names = ["moments", "cfield", "mufield", "collide"]
    bytes_per_val = 4 if w["precision"] == "f32" else 8
    kern = {}
    for k, t in zip(names, ms):
        alg = ALG_VALUES[k] * bytes_per_val * n
        kern[k] = {"ms": t, "alg_GBs": alg / (t * 1e-3) / 1e9 if t > 0 else None}
    coll = kern["collide"]
    copy_gbs = copy_probe(1024, 10)
    stream_gbs = stream_probe(1024, 10)
    roof = {"bound": "hbm", "achieved": coll["alg_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": coll["alg_GBs"] / HBM_PEAK_GBS, "traffic": _profile_traffic(args.workload),
            "kernel": "k_collide (fused MRT collision, gather streaming)",
            "alg_bytes_per_launch": ALG_VALUES["collide"] * bytes_per_val * n,
            "copy_probe_GBs": copy_gbs, "stream_probe_GBs": stream_gbs}
    mass_drift = abs(eng.current_mass() - eng.get_target_mass())

    # the same lattice carrying 12,000 tracers on the two slab interfaces, six
    # diffusivities x 2000, each plane carrying every diffusivity
    tr_mlups = None
    if args.workload == "porous_300":
        edge = w["solid"].shape[2]
        pos, _ = W.seed_planes(w["solid"], 6000, (0.0, edge / 2.0), axis=0, seed=12)
        ladder = "1e-4:1000, 3.98e-4:1000, 1.58e-3:1000, 6.31e-3:1000, 2.51e-2:1000, 1e-1:1000"
        dm = np.concatenate([W.dm_groups(ladder), W.dm_groups(ladder)])
        eng.tracers_init(pos, dm, seed=3, substeps=1, velocity_skip=1)
        eng.step(5)
        eng.batch.synchronize()
        tr_steps = max(args.steps, 200)
        t1 = time.perf_counter()
        eng.step(tr_steps)
        eng.batch.synchronize()
        tr_mlups = n * tr_steps / (time.perf_counter() - t1) / 1e6
        eng.tracers_init(np.zeros((0, 3)), np.zeros(0))
        print("[bench] with 12,000 tracers: %.1f MLUPS" % tr_mlups, file=sys.stderr, flush=True)
    print("[bench] %.1f MLUPS, kernels %s" % (mlups, json.dumps(kern)), file=sys.stderr, flush=True)

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    # FP32 error at full size: the FP32 engine against the FP64 engine (bit-identical
    # to the FP64 oracle wherever the two are compared) from the same initial state,
    # max |difference| of each downloaded field after err_steps steps
    f32_err = None
    if args.err_steps > 0 and w["precision"] == "f32":
        eng.free()
        fields, rates = {}, {}
        for prec in ("f64", "f32"):
            e2 = MultiPhaseGPU(precision=prec).init(w["solid"], w["params"], w["switches"])
            e2.init_rest(w["c0"], w["p0"])
            e2.record_target_mass()
            e2.batch.synchronize()
            t_e = time.perf_counter()
            for _ in range(args.err_steps // 1000):     # progress between chunks
                e2.step(1000)
                e2.batch.synchronize()
            e2.step(args.err_steps % 1000)
            e2.batch.synchronize()
            rates[prec] = e2.n * args.err_steps / (time.perf_counter() - t_e) / 1e6
            fields[prec] = e2.download()
            e2.free()
        f32_err = {"steps": args.err_steps, "MLUPS_over_run": rates}
        for k in ("rho", "c", "ux", "uy", "uz", "p"):
            f32_err[k] = float(np.abs(fields["f32"][k] - fields["f64"][k]).max())
        for k in ("c", "ux"):
            d = np.abs(fields["f32"][k] - fields["f64"][k])
            f32_err[k + "_p99"], f32_err[k + "_median"] = float(np.percentile(d, 99)), float(np.median(d))
        del fields
        log("fp32 vs fp64 after %d steps: %s" % (args.err_steps, json.dumps(f32_err)))

    sweep = None
    if args.sweep_steps > 0:
        eng.free()
        from felbm_gpu_amd import serving
        sweep = serving.run_sweep(W.sweep_instances(48, 1000), args.sweep_steps, "f32", log=log, check_fp64=True)
        log("sweep: %s" % json.dumps(sweep))
        # the same sweep with every instance at edge 64 (about 1e5 sites each):
        # the regime where one launch per pass over the whole batch matters
        small = serving.run_sweep(W.sweep_instances(48, 1000, edges=(64,)), args.sweep_steps, "f32", log=log)
        sweep["small_64"] = {k: small[k] for k in ("instances", "sites", "steps", "batched_MLUPS", "sequential_MLUPS",
                                                  "speedup", "retired_nonfinite")}
        log("sweep 48 x 64^3: %s" % json.dumps(sweep["small_64"]))
    served = None
    if args.serve > 0:
        eng.free()
        from felbm_gpu_amd import serving
        cap = int(args.serve_max_sites) or None
        res = serving.run_serve(args.serve, rate=4.0, seed=2024, precision="f32", log=log, check_fp64_steps=2000,
                                max_sites=cap)
        served = {k: v for k, v in res.items() if k != "done"}
        served["requests_steps"] = [d["steps"] for d in res["done"]]
        log("serve: %s" % json.dumps(served))
        if cap is not None:
            # the same stream admitted by the HBM budget alone, for comparison
            alt = serving.run_serve(args.serve, rate=4.0, seed=2024, precision="f32", log=log)
            served["hbm_budget_only"] = {k: alt[k] for k in ("p50_latency_s", "p99_latency_s", "aggregate_MLUPS",
                                                             "makespan_s", "converged", "step_cap")}
            log("serve, HBM budget alone: %s" % json.dumps(served["hbm_budget_only"]))

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (%d steps)" % args.cpu_steps, file=sys.stderr, flush=True)
        eng.free()
        cpu = cpu_baseline(args.workload, w, args.cpu_steps)
        # the GPU rates on the same lattice beside it: FP64 (like for like with
        # the FP64 oracle) over the FP32-vs-FP64 run, FP32 the headline
        cpu["gpu_same_lattice_MLUPS"] = {"f32": mlups, "f64": f32_err["MLUPS_over_run"]["f64"] if f32_err else None}

    out = {
        "metric": "MLUPS (300^3 fused MRT) + batched-sweep MLUPS and p50/p99 solve latency",
        "value": mlups, "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if w["precision"] == "f32" else "f64",
        "data": "synthetic (seeded random-sphere pack, slab initial condition)",
        "config": {"workload": args.workload, "sites": n, "collision": "MRT" if w["switches"]["use_mrt"] else "BGK",
                   "correct_op_mass": bool(w["switches"]["correct_op_mass"]),
                   "hbm_frac_total": mlups * 1e6 * B_ALG_PER_LU / (HBM_PEAK_GBS * 1e9)},
        "kernels_ms": {k: v["ms"] for k, v in kern.items()},
        "autotune": tune,
        "mass_drift": mass_drift,
        "with_tracers_MLUPS": tr_mlups,
        "fp32_max_abs_diff_vs_fp64": f32_err,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_sweep": sweep,
        "served": served,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
