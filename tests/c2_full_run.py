"""The 150^3 porous FP64 case over its full 5000 steps: the HIP engine against
the FP64 oracle (CPU, OpenMP), distributions and fields compared at the end.

Not collected by pytest (the oracle needs several minutes on 16 host cores);
run it on an MI355X host as

    python -u tests/c2_full_run.py [steps] [mass_mode]

mass_mode 2 (default) is the oracle with the correction folded into the next
step as the engine folds it (bit-identical); mass_mode 1 is the reference's
per-population injection after every step (include/felbm_gpu/
device_engine.cuh:1128-1139), which the engine matches to rounding.  It
prints progress every 250 steps and exits non-zero if the engine's state
differs from the oracle's by more than 1e-13 (the FP64 bar) in any value."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from felbm_gpu_amd import MultiPhaseGPU  # noqa: E402
from felbm_gpu_amd import workloads as W  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 5000
    mode = int(sys.argv[2]) if len(sys.argv) > 2 else 2
    w = W.workload("porous_150")
    o = Oracle(w["solid"], w["params"], use_mrt=True, fuse_collision=True, mass_mode=mode)
    g = MultiPhaseGPU("f64").init(w["solid"], w["params"], {"use_mrt": 1, "correct_op_mass": 1})
    o.init_rest(w["c0"], 1.0)
    h, gg = o.download_state()
    g.upload_state(h, gg)
    o.record_target_mass()
    g.record_target_mass()
    print("150^3 porous FP64, %d sites, %d steps, oracle mass_mode %d, threads %s"
          % (g.n, steps, mode, os.environ.get("OMP_NUM_THREADS", "all")), flush=True)
    t0 = time.time()
    done = 0
    while done < steps:
        k = min(250, steps - done)
        o.run(k)
        g.step(k)
        done += k
        print("step %d, %.0f s" % (done, time.time() - t0), flush=True)
    ho, go = o.download_state()
    hg, gg = g.download_state()
    dh, dg = float(np.abs(ho - hg).max()), float(np.abs(go - gg).max())
    fo, fg = o.download(), g.download()
    df = {k: float(np.abs(fo[k] - fg[k]).max()) for k in ("rho", "c", "ux", "uy", "uz", "p")}
    dm = abs(o.current_mass() - g.current_mass())
    print("max|dh| %.3e max|dg| %.3e fields %s |dM| %.3e" % (dh, dg, df, dm), flush=True)
    ok = dh <= 1e-13 and dg <= 1e-13 and max(df.values()) <= 1e-13
    print("PASS" if ok else "FAIL", flush=True)
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main()
