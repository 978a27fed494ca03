"""Many lattices on one GPU: the batched parameter sweep and the served
request stream.

sweep     N independent instances (edge, porosity, contact angle, mobility
          drawn per instance) advanced together -- one launch per pass serves
          all of them, the workgroup shape of each pass autotuned on the batch,
          the step loop captured in a HIP graph -- against the same instances
          run one after another through a single-instance engine.
          Batched MLUPS = sum_i n_i steps / wall seconds.

serve     a stream of solve requests with Poisson arrivals.  Each request is an
          open-boundary channel (inlet/outlet along y, two cleared layers and
          buffer nodes at each end, copy_to_buffers, no mass correction), run
          until the mean order parameter changes by less than `tol` (relative)
          over one check interval or `max_steps` are done.  The server admits
          arrived requests while the instances fit the HBM budget (and the
          optional budget of lattice sites in flight), advances all
          active instances together `admit_every` steps at a time, checks each
          request every `check_every` of its own steps, and retires converged
          ones.  Latency = completion - arrival; aggregate MLUPS =
          sum_i n_i steps_i / makespan.

The scheduler (Server) talks to an engine through four calls -- admit, step,
mean_c, retire -- so its policy is tested on the CPU with a stand-in engine.
"""
import math
import time

import numpy as np

from . import workloads as W
from ._lib import FELBM_ENONFINITE, FelbmError


def instance_bytes(n_sites, precision, n_open=0):
    """Device bytes of one instance: h, g0, g1 (19 each), 13 fields, coords,
    and the row words and accumulators (rounded up); with open boundaries also
    the per-site entry codes (4 B per site) and 8 values per boundary node
    (engine_impl.hpp set_open_bnd)."""
    R = 8 if precision in ("f64", 8) else 4
    ob = (4 * n_sites + 8 * R * max(1, n_open) + 512) if n_open else 0
    return int(n_sites * (70 * R + 4) + ob + (1 << 20))


# ---- parameter sweep ----------------------------------------------------------------

def sweep_problem(spec):
    edge = int(spec["edge"])
    solid = W.spheres(edge, edge, edge, 5.0, 9.0, spec["porosity"], spec["seed"])
    prm = W.derive_params(sigma=0.01, width=4.0, mobility_coeff=spec["mobility_coeff"], theta_deg=spec["theta"],
                          acceleration=(1e-6, 0.0, 0.0))
    return solid, prm, W.slab_c(solid, 0, 0, edge // 2, 4.0)


def _batch_fields(problems, precision, steps, graph_chunk):
    """Fields of every instance after 2 + steps batched steps (graph replay)."""
    from .engine import Batch, MultiPhaseGPU
    sw = {"use_mrt": 1, "correct_op_mass": 1}
    b = Batch(precision=precision)
    insts = []
    for solid, prm, c0 in problems:
        g = MultiPhaseGPU(batch=b).init(solid, prm, sw)
        g.init_rest(c0, 1.0)
        g.record_target_mass()
        insts.append(g)
    b.step(2)
    b.step_graph(graph_chunk, steps // graph_chunk)
    b.synchronize()
    out = [g.download() for g in insts]
    for g in insts:
        g.free()
    b.free()
    return out


def run_sweep(specs, steps, precision="f32", graph_chunk=50, log=None, check_fp64=False):
    """Batched vs sequential MLUPS over the same instances (MRT, mass
    correction on, FP32 distributions with FP64 sums).  check_fp64 also runs
    the batch in FP64 and reports the largest FP32 - FP64 difference of c and
    u over all instances after the same steps."""
    from .engine import Batch, MultiPhaseGPU
    sw = {"use_mrt": 1, "correct_op_mass": 1}
    t_gen = time.time()
    problems = [sweep_problem(s) for s in specs]
    if log:
        log("sweep: %d instances generated in %.1fs" % (len(problems), time.time() - t_gen))
    steps = max(graph_chunk, (int(steps) // graph_chunk) * graph_chunk)

    # batched: one handle, every instance in the same launches, graph replay
    b = Batch(precision=precision)
    insts = []
    for solid, prm, c0 in problems:
        g = MultiPhaseGPU(batch=b).init(solid, prm, sw)
        g.init_rest(c0, 1.0)
        g.record_target_mass()
        insts.append(g)
    total = sum(g.n for g in insts)
    # workgroup shape per pass, timed on this batch (the state is left as it was)
    tune = b.autotune(10)
    if log:
        log("sweep: autotuned sites per workgroup %s" % tune["block_sizes"])
    b.step(2)
    b.synchronize()
    # each check interval (2 graph chunks): graph replays, then two eager
    # steps the last of which fills the exact monitoring sums, copied back for
    # every instance at once; an instance whose sums report non-finite fields
    # is retired (its parameter point left the method's stable range) and the
    # batch goes on without it
    check = 2 * graph_chunk
    steps_of = [0] * len(insts)
    live = list(range(len(insts)))
    nonfinite = []
    t0 = time.perf_counter()
    done = 0
    while done < steps and live:
        k = min(check, steps - done)
        if k > 2:
            b.step_graph(k - 2, 1)
        b.step(2, stats_last=True)
        done += k
        for j in live:
            steps_of[j] += k
        insts[live[0]].flow_stats_launch()
        for j in list(live):
            try:
                insts[j].flow_stats_collect()
            except FelbmError as err:
                if err.code != FELBM_ENONFINITE:
                    raise
                insts[j].free()
                live.remove(j)
                nonfinite.append(j)
    b.synchronize()
    t_batched = time.perf_counter() - t0
    for j in live:
        insts[j].free()
    b.free()
    lu = float(sum(insts[j].n * steps_of[j] for j in range(len(insts))))
    if log:
        log("sweep: batched %d sites x %d steps in %.3fs, %d instances retired non-finite"
            % (total, steps, t_batched, len(nonfinite)))

    # sequential: the same instances one at a time through a one-instance
    # engine, each for the steps it ran in the batch
    t_seq = 0.0
    for j, (solid, prm, c0) in enumerate(problems):
        g = MultiPhaseGPU(precision).init(solid, prm, sw)
        g.init_rest(c0, 1.0)
        g.record_target_mass()
        g.step(2)
        g.batch.synchronize()
        t0 = time.perf_counter()
        g.step(steps_of[j])
        g.batch.synchronize()
        t_seq += time.perf_counter() - t0
        g.free()
        g.batch.free()
    if log:
        log("sweep: sequential %.3fs" % t_seq)
    res = {"instances": len(problems), "sites": int(total), "steps": steps,
           "batched_MLUPS": lu / t_batched / 1e6, "sequential_MLUPS": lu / t_seq / 1e6,
           "speedup": t_seq / t_batched, "batched_s": t_batched, "sequential_s": t_seq,
           "lattice_updates": lu, "retired_nonfinite": len(nonfinite),
           "retired": [{"instance": j, "after_steps": steps_of[j], "theta": specs[j]["theta"],
                        "mobility_coeff": specs[j]["mobility_coeff"]} for j in nonfinite],
           "autotune": tune}
    if check_fp64 and precision == "f32":
        f32 = _batch_fields(problems, "f32", steps, graph_chunk)
        f64 = _batch_fields(problems, "f64", steps, graph_chunk)
        finite = lambda f: all(np.isfinite(f[k]).all() for k in f)
        # an instance whose FP64 run is not finite (the parameters leave the
        # method's stable range) has no reference to compare against
        pairs = [(a, b) for a, b in zip(f32, f64) if finite(b)]
        dc = [float(np.abs(a["c"] - b["c"]).max()) for a, b in pairs]
        du = [max(float(np.abs(a[k] - b[k]).max()) for k in ("ux", "uy", "uz")) for a, b in pairs]
        res["fp32_vs_fp64"] = {"steps": steps + 2, "compared": len(pairs),
                               "fp64_nonfinite": len(problems) - len(pairs),
                               "fp32_nonfinite_where_fp64_finite": int(sum(not finite(a) for a, _ in pairs)),
                               "max_dc": max(dc) if dc else None, "max_du": max(du) if du else None,
                               "median_instance_dc": float(np.median(dc)) if dc else None,
                               "instances_dc_above_1e-5": int(sum(v > 1e-5 for v in dc))}
        if log:
            log("sweep: fp32 vs fp64 %s" % res["fp32_vs_fp64"])
    return res


# ---- served request stream ----------------------------------------------------------

def make_requests(count, rate=4.0, seed=2024, edge_range=(48, 150), dp_range=(1e-4, 1e-3),
                  theta_range=(30.0, 150.0), porosity=0.45):
    """Seeded request stream: arrival offsets (s) of a Poisson process and the
    per-request edge, pressure drop and contact angle."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        out.append({"id": i, "arrival": t, "edge": int(rng.integers(edge_range[0], edge_range[1] + 1)),
                    "dp": float(rng.uniform(*dp_range)), "theta": float(rng.uniform(*theta_range)),
                    "porosity": porosity, "seed": int(seed + 1 + i)})
    return out


def request_problem(req):
    """Open channel for one request: sphere pack, flow along y, inlet phase
    injected (c = 1), outlet drains at p = 1, uniform initial c = 0."""
    e = int(req["edge"])
    rmin, rmax = 3.0, 6.0
    solid = W.spheres(e, e, e, rmin, rmax, req["porosity"], req["seed"], keep_largest=True, periodic=(1, 0, 1))
    s, inl, out, ic, bv, br = W.open_channel(solid, 2, copy_to_buffers=True)
    prm = W.derive_params(sigma=0.01, width=4.0, mobility_coeff=0.02, theta_deg=req["theta"])
    ob = {"use_in_prs": 1, "use_in_fluid": 1, "use_out_prs": 1, "p_in": 1.0 + req["dp"], "p_out": 1.0,
          "inlet_c_a": 1.0}
    return {"solid": s, "params": prm, "open_bnd": ob, "inlet": inl, "outlet": out, "icoord": ic,
            "buf_verts": bv, "buf_refs": br}


class GpuBackend:
    """The engine side of the server: one Batch, one slot per request."""

    def __init__(self, precision="f32"):
        from .engine import Batch
        self.precision = precision
        self.batch = Batch(precision=precision)

    def admit(self, prob):
        from .engine import MultiPhaseGPU
        g = MultiPhaseGPU(batch=self.batch).init(prob["solid"], prob["params"], {"use_mrt": 1, "correct_op_mass": 0},
                                                 periodic=(1, 0, 1))
        g.set_open_bnd(prob["open_bnd"], prob["inlet"], prob["outlet"], prob["icoord"], prob["buf_verts"],
                       prob["buf_refs"], True)
        g.init_rest(np.zeros(g.n), 1.0)
        return g, g.n

    def step(self, nsteps):
        self.batch.step(nsteps, stats_last=True)

    def mean_c(self, g):
        """Mean c over the last step; raises FelbmError(FELBM_ENONFINITE) when
        the instance's fields are no longer finite."""
        return float(g.flow_stats()[4]) / g.n

    def retire(self, g):
        g.free()

    def close(self):
        self.batch.free()


class Server:
    """Continuous batching with HBM admission control.

    Each tick: admit every arrived request that fits the budgets (FIFO) and
    advance all active instances by `admit_every` steps.  Two budgets bound
    what runs at once: device memory (`hbm_budget`, bytes) and, optionally,
    compute (`max_sites`, lattice sites).  Every active instance advances at
    the batch's step rate, which falls as the batch grows, so with memory as
    the only limit every request of a burst shares the GPU until nearly the
    end; a site budget large enough to keep the GPU full (a few million
    sites) lets the first arrivals finish early at about the same aggregate
    rate.  A request larger than either budget still runs, alone.  A request is checked
    at every multiple of `check_every` of its own steps: it is retired when its
    mean c moved by less than `tol` relative to its previous check, or when it
    reached `max_steps`.  The tick sets how long an arrival waits for a running
    batch (at most one tick) independently of the convergence window.
    `clock` returns seconds since the stream started."""

    def __init__(self, backend, hbm_budget=200e9, check_every=500, tol=1e-6, max_steps=20000, precision="f32",
                 clock=None, sleep=None, admit_every=100, max_sites=None):
        self.be = backend
        self.budget = float(hbm_budget)
        self.max_sites = int(max_sites) if max_sites else None
        self.check_every = int(check_every)
        self.tick = int(admit_every) if admit_every else self.check_every
        self.tol = float(tol)
        self.max_steps = int(max_steps)
        if self.tick <= 0 or self.check_every % self.tick or self.max_steps % self.tick:
            raise ValueError("check_every and max_steps must be multiples of admit_every")
        self.precision = precision
        self.clock = clock
        self.sleep = sleep or time.sleep

    def run(self, requests, problems, log=None):
        t0 = time.perf_counter()
        clock = self.clock or (lambda: time.perf_counter() - t0)
        pending = sorted(requests, key=lambda r: r["arrival"])
        active, done = [], []
        used = 0.0
        sites = 0
        lu = 0.0
        while pending or active:
            now = clock()
            # admission (FIFO, budgets permitting)
            while pending and pending[0]["arrival"] <= now:
                r = pending[0]
                pr = problems[r["id"]]
                n_open = sum(len(pr.get(k, ())) for k in ("inlet", "outlet", "buf_verts"))
                n_sites = int(np.count_nonzero(pr["solid"] == 0))
                need = instance_bytes(n_sites, self.precision, n_open)
                if active and (used + need > self.budget or
                               (self.max_sites is not None and sites + n_sites > self.max_sites)):
                    break
                pending.pop(0)
                h, n = self.be.admit(problems[r["id"]])
                active.append({"req": r, "h": h, "n": n, "steps": 0, "last": None, "bytes": need,
                               "sites": n_sites, "admitted": clock()})
                used += need
                sites += n_sites
            if not active:
                if pending:
                    self.sleep(max(0.0, pending[0]["arrival"] - clock()))
                continue
            self.be.step(self.tick)
            lu += sum(a["n"] for a in active) * float(self.tick)
            still = []
            for a in active:
                a["steps"] += self.tick
                if a["steps"] % self.check_every and a["steps"] < self.max_steps:
                    still.append(a)
                    continue
                try:
                    m = self.be.mean_c(a["h"])
                except FelbmError as err:
                    if err.code != FELBM_ENONFINITE:
                        raise
                    m = float("nan")
                finite = m == m
                rel = abs(m - a["last"]) / max(abs(m), 1e-300) if (finite and a["last"] is not None) else None
                conv = rel is not None and rel <= self.tol
                a["last"] = m
                if conv or not finite or a["steps"] >= self.max_steps:
                    t = clock()
                    self.be.retire(a["h"])
                    used -= a["bytes"]
                    sites -= a["sites"]
                    status = "converged" if conv else ("nonfinite" if not finite else "step cap")
                    done.append({"id": a["req"]["id"], "arrival": a["req"]["arrival"], "done": t,
                                 "latency": t - a["req"]["arrival"], "steps": a["steps"], "sites": a["n"],
                                 "converged": bool(conv), "status": status, "mean_c": m, "rel_change": rel})
                    if log:
                        log("serve: request %d done after %d steps (%s, last relative change of mean c %s), "
                            "latency %.2fs" % (a["req"]["id"], a["steps"], status,
                                               "%.2e" % rel if rel is not None else "-", t - a["req"]["arrival"]))
                else:
                    still.append(a)
            active = still
        makespan = max(d["done"] for d in done) - min(d["arrival"] for d in done) if done else 0.0
        lat = np.array([d["latency"] for d in done])
        rels = [d["rel_change"] for d in done if d["rel_change"] is not None]
        return {"requests": len(done), "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
                "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
                "aggregate_MLUPS": lu / makespan / 1e6 if makespan > 0 else None,
                "makespan_s": makespan, "converged": int(sum(d["converged"] for d in done)),
                "nonfinite": int(sum(d["status"] == "nonfinite" for d in done)),
                "step_cap": int(sum(d["status"] == "step cap" for d in done)),
                "rel_change_at_retirement": {"min": min(rels), "median": float(np.median(rels)), "max": max(rels)}
                if rels else None,
                "lattice_updates": lu, "hbm_budget": self.budget, "max_sites": self.max_sites, "done": done}


def _served_fields(problems, precision, steps):
    """Fields of every request problem after `steps` steps, all in one batch."""
    be = GpuBackend(precision)
    try:
        hs = [be.admit(pr)[0] for pr in problems]
        be.batch.step(steps)
        be.batch.synchronize()
        return [h.download() for h in hs]
    finally:
        be.close()


def run_serve(count, rate=4.0, seed=2024, precision="f32", check_every=500, tol=1e-6, max_steps=20000,
              hbm_budget=200e9, log=None, check_fp64_steps=0, admit_every=100, max_sites=None):
    """The served stream; check_fp64_steps > 0 also runs every request problem
    that many steps in FP32 and in FP64 and reports the largest FP32 - FP64
    difference of c and u over them."""
    reqs = make_requests(count, rate, seed)
    t_gen = time.time()
    problems = {r["id"]: request_problem(r) for r in reqs}
    if log:
        log("serve: %d requests generated in %.1fs" % (len(reqs), time.time() - t_gen))
    be = GpuBackend(precision)
    try:
        res = Server(be, hbm_budget, check_every, tol, max_steps, precision,
                     admit_every=admit_every, max_sites=max_sites).run(reqs, problems, log)
    finally:
        be.close()
    res["rate_per_s"] = rate
    if check_fp64_steps > 0 and precision == "f32":
        probs = [problems[r["id"]] for r in reqs]
        f32 = _served_fields(probs, "f32", check_fp64_steps)
        f64 = _served_fields(probs, "f64", check_fp64_steps)
        finite = lambda f: all(np.isfinite(f[k]).all() for k in f)
        pairs = [(a, b) for a, b in zip(f32, f64) if finite(b)]
        dc = [float(np.abs(a["c"] - b["c"]).max()) for a, b in pairs]
        du = [max(float(np.abs(a[k] - b[k]).max()) for k in ("ux", "uy", "uz")) for a, b in pairs]
        res["fp32_vs_fp64"] = {"steps": int(check_fp64_steps), "compared": len(pairs),
                               "fp64_nonfinite": len(probs) - len(pairs),
                               "max_dc": max(dc) if dc else None, "max_du": max(du) if du else None,
                               "median_request_dc": float(np.median(dc)) if dc else None}
        if log:
            log("serve: fp32 vs fp64 %s" % res["fp32_vs_fp64"])
    return res
