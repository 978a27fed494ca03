"""Scheduler policy of the served request stream, on the CPU with a stand-in
engine and a simulated clock: every request completes, convergence and the
step cap retire instances, admission is FIFO and never exceeds the HBM
budget while another instance is active, and the latency and throughput
figures follow from the clock."""
import numpy as np

from felbm_gpu_amd import serving as S


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class FakeBackend:
    """mean c of request i approaches 1 as 1 - r_i^k after k check intervals
    (500 steps) of its own; one step of one site costs 1e-9 s of simulated
    time."""

    def __init__(self, clock, rates, sizes):
        self.clock, self.rates, self.sizes = clock, rates, sizes
        self.active = {}
        self.log = []

    def admit(self, prob):
        rid = prob["id"]
        self.active[rid] = 0
        self.log.append(("admit", rid, self.clock()))
        return rid, self.sizes[rid]

    def step(self, nsteps):
        self.clock.t += 1e-9 * nsteps * sum(self.sizes[r] for r in self.active)
        for r in self.active:
            self.active[r] += nsteps

    def mean_c(self, rid):
        return 1.0 - self.rates[rid] ** (self.active[rid] // 500)

    def retire(self, rid):
        del self.active[rid]
        self.log.append(("retire", rid, self.clock()))


def make(count, sizes):
    reqs = S.make_requests(count, rate=4.0, seed=5)
    probs = {r["id"]: {"id": r["id"], "solid": np.zeros(sizes[r["id"]], np.uint8)} for r in reqs}
    return reqs, probs


def test_requests_arrive_as_a_seeded_poisson_stream():
    a = S.make_requests(2000, rate=4.0, seed=1)
    b = S.make_requests(2000, rate=4.0, seed=1)
    assert a == b
    gaps = np.diff([0.0] + [r["arrival"] for r in a])
    assert abs(gaps.mean() - 0.25) < 0.02 and np.all(gaps > 0)
    edges = np.array([r["edge"] for r in a])
    assert edges.min() >= 48 and edges.max() <= 150
    dps = np.array([r["dp"] for r in a])
    assert dps.min() >= 1e-4 and dps.max() <= 1e-3


def test_every_request_completes_and_retires_on_convergence_or_cap():
    n = 12
    sizes = {i: 100000 + 5000 * i for i in range(n)}
    rates = {i: (0.001 if i % 3 else 0.999999) for i in range(n)}   # every third never converges
    clk = FakeClock()
    be = FakeBackend(clk, rates, sizes)
    reqs, probs = make(n, sizes)
    res = S.Server(be, hbm_budget=1e12, check_every=500, tol=1e-6, max_steps=5000, clock=clk,
                   sleep=clk.sleep).run(reqs, probs)
    assert res["requests"] == n
    by_id = {d["id"]: d for d in res["done"]}
    for i in range(n):
        d = by_id[i]
        if i % 3:
            assert d["converged"] and d["steps"] < 5000
        else:
            assert not d["converged"] and d["steps"] == 5000
        assert d["latency"] >= 0.0 and d["done"] >= d["arrival"]
    assert res["p50_latency_s"] <= res["p99_latency_s"]
    assert res["lattice_updates"] == sum(d["sites"] * d["steps"] for d in res["done"])
    assert abs(res["aggregate_MLUPS"] - res["lattice_updates"] / res["makespan_s"] / 1e6) < 1e-9


def test_admission_is_fifo_and_respects_the_hbm_budget():
    n = 10
    sizes = {i: 400000 for i in range(n)}
    per = S.instance_bytes(400000, "f32")
    budget = 3.5 * per                     # at most three at a time
    clk = FakeClock()
    be = FakeBackend(clk, {i: 0.5 for i in range(n)}, sizes)
    reqs, probs = make(n, sizes)
    S.Server(be, hbm_budget=budget, check_every=500, tol=1e-6, max_steps=20000, clock=clk,
             sleep=clk.sleep).run(reqs, probs)
    live, peak, order = set(), 0, []
    for ev, rid, t in be.log:
        if ev == "admit":
            live.add(rid)
            order.append(rid)
        else:
            live.discard(rid)
        peak = max(peak, len(live))
    assert peak == 3
    assert order == sorted(order)


class BlowUpBackend(FakeBackend):
    """request `bad` raises the engine's non-finite error from its third check on."""

    def __init__(self, clock, rates, sizes, bad):
        super().__init__(clock, rates, sizes)
        self.bad = bad

    def mean_c(self, rid):
        if rid == self.bad and self.active[rid] >= 3 * 500:
            raise S.FelbmError(S.FELBM_ENONFINITE, "non-finite or out-of-range value in the monitored fields")
        return super().mean_c(rid)


def test_nonfinite_request_is_retired_and_frees_its_budget():
    n = 6
    sizes = {i: 400000 for i in range(n)}
    per = S.instance_bytes(400000, "f32")
    clk = FakeClock()
    be = BlowUpBackend(clk, {i: 0.999999 for i in range(n)}, sizes, bad=1)
    reqs, probs = make(n, sizes)
    res = S.Server(be, hbm_budget=2.5 * per, check_every=500, tol=1e-6, max_steps=4000, clock=clk,
                   sleep=clk.sleep).run(reqs, probs)
    by_id = {d["id"]: d for d in res["done"]}
    assert res["requests"] == n and res["nonfinite"] == 1 and res["step_cap"] == n - 1
    assert by_id[1]["status"] == "nonfinite" and by_id[1]["steps"] == 1500 and not by_id[1]["converged"]
    assert all(by_id[i]["status"] == "step cap" and by_id[i]["steps"] == 4000 for i in range(n) if i != 1)
    # the retired request's slot went to the next one in line at once
    admit_t = {rid: t for ev, rid, t in be.log if ev == "admit"}
    retire_t = {rid: t for ev, rid, t in be.log if ev == "retire"}
    assert admit_t[2] == retire_t[1]


def test_arrivals_join_a_running_batch_within_one_tick():
    """Admission is checked every `admit_every` steps, not only at the
    convergence checks: a request arriving while others run waits at most one
    tick (100 steps of the running batch), and every request is still checked
    at multiples of 500 of its own steps."""
    n = 8
    sizes = {i: 200000 for i in range(n)}
    clk = FakeClock()
    be = FakeBackend(clk, {i: 0.999999 for i in range(n)}, sizes)
    reqs, probs = make(n, sizes)
    res = S.Server(be, hbm_budget=1e12, check_every=500, tol=1e-6, max_steps=3000, clock=clk,
                   sleep=clk.sleep, admit_every=100).run(reqs, probs)
    admit_t = {rid: t for ev, rid, t in be.log if ev == "admit"}
    for r in reqs:
        # one tick of the largest batch: 100 steps x every request's sites
        assert admit_t[r["id"]] - r["arrival"] <= 1e-9 * 100 * sum(sizes.values()) + 1e-12
    assert all(d["steps"] == 3000 and d["status"] == "step cap" for d in res["done"])
    with np.testing.assert_raises(ValueError):
        S.Server(be, check_every=500, admit_every=300)


def test_open_boundary_allocation_is_in_the_budget():
    assert S.instance_bytes(10 ** 6, "f32", n_open=20000) - S.instance_bytes(10 ** 6, "f32") == \
        4 * 10 ** 6 + 8 * 4 * 20000 + 512


def test_site_budget_bounds_the_batch_and_lets_early_arrivals_finish_first():
    """With a budget of sites in flight the batch never holds more than the
    budget (a request larger than it runs alone), admission stays FIFO, and
    on a burst of equal requests the median latency falls against memory-only
    admission while the last request finishes no later than the step rate
    allows."""
    n = 12
    sizes = {i: 300000 if i != 5 else 900000 for i in range(n)}
    runs = {}
    for cap in (None, 800000):
        clk = FakeClock()
        be = FakeBackend(clk, {i: 0.999999 for i in range(n)}, sizes)
        reqs, probs = make(n, sizes)
        res = S.Server(be, hbm_budget=1e12, check_every=500, tol=1e-6, max_steps=2000, clock=clk,
                       sleep=clk.sleep, max_sites=cap).run(reqs, probs)
        live, peak, order, alone = set(), 0, [], True
        for ev, rid, t in be.log:
            if ev == "admit":
                live.add(rid)
                order.append(rid)
                if rid == 5:
                    alone = live == {5}
            else:
                live.discard(rid)
            if cap is not None and 5 not in live:
                peak = max(peak, sum(sizes[r] for r in live))
        assert order == sorted(order) and res["requests"] == n
        assert res["max_sites"] == cap
        if cap is not None:
            assert peak <= cap and alone
        runs[cap] = res
    assert runs[800000]["p50_latency_s"] < runs[None]["p50_latency_s"]
