"""Continuous batching on the GPU: a request's result does not depend on what
else shares the device.  Requests are admitted into a running batch while
others are mid-run, one is retired mid-run, and each one's state must equal
the state of the same request run alone for the same number of steps, bit for
bit (every pass is site-parallel, the open-boundary requests carry no mass
correction, and the monitoring sums are exact fixed-point sums)."""
import sys

import numpy as np
import pytest

from felbm_gpu_amd import serving

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_requests_admitted_mid_run_equal_alone(precision):
    reqs = serving.make_requests(3, rate=4.0, seed=77, edge_range=(24, 32))
    probs = [serving.request_problem(r) for r in reqs]
    be = serving.GpuBackend(precision)
    a, _ = be.admit(probs[0])
    be.step(30)
    b, _ = be.admit(probs[1])               # admitted while request 0 is running
    be.step(20)
    mean_a = be.mean_c(a)
    state_a = a.download_state()            # request 0 after 50 steps
    be.retire(a)                            # retired while 1 is running
    c, _ = be.admit(probs[2])
    be.step(15)
    mean_b = be.mean_c(b)
    state_b, state_c = b.download_state(), c.download_state()   # 35 and 15 steps
    mean_c = be.mean_c(c)
    be.retire(b)
    be.retire(c)
    be.close()
    for idx, steps, mixed, mean_mixed in ((0, 50, state_a, mean_a), (1, 35, state_b, mean_b), (2, 15, state_c, mean_c)):
        alone = serving.GpuBackend(precision)
        g, n = alone.admit(probs[idx])
        alone.step(steps)
        h, gg = g.download_state()
        m = alone.mean_c(g)
        alone.retire(g)
        alone.close()
        _say("request %d (%d sites), %d steps: mean c alone %.17g, in the mix %.17g" % (idx, n, steps, m, mean_mixed))
        assert np.array_equal(h, mixed[0]) and np.array_equal(gg, mixed[1]), idx
        assert m == mean_mixed, idx


def test_converging_request_retires_on_engine_output():
    """The served stream's retire-on-convergence path on real engine output:
    a request whose inlet injects the phase already in the medium
    (single-phase pressure-driven flow, inlet_c_a = 0 into c = 0) keeps its
    mean c and is retired as converged at its second check, while a
    two-phase request admitted with it is still invading when it reaches the
    step cap."""
    reqs = serving.make_requests(2, rate=4.0, seed=77, edge_range=(24, 32))
    probs = {r["id"]: serving.request_problem(r) for r in reqs}
    probs[0]["open_bnd"] = dict(probs[0]["open_bnd"], inlet_c_a=0.0)
    for r in reqs:
        r["arrival"] = 0.0
    be = serving.GpuBackend("f32")
    try:
        res = serving.Server(be, hbm_budget=200e9, check_every=500, tol=1e-6, max_steps=3000).run(reqs, probs)
    finally:
        be.close()
    by_id = {d["id"]: d for d in res["done"]}
    _say("served: %s" % [(d["id"], d["status"], d["steps"], d["rel_change"]) for d in res["done"]])
    assert by_id[0]["status"] == "converged" and by_id[0]["steps"] == 1000
    assert by_id[1]["status"] == "step cap" and by_id[1]["rel_change"] > 1e-6
    assert res["converged"] == 1 and res["nonfinite"] == 0
