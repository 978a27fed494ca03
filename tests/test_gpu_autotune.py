"""Workgroup shapes and the tile-shape autotuner (GPU).

Results must not depend on the sites per workgroup of any pass (the passes are
site-parallel and every reduction is an exact fixed-point sum), and the
autotuner's timing steps must leave the state of every instance as it was.
"""
import sys

import numpy as np
import pytest

from felbm_gpu_amd import Batch, MultiPhaseGPU
from felbm_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _batch(precision="f64"):
    prm = W.derive_params(theta_deg=60.0, acceleration=(1e-5, 0, 0))
    solids = [W.spheres(18, 18, 18, 3.0, 5.0, 0.55, 31), W.spheres(22, 20, 16, 3.0, 5.0, 0.5, 32)]
    b = Batch(precision=precision)
    insts = []
    for s in solids:
        g = MultiPhaseGPU(batch=b).init(s, prm, {"use_mrt": 1, "correct_op_mass": 1})
        g.init_rest(W.slab_c(s, 0, 0, s.shape[2] // 2, 4.0), 1.0)
        g.record_target_mass()
        insts.append(g)
    return b, insts


def _run(shapes, steps=12):
    b, insts = _batch()
    if shapes is not None:
        b.set_block_sizes(shapes)
    b.step(steps)
    out = [g.download_state() for g in insts]
    for g in insts:
        g.free()
    b.free()
    return out


def test_workgroup_shapes_give_identical_states():
    ref = _run(None)
    for shapes in ([64, 64, 64, 64], [128, 128, 128, 128], [64, 256, 128, 64], [256, 64, 64, 128]):
        _say("shapes", shapes)
        got = _run(shapes)
        for (h0, g0), (h1, g1) in zip(ref, got):
            assert np.array_equal(h0, h1) and np.array_equal(g0, g1), shapes


def test_autotune_leaves_state_and_steps_unchanged():
    b, insts = _batch()
    b.step(3)
    before = [g.download_state() for g in insts]
    res = b.autotune(4)
    _say("autotune", res)
    assert set(res["block_sizes"].values()) <= {64, 128, 256}
    after = [g.download_state() for g in insts]
    for (h0, g0), (h1, g1) in zip(before, after):
        assert np.array_equal(h0, h1) and np.array_equal(g0, g1)
    b.step(9)
    tuned = [g.download_state() for g in insts]
    for g in insts:
        g.free()
    b.free()
    plain = _run(None, steps=12)
    for (h0, g0), (h1, g1) in zip(plain, tuned):
        assert np.array_equal(h0, h1) and np.array_equal(g0, g1)
    with pytest.raises(Exception):
        b2, i2 = _batch()
        try:
            b2.set_block_sizes([32, 64, 64, 64])
        finally:
            for g in i2:
                g.free()
            b2.free()


def _open_batch(precision):
    from felbm_gpu_amd import serving
    reqs = serving.make_requests(2, rate=4.0, seed=91, edge_range=(24, 30))
    be = serving.GpuBackend(precision)
    hs = [be.admit(serving.request_problem(r))[0] for r in reqs]
    prm = W.derive_params(theta_deg=60.0, acceleration=(1e-5, 0, 0))
    s = W.spheres(18, 18, 18, 3.0, 5.0, 0.55, 31)
    g = MultiPhaseGPU(batch=be.batch).init(s, prm, {"use_mrt": 1, "correct_op_mass": 1})
    g.init_rest(W.slab_c(s, 0, 0, 9, 4.0), 1.0)
    g.record_target_mass()
    return be, hs + [g]


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_shapes_identical_with_open_boundaries(precision):
    """Open-boundary and periodic instances in one batch, FP32 and FP64: every
    shape gives the same state."""
    ref = None
    for shapes in (None, [64, 128, 64, 128], [128, 64, 256, 64]):
        be, insts = _open_batch(precision)
        if shapes is not None:
            be.batch.set_block_sizes(shapes)
        be.batch.step(15)
        got = [g.download_state() for g in insts]
        for g in insts:
            g.free()
        be.close()
        if ref is None:
            ref = got
            continue
        _say(precision, "shapes", shapes)
        for (h0, g0), (h1, g1) in zip(ref, got):
            assert np.array_equal(h0, h1) and np.array_equal(g0, g1), shapes
