"""The caller's site numbering at the boundary (felbm_geometry.site_xyz).

The reference numbers the compressed sites as its SubDomain does
(MultiPhaseGPU::init takes the SubDomain, include/felbm_gpu/multiphase_gpu.cuh:
75-79; the order is sd.idx_to_coords, src/felbm_gpu_main.cu:124-143).  An
engine given a permuted numbering must hold the same physics: after the same
steps its state, fields, stats, mass and exported tables equal those of the
engine in its own (x-fastest) order, permuted, bit for bit -- with the mass
correction, with open-boundary node lists given in the caller's numbering,
and in both precisions.  A site list with a solid voxel, a repeated voxel or
the wrong length is refused."""
import sys

import numpy as np
import pytest

from felbm_gpu_amd import MultiPhaseGPU
from felbm_gpu_amd import workloads as W
from felbm_gpu_amd._lib import FelbmError

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _perm(n, seed):
    return np.random.default_rng(seed).permutation(n)


def _xyz(solid):
    x, y, z = W.site_coords(solid)
    return np.stack([x, y, z], 1).astype(np.int32)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_permuted_numbering_steps_identically(precision):
    solid = W.spheres(24, 24, 24, 3.0, 5.0, 0.55, 21)
    prm = W.derive_params(theta_deg=50.0, acceleration=(1e-5, 0.0, 0.0), rho0=3.0)
    sw = {"use_mrt": 1, "correct_op_mass": 1}
    xyz = _xyz(solid)
    n = len(xyz)
    P = _perm(n, 5)                      # caller site i is engine-order site P[i]
    a = MultiPhaseGPU(precision).init(solid, prm, sw)
    b = MultiPhaseGPU(precision).init(solid, prm, sw, site_xyz=xyz[P])
    assert a.n == b.n == n
    c0 = W.slab_c(solid, 0, 0, 12, 4.0)
    a.init_rest(c0, 1.0)
    b.init_rest(c0[P], 1.0)
    a.record_target_mass()
    b.record_target_mass()
    _say("numbering %s: %d sites, 30 steps" % (precision, n))
    a.step(30, stats_last=True)
    b.step(30, stats_last=True)
    ha, ga = (s.reshape(19, n) for s in a.download_state())
    hb, gb = (s.reshape(19, n) for s in b.download_state())
    assert np.array_equal(hb, ha[:, P]) and np.array_equal(gb, ga[:, P])
    fa, fb = a.download(), b.download()
    for k in fa:
        assert np.array_equal(fb[k], fa[k][P]), k
    assert np.array_equal(a.flow_stats(), b.flow_stats())
    assert a.current_mass() == b.current_mass()
    # a state uploaded in the caller's numbering goes back out unchanged
    b.upload_state(hb.ravel(), gb.ravel())
    h2, g2 = b.download_state()
    assert np.array_equal(h2, hb.ravel()) and np.array_equal(g2, gb.ravel())
    a.batch.free()
    b.batch.free()


def test_permuted_numbering_tables_and_open_boundary():
    solid0 = W.spheres(20, 24, 20, 3.0, 5.0, 0.55, 23, keep_largest=True, periodic=(1, 0, 1))
    s, inl, out, ic, bv, br = W.open_channel(solid0, 2, copy_to_buffers=True)
    prm = W.derive_params(theta_deg=70.0)
    sw = {"use_mrt": 1, "correct_op_mass": 0}
    ob = {"use_in_prs": 1, "use_in_fluid": 1, "use_out_prs": 1, "p_in": 1.0005, "p_out": 1.0, "inlet_c_a": 1.0}
    xyz = _xyz(s)
    n = len(xyz)
    P = _perm(n, 9)
    Pinv = np.empty(n, np.int64)
    Pinv[P] = np.arange(n)
    a = MultiPhaseGPU("f64").init(s, prm, sw, periodic=(1, 0, 1))
    b = MultiPhaseGPU("f64").init(s, prm, sw, periodic=(1, 0, 1), site_xyz=xyz[P])
    a.set_open_bnd(ob, inl, out, ic, bv, br, True)
    b.set_open_bnd(ob, Pinv[inl], Pinv[out], ic, Pinv[bv], Pinv[br], True)
    a.init_rest(np.zeros(n), 1.0)
    b.init_rest(np.zeros(n), 1.0)
    _say("numbering with open boundaries: %d sites, %d inlet nodes, 40 steps" % (n, len(inl)))
    a.step(40)
    b.step(40)
    ha, ga = (x.reshape(19, n) for x in a.download_state())
    hb, gb = (x.reshape(19, n) for x in b.download_state())
    assert np.array_equal(hb, ha[:, P]) and np.array_equal(gb, ga[:, P])
    ta, tb = a.export_tables(), b.export_tables()
    tn = b.export_tables(noalloc=True)
    for k in ta:
        rows = ta[k].size // n
        va = ta[k].reshape(rows, n)[:, P].copy()
        if k == "src":
            kk, ss = va // n, va % n
            va = np.where(va >= 0, kk * n + Pinv[ss], va)
        elif k not in ("bcase", "lcase"):
            va = np.where(va >= 0, Pinv[np.maximum(va, 0)], va)
        assert np.array_equal(tb[k].reshape(rows, n), va), k
        assert np.array_equal(tn[k], tb[k]), k
    a.batch.free()
    b.batch.free()


def test_site_list_is_validated():
    solid = W.spheres(16, 16, 16, 3.0, 4.0, 0.55, 3)
    prm = W.derive_params()
    xyz = _xyz(solid)
    sz, sy, sx = np.nonzero(solid)
    bad_solid = xyz.copy()
    bad_solid[3] = (sx[0], sy[0], sz[0])
    bad_twice = xyz.copy()
    bad_twice[5] = bad_twice[6]
    for lst, word in ((bad_solid, "solid"), (bad_twice, "twice"), (xyz[:-1], "entries")):
        with pytest.raises(FelbmError) as ei:
            MultiPhaseGPU("f64").init(solid, prm, {}, site_xyz=lst)
        assert ei.value.code == -1 and word in str(ei.value), str(ei.value)
    g = MultiPhaseGPU("f64").init(solid, prm, {}, site_xyz=xyz)      # the engine's own order
    assert g.n == len(xyz)
    g.batch.free()


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("permuted", [False, True])
def test_init_rest_equals_uploaded_rest_state(precision, permuted):
    """init_rest builds the rest state h_k = c w_k, g_k = w_k p0 on the device
    from c alone; it must give the same bits as that state built on the host
    and passed through upload_state, before and after steps, in the engine's
    own numbering and a permuted one, with an off-unity rest pressure."""
    solid = W.spheres(20, 18, 16, 2.5, 4.0, 0.5, 33)
    prm = W.derive_params(theta_deg=70.0, acceleration=(1e-5, 0.0, 0.0))
    sw = {"use_mrt": 1, "correct_op_mass": 1}
    xyz = _xyz(solid)
    n = len(xyz)
    P = _perm(n, 9) if permuted else np.arange(n)
    kw = {"site_xyz": xyz[P]} if permuted else {}
    wk = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    c = W.slab_c(solid, 0, 0, 10, 4.0)[P]
    p0 = 1.0 + 3.7e-4
    a = MultiPhaseGPU(precision).init(solid, prm, sw, **kw)
    b = MultiPhaseGPU(precision).init(solid, prm, sw, **kw)
    a.init_rest(c, p0)
    b.upload_state((wk[:, None] * c[None, :]).ravel(), (wk[:, None] * np.full(n, p0)[None, :]).ravel())
    a.record_target_mass()
    b.record_target_mass()
    assert a.get_target_mass() == b.get_target_mass()
    _say("init_rest %s permuted=%s: %d sites" % (precision, permuted, n))
    for steps in (0, 12):
        a.step(steps)
        b.step(steps)
        ha, ga = a.download_state()
        hb, gb = b.download_state()
        assert np.array_equal(ha, hb) and np.array_equal(ga, gb), steps
    a.free()
    b.free()
