"""GPU tracers against the NumPy restatement (oracle/tracers.py), fed with the
FP64 oracle's velocity of each step's field pass -- the u the reference's
driver copies to its ParticleManager after gpu.step()
(src/felbm_gpu_main.cu:744-777).  With Dm = 0 the positions agree bit for bit;
Brownian tracers within 1e-12 (log/cos/sin of two libraries)."""
import sys

import numpy as np
import pytest

from felbm_gpu_amd import MultiPhaseGPU
from felbm_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(precision="f64", substeps=2, vskip=1, seed=11, n_per_plane=300):
    from oracle.pyoracle import Oracle
    from oracle.tracers import Tracers
    solid = W.spheres(24, 24, 24, 3.0, 5.0, 0.55, 6)
    prm = W.derive_params(theta_deg=60.0, acceleration=(2e-5, 0.0, 0.0))
    c0 = W.slab_c(solid, 0, 0, 12, 4.0)
    o = Oracle(solid, prm, use_mrt=True, mass_mode=2)
    o.init_rest(c0, 1.0)
    o.record_target_mass()
    g = MultiPhaseGPU(precision).init(solid, prm, {"use_mrt": 1, "correct_op_mass": 1})
    h, gg = o.download_state()
    g.upload_state(h, gg)
    g.record_target_mass()
    pos, pid = W.seed_planes(solid, n_per_plane, (5.5, 17.5), axis=0, seed=3)
    dm = W.dm_groups("0:%d, 1e-3:%d, 2.5e-2:%d" % (n_per_plane // 3, n_per_plane // 3,
                                                   n_per_plane - 2 * (n_per_plane // 3)))
    dm = np.concatenate([dm, dm])
    g.tracers_init(pos, dm, seed=seed, substeps=substeps, velocity_skip=vskip)
    t = Tracers(solid, (1, 1, 1), pos, dm, seed=seed, substeps=substeps, velocity_skip=vskip)
    return o, g, t, dm


def advance(o, g, t, steps):
    for _ in range(steps):
        o.step()
        o.apply_mass_correction()
        t.step(o.field("ux"), o.field("uy"), o.field("uz"))
    g.step(steps)


@pytest.mark.parametrize("vskip", [1, 3])
def test_tracers_match_restatement_fp64(vskip):
    o, g, t, dm = setup(vskip=vskip)
    _say("tracers fp64, velocity_skip", vskip, "n", len(dm))
    advance(o, g, t, 24)
    xyz, uvw, blk = g.tracers_download()
    det = dm == 0.0
    d_det = float(np.abs(xyz[det] - t.pos[det]).max())
    d_rnd = float(np.abs(xyz[~det] - t.pos[~det]).max())
    _say("  max|dx| Dm=0 %.3e, Dm>0 %.3e, blocked %d/%d" % (d_det, d_rnd, int(blk.sum()), int(t.blocked.sum())))
    assert d_det == 0.0
    assert d_rnd <= 1e-12
    assert np.array_equal(blk, t.blocked)
    u_ref = t.interp(t.pos, t.u)
    assert float(np.abs(uvw[det] - u_ref[det]).max()) == 0.0
    # tracers moved: the flow and the kicks are not trivial
    assert float(np.abs(xyz - t.pos).max()) <= 1e-12 and float(np.abs(t.pos[det][:, 0]).std()) > 0


def test_tracers_fp32_engine_within_tolerance():
    o, g, t, dm = setup(precision="f32", substeps=1)
    advance(o, g, t, 12)
    xyz, uvw, blk = g.tracers_download()
    det = dm == 0.0
    d = float(np.abs(xyz[det] - t.pos[det]).max())
    _say("tracers fp32 engine: max|dx| (Dm=0) %.3e" % d)
    assert d <= 1e-5


def test_stretching_matches_restatement_fp64():
    """Eq. (12) stretching with the phase split on the real flow: positions of
    the Dm = 0 tracers bit-identical, rho_hat and log_rho within 1e-12 (sqrt
    and log of two libraries), rows within 1e-10."""
    from oracle.tracers import Tracers
    o, g, t, dm = setup(substeps=2, vskip=1)
    n = len(dm)
    r0 = np.random.default_rng(4).normal(size=(n, 3))
    r0 /= np.linalg.norm(r0, axis=1)[:, None]
    g.tracers_stretching(True, rho_hat=r0, phase_split=True, capacity=64)
    t.rho_hat, t.log_rho = r0.copy(), np.zeros(n)
    for _ in range(16):
        o.step()
        o.apply_mass_correction()
        t.step(o.field("ux"), o.field("uy"), o.field("uz"), o.field("c"))
    g.step(16)
    xyz, uvw, blk = g.tracers_download()
    rh, lr = g.tracers_stretch_state()
    rows, lost = g.tracers_stretch_rows()
    det = dm == 0.0
    _say("stretching: max|drho_hat| %.3e max|dlog_rho| %.3e rows %d"
         % (np.abs(rh - t.rho_hat).max(), np.abs(lr - t.log_rho).max(), len(rows)))
    assert float(np.abs(xyz[det] - t.pos[det]).max()) == 0.0
    assert np.abs(rh - t.rho_hat).max() <= 1e-12 and np.abs(lr - t.log_rho).max() <= 1e-12
    assert lost == 0 and len(rows) == 16
    ref = np.array(t.rows)
    assert np.array_equal(rows[:, 0], ref[:, 0]) and np.array_equal(rows[:, [4, 7, 8]], ref[:, [4, 7, 8]])
    for col in (1, 2, 3, 5, 6):
        assert np.allclose(rows[:, col], ref[:, col], rtol=1e-10, atol=1e-18), col
    # a second drain returns nothing new; more steps than the ring holds are counted as lost
    assert len(g.tracers_stretch_rows()[0]) == 0
    g.step(70)
    rows2, lost2 = g.tracers_stretch_rows()
    assert len(rows2) == 64 and lost2 == 6 and rows2[-1, 0] == 16 + 70


def test_tracers_in_graph_replays_equal_eager():
    """The tracer kernels run on the engine stream between the mu pass and the
    collision pass; captured into a graph and replayed, they move exactly as
    in eager steps (positions, stretching state and rows)."""
    runs = []
    for graph in (False, True):
        o, g, t, dm = setup(vskip=2)
        g.tracers_stretching(True, phase_split=True, capacity=64)
        if graph:
            g.batch.step_graph(4, 3)
        else:
            g.step(12)
        xyz, _, blk = g.tracers_download()
        rh, lr = g.tracers_stretch_state()
        rows, _ = g.tracers_stretch_rows()
        runs.append((xyz, blk, rh, lr, rows))
    _say("graph vs eager tracers: max|dx| %.3e" % float(np.abs(runs[0][0] - runs[1][0]).max()))
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def test_graph_recaptured_after_tracers_reallocate():
    """A step graph captured before tracers_init / tracers_stretching holds
    the old tracer buffers as kernel arguments; the engine drops it when they
    are reallocated.  step_graph, tracers_init (new tracers), stretching on,
    step_graph again: positions, stretching state and fields equal the same
    sequence run with eager steps."""
    runs = []
    for graph in (False, True):
        o, g, t, dm = setup(vskip=1)
        adv = (lambda k: g.batch.step_graph(4, k)) if graph else (lambda k: g.step(4 * k))
        adv(1)
        pos2, _ = W.seed_planes(g.solid, 150, (8.5,), axis=0, seed=9)
        g.tracers_init(pos2, np.full(len(pos2), 1e-3), seed=5, substeps=1, velocity_skip=1)
        adv(2)
        g.tracers_stretching(True, phase_split=True, capacity=32)
        adv(1)
        xyz, _, blk = g.tracers_download()
        rh, lr = g.tracers_stretch_state()
        f = g.download()
        runs.append((xyz, blk, rh, lr, f["c"], f["ux"]))
    _say("graph after tracer reallocation vs eager: max|dx| %.3e" % float(np.abs(runs[0][0] - runs[1][0]).max()))
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)
