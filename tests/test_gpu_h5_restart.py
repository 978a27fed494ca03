"""Checkpoint / restart and the field and particle files through the engine
(src/felbm_gpu_main.cu:395-410, :566-575, :701-770; README.md:397-416).

  - a run checkpointed at step 6 and resumed in a new engine from
    checkpoint_6.h5 ends in the state of the uninterrupted run, bit for bit
    (FP64 and FP32; the checkpoint stores doubles, README.md:409-410), and
    writes a byte-identical fields file at the end (times masked);
  - with the mass correction on, the resumed state agrees to rounding: the
    checkpoint holds the corrected populations, the running engine keeps the
    correction folded into the next step;
  - tracers restored from particles_<t>.h5 (positions, Dm, rho_hat, log_rho)
    with their step counter continue the uninterrupted trajectories bit for bit
    in FP64, Brownian groups included, and number the stretching rows by the
    lattice step;
  - the fields file holds what download() returns.
"""
import sys
import time

import numpy as np
import pytest

from felbm_gpu_amd import MultiPhaseGPU, h5io
from felbm_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(solid, prm, precision, mass):
    g = MultiPhaseGPU(precision).init(solid, prm, {"use_mrt": 1, "correct_op_mass": mass})
    g.init_rest(W.slab_c(solid, 0, 0, solid.shape[2] // 2, 4.0), 1.0)
    g.record_target_mass()
    return g


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_checkpoint_restart_is_exact(tmp_path, precision):
    from tests.test_h5io import _same_bytes
    t0 = int(time.time())
    solid = W.spheres(20, 20, 20, 3.0, 5.0, 0.55, 7)
    prm = W.derive_params(theta_deg=60.0, forcing=(1e-5, 0.0, 0.0))
    _say("checkpoint restart", precision, "sites", int((solid == 0).sum()))
    a = _engine(solid, prm, precision, 0)
    out_a = h5io.OutputSchedule(tmp_path / "a", "run", file_skip=12, checkpoint_skip=6)
    (tmp_path / "a").mkdir()
    a.run(0, 12, "sinusoidal", 10.0, output=out_a)
    assert [p.split("/")[-1] for p in out_a.written] == ["run_0.h5", "checkpoint_6.h5", "checkpoint_12.h5",
                                                         "run_12.h5"]
    b = MultiPhaseGPU(precision).init(solid, prm, {"use_mrt": 1, "correct_op_mass": 0})
    step = b.read_checkpoint(tmp_path / "a" / "checkpoint_6.h5")
    assert step == 6 and b.get_target_mass() == a.get_target_mass()
    (tmp_path / "b").mkdir()
    out_b = h5io.OutputSchedule(tmp_path / "b", "run", file_skip=12, checkpoint_skip=6, restart_step=6)
    b.run(6, 12, "sinusoidal", 10.0, output=out_b)
    ha, ga = a.download_state()
    hb, gb = b.download_state()
    assert np.array_equal(ha, hb) and np.array_equal(ga, gb)
    assert _same_bytes(tmp_path / "a" / "run_12.h5", tmp_path / "b" / "run_12.h5", t0)
    assert _same_bytes(tmp_path / "a" / "checkpoint_12.h5", tmp_path / "b" / "checkpoint_12.h5", t0)
    with pytest.raises(h5io.H5Error):
        MultiPhaseGPU(precision).init(solid[:, :, :-1].copy(), prm, {}).read_checkpoint(
            tmp_path / "a" / "checkpoint_6.h5")


def test_checkpoint_restart_with_mass_correction_to_rounding(tmp_path):
    solid = W.spheres(20, 20, 20, 3.0, 5.0, 0.55, 8)
    prm = W.derive_params(theta_deg=60.0)
    a = _engine(solid, prm, "f64", 1)
    a.step(6)
    a.write_checkpoint(tmp_path / "c.h5", 6)
    a.step(10)
    b = MultiPhaseGPU("f64").init(solid, prm, {"use_mrt": 1, "correct_op_mass": 1})
    assert b.read_checkpoint(tmp_path / "c.h5") == 6
    b.step(10)
    ha, ga = a.download_state()
    hb, gb = b.download_state()
    d = max(np.abs(ha - hb).max(), np.abs(ga - gb).max())
    _say("mass-corrected restart: max |dh|, |dg| = %.3e" % d)
    assert d <= 1e-13
    assert abs(a.current_mass() - b.current_mass()) <= 1e-9 * abs(a.current_mass())


def test_fields_file_holds_download(tmp_path):
    solid = W.spheres(16, 16, 16, 2.0, 4.0, 0.6, 9)
    g = _engine(solid, W.derive_params(), "f32", 1)
    g.step(5)
    g.write_fields_h5(tmp_path / "f.h5", deflate=3, f32=True)
    g.write_geometry_h5(tmp_path / "geometry.h5")
    f = g.download()
    for name, key in zip(h5io.FIELD_NAMES, h5io.FIELD_KEYS):
        assert np.array_equal(h5io.read_dataset(tmp_path / "f.h5", name, np.float32), f[key].astype(np.float32))
    co = h5io.read_dataset(tmp_path / "geometry.h5", "coords", np.int32)
    assert co.shape == (g.n, 3)


def test_tracer_restore_continues_exactly(tmp_path):
    solid = W.spheres(24, 24, 24, 3.0, 5.0, 0.55, 6)
    prm = W.derive_params(theta_deg=60.0, acceleration=(2e-5, 0.0, 0.0))
    pos, pid = W.seed_planes(solid, 200, (5.5, 17.5), axis=0, seed=3)
    dm = np.concatenate([W.dm_groups("0:100, 1e-3:50, 2.5e-2:50")] * 2)
    kw = dict(seed=11, substeps=2, velocity_skip=2)

    a = _engine(solid, prm, "f64", 0)
    a.tracers_init(pos, dm, plane_id=pid, **kw)
    a.tracers_stretching(True, phase_split=True, capacity=64)
    (tmp_path / "a").mkdir()
    out = h5io.OutputSchedule(tmp_path / "a", checkpoint_skip=4)
    a.run(0, 10, output=out)
    rows_a, _ = a.tracers_stretch_rows()
    _say("tracer restore: rows", len(rows_a), "files", [p.split("/")[-1] for p in out.written])

    b = MultiPhaseGPU("f64").init(solid, prm, {"use_mrt": 1, "correct_op_mass": 0})
    t = b.read_checkpoint(tmp_path / "a" / "checkpoint_8.h5")
    b.restore_particles_h5(tmp_path / "a" / "particles_8.h5", t, capacity=64, **kw)
    f = h5io.read_particles(tmp_path / "a" / "particles_8.h5")
    assert np.array_equal(f["Dm"], dm) and np.array_equal(f["plane_id"], pid)
    b.run(t, 10)
    xa, _, _ = a.tracers_download()
    xb, _, _ = b.tracers_download()
    ra, la = a.tracers_stretch_state()
    rb, lb = b.tracers_stretch_state()
    rows_b, _ = b.tracers_stretch_rows()
    d = float(np.abs(xa - xb).max())
    _say("  max |dx| %.3e, max |dlog_rho| %.3e" % (d, float(np.abs(la - lb).max())))
    assert d == 0.0 and np.array_equal(ra, rb) and np.array_equal(la, lb)
    assert rows_b[:, 0].tolist() == [9.0, 10.0] and rows_a[-2:, 0].tolist() == [9.0, 10.0]
    assert np.array_equal(rows_b[:, 1:], rows_a[-2:, 1:])
