"""GPU runs of the workload configurations at their real sizes, through the C ABI.

  reference injection  the FP64 engine against the oracle in mass_mode 1, the
                       reference's per-population injection h_k += (real)(w_k dc)
                       after each step (include/felbm_gpu/device_engine.cuh:
                       1128-1139, multiphase_gpu.cuh:689-700), <= 1e-13: the
                       128x128x1 BGK case over its 2000 steps and the 150^3 MRT
                       pack over 200 steps (the engine folds the same correction
                       into the next step's c, which agrees to rounding)
  300^3 FP32           the throughput lattice (porosity 0.37, 9,988,968 sites,
                       MRT, mass correction) with the 12,000-tracer ladder
                       (6 diffusivities 1e-4 .. 1e-1 on the two slab
                       interfaces), FP32 engine against the FP64 engine from the
                       same initial state: c and u within 1e-5, tracers finite
                       and in pore space
  48-instance sweep    the batch exactly as the sweep benchmark builds it (graph
                       replay), per instance FP32 against FP64 within 1e-5 in c
                       and u; three instances batched equal the same instance
                       run alone, bit for bit
  served requests      requests drawn from the served stream's distribution
                       (edge 48-150, pressure drop 1e-4 .. 1e-3, contact angle
                       30-150 degrees) admitted into a running batch equal the
                       same request run alone, FP32 and FP64, bit for bit
  3d_image ingest      a TIFF stack written in the reference's format with the
                       cfg/3d_image/domain.cfg:19-42 keys (interface_solid,
                       invert_image, coarsening_levels) ingested and run on the
                       engine: bit-identical to the FP64 oracle
"""
import sys

import numpy as np
import pytest

from felbm_gpu_amd import MultiPhaseGPU, Batch
from felbm_gpu_amd import serving
from felbm_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair_mode1(w, use_mrt):
    from oracle.pyoracle import Oracle
    o = Oracle(w["solid"], w["params"], use_mrt=use_mrt, fuse_collision=True, mass_mode=1)
    g = MultiPhaseGPU("f64").init(w["solid"], w["params"], {"use_mrt": use_mrt, "correct_op_mass": 1})
    o.init_rest(w["c0"], 1.0)
    h, gg = o.download_state()
    g.upload_state(h, gg)
    o.record_target_mass()
    g.record_target_mass()
    return o, g


@pytest.mark.parametrize("name,steps,use_mrt", [("2d_cylinders_128", 2000, 0), ("porous_150", 200, 1)])
def test_fp64_engine_vs_reference_injection(name, steps, use_mrt):
    w = W.workload(name)
    o, g = _pair_mode1(w, use_mrt)
    _say("%s: FP64 engine vs oracle mass_mode 1 (per-population injection), %d steps, n = %d" % (name, steps, g.n))
    o.run(steps)                    # step, then the per-population injection
    g.step(steps)
    ho, go = o.download_state()
    hg, gg = g.download_state()
    dh, dg = float(np.abs(ho - hg).max()), float(np.abs(go - gg).max())
    fo, fg = o.download(), g.download()
    df = {k: float(np.abs(fo[k] - fg[k]).max()) for k in ("rho", "c", "ux", "uy", "uz", "p")}
    dm = abs(o.current_mass() - g.current_mass())
    _say("  max|dh| %.3e max|dg| %.3e fields %s |dM| %.3e" % (dh, dg, df, dm))
    assert dh <= 1e-13 and dg <= 1e-13
    assert max(df.values()) <= 1e-13
    assert dm <= 1e-9 * abs(o.current_mass())


def _porous_300_tracers(eng, solid):
    edge = solid.shape[2]
    pos, _ = W.seed_planes(solid, 6000, (0.0, edge / 2.0), axis=0, seed=12)
    ladder = "1e-4:1000, 3.98e-4:1000, 1.58e-3:1000, 6.31e-3:1000, 2.51e-2:1000, 1e-1:1000"
    dm = np.concatenate([W.dm_groups(ladder), W.dm_groups(ladder)])
    eng.tracers_init(pos, dm, seed=3, substeps=1, velocity_skip=1)
    return dm


def test_porous_300_fp32_with_tracers_vs_fp64():
    w = W.workload("porous_300")
    steps = 500
    out = {}
    for prec in ("f64", "f32"):
        g = MultiPhaseGPU(prec).init(w["solid"], w["params"], w["switches"])
        g.init_rest(w["c0"], w["p0"])
        g.record_target_mass()
        _porous_300_tracers(g, w["solid"])
        _say("porous_300 %s: %d sites, 12,000 tracers, %d steps" % (prec, g.n, steps))
        g.step(steps)
        f = g.download()
        xyz, uvw, blk = g.tracers_download()
        out[prec] = ({k: f[k] for k in ("c", "ux", "uy", "uz")}, xyz, blk,
                     abs(g.current_mass() - g.get_target_mass()))
        g.free()
        g.batch.free()
    f32, f64 = out["f32"][0], out["f64"][0]
    dc = float(np.abs(f32["c"] - f64["c"]).max())
    du = max(float(np.abs(f32[k] - f64[k]).max()) for k in ("ux", "uy", "uz"))
    nz, ny, nx = w["solid"].shape
    for prec in ("f32", "f64"):
        xyz = out[prec][1]
        assert np.isfinite(xyz).all(), prec
        v = np.rint(xyz).astype(np.int64)
        v[:, 0] %= nx
        v[:, 1] %= ny
        v[:, 2] %= nz
        assert not w["solid"][v[:, 2], v[:, 1], v[:, 0]].any(), prec       # every tracer in pore space
    dx = float(np.abs(out["f32"][1] - out["f64"][1]).max())
    _say("  FP32 vs FP64 after %d steps: max|dc| %.3e max|du| %.3e; tracers max|dx| %.3e, blocked %d / %d;"
         " mass drift %.2e / %.2e" % (steps, dc, du, dx, int(out["f32"][2].sum()), int(out["f64"][2].sum()),
                                      out["f32"][3], out["f64"][3]))
    assert dc <= 1e-5 and du <= 1e-5
    assert dx <= 1e-2


def test_sweep_48_batch_fp32_vs_fp64_and_alone():
    specs = W.sweep_instances(48, 1000)
    problems = [serving.sweep_problem(s) for s in specs]
    steps, chunk = 100, 50
    _say("48-instance sweep: %d sites, %d steps (graph chunks of %d)"
         % (sum(int((p[0] == 0).sum()) for p in problems), steps + 2, chunk))
    f32 = serving._batch_fields(problems, "f32", steps, chunk)
    f64 = serving._batch_fields(problems, "f64", steps, chunk)
    finite = lambda f: all(np.isfinite(f[k]).all() for k in f)
    worst_c = worst_u = 0.0
    compared = 0
    for i, (a, b) in enumerate(zip(f32, f64)):
        if not finite(b):
            _say("  instance %d: FP64 run not finite (theta %.1f, mobility coeff %.4f)"
                 % (i, specs[i]["theta"], specs[i]["mobility_coeff"]))
            continue
        assert finite(a), i
        dc = float(np.abs(a["c"] - b["c"]).max())
        du = max(float(np.abs(a[k] - b[k]).max()) for k in ("ux", "uy", "uz"))
        worst_c, worst_u = max(worst_c, dc), max(worst_u, du)
        compared += 1
        assert dc <= 1e-5 and du <= 1e-5, (i, dc, du)
    _say("  FP32 vs FP64 over %d instances: max|dc| %.3e max|du| %.3e" % (compared, worst_c, worst_u))
    assert compared >= 44
    for i in (0, 17, 47):
        alone = serving._batch_fields([problems[i]], "f32", steps, chunk)[0]
        for k in alone:
            assert np.array_equal(alone[k], f32[i][k]), (i, k)
    _say("  instances 0, 17, 47 alone equal the batch bit for bit")


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_served_requests_full_size_mix_equals_alone(precision):
    reqs = serving.make_requests(3, rate=4.0, seed=2024)       # edges 57, 70, 75
    big = serving.make_requests(16, rate=4.0, seed=2024)[5]     # edge 148
    reqs = [reqs[0], big, reqs[2]]
    probs = [serving.request_problem(r) for r in reqs]
    _say("served %s: request edges %s" % (precision, [r["edge"] for r in reqs]))
    be = serving.GpuBackend(precision)
    a, _ = be.admit(probs[0])
    be.step(40)
    b, _ = be.admit(probs[1])
    be.step(30)
    sa, ma = a.download_state(), be.mean_c(a)
    be.retire(a)
    c, _ = be.admit(probs[2])
    be.step(20)
    sb, mb = b.download_state(), be.mean_c(b)
    sc, mc = c.download_state(), be.mean_c(c)
    be.retire(b)
    be.retire(c)
    be.close()
    for idx, steps, st, m in ((0, 70, sa, ma), (1, 50, sb, mb), (2, 20, sc, mc)):
        alone = serving.GpuBackend(precision)
        g, n = alone.admit(probs[idx])
        alone.step(steps)
        h, gg = g.download_state()
        m1 = alone.mean_c(g)
        alone.retire(g)
        alone.close()
        _say("  request edge %d (%d sites), %d steps: mean c %.17g alone, %.17g in the mix"
             % (reqs[idx]["edge"], n, steps, m1, m))
        assert np.isfinite(h).all()
        assert np.array_equal(h, st[0]) and np.array_equal(gg, st[1]), idx
        assert m1 == m, idx


def test_3d_image_stack_ingested_runs_bitexact(tmp_path):
    from felbm_gpu_amd import geometry_io as G
    from oracle.pyoracle import Oracle
    # a 64^3 stack whose non-zero voxels are pore (invert_image with
    # interface_solid), coarsened once to 32^3
    grains = W.spheres(64, 64, 64, 4.0, 7.0, 0.45, 31, keep_largest=False)
    img = np.where(grains == 0, 255, 0).astype(np.uint8)
    d = tmp_path / "image"
    d.mkdir()
    G.write_tiff(str(d / "stack_000.tif"), img[:32])
    G.write_tiff(str(d / "stack_001.tif"), img[32:])
    keys = {"image_dir": "./image/", "extension": "tif", "interface_solid": "true", "invert_image": "true",
            "coarsening_levels": "1", "use_subvolume": "false", "add_inlet_pores": "false",
            "add_outlet_pores": "false"}
    cfg = tmp_path / "domain.cfg"
    cfg.write_text("".join("%s = %s   # key\n" % kv for kv in keys.items()))
    solid = G.load_3d_image(str(cfg))
    assert solid.shape == (32, 32, 32)
    assert np.array_equal(solid, G.coarsen(grains, 1))
    nx = solid.shape[2]
    prm = W.derive_params(theta_deg=45.0, acceleration=(1e-5, 0.0, 0.0))
    c0 = W.slab_c(solid, 0, 0, nx // 2, 4.0)
    o = Oracle(solid, prm, use_mrt=True, mass_mode=2)
    g = MultiPhaseGPU("f64").init(solid, prm, {"use_mrt": 1, "correct_op_mass": 1})
    o.init_rest(c0, 1.0)
    h, gg = o.download_state()
    g.upload_state(h, gg)
    o.record_target_mass()
    g.record_target_mass()
    _say("3d_image stack 64^3 -> coarsened %s, %d sites" % (solid.shape, g.n))
    o.run(30)
    g.step(30)
    ho, go = o.download_state()
    hg, ggg = g.download_state()
    assert np.array_equal(ho, hg) and np.array_equal(go, ggg)
    fo, fg = o.download(), g.download()
    for k in fo:
        assert np.array_equal(fo[k], fg[k]), k


@pytest.mark.parametrize("precision,edge", [("f32", 390), ("f64", 310)])
def test_instance_beyond_32bit_offsets_is_refused(precision, edge):
    """An all-fluid box whose distribution array would reach 2^32 bytes
    (19 * sites * sizeof(real)) is refused before anything is allocated, and
    the engine keeps working for the next instance."""
    from felbm_gpu_amd import FelbmError
    prm = W.derive_params()
    sw = {"use_mrt": 1, "correct_op_mass": 0}
    b = Batch(precision=precision)
    with pytest.raises(FelbmError) as err:
        MultiPhaseGPU(batch=b).init(np.zeros((edge, edge, edge), np.uint8), prm, sw)
    assert err.value.code == -5 and "4 GiB" in str(err.value)
    _say("refused %d^3 %s: %s" % (edge, precision, err.value))
    small = MultiPhaseGPU(batch=b).init(W.spheres(16, 16, 16, 2.0, 3.0, 0.6, 9), prm, sw)
    small.init_rest(np.full(small.n, 0.3), 1.0)
    b.step(2)
    assert np.isfinite(small.download()["c"]).all()
    b.free()
