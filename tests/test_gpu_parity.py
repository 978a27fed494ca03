"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

The bar: FP64 results bit-identical to the oracle (its folded mass-correction
mode, mass_mode=2, which applies the reference's correction inside the next
moments exactly as the engine does); FP32 within the stated tolerance of the
FP64 oracle.  Mirrors the reference's compare_cpu_gpu harness
(src/compare_cpu_gpu.cu:241-294): same initial distributions on both sides,
distributions compared after N steps, then the fields.
"""
import sys

import numpy as np
import pytest

from felbm_gpu_amd import MultiPhaseGPU, Batch
from felbm_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_pack(edge=24, por=0.55, seed=11):
    return W.spheres(edge, edge, edge, 3.0, 5.0, por, seed)


def make_pair(solid, params, use_mrt, mass, precision="f64", periodic=(1, 1, 1), hbb=True, c0=None):
    from oracle.pyoracle import Oracle
    o = Oracle(solid, params, use_mrt=use_mrt, fuse_collision=True, mass_mode=2 if mass else 0,
               periodic=periodic, use_halfway_bb=hbb)
    g = MultiPhaseGPU(precision=precision).init(solid, params, {"use_mrt": use_mrt, "correct_op_mass": mass},
                                                 periodic=periodic, use_halfway_bb=hbb)
    assert g.n == o.n
    if c0 is None:
        c0 = W.slab_c(solid, 0, 0, solid.shape[2] // 2, 4.0)
    o.init_rest(c0, 1.0)
    h, gg = o.download_state()
    g.upload_state(h, gg)
    o.record_target_mass()
    g.record_target_mass()
    return o, g


def test_tables_bitwise_match_oracle():
    from oracle.pyoracle import Oracle
    cases = [("2d", W.cylinders(64, 48, 1, 6, 4.0, 7.0, seed=5)[0], True),
             ("3d", small_pack(20, 0.5, 3), True),
             ("fluid_nohbb", np.zeros((8, 9, 10), np.uint8), False)]
    prm = W.derive_params()
    for name, solid, hbb in cases:
        _say("tables case", name, solid.shape)
        o = Oracle(solid, prm, use_halfway_bb=hbb)
        g = MultiPhaseGPU("f64").init(solid, prm, {}, use_halfway_bb=hbb)
        to, tg = o.tables(), g.export_tables()
        for k in to:
            assert np.array_equal(to[k], tg[k]), (name, k)


@pytest.mark.parametrize("use_mrt", [0, 1])
def test_one_step_fields_and_state_bitexact(use_mrt):
    w = W.workload("2d_cylinders_128")
    _say("one step, mrt", use_mrt)
    o, g = make_pair(w["solid"], w["params"], use_mrt, 1, c0=w["c0"])
    o.run(1)
    g.step(1)
    ho, go = o.download_state()
    hg, gg = g.download_state()
    assert np.array_equal(ho, hg)
    assert np.array_equal(go, gg)
    fo, fg = o.download(), g.download()
    for k in fo:
        assert np.array_equal(fo[k], fg[k]), k


@pytest.mark.parametrize("use_mrt", [0, 1])
def test_20_steps_porous_3d_bitexact(use_mrt):
    solid = small_pack()
    prm = W.derive_params(theta_deg=45.0, acceleration=(1e-5, 0, 0), rho0=1.0, rho1=1.0)
    _say("20 steps 3d, mrt", use_mrt, int((solid == 0).sum()))
    o, g = make_pair(solid, prm, use_mrt, 1)
    o.run(20)
    g.step(20)
    ho, go = o.download_state()
    hg, gg = g.download_state()
    assert np.isfinite(hg).all()
    assert np.abs(ho - hg).max() == 0.0
    assert np.abs(go - gg).max() == 0.0


def test_density_ratio_bitexact():
    """density ratio 5 (the reference's validate.sh cases) with MRT"""
    solid = small_pack(20, 0.6, 4)
    prm = W.derive_params(theta_deg=70.0, rho0=5.0, rho1=1.0, tau0=0.8, tau1=1.0, acceleration=(0, 1e-5, 0))
    o, g = make_pair(solid, prm, 1, 1)
    o.run(20)
    g.step(20)
    assert np.array_equal(o.download_state()[0], g.download_state()[0])
    assert np.array_equal(o.download_state()[1], g.download_state()[1])


def test_stats_and_mass_match():
    solid = small_pack()
    prm = W.derive_params(theta_deg=45.0, acceleration=(1e-5, 0, 0))
    o, g = make_pair(solid, prm, 1, 1)
    o.run(10)
    g.step(10, stats_last=True)
    so, sg = o.flow_stats(), g.flow_stats()
    assert np.array_equal(so, sg), (so, sg)
    assert o.current_mass() == g.current_mass()
    assert abs(g.current_mass() - g.get_target_mass()) <= 1e-9 * abs(g.get_target_mass())


def test_c1_full_2000_steps():
    w = W.workload("2d_cylinders_128")
    o, g = make_pair(w["solid"], w["params"], 0, 1, c0=w["c0"])
    _say("C1 2000 steps, n =", g.n)
    o.run(2000)
    g.step(2000)
    ho, go = o.download_state()
    hg, gg = g.download_state()
    dh, dg = np.abs(ho - hg).max(), np.abs(go - gg).max()
    _say("C1 max|dh| max|dg|", dh, dg)
    assert dh <= 1e-13 and dg <= 1e-13
    fo, fg = o.download(), g.download()
    for k in ("c", "rho", "ux", "uy", "uz", "p"):
        assert np.abs(fo[k] - fg[k]).max() <= 1e-13, k


def test_c2_porous_150_fp64_200_steps():
    """The 150^3 sphere pack at porosity 0.40 (1.35e6 sites), MRT, all switches,
    mass correction on: FP64 bit-identical to the oracle after 200 steps."""
    w = W.workload("porous_150")
    o, g = make_pair(w["solid"], w["params"], 1, 1, c0=w["c0"])
    _say("C2 200 steps, n =", g.n)
    o.run(200)
    g.step(200)
    ho, go = o.download_state()
    hg, gg = g.download_state()
    dh, dg = float(np.abs(ho - hg).max()), float(np.abs(go - gg).max())
    _say("C2 max|dh| max|dg|", dh, dg)
    assert dh == 0.0 and dg == 0.0
    assert abs(o.current_mass() - g.current_mass()) <= 1e-9


def _fp32_errors(o, g, steps):
    o.run(steps)
    g.step(steps)
    fo, fg = o.download(), g.download()
    dc = np.abs(fo["c"] - fg["c"]).max()
    du = max(np.abs(fo[k] - fg[k]).max() for k in ("ux", "uy", "uz"))
    return dc, du


# FP32 tolerance against the FP64 oracle: 1e-6 in c over 20 and over 200
# steps, and 2e-8 in u (measured: 1.3e-7 and 1.6e-7 in c, 2.4e-9 and 2.8e-9 in
# u; 1.0e-6 in c before the populations carried the moments pass's rounding
# remainder, 3.1e-6 in c and 1.2e-7 in u with absolute FP32 pressures).
# The u bound holds only with the pressure stored as a deviation from the rest
# pressure (p - 1, g_k - w_k): with absolute FP32 pressures u was off by 1.2e-7
# after 200 steps (and c by 3.1e-6).
def test_fp32_within_tolerance_of_fp64_oracle():
    solid = small_pack(24, 0.55, 11)
    prm = W.derive_params(theta_deg=45.0, acceleration=(1e-5, 0, 0))
    o, g = make_pair(solid, prm, 1, 1, precision="f32")
    dc, du = _fp32_errors(o, g, 20)
    _say("fp32 mrt 20 steps max|dc|", dc, "max|du|", du)
    assert dc <= 1e-6 and du <= 2e-8
    dc, du = _fp32_errors(o, g, 180)
    _say("fp32 mrt 200 steps max|dc|", dc, "max|du|", du)
    assert du <= 2e-8 and dc <= 1e-6


# Without the mass correction, the FP32 run's total c may differ from the FP64
# run's only by rounding: lattice weights whose FP32 values sum to 1 + 2^-26
# (1/3, 1/18, 1/36 rounded to nearest) add 1.5e-8 of the mass every step
# (measured 1.6e-8 per step here); the engine's weights sum to exactly 1
# (measured 1.6e-9 per step).  What remained came from rounding
# G_0 = w_0 (1 - 1.5 u^2) to FP32 before it multiplies c: at a given |u| the
# rest population's offset from w_0 is the same fraction of an ulp at every
# site, so every bulk site gained or lost the same mass (measured 1.3e-9 per
# step); h_k = w_k c + (w_k u_term_k) c + G_k s_k, rounded once, gives 2.4e-11,
# and carrying the moments pass's rounding remainder into the populations 6e-12.
def test_fp32_mass_drift_below_weight_rounding():
    edge, steps = 64, 400
    sol = W.spheres(edge, edge, edge, 5.0, 9.0, 0.4, 1000)
    prm = W.derive_params(theta_deg=60.0, acceleration=(1e-6, 0.0, 0.0))
    gain = {}
    for prec in ("f64", "f32"):
        e = MultiPhaseGPU(prec).init(sol, prm, {"use_mrt": 1, "correct_op_mass": 0})
        e.init_rest(W.slab_c(sol, 0, 0, edge // 2, 4.0), 1.0)
        m0 = e.current_mass()
        e.step(steps)
        gain[prec] = (e.current_mass() - m0, m0)
        e.free()
    rel = abs(gain["f32"][0] - gain["f64"][0]) / gain["f64"][1] / steps
    _say("fp32 - fp64 mass gain per step, relative: %.3e" % rel)
    assert rel < 5e-11


# The mass correction at an obtuse contact angle: sweep instance 42 (96^3 at
# porosity 0.46, 139 degrees, mobility coefficient 0.012).  Near the interface
# the correction per step is below half an ulp of the FP32 c; unless what the
# rounding drops is carried into the populations, the same rounding at the
# same sites every step left c off by 2.1e-4 after 1000 steps (1.1e-5 with it;
# measured 6.6e-7, p99 1.7e-7, once the moments pass's rounding remainder is
# carried too; FP64 runs from initial states 1e-7 apart stay within 1.2e-7 of
# each other, so the dynamics do not amplify).
def test_fp32_mass_correction_obtuse_angle_sweep_instance():
    from felbm_gpu_amd import serving
    spec = W.sweep_instances(48, 1000)[42]
    solid, prm, c0 = serving.sweep_problem(spec)
    _say("fp32 sweep instance 42: edge", spec["edge"], "theta %.1f" % spec["theta"])
    f = {}
    for prec in ("f64", "f32"):
        e = MultiPhaseGPU(prec).init(solid, prm, {"use_mrt": 1, "correct_op_mass": 1})
        e.init_rest(c0, 1.0)
        e.record_target_mass()
        e.step(1000)
        f[prec] = e.download()
        e.free()
    dc = np.abs(f["f32"]["c"] - f["f64"]["c"])
    du = max(np.abs(f["f32"][k] - f["f64"][k]).max() for k in ("ux", "uy", "uz"))
    _say("fp32 sweep instance 42, 1000 steps: max|dc| %.3e p99 %.3e max|du| %.3e" % (dc.max(), np.percentile(dc, 99), du))
    assert dc.max() <= 5e-6 and np.percentile(dc, 99) <= 1e-6 and du <= 1e-7


def test_batch_equals_alone_and_graph_equals_eager():
    prm = W.derive_params(theta_deg=60.0, acceleration=(1e-5, 0, 0))
    solids = [small_pack(16, 0.6, 21), small_pack(20, 0.5, 22), W.cylinders(40, 30, 1, 4, 3.0, 6.0, seed=9)[0]]
    sw = {"use_mrt": 1, "correct_op_mass": 1}
    b = Batch(precision="f64")
    mixed = [MultiPhaseGPU(batch=b).init(s, prm, sw) for s in solids]
    alone = [MultiPhaseGPU("f64").init(s, prm, sw) for s in solids]
    for s, m, a in zip(solids, mixed, alone):
        c0 = W.slab_c(s, 0, 0, s.shape[2] // 2, 4.0)
        m.init_rest(c0); a.init_rest(c0)
        m.record_target_mass(); a.record_target_mass()
    b.step_graph(4, 3)          # 12 steps through a captured graph
    for a in alone:
        a.step(12)
    for m, a in zip(mixed, alone):
        hm, gm = m.download_state()
        ha, ga = a.download_state()
        assert np.array_equal(hm, ha) and np.array_equal(gm, ga)


def test_upload_download_roundtrip_and_restart():
    solid = small_pack(16, 0.6, 31)
    prm = W.derive_params()
    o, g = make_pair(solid, prm, 1, 0)
    h, gg = g.download_state()
    ho, go = o.download_state()
    assert np.array_equal(h, ho) and np.array_equal(gg, go)
    g.step(5)
    h1, g1 = g.download_state()
    g2 = MultiPhaseGPU("f64").init(solid, prm, {"use_mrt": 1})
    g2.upload_state(h1, g1)
    g.step(5); g2.step(5)
    assert np.array_equal(g.download_state()[0], g2.download_state()[0])


@pytest.mark.parametrize("timedep", ["sinusoidal", "square"])
def test_time_dependent_forcing_bitexact(timedep):
    """The driver loop (src/felbm_gpu_main.cu:701-749): forcing_factor_at(t)
    set before every step with a body force in x; FP64 bit-identical."""
    from oracle.pyoracle import forcing_factor_at
    solid = small_pack(20, 0.55, 5)
    prm = W.derive_params(theta_deg=60.0, forcing=(2e-5, 0.0, 0.0))
    o, g = make_pair(solid, prm, True, True)
    T = 8.0
    _say("forcing schedule", timedep, "n", g.n)
    for t in range(3, 15):
        o.set_forcing_factor(forcing_factor_at(timedep, T, t))
        o.step()
        o.apply_mass_correction()
    g.run(3, 15, timedep, T)
    ho, go = o.download_state()
    hg, gg = g.download_state()
    assert np.array_equal(ho, hg) and np.array_equal(go, gg)
    fo, fg = o.download(), g.download()
    for name in ("rho", "c", "ux", "uy", "uz", "p"):
        assert np.array_equal(fo[name], fg[name]), name


def test_fp32_bgk_within_tolerance():
    w = W.workload("2d_cylinders_128")
    o, g = make_pair(w["solid"], w["params"], 0, 1, precision="f32", c0=w["c0"])
    dc, du = _fp32_errors(o, g, 20)
    _say("fp32 bgk 20 steps max|dc|", dc, "max|du|", du)
    assert dc <= 2e-6 and du <= 1e-7
    dc, du = _fp32_errors(o, g, 180)
    _say("fp32 bgk 200 steps max|dc|", dc, "max|du|", du)
    assert du <= 1e-7 and dc <= 3e-6


@pytest.mark.parametrize("use_mrt", [0, 1])
def test_nonperiodic_faces_bitexact(use_mrt):
    """a box closed along z (populations arriving from outside are zero,
    the streaming source -2 of the reference's tables)"""
    solid = small_pack(20, 0.6, 41)
    prm = W.derive_params(theta_deg=50.0, acceleration=(1e-5, 0, 0))
    o, g = make_pair(solid, prm, use_mrt, 1, periodic=(1, 1, 0))
    o.run(15)
    g.step(15)
    assert np.array_equal(o.download_state()[0], g.download_state()[0])
    assert np.array_equal(o.download_state()[1], g.download_state()[1])


def test_no_bounce_back_all_fluid_bitexact():
    solid = np.zeros((10, 12, 14), np.uint8)
    prm = W.derive_params(acceleration=(1e-5, 0, 0))
    c0 = W.slab_c(solid, 0, 0, 7, 4.0)
    o, g = make_pair(solid, prm, 1, 1, hbb=False, c0=c0)
    o.run(15)
    g.step(15)
    assert np.array_equal(o.download_state()[0], g.download_state()[0])
    assert np.array_equal(o.download_state()[1], g.download_state()[1])


@pytest.mark.parametrize("nx,per_x", [(3, 1), (33, 1), (65, 1), (34, 0), (64, 0)])
def test_row_word_windows_bitexact(nx, per_x):
    """Row lengths around the 32-voxel words of the neighbour table: rows of
    one, two and three words, x + dx crossing a word boundary and wrapping
    around a periodic x axis (or leaving a closed one), against the oracle."""
    rng = np.random.default_rng(nx)
    solid = (rng.random((8, 7, nx)) < 0.3).astype(np.uint8)
    solid[:, :, 0] = 0                       # fluid on both x faces
    solid[:, :, -1] = 0
    prm = W.derive_params(theta_deg=70.0, acceleration=(1e-5, 0, 0))
    c0 = W.slab_c(solid, 0, 0, max(nx // 2, 1), 2.0)
    _say("row windows nx", nx, "periodic x", per_x)
    o, g = make_pair(solid, prm, 1, 1, periodic=(per_x, 1, 1), c0=c0)
    o.run(6)
    g.step(6)
    assert np.array_equal(o.download_state()[0], g.download_state()[0])
    assert np.array_equal(o.download_state()[1], g.download_state()[1])
