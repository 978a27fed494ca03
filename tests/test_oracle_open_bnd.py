"""CPU pins of the oracle's open inlet/outlet boundaries
(multiphase_gpu.cuh:369-446, device_engine.cuh:699-846).

The reference's boundary node lists come from felbm_local (absent), so the
lists here are the build's own (workloads.open_channel); what is pinned is
what the boundary does with them:
  - the injection schedule imposes c = (1-a) + (2a-1) gate on the inlet, with
    gate a square wave in time (alternate), a tanh profile across the inlet
    (split) or periodic tanh stripes (stripes), in closed form;
  - an imposed inlet velocity and the imposed pressures appear unchanged in
    the fields;
  - a uniform rest state with equal end pressures stays at rest;
  - a pressure-driven single-phase channel between walls reaches the plane
    Poiseuille mean velocity G H^2 / (12 nu) for the pressure gradient G it
    sets up (nu = tau cs^2), with a linear pressure profile.
"""
import math

import numpy as np
import pytest

from felbm_gpu_amd import workloads as W
from oracle.pyoracle import Oracle

PER = (1, 0, 1)


def walled_channel(nx=10, ny=40, copy=True):
    s = np.zeros((1, ny, nx), np.uint8)
    s[:, :, 0] = 1
    s[:, :, nx - 1] = 1
    return W.open_channel(s, 0, copy_to_buffers=copy, split_axis=0)


def test_alternate_injection_is_a_square_wave_in_time():
    s, inl, out, ic, bv, br = walled_channel(copy=False)
    ob = {"use_in_fluid": 1, "inlet_c_a": 1.0, "inlet_mode": 1, "inlet_period": 12.0, "inlet_duty": 0.5,
          "inlet_ramp": 0.0}
    o = Oracle(s, W.derive_params(theta_deg=90.0), use_mrt=True, periodic=PER, open_bnd=ob)
    o.set_open_bnd(inl, out, ic)
    o.init_rest(np.zeros(o.n), 1.0)
    for t in range(30):
        c_in = o.download()["c"][inl]
        expect = 1.0 if 3 <= (t % 12) < 9 else 0.0      # window [P/2 - dP/2, P/2 + dP/2)
        assert np.all(c_in == expect), t
        o.step()


def test_inlet_fluid_fraction_maps_to_c():
    s, inl, out, ic, bv, br = walled_channel(copy=False)
    for a in (0.0, 0.25, 1.0):
        ob = {"use_in_fluid": 1, "inlet_c_a": a, "inlet_mode": 0}
        o = Oracle(s, W.derive_params(theta_deg=90.0), periodic=PER, open_bnd=ob)
        o.set_open_bnd(inl, out, ic)
        o.init_rest(np.full(o.n, 0.5), 1.0)
        f = o.download()
        assert np.all(f["c"][inl] == (1.0 - a) + (2.0 * a - 1.0))
        rho = f["c"][inl] * 1.0 + (1.0 - f["c"][inl]) * 1.0
        assert np.all(f["rho"][inl] == rho)


@pytest.mark.parametrize("mode", [2, 3])
def test_split_and_stripes_profiles(mode):
    nx = 24
    s, inl, out, ic, bv, br = walled_channel(nx=nx, copy=True)
    w, pos, per, duty = 2.0, 9.0, 8.0, 0.25
    ob = {"use_in_fluid": 1, "inlet_c_a": 1.0, "inlet_mode": mode, "inlet_ramp": w, "split_pos": pos,
          "inlet_period": per, "inlet_duty": duty}
    o = Oracle(s, W.derive_params(theta_deg=90.0), periodic=PER, open_bnd=ob)
    o.set_open_bnd(inl, out, ic, bv, br, True)
    o.init_rest(np.zeros(o.n), 1.0)
    c_in = o.download()["c"][inl]
    if mode == 2:
        expect = 0.5 * (1.0 - np.tanh(2.0 * (ic - pos) / w))
    else:
        t = (ic - pos) - per * np.floor((ic - pos) / per)
        lo, hi = 0.5 * per - 0.5 * duty * per, 0.5 * per + 0.5 * duty * per
        expect = 0.5 * (np.tanh(2.0 * (t - lo) / w) - np.tanh(2.0 * (t - hi) / w))
    assert np.max(np.abs(c_in - expect)) < 1e-15
    # buffer nodes mirror their inlet vertex
    f = o.download()
    for k in ("c", "rho", "p", "ux", "uy", "uz"):
        assert np.array_equal(f[k][bv], f[k][br]), k


def test_imposed_velocity_and_pressures_survive_the_correction():
    s, inl, out, ic, bv, br = walled_channel(copy=True)
    ob = {"use_in_vel": 1, "use_in_fluid": 1, "use_out_prs": 1, "u_in": (0.0, 1e-3, 0.0), "p_out": 0.999,
          "inlet_c_a": 1.0}
    o = Oracle(s, W.derive_params(sigma=0.0), use_mrt=True, periodic=PER, open_bnd=ob)
    o.set_open_bnd(inl, out, ic, bv, br, True)
    o.init_rest(np.ones(o.n), 1.0)
    o.run(20)
    f = o.download()
    assert np.all(f["uy"][inl] == 1e-3) and np.all(f["ux"][inl] == 0.0)
    assert np.all(f["p"][out] == 0.999)


def test_rest_state_with_equal_pressures_is_a_fixed_point():
    s, inl, out, ic, bv, br = walled_channel(copy=True)
    ob = {"use_in_prs": 1, "use_in_fluid": 1, "use_out_prs": 1, "use_out_fluid": 1, "p_in": 1.0, "p_out": 1.0,
          "inlet_c_a": 1.0, "c_out_fixed": 1.0, "rho_out_fixed": 1.0}
    o = Oracle(s, W.derive_params(sigma=0.0), use_mrt=True, periodic=PER, open_bnd=ob)
    o.set_open_bnd(inl, out, ic, bv, br, True)
    o.init_rest(np.ones(o.n), 1.0)
    o.run(200)
    f = o.download()
    u = max(float(np.abs(f[k]).max()) for k in ("ux", "uy", "uz"))
    assert u < 1e-15 and float(np.abs(f["p"] - 1.0).max()) < 1e-14


def test_pressure_driven_channel_is_plane_poiseuille():
    nx, ny = 10, 40
    s, inl, out, ic, bv, br = walled_channel(nx, ny, copy=True)
    dp = 1e-4
    ob = {"use_in_prs": 1, "use_in_fluid": 1, "use_out_prs": 1, "use_out_fluid": 1, "p_in": 1.0 + dp,
          "p_out": 1.0, "inlet_c_a": 1.0, "c_out_fixed": 1.0, "rho_out_fixed": 1.0}
    o = Oracle(s, W.derive_params(sigma=0.0, tau0=1.0, tau1=1.0), use_mrt=True, periodic=PER, open_bnd=ob)
    o.set_open_bnd(inl, out, ic, bv, br, True)
    o.init_rest(np.ones(o.n), 1.0)
    o.run(8000)
    f = o.download()
    x, y, z = W.site_coords(s)
    pm = np.array([f["p"][y == j].mean() for j in range(ny)])
    assert pm[1] == 1.0 + dp and pm[ny - 2] == 1.0
    # interior pressure is linear: fit and check the residual
    js = np.arange(4, ny - 4)
    G, b = np.polyfit(js, pm[js], 1)
    assert np.max(np.abs(pm[js] - (G * js + b))) < 1e-3 * dp
    H, nu = nx - 2, 1.0 / 3.0
    u_mean = f["uy"][y == ny // 2].mean()
    u_theory = -G * H * H / (12.0 * nu)
    assert abs(u_mean / u_theory - 1.0) < 0.01, (u_mean, u_theory)
