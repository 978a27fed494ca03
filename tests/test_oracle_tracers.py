"""CPU pins of the tracer restatement (oracle/tracers.py).

ParticleManager is not available, so the tracer update is pinned by what the
mathematics fixes (the analytic cases of docs/HANDOFF_stretching.md:252-289
and README.md:428-446):
  - a uniform velocity moves every tracer by U t (trilinear weights sum to 1);
  - a linear shear u_x = G y is interpolated exactly, so x(t) = x0 + G y0 t;
  - with u = 0 the Brownian displacement has mean 0 and variance 2 Dm t per
    axis, and no z displacement in 2D;
  - Dm = 0 tracers draw nothing and are deterministic;
  - no tracer ever ends in a solid voxel (wall blocking), and blocked moves
    are counted;
  - the normal generator is deterministic per (seed, tracer, draw) with mean 0
    and variance 1;
  - stretching (Eq. 12): in a uniform shear log|rho| follows the closed form of
    the composed shear and the first row's lambda is the substep mean of
    G rho_x rho_y; in solid-body rotation lambda vanishes while explicit Euler
    grows |rho| by sqrt(1 + (Omega dt)^2) per substep; the wetting and
    non-wetting rows recombine to the total.
"""
import numpy as np

from felbm_gpu_amd import workloads as W
from oracle.tracers import Tracers, normals3


def site_u(solid, fn):
    x, y, z = W.site_coords(solid)
    return fn(x.astype(float), y.astype(float), z.astype(float))


def test_uniform_velocity_translates_exactly():
    solid = np.zeros((6, 7, 8), np.uint8)
    rng = np.random.default_rng(1)
    p0 = rng.uniform(0, 5.5, (200, 3))
    U = (0.013, -0.007, 0.002)
    t = Tracers(solid, (1, 1, 1), p0, np.zeros(200), substeps=3)
    n = site_u(solid, lambda x, y, z: x * 0.0)
    for _ in range(40):
        t.step(n + U[0], n + U[1], n + U[2])
    L = np.array([8.0, 7.0, 6.0])
    d = np.mod(t.pos - p0 - 40 * np.array(U) + L / 2, L) - L / 2
    assert np.abs(d).max() < 1e-12


def test_linear_shear_is_interpolated_exactly():
    solid = np.zeros((4, 20, 10), np.uint8)        # y non-periodic, no solid
    G = 1e-3
    rng = np.random.default_rng(2)
    p0 = np.column_stack([rng.uniform(0, 9, 100), rng.uniform(1, 18, 100), rng.uniform(0, 3, 100)])
    t = Tracers(solid, (1, 0, 1), p0, np.zeros(100))
    z0 = site_u(solid, lambda x, y, z: 0.0 * x)
    ux = site_u(solid, lambda x, y, z: G * y)
    for _ in range(50):
        t.step(ux, z0, z0)
    dx = np.mod(t.pos[:, 0] - p0[:, 0] + 5.0, 10.0) - 5.0
    assert np.abs(dx - G * p0[:, 1] * 50).max() < 1e-12
    assert np.array_equal(t.pos[:, 1:], p0[:, 1:])


def test_brownian_variance_and_2d_suppression():
    solid = np.zeros((1, 64, 64), np.uint8)
    n, steps, dm = 20000, 50, 0.02
    p0 = np.column_stack([np.full(n, 32.0), np.full(n, 32.0), np.zeros(n)])
    t = Tracers(solid, (1, 1, 1), p0, np.full(n, dm), seed=7, substeps=2)
    z = site_u(solid, lambda x, y, z: 0.0 * x)
    for _ in range(steps):
        t.step(z, z, z)
    d = t.pos - p0
    var = 2 * dm * steps
    for a in (0, 1):
        assert abs(d[:, a].mean()) < 4 * np.sqrt(var / n)
        assert abs(d[:, a].var() / var - 1.0) < 4 * np.sqrt(2.0 / n)
    assert np.all(d[:, 2] == 0.0)


def test_zero_diffusivity_is_deterministic_and_draws_nothing():
    solid = np.zeros((3, 8, 8), np.uint8)
    p0 = np.full((10, 3), 2.25)
    u = site_u(solid, lambda x, y, z: 1e-3 + 0.0 * x)
    a = Tracers(solid, (1, 1, 1), p0, np.zeros(10), seed=1)
    b = Tracers(solid, (1, 1, 1), p0, np.zeros(10), seed=99)
    for _ in range(5):
        a.step(u, u, u)
        b.step(u, u, u)
    assert np.array_equal(a.pos, b.pos)


def test_wall_blocking_keeps_tracers_in_the_pore_space():
    solid = W.spheres(24, 24, 24, 3.0, 5.0, 0.5, 4)
    rng = np.random.default_rng(3)
    x, y, z = W.site_coords(solid)
    pick = rng.choice(len(x), 500)
    p0 = np.column_stack([x[pick], y[pick], z[pick]]).astype(float)
    t = Tracers(solid, (1, 1, 1), p0, np.full(500, 0.3), seed=5)
    zero = 0.0 * x
    for _ in range(30):
        t.step(zero, zero, zero)
    v = np.mod(np.rint(t.pos).astype(int), 24)
    assert np.all(solid[v[:, 2], v[:, 1], v[:, 0]] == 0)
    assert t.blocked.sum() > 0


def test_normals_are_deterministic_standard_normal():
    ids = np.arange(100000, dtype=np.uint64)
    a = normals3(42, ids, 3)
    b = normals3(42, ids, 3)
    c = normals3(42, ids, 4)
    for q in range(3):
        assert np.array_equal(a[q], b[q])
        assert not np.array_equal(a[q], c[q])
        assert abs(a[q].mean()) < 0.02 and abs(a[q].var() - 1.0) < 0.02
    assert abs(np.corrcoef(a[0], a[1])[0, 1]) < 0.02


# ---- Lagrangian stretching (docs/HANDOFF_stretching.md:165-210, :252-289) ----

def _unit(n, seed):
    v = np.random.default_rng(seed).normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def test_stretching_uniform_shear_is_exact():
    # u_x = G y: J = [[0, G, 0], 0, 0] exactly; S substeps of rho* = rho + dt J rho
    # compose to one shear by G t, so log|rho| is known in closed form
    solid = np.zeros((4, 24, 10), np.uint8)
    G, S, steps = 2e-3, 3, 40
    rng = np.random.default_rng(5)
    p0 = np.column_stack([rng.uniform(0, 9, 200), rng.uniform(6, 17, 200), rng.uniform(0, 3, 200)])
    r0 = _unit(200, 6)
    t = Tracers(solid, (1, 0, 1), p0, np.zeros(200), substeps=S, rho_hat=r0)
    z0 = site_u(solid, lambda x, y, z: 0.0 * x)
    ux = site_u(solid, lambda x, y, z: G * y)
    for _ in range(steps):
        t.step(ux, z0, z0)
    T = float(steps)
    expect = 0.5 * np.log((r0[:, 0] + T * G * r0[:, 1]) ** 2 + r0[:, 1] ** 2 + r0[:, 2] ** 2)
    assert np.abs(t.log_rho - expect).max() < 1e-13
    assert np.abs(np.linalg.norm(t.rho_hat, axis=1) - 1.0).max() < 1e-15
    # first row: lambda = < rho_x rho_y G > averaged over the substeps of step 1
    lam = np.zeros(200)
    rh = r0.copy()
    for _ in range(S):
        lam += G * rh[:, 0] * rh[:, 1]
        rh = rh + (1.0 / S) * np.column_stack([G * rh[:, 1], 0 * rh[:, 1], 0 * rh[:, 1]])
        rh /= np.linalg.norm(rh, axis=1)[:, None]
    assert abs(t.rows[0][1] - (lam / S).mean()) < 1e-15
    assert t.rows[0][0] == 1 and t.rows[-1][0] == steps and t.rows[0][4] == 200


def test_stretching_rotation_has_zero_lambda_and_euler_growth():
    # solid-body rotation: J antisymmetric, so rho.J rho vanishes; explicit Euler
    # still inflates |rho| by sqrt(1 + (Omega dt)^2) per substep for rho in the plane
    solid = np.zeros((3, 40, 40), np.uint8)
    Om, S, steps = 1e-3, 2, 30
    p0 = np.column_stack([np.full(50, 19.3), np.full(50, 20.6), np.ones(50)])
    r0 = _unit(50, 7)
    r0[:, 2] = 0.0
    r0 /= np.linalg.norm(r0, axis=1)[:, None]
    t = Tracers(solid, (0, 0, 1), p0, np.zeros(50), substeps=S, rho_hat=r0)
    ux = site_u(solid, lambda x, y, z: -Om * (y - 20.0))
    uy = site_u(solid, lambda x, y, z: Om * (x - 20.0))
    z0 = site_u(solid, lambda x, y, z: 0.0 * x)
    for _ in range(steps):
        t.step(ux, uy, z0)
    assert max(abs(r[1]) for r in t.rows) < 1e-15
    expect = steps * S * 0.5 * np.log1p((Om / S) ** 2)
    assert np.abs(t.log_rho - expect).max() < 1e-12


def test_stretching_phase_split_recombines():
    solid = np.zeros((3, 16, 16), np.uint8)
    rng = np.random.default_rng(8)
    n = 400
    p0 = np.column_stack([rng.uniform(0, 15, n), rng.uniform(0, 15, n), rng.uniform(0, 2, n)])
    t = Tracers(solid, (1, 1, 1), p0, np.full(n, 1e-3), seed=3, rho_hat=_unit(n, 9))
    x, y, z = W.site_coords(solid)
    ux = 1e-3 * np.sin(2 * np.pi * y / 16.0)
    uy = 1e-3 * np.cos(2 * np.pi * x / 16.0)
    c = (x < 8).astype(float)
    for _ in range(10):
        t.step(ux, uy, 0.0 * x, c)
    for r in t.rows:
        step, lam, mean, var, na, lw, lnw, nw, nnw = r
        assert nw + nnw == na and nw > 0 and nnw > 0
        assert abs((nw * lw + nnw * lnw) / na - lam) < 1e-15
