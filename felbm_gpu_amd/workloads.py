"""Physical parameters, synthetic media and initial conditions of the solver
workloads.

Parameter derivation follows the reference's harness
(src/compare_cpu_gpu.cu:111-118) and phi = 6 sigma cos(theta)
(scripts/run_fig5_2d.sh:23):

    kappa = 1.5 W sigma,  beta = 12 sigma / W,  mobility = mobility_coeff / beta,
    boundary_coefficient = phi / kappa  (0 when kappa ~ 0)

Workloads (each generated from a fixed seed by the engine's own generators):
  2d_cylinders_128  128x128x1 periodic, 24 non-overlapping cylinders r~U[6,10],
                    BGK FP64, theta 60 deg, slab c=1 for x<64, 2000 steps
  porous_150        150^3 periodic overlapping spheres r~U[5,9] at porosity
                    0.40, largest pore cluster kept, MRT FP64, theta 45 deg
  porous_300        300^3 spheres at porosity 0.37, MRT FP32
  sweep_48          48 independent instances of edge 64/96/128/160
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check


def derive_params(sigma=0.01, width=4.0, mobility_coeff=0.02, theta_deg=60.0, rho0=1.0, rho1=1.0,
                  tau0=1.0, tau1=1.0, acceleration=(0.0, 0.0, 0.0), forcing=(0.0, 0.0, 0.0),
                  mrt_lambda=0.1875, forcing_factor=1.0):
    kappa = 1.5 * width * sigma
    beta = 12.0 * sigma / width
    phi = 6.0 * sigma * math.cos(math.radians(theta_deg))
    return {
        "rho0": rho0, "rho1": rho1, "kappa": kappa, "beta": beta,
        "boundary_coefficient": 0.0 if abs(kappa) < 1e-16 else phi / kappa,
        "tau0": tau0, "tau1": tau1,
        # with sigma = 0 (single phase) beta is 0 and the mobility term vanishes
        "mobility": mobility_coeff / beta if beta != 0.0 else 0.0,
        "forcing_factor": forcing_factor, "acceleration": tuple(acceleration), "forcing": tuple(forcing),
        "mrt_lambda": mrt_lambda,
    }


def cylinders(nx, ny, nz, count, rmin, rmax, seed, gap=0.0, max_tries=200000):
    lib = _lib.load()
    solid = np.zeros((nz, ny, nx), dtype=np.uint8)
    placed = check(lib.felbm_geom_cylinders(nx, ny, nz, count, rmin, rmax, gap, seed, max_tries,
                                            solid.ctypes.data_as(C.POINTER(C.c_uint8))))
    return solid, placed


def spheres(nx, ny, nz, rmin, rmax, porosity, seed, keep_largest=True, periodic=(1, 1, 1)):
    lib = _lib.load()
    solid = np.zeros((nz, ny, nx), dtype=np.uint8)
    check(lib.felbm_geom_spheres(nx, ny, nz, rmin, rmax, porosity, seed,
                                 solid.ctypes.data_as(C.POINTER(C.c_uint8))))
    if keep_largest:
        check(lib.felbm_keep_largest_cluster(nx, ny, nz, int(periodic[0]), int(periodic[1]), int(periodic[2]),
                                             solid.ctypes.data_as(C.POINTER(C.c_uint8))))
    return solid


def site_coords(solid):
    """(x, y, z) of every site in the engine's x-fastest order."""
    z, y, x = np.nonzero(solid == 0)
    return x, y, z


def slab_c(solid, axis=0, lo=0.0, hi=None, width=4.0):
    """Two-interface periodic slab: c = 1/2 (1 + tanh(2 d / W)), d the signed
    distance to the slab faces lo-1/2 and hi-1/2 (positive inside)."""
    nz, ny, nx = solid.shape
    L = (nx, ny, nz)[axis]
    hi = L / 2.0 if hi is None else hi
    coords = site_coords(solid)[axis].astype(np.float64)
    centre = 0.5 * (lo + hi) - 0.5
    half = 0.5 * (hi - lo)
    dd = np.abs(coords - centre)
    dd = np.minimum(dd, L - dd)
    d = half - dd
    return 0.5 * (1.0 + np.tanh(2.0 * d / width))


def open_channel(solid, empty_layers=2, copy_to_buffers=True, split_axis=0):
    """Open inlet/outlet lists for flow along y (the reference's SubDomain
    builds them in felbm_local, which is not available; this is the build's own
    rule).  The first and last `empty_layers` y-planes are cleared of solid.
    Without buffers the inlet is the plane y=0 and the outlet y=Ny-1; with
    copy_to_buffers the planes y=0 and y=Ny-1 are buffer nodes mirroring the
    same (x, z) node of y=1 and y=Ny-2, which are the inlet and outlet.
    Returns (solid, inlet, outlet, icoord, buf_verts, buf_refs); icoord is each
    inlet node's coordinate along split_axis."""
    s = np.array(solid, dtype=np.uint8, copy=True)
    nz, ny, nx = s.shape
    e = int(empty_layers)
    s[:, :e, :] = 0
    s[:, ny - e:, :] = 0
    x, y, z = site_coords(s)
    idx = -np.ones((nz, ny, nx), dtype=np.int64)
    idx[z, y, x] = np.arange(len(x))
    coords = (x, y, z)
    yin, yout = (1, ny - 2) if copy_to_buffers else (0, ny - 1)
    inlet = np.nonzero(y == yin)[0].astype(np.int32)
    outlet = np.nonzero(y == yout)[0].astype(np.int32)
    icoord = coords[split_axis][inlet].astype(np.float64)
    bv, br = np.zeros(0, np.int32), np.zeros(0, np.int32)
    if copy_to_buffers:
        b_in = np.nonzero(y == 0)[0]
        b_out = np.nonzero(y == ny - 1)[0]
        bv = np.concatenate([b_in, b_out]).astype(np.int32)
        br = np.concatenate([idx[z[b_in], 1, x[b_in]], idx[z[b_out], ny - 2, x[b_out]]]).astype(np.int32)
    return s, inlet, outlet, icoord, bv, br


def dm_groups(spec):
    """Per-tracer diffusivities from particles_dm_groups = "Dm:count, ..."
    (README.md:428-433)."""
    out = []
    for part in str(spec).split(","):
        part = part.strip()
        if not part:
            continue
        dm, cnt = part.split(":")
        out += [float(dm)] * int(cnt)
    return np.array(out, dtype=np.float64)


def seed_planes(solid, per_plane, offsets, axis=0, seed=0):
    """Tracers seeded uniformly on planes normal to `axis` at the given offsets
    (particles_init_mode = plane_*, README.md:435-442): each plane resamples
    until its quota of points whose nearest voxel is fluid is reached.
    Returns positions (n, 3) as (x, y, z) and the plane id of each."""
    nz, ny, nx = solid.shape
    dims = (nx, ny, nz)
    rng = np.random.default_rng(seed)
    pos, pid = [], []
    for k, off in enumerate(offsets):
        got = []
        need = int(per_plane)
        while need > 0:
            cand = np.empty((max(2 * need, 64), 3))
            for a in range(3):
                cand[:, a] = rng.uniform(-0.5, dims[a] - 0.5, len(cand))
            cand[:, axis] = float(off)
            v = np.rint(cand).astype(np.int64)
            for a in range(3):
                v[:, a] = np.clip(v[:, a], 0, dims[a] - 1)
            ok = solid[v[:, 2], v[:, 1], v[:, 0]] == 0
            take = cand[ok][:need]
            got.append(take)
            need -= len(take)
        pos.append(np.concatenate(got))
        pid += [k] * int(per_plane)
    return np.concatenate(pos), np.array(pid, dtype=np.int32)


def workload(name):
    """Geometry, params, switches and IC for a named workload."""
    if name == "2d_cylinders_128":
        solid, _ = cylinders(128, 128, 1, 24, 6.0, 10.0, seed=1)
        prm = derive_params(sigma=0.01, width=4.0, mobility_coeff=0.02, theta_deg=60.0,
                            acceleration=(1e-6, 0.0, 0.0))
        sw = {"use_mrt": 0, "correct_op_mass": 1}
        return dict(solid=solid, params=prm, switches=sw, c0=slab_c(solid, 0, 0, 64, 4.0), p0=1.0,
                    precision="f64", steps=2000)
    if name.startswith("porous_"):
        edge = int(name.split("_")[1])
        por, seed, prec, theta = {150: (0.40, 2, "f64", 45.0), 300: (0.37, 3, "f32", 45.0)}.get(
            edge, (0.40, 7, "f64", 45.0))
        solid = spheres(edge, edge, edge, 5.0, 9.0, por, seed)
        prm = derive_params(sigma=0.01, width=4.0, mobility_coeff=0.02, theta_deg=theta,
                            acceleration=(1e-6, 0.0, 0.0))
        sw = {"use_mrt": 1, "correct_op_mass": 1}
        return dict(solid=solid, params=prm, switches=sw, c0=slab_c(solid, 0, 0, edge // 2, 4.0), p0=1.0,
                    precision=prec, steps=5000 if edge == 150 else 10000)
    raise KeyError(name)


def sweep_instances(count=48, seed=1000, edges=(64, 96, 128, 160)):
    """The parameter sweep: per-instance edge, porosity, contact angle and
    mobility drawn from a seeded stream; instance i uses geometry seed+i.
    edges=(64,) gives the many-small-instances form of the same sweep."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        edge = int(rng.choice(list(edges)))
        por = float(rng.uniform(0.30, 0.50))
        theta = float(rng.uniform(30.0, 150.0))
        mob = float(math.exp(rng.uniform(math.log(0.005), math.log(0.05))))
        out.append(dict(edge=edge, porosity=por, theta=theta, mobility_coeff=mob, seed=seed + i))
    return out
