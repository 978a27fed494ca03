"""Python mirror of ``felbm_gpu::MultiPhaseGPU`` over the C ABI.

Names and argument meanings follow ``include/felbm_gpu/multiphase_gpu.cuh``
of the reference (init, upload_state, step, apply_mass_correction,
record_target_mass, current_mass, flow_stats*, download, download_state,
free), so a driver or test written against the reference reads the same.
One ``MultiPhaseGPU`` is one slot of an engine handle; ``Batch`` exposes the
handle with many slots (the batched form).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

D3Q19_W = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)

# open-boundary settings (felbm_gpu_main.cu:340-364); keys of felbm_open_bnd
OPEN_BND_DEFAULTS = {"use_in_vel": 0, "use_in_prs": 0, "use_in_fluid": 0, "use_out_prs": 0, "use_out_fluid": 0,
                     "u_in": (0.0, 0.0, 0.0), "p_in": 1.0, "p_out": 1.0, "c_out_fixed": 0.0, "rho_out_fixed": 1.0,
                     "inlet_c_a": 1.0, "inlet_mode": 0, "split_axis": 0, "inlet_period": 0.0, "inlet_duty": 0.5,
                     "inlet_ramp": 0.0, "split_pos": 0.0}


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _geom(solid, periodic=(1, 1, 1), use_halfway_bb=True, site_xyz=None):
    """felbm_geometry of a solid mask; site_xyz [n][3] (x, y, z) is the caller's
    site numbering (None: x-fastest order).  Returns the struct and the arrays
    it points at (to be kept alive for the call)."""
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    nz, ny, nx = solid.shape
    g = _lib.Geometry()
    g.nx, g.ny, g.nz = nx, ny, nz
    for a in range(3):
        g.periodic[a] = int(bool(periodic[a]))
    g.use_halfway_bb = int(bool(use_halfway_bb))
    g.solid = solid.ctypes.data_as(C.POINTER(C.c_uint8))
    keep = [solid]
    if site_xyz is not None:
        xyz = np.ascontiguousarray(site_xyz, dtype=np.int32).reshape(-1, 3)
        g.site_xyz = xyz.ctypes.data_as(C.POINTER(C.c_int32))
        g.n_sites = xyz.shape[0]
        keep.append(xyz)
    return g, keep


def _params(p):
    q = _lib.Params()
    for k in ("rho0", "rho1", "kappa", "beta", "boundary_coefficient", "tau0", "tau1", "mobility",
              "forcing_factor", "mrt_lambda"):
        setattr(q, k, float(p[k]))
    for a in range(3):
        q.acceleration[a] = float(p["acceleration"][a])
        q.forcing[a] = float(p["forcing"][a])
    return q


def _switches(s):
    w = _lib.Switches()
    for k in ("use_mrt", "correct_op_mass", "stream_matrix_free", "grad_matrix_free", "fused",
              "fuse_collision", "mrt_fast_transform", "stream_inplace"):
        setattr(w, k, int(bool(s.get(k, 1 if k not in ("use_mrt", "correct_op_mass") else 0))))
    return w


class Batch:
    """An engine handle on one GPU holding any number of lattice instances."""

    def __init__(self, device=0, precision="f64"):
        self.lib = _lib.load()
        self.precision = {"f64": 8, "f32": 4, 8: 8, 4: 4}[precision]
        h = C.c_void_p()
        check(self.lib.felbm_create(C.byref(h), int(device), self.precision))
        self.h = h
        self._keep = {}

    def add(self, solid, params, switches, periodic=(1, 1, 1), use_halfway_bb=True, site_xyz=None):
        g, arr = _geom(solid, periodic, use_halfway_bb, site_xyz)
        slot = C.c_int()
        check(self.lib.felbm_add_instance(self.h, C.byref(g), C.byref(_params(params)),
                                          C.byref(_switches(switches)), C.byref(slot)))
        return slot.value

    def remove(self, slot):
        check(self.lib.felbm_remove_instance(self.h, int(slot)))

    def n_sites(self, slot):
        n = C.c_int64()
        check(self.lib.felbm_num_sites(self.h, int(slot), C.byref(n)))
        return n.value

    def step(self, nsteps=1, stats_last=False):
        check(self.lib.felbm_step(self.h, int(nsteps), int(bool(stats_last))))

    def step_graph(self, nsteps, replays):
        check(self.lib.felbm_step_graph(self.h, int(nsteps), int(replays)))

    def synchronize(self):
        check(self.lib.felbm_synchronize(self.h))

    def kernel_times(self, nsteps):
        ms = (C.c_float * 4)()
        check(self.lib.felbm_kernel_times(self.h, int(nsteps), ms))
        return list(ms)

    def autotune(self, nsteps=10):
        """Pick the sites per workgroup of each pass (64, 128, 256) by timing
        them on the instances present; the state is left as it was."""
        bs = (C.c_int * 4)()
        ms = (C.c_float * 12)()
        check(self.lib.felbm_autotune(self.h, int(nsteps), bs, ms))
        passes = ("moments", "cfield", "mufield", "collide")
        return {"block_sizes": dict(zip(passes, list(bs))),
                "ms": {str(b): dict(zip(passes, list(ms)[4 * q:4 * q + 4])) for q, b in enumerate((64, 128, 256))}}

    def set_block_sizes(self, sizes):
        """Sites per workgroup of moments, c-stencil, mu-stencil, collide."""
        bs = (C.c_int * 4)(*[int(v) for v in sizes])
        check(self.lib.felbm_set_block_sizes(self.h, bs))

    def free(self):
        if self.h:
            self.lib.felbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MultiPhaseGPU:
    """One lattice on one GPU: the drop-in for the reference's MultiPhaseGPU."""

    def __init__(self, precision="f64", device=0, batch=None):
        self.batch = batch if batch is not None else Batch(device, precision)
        self.lib = self.batch.lib
        self.slot = None
        self.n = 0
        self.Vn = 0
        self.correct_mass = False
        self.solid = None
        self._tr = None          # host copy of the tracer set-up (dm, plane_id) for particle files
        self.site_xyz = None
        self._stretch = False

    # multiphase_gpu.cuh:75-351; site_xyz [n][3] is the caller's site numbering
    # (SubDomain::idx_to_coords order, src/felbm_gpu_main.cu:124-143)
    def init(self, solid, params, switches, periodic=(1, 1, 1), use_halfway_bb=True, site_xyz=None):
        self.slot = self.batch.add(solid, params, switches, periodic, use_halfway_bb, site_xyz)
        self.n = self.batch.n_sites(self.slot)
        self.Vn = 19 * self.n
        self.correct_mass = bool(switches.get("correct_op_mass", False))
        self.solid = np.asarray(solid)
        self.site_xyz = None if site_xyz is None else np.asarray(site_xyz, dtype=np.int32).reshape(-1, 3)
        return self

    @property
    def h(self):
        return self.batch.h

    # :354-359
    def upload_state(self, h, g):
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        if h.size != self.Vn or g.size != self.Vn:
            raise ValueError("state arrays must hold 19*n values")
        check(self.lib.felbm_upload_state(self.h, self.slot, _dptr(h), _dptr(g)))

    def init_rest(self, c_site, p0=1.0):
        c = np.ascontiguousarray(c_site, dtype=np.float64)
        if c.size != self.n:
            raise ValueError("c must hold n values")
        check(self.lib.felbm_init_rest(self.h, self.slot, _dptr(c), float(p0)))

    # :449-455
    def download_state(self):
        h = np.empty(self.Vn)
        g = np.empty(self.Vn)
        check(self.lib.felbm_download_state(self.h, self.slot, _dptr(h), _dptr(g)))
        return h, g

    def state_size(self):
        return self.Vn

    # :522 + :689 (the correction runs inside when correct_op_mass is on)
    def step(self, nsteps=1, stats_last=False):
        self.batch.step(nsteps, stats_last)

    def apply_mass_correction(self):
        """Kept for interface parity: the correction is applied by step()."""
        return 0.0

    def set_forcing_factor(self, ff):
        check(self.lib.felbm_set_forcing_factor(self.h, self.slot, float(ff)))

    # :369-400, with the DevParams open-BC block of felbm_gpu_main.cu:340-364
    def set_open_bnd(self, ob, inlet, outlet, icoord=None, buf_verts=(), buf_refs=(), copy_to_buffers=False):
        f = dict(OPEN_BND_DEFAULTS)
        f.update(ob)
        s = _lib.OpenBnd()
        for key in ("use_in_vel", "use_in_prs", "use_in_fluid", "use_out_prs", "use_out_fluid", "inlet_mode",
                    "split_axis"):
            setattr(s, key, int(f[key]))
        for a in range(3):
            s.u_in[a] = float(f["u_in"][a])
        for key in ("p_in", "p_out", "c_out_fixed", "rho_out_fixed", "inlet_c_a", "inlet_period", "inlet_duty",
                    "inlet_ramp", "split_pos"):
            setattr(s, key, float(f[key]))
        inl = np.ascontiguousarray(inlet, dtype=np.int32)
        out = np.ascontiguousarray(outlet, dtype=np.int32)
        bv = np.ascontiguousarray(buf_verts, dtype=np.int32)
        br = np.ascontiguousarray(buf_refs, dtype=np.int32)
        ic = None if icoord is None else np.ascontiguousarray(icoord, dtype=np.float64)
        pi = C.POINTER(C.c_int32)
        check(self.lib.felbm_set_open_bnd(self.h, self.slot, C.byref(s), inl.ctypes.data_as(pi), len(inl),
                                          out.ctypes.data_as(pi), len(out), None if ic is None else _dptr(ic),
                                          bv.ctypes.data_as(pi), br.ctypes.data_as(pi), len(bv),
                                          int(bool(copy_to_buffers))))

    # tracers: ParticleManager (felbm_gpu_main.cu:426-438, :752-777)
    def tracers_init(self, xyz, dm, seed=0, substeps=1, velocity_skip=1, plane_id=None):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        dm = np.ascontiguousarray(dm, dtype=np.float64).reshape(-1)
        assert len(dm) == len(xyz)
        check(self.lib.felbm_tracers_init(self.h, self.slot, len(dm), _dptr(xyz), _dptr(dm), int(seed),
                                          int(substeps), int(velocity_skip)))
        self._stretch = False
        self._tr = None if len(dm) == 0 else {
            "dm": dm.copy(), "plane_id": None if plane_id is None else np.asarray(plane_id, dtype=np.int32).copy()}

    def has_tracers(self):
        return self._tr is not None

    def tracers_set_step(self, step):
        check(self.lib.felbm_tracers_set_step(self.h, self.slot, int(step)))

    def tracers_download(self):
        np_ = C.c_int64()
        check(self.lib.felbm_tracers_count(self.h, self.slot, C.byref(np_)))
        n = np_.value
        xyz, uvw = np.empty((n, 3)), np.empty((n, 3))
        blk = np.empty(n, dtype=np.int32)
        check(self.lib.felbm_tracers_download(self.h, self.slot, _dptr(xyz), _dptr(uvw),
                                              blk.ctypes.data_as(C.POINTER(C.c_int32))))
        return xyz, uvw, blk

    # particles_stretching (docs/HANDOFF_stretching.md:165-210)
    def tracers_stretching(self, enable=True, rho_hat=None, log_rho=None, seed=20260801, phase_split=True,
                           capacity=4096):
        rh = None if rho_hat is None else np.ascontiguousarray(rho_hat, dtype=np.float64).reshape(-1)
        lr = None if log_rho is None else np.ascontiguousarray(log_rho, dtype=np.float64).reshape(-1)
        check(self.lib.felbm_tracers_stretching(self.h, self.slot, int(bool(enable)),
                                                None if rh is None else _dptr(rh), None if lr is None else _dptr(lr),
                                                int(seed), int(bool(phase_split)), int(capacity)))
        self._stretch = bool(enable)

    def tracers_stretch_rows(self, max_rows=100000):
        """take_stretch_rows: (rows [k][9], lost); columns step, lambda, mean_log_rho,
        var_log_rho, n_active, lambda_w, lambda_nw, n_w, n_nw."""
        buf = np.empty((int(max_rows), 9))
        cnt, lost = C.c_int(), C.c_int64()
        check(self.lib.felbm_tracers_stretch_rows(self.h, self.slot, _dptr(buf), int(max_rows), C.byref(cnt),
                                                  C.byref(lost)))
        return buf[:cnt.value].copy(), lost.value

    def tracers_stretch_state(self):
        np_ = C.c_int64()
        check(self.lib.felbm_tracers_count(self.h, self.slot, C.byref(np_)))
        rh, lr = np.empty((np_.value, 3)), np.empty(np_.value)
        check(self.lib.felbm_tracers_stretch_state(self.h, self.slot, _dptr(rh), _dptr(lr)))
        return rh, lr

    def run(self, t_begin, t_end, timedep="constant", period=1.0, output=None):
        """The driver's stepping loop (src/felbm_gpu_main.cu:701-749): for t in
        [t_begin, t_end) set forcing_factor_at(t), step, apply the mass
        correction.  With an ``h5io.OutputSchedule`` its files for step t are
        written before that step, and for t_end after the last one."""
        for t in range(int(t_begin), int(t_end) + (1 if output is not None else 0)):
            if output is not None:
                output.before_step(self, t)
            if t < int(t_end):
                self.set_forcing_factor(forcing_factor_at(timedep, period, t))
                self.step()
                self.apply_mass_correction()
        if output is not None:
            output.finish(self)

    # ---- HDF5 files (src/felbm_gpu_main.cu:47-189, :276, :395-410, :566-575, :706-770)
    def write_fields_h5(self, path, deflate=0, f32=False):
        """<output_name>_<t>.h5 of the current state (download_raw + write_fields_h5)."""
        from . import h5io
        h5io.write_fields(path, self.download(), deflate, f32)

    def write_geometry_h5(self, path, deflate=0):
        from . import h5io
        coords = h5io.geometry_coords(self.solid) if self.site_xyz is None else self.site_xyz
        h5io.write_geometry(path, coords, deflate)

    def write_checkpoint(self, path, step):
        from . import h5io
        h, g = self.download_state()
        h5io.write_checkpoint(path, h, g, step, self.get_target_mass())

    def read_checkpoint(self, path):
        """Restores h, g and the original target mass; returns the step."""
        from . import h5io
        h, g, step, tm = h5io.read_checkpoint(path, self.Vn)
        self.upload_state(h, g)
        self.set_target_mass(tm)
        return step

    def write_particles_h5(self, path):
        from . import h5io
        if self._tr is None:
            raise RuntimeError("no tracers on this lattice")
        xyz, uvw, _ = self.tracers_download()
        rh, lr = self.tracers_stretch_state() if self._stretch else (None, None)
        h5io.write_particles(path, xyz, uvw, None, self._tr["dm"], self._tr["plane_id"], rh, lr)

    def restore_particles_h5(self, path, step, seed=0, substeps=1, velocity_skip=1, dm=None, phase_split=True,
                             capacity=4096):
        """Tracers from particles_<step>.h5, in id order: positions, Dm (the
        file's, else ``dm``), plane_id, and -- when the file holds them --
        rho_hat/log_rho with stretching re-enabled.  The tracer step counter is
        set to ``step``."""
        from . import h5io
        f = h5io.read_particles(path)
        order = np.argsort(f["id"], kind="stable") if "id" in f else np.arange(len(f["position"]))
        dmv = f["Dm"][order] if "Dm" in f else np.asarray(dm, dtype=np.float64)
        pl = f["plane_id"][order] if "plane_id" in f else None
        self.tracers_init(f["position"][order], dmv, seed, substeps, velocity_skip, plane_id=pl)
        self.tracers_set_step(step)
        if "rho_hat" in f:
            self.tracers_stretching(True, f["rho_hat"][order], f["log_rho"][order], phase_split=phase_split,
                                    capacity=capacity)

    def record_target_mass(self):
        check(self.lib.felbm_record_target_mass(self.h, self.slot))

    def get_target_mass(self):
        m = C.c_double()
        check(self.lib.felbm_get_target_mass(self.h, self.slot, C.byref(m)))
        return m.value

    def set_target_mass(self, m):
        check(self.lib.felbm_set_target_mass(self.h, self.slot, float(m)))

    def current_mass(self):
        m = C.c_double()
        check(self.lib.felbm_current_mass(self.h, self.slot, C.byref(m)))
        return m.value

    # :645-683
    def flow_stats(self):
        out = np.empty(6)
        check(self.lib.felbm_flow_stats(self.h, self.slot, _dptr(out)))
        return out

    def flow_stats_launch(self):
        check(self.lib.felbm_flow_stats_launch(self.h))

    def flow_stats_ready(self):
        r = C.c_int()
        check(self.lib.felbm_flow_stats_ready(self.h, C.byref(r)))
        return bool(r.value)

    def flow_stats_collect(self):
        out = np.empty(6)
        check(self.lib.felbm_flow_stats_collect(self.h, self.slot, _dptr(out)))
        return out

    # :704-757 -> dict of rho, c, ux, uy, uz, p
    def download(self):
        out = np.empty(6 * self.n)
        check(self.lib.felbm_download_fields(self.h, self.slot, _dptr(out)))
        f = out.reshape(6, self.n)
        return {"rho": f[0], "c": f[1], "ux": f[2], "uy": f[3], "uz": f[4], "p": f[5]}

    def export_tables(self, noalloc=False):
        """Reference-format tables; noalloc stages them in the library's static
        device workspace (felbm_export_tables_noalloc) instead of a device
        allocation."""
        n = self.n
        names = ("cdm", "cdp", "bcase", "bc1", "bc2", "avgnext", "lcase", "lc1", "lc2", "src")
        sizes = (18, 18, 18, 18, 18, 19, 18, 18, 18, 19)
        arrs = [np.empty(s * n, dtype=np.int32) for s in sizes]
        fn = self.lib.felbm_export_tables_noalloc if noalloc else self.lib.felbm_export_tables
        check(fn(self.h, self.slot, *[a.ctypes.data_as(C.POINTER(C.c_int32)) for a in arrs]))
        return dict(zip(names, arrs))

    def free(self):
        if self.slot is not None:
            self.batch.remove(self.slot)
            self.slot = None


def forcing_factor_at(timedep, period, t):
    """forcing_factor_at(t) of the driver (src/felbm_gpu_main.cu:587-594)."""
    ff = C.c_double()
    check(_lib.load().felbm_forcing_factor_at(str(timedep).encode(), float(period), int(t), C.byref(ff)))
    return ff.value


def mrt_matrices():
    lib = _lib.load()
    M = np.empty((19, 19))
    Mi = np.empty((19, 19))
    check(lib.felbm_mrt_matrices(_dptr(M), _dptr(Mi)))
    return M, Mi
