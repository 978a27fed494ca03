"""ctypes wrapper of the FP64 CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the oracle.  The product never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")


class OrcParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("rho0", "rho1", "kappa", "beta", "bnd_coeff", "tau0", "tau1",
                                          "mobility", "forcing_factor", "gx", "gy", "gz", "fx", "fy", "fz",
                                          "mrt_lambda")] + \
               [(k, C.c_int) for k in ("open_bnd", "n_inlet", "n_outlet", "n_buf", "copy_to_buf", "use_in_vel",
                                       "use_in_prs", "use_in_fluid", "use_out_prs", "use_out_fluid")] + \
               [(k, C.c_double) for k in ("u_in_x", "u_in_y", "u_in_z", "p_in", "p_out", "c_out_fixed",
                                          "rho_out_fixed")] + \
               [("inlet_mode", C.c_int), ("split_axis", C.c_int)] + \
               [(k, C.c_double) for k in ("inlet_c_a", "inlet_period", "inlet_duty", "inlet_ramp", "split_pos")]


class OrcSwitches(C.Structure):
    _fields_ = [("use_mrt", C.c_int), ("fuse_collision", C.c_int), ("stream_inplace", C.c_int),
                ("mass_mode", C.c_int)]


_lib = None
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int32)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    sig = {
        "orc_new": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.POINTER(OrcParams), C.POINTER(OrcSwitches)]),
        "orc_delete": (None, [C.c_void_p]),
        "orc_n": (C.c_int, [C.c_void_p]),
        "orc_inplace_active": (C.c_int, [C.c_void_p]),
        "orc_upload_state": (None, [C.c_void_p, _PD, _PD]),
        "orc_download_state": (None, [C.c_void_p, _PD, _PD]),
        "orc_init_rest": (None, [C.c_void_p, _PD, C.c_double]),
        "orc_step": (None, [C.c_void_p]),
        "orc_run": (None, [C.c_void_p, C.c_int]),
        "orc_record_target_mass": (None, [C.c_void_p]),
        "orc_current_mass": (C.c_double, [C.c_void_p]),
        "orc_apply_mass_correction": (C.c_double, [C.c_void_p]),
        "orc_mass_weight": (None, [C.c_void_p, _PD, _PD]),
        "orc_flow_stats": (None, [C.c_void_p, _PD]),
        "orc_download_fields": (None, [C.c_void_p, _PD]),
        "orc_get_tables": (None, [C.c_void_p] + [_PI] * 10),
        "orc_num_ops": (C.c_int64, [C.c_void_p]),
        "orc_verify_inplace": (C.c_int, [C.c_void_p]),
        "orc_field": (_PD, [C.c_void_p, C.c_char_p]),
        "orc_fixed_sum": (C.c_double, [_PD, C.c_int64]),
        "orc_d3q19": (None, [_PI, _PI, _PI, _PI, _PD]),
        "orc_mrt_matrices": (None, [_PD, _PD]),
        "orc_set_forcing_factor": (None, [C.c_void_p, C.c_double]),
        "orc_forcing_factor_at": (C.c_double, [C.c_char_p, C.c_double, C.c_longlong]),
        "orc_get_target_mass": (C.c_double, [C.c_void_p]),
        "orc_set_target_mass": (None, [C.c_void_p, C.c_double]),
        "orc_get_coords": (None, [C.c_void_p, _PI]),
        "orc_set_open_bnd": (C.c_int, [C.c_void_p, _PI, C.c_int, _PI, C.c_int, _PD, _PI, _PI, C.c_int, C.c_int]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


OPEN_BND_DEFAULTS = {"use_in_vel": 0, "use_in_prs": 0, "use_in_fluid": 0, "use_out_prs": 0, "use_out_fluid": 0,
                     "u_in": (0.0, 0.0, 0.0), "p_in": 1.0, "p_out": 1.0, "c_out_fixed": 0.0, "rho_out_fixed": 1.0,
                     "inlet_c_a": 1.0, "inlet_mode": 0, "split_axis": 0, "inlet_period": 0.0, "inlet_duty": 0.5,
                     "inlet_ramp": 0.0, "split_pos": 0.0}


def open_bnd_fields(ob):
    out = dict(OPEN_BND_DEFAULTS)
    out.update(ob)
    return out


def forcing_factor_at(timedep, period, t):
    return load().orc_forcing_factor_at(timedep.encode(), float(period), int(t))


def _d(a):
    return a.ctypes.data_as(_PD)


def to_orc_params(p):
    q = OrcParams()
    q.rho0, q.rho1, q.kappa, q.beta = p["rho0"], p["rho1"], p["kappa"], p["beta"]
    q.bnd_coeff, q.tau0, q.tau1, q.mobility = p["boundary_coefficient"], p["tau0"], p["tau1"], p["mobility"]
    q.forcing_factor, q.mrt_lambda = p["forcing_factor"], p["mrt_lambda"]
    q.gx, q.gy, q.gz = p["acceleration"]
    q.fx, q.fy, q.fz = p["forcing"]
    return q


class Oracle:
    """FP64 CPU restatement of MultiPhaseGPU (the oracle)."""

    def __init__(self, solid, params, use_mrt=False, fuse_collision=True, stream_inplace=False, mass_mode=0,
                 periodic=(1, 1, 1), use_halfway_bb=True, open_bnd=None):
        self.lib = load()
        self.solid = np.ascontiguousarray(solid, dtype=np.uint8)
        nz, ny, nx = self.solid.shape
        sw = OrcSwitches(int(use_mrt), int(fuse_collision), int(stream_inplace), int(mass_mode))
        self._p = to_orc_params(params)
        if open_bnd is not None:
            ob = open_bnd_fields(open_bnd)
            self._p.open_bnd = 1
            for key in ("use_in_vel", "use_in_prs", "use_in_fluid", "use_out_prs", "use_out_fluid", "inlet_mode",
                        "split_axis"):
                setattr(self._p, key, int(ob[key]))
            self._p.u_in_x, self._p.u_in_y, self._p.u_in_z = ob["u_in"]
            for key in ("p_in", "p_out", "c_out_fixed", "rho_out_fixed", "inlet_c_a", "inlet_period", "inlet_duty",
                        "inlet_ramp", "split_pos"):
                setattr(self._p, key, float(ob[key]))
        self.E = self.lib.orc_new(nx, ny, nz, self.solid.ctypes.data_as(C.POINTER(C.c_uint8)), int(use_halfway_bb),
                                  int(periodic[0]), int(periodic[1]), int(periodic[2]), C.byref(self._p), C.byref(sw))
        self.n = self.lib.orc_n(self.E)
        self.Vn = 19 * self.n

    def __del__(self):
        try:
            if self.E:
                self.lib.orc_delete(self.E)
                self.E = None
        except Exception:
            pass

    def init_rest(self, c, p0=1.0):
        c = np.ascontiguousarray(c, dtype=np.float64)
        self.lib.orc_init_rest(self.E, _d(c), float(p0))

    def upload_state(self, h, g):
        h = np.ascontiguousarray(h, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        self.lib.orc_upload_state(self.E, _d(h), _d(g))

    def download_state(self):
        h = np.empty(self.Vn)
        g = np.empty(self.Vn)
        self.lib.orc_download_state(self.E, _d(h), _d(g))
        return h, g

    def step(self):
        self.lib.orc_step(self.E)

    def run(self, nsteps):
        self.lib.orc_run(self.E, int(nsteps))

    def record_target_mass(self):
        self.lib.orc_record_target_mass(self.E)

    def current_mass(self):
        return self.lib.orc_current_mass(self.E)

    def apply_mass_correction(self):
        return self.lib.orc_apply_mass_correction(self.E)

    def mass_weight(self):
        M, W = C.c_double(), C.c_double()
        self.lib.orc_mass_weight(self.E, C.byref(M), C.byref(W))
        return M.value, W.value

    def flow_stats(self):
        out = np.empty(6)
        self.lib.orc_flow_stats(self.E, _d(out))
        return out

    def download(self):
        out = np.empty(6 * self.n)
        self.lib.orc_download_fields(self.E, _d(out))
        f = out.reshape(6, self.n)
        return {"rho": f[0], "c": f[1], "ux": f[2], "uy": f[3], "uz": f[4], "p": f[5]}

    def field(self, name):
        p = self.lib.orc_field(self.E, name.encode())
        size = self.Vn if name in ("h", "g") else self.n
        return np.ctypeslib.as_array(p, shape=(size,)).copy()

    def tables(self):
        n = self.n
        names = ("cdm", "cdp", "bcase", "bc1", "bc2", "avgnext", "lcase", "lc1", "lc2", "src")
        sizes = (18, 18, 18, 18, 18, 19, 18, 18, 18, 19)
        arrs = [np.empty(s * n, dtype=np.int32) for s in sizes]
        self.lib.orc_get_tables(self.E, *[a.ctypes.data_as(_PI) for a in arrs])
        return dict(zip(names, arrs))

    def set_open_bnd(self, inlet, outlet, icoord=None, buf_verts=(), buf_refs=(), copy_to_buffers=False):
        inl = np.ascontiguousarray(inlet, dtype=np.int32)
        out = np.ascontiguousarray(outlet, dtype=np.int32)
        bv = np.ascontiguousarray(buf_verts, dtype=np.int32)
        br = np.ascontiguousarray(buf_refs, dtype=np.int32)
        ic = None if icoord is None else np.ascontiguousarray(icoord, dtype=np.float64)
        self._keep = (inl, out, bv, br, ic)
        pi = C.POINTER(C.c_int32)
        rc = self.lib.orc_set_open_bnd(self.E, inl.ctypes.data_as(pi), len(inl), out.ctypes.data_as(pi), len(out),
                                       None if ic is None else _d(ic), bv.ctypes.data_as(pi),
                                       br.ctypes.data_as(pi), len(bv), int(copy_to_buffers))
        assert rc == 0

    def set_forcing_factor(self, ff):
        self.lib.orc_set_forcing_factor(self.E, float(ff))

    def coords(self):
        xyz = np.empty(3 * self.n, dtype=np.int32)
        self.lib.orc_get_coords(self.E, xyz.ctypes.data_as(_PI))
        return xyz.reshape(self.n, 3)


def d3q19():
    lib = load()
    cx, cy, cz, opp = (np.empty(19, dtype=np.int32) for _ in range(4))
    w = np.empty(19)
    lib.orc_d3q19(cx.ctypes.data_as(_PI), cy.ctypes.data_as(_PI), cz.ctypes.data_as(_PI),
                  opp.ctypes.data_as(_PI), _d(w))
    return cx, cy, cz, opp, w


def mrt_matrices():
    lib = load()
    M = np.empty((19, 19))
    Mi = np.empty((19, 19))
    lib.orc_mrt_matrices(_d(M), _d(Mi))
    return M, Mi


def fixed_sum(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return load().orc_fixed_sum(_d(v), v.size)
