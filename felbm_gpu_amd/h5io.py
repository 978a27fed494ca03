"""HDF5 output and restart files in felbm_gpu's layout.

ctypes binding of ``include/felbm_h5.h`` (libfelbm_h5.so, built by
``felbm_gpu_amd/build.py`` against the HDF5 1.10 C library), plus the output
schedule of the reference's driver loop (src/felbm_gpu_main.cu:701-770):

  fields      ``<output_name>_<t>.h5``: density, concentration, u_x, u_y, u_z,
              pressure (write_fields_h5, :47-118; output_deflate / output_float32)
  geometry    ``geometry.h5``: coords n x 3 int32, size 3 int32 (:124-143, :276)
  checkpoint  ``checkpoint_<t>.h5``: h, g, step, target_mass, Vn (:151-189, :706-714)
  particles   ``particles_<t>.h5``: position, velocity, id [, Dm, plane_id,
              rho_hat, log_rho] (:712, :765-770; README.md:421-455)

Loading fails loudly when the library has not been built.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfelbm_h5.so")

F64, F32, I32 = 0, 1, 2
EMISSING = -4
FIELD_NAMES = ("density", "concentration", "u_x", "u_y", "u_z", "pressure")
# engine field keys in the order of FIELD_NAMES
FIELD_KEYS = ("rho", "c", "ux", "uy", "uz", "p")

_P = C.c_void_p
_S = C.c_char_p
_I = C.c_int
_I64 = C.c_int64
_PD = C.POINTER(C.c_double)
_PI32 = C.POINTER(C.c_int32)

SIGNATURES = {
    "felbm_h5_last_error": (_S, []),
    "felbm_h5_write_fields": (_I, [_S, _I64, _P, _P, _P, _P, _P, _P, _I, _I, _I]),
    "felbm_h5_write_geometry": (_I, [_S, _I64, _PI32, _I]),
    "felbm_h5_write_checkpoint": (_I, [_S, _I64, _PD, _PD, C.c_uint32, C.c_double]),
    "felbm_h5_read_checkpoint": (_I, [_S, _I64, _PD, _PD, C.POINTER(C.c_uint32), _PD]),
    "felbm_h5_write_particles": (_I, [_S, _I64, _PD, _PD, _PI32, _PD, _PI32, _PD, _PD]),
    "felbm_h5_dataset_shape": (_I, [_S, _S, C.POINTER(_I64), C.POINTER(_I)]),
    "felbm_h5_read_dataset": (_I, [_S, _S, _I, _P, _I64]),
}

_lib = None


class H5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("felbm_h5 error %d: %s" % (code, msg))
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is not built (run python -m felbm_gpu_amd.build)" % LIB_PATH)
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype, f.argtypes = res, args
        _lib = h
    return _lib


def _check(rc):
    if rc != 0:
        raise H5Error(rc, lib().felbm_h5_last_error().decode())


def _path(p):
    return os.fsencode(os.fspath(p))


def _d(a):
    return None if a is None else a.ctypes.data_as(_PD)


def _i(a):
    return None if a is None else a.ctypes.data_as(_PI32)


def write_fields(path, fields, deflate=0, f32=False):
    """fields: dict with rho, c, ux, uy, uz, p (all float64 or all float32)."""
    arrs = [np.ascontiguousarray(fields[k]).reshape(-1) for k in FIELD_KEYS]
    src_f32 = arrs[0].dtype == np.float32
    want = np.float32 if src_f32 else np.float64
    arrs = [np.ascontiguousarray(a, dtype=want) for a in arrs]
    n = arrs[0].size
    if any(a.size != n for a in arrs):
        raise ValueError("the six fields must have the same length")
    _check(lib().felbm_h5_write_fields(_path(path), n, *[a.ctypes.data_as(_P) for a in arrs], int(src_f32),
                                       int(deflate), int(bool(f32))))


def geometry_coords(solid):
    """coords [n][3] int32 (x, y, z) of the sites of a solid mask [nz][ny][nx],
    in the engine's site order (fluid voxels, x fastest)."""
    z, y, x = np.nonzero(np.asarray(solid) == 0)
    return np.ascontiguousarray(np.stack([x, y, z], axis=1), dtype=np.int32)


def write_geometry(path, coords, deflate=0):
    co = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1, 3)
    _check(lib().felbm_h5_write_geometry(_path(path), len(co), _i(co), int(deflate)))


def write_checkpoint(path, h, g, step, target_mass):
    h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
    g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
    if h.size != g.size:
        raise ValueError("h and g must have the same length")
    _check(lib().felbm_h5_write_checkpoint(_path(path), h.size, _d(h), _d(g), int(step), float(target_mass)))


def read_checkpoint(path, vn):
    """(h, g, step, target_mass); H5Error with code -3 when the file's Vn != vn."""
    h, g = np.empty(int(vn)), np.empty(int(vn))
    st, tm = C.c_uint32(), C.c_double()
    _check(lib().felbm_h5_read_checkpoint(_path(path), int(vn), _d(h), _d(g), C.byref(st), C.byref(tm)))
    return h, g, int(st.value), float(tm.value)


def write_particles(path, position, velocity, ids=None, dm=None, plane_id=None, rho_hat=None, log_rho=None):
    pos = np.ascontiguousarray(position, dtype=np.float64).reshape(-1, 3)
    m = len(pos)
    vel = np.ascontiguousarray(velocity, dtype=np.float64).reshape(m, 3)
    ids = np.arange(m, dtype=np.int32) if ids is None else np.ascontiguousarray(ids, dtype=np.int32).reshape(m)
    dm = None if dm is None else np.ascontiguousarray(dm, dtype=np.float64).reshape(m)
    pl = None if plane_id is None else np.ascontiguousarray(plane_id, dtype=np.int32).reshape(m)
    rh = None if rho_hat is None else np.ascontiguousarray(rho_hat, dtype=np.float64).reshape(m, 3)
    lr = None if log_rho is None else np.ascontiguousarray(log_rho, dtype=np.float64).reshape(m)
    _check(lib().felbm_h5_write_particles(_path(path), m, _d(pos), _d(vel), _i(ids), _d(dm), _i(pl), _d(rh),
                                          _d(lr)))


def write_particles_csv(path, position, ids=None):
    """particles_format = csv (felbm_gpu_main.cu:547-555): one plain-text line
    id,x,y,z per tracer.  The reference's writer (ParticleManager) is not
    available, so the header line and the digits are this engine's choice:
    a '# id,x,y,z' line, then 17 significant digits (round-trip exact)."""
    pos = np.asarray(position, dtype=np.float64).reshape(-1, 3)
    ids = np.arange(len(pos)) if ids is None else np.asarray(ids).reshape(-1)
    with open(path, "w") as f:
        f.write("# id,x,y,z\n")
        f.write("".join("%d,%.17g,%.17g,%.17g\n" % (int(i), p[0], p[1], p[2]) for i, p in zip(ids, pos)))


def read_particles_csv(path):
    """(ids, positions) of a file written by write_particles_csv."""
    a = np.loadtxt(path, delimiter=",", comments="#", ndmin=2)
    return a[:, 0].astype(np.int64), a[:, 1:4].copy()


def dataset_shape(path, name):
    """shape tuple, or None when the file has no such dataset."""
    dims, nd = (_I64 * 2)(), _I()
    rc = lib().felbm_h5_dataset_shape(_path(path), name.encode(), dims, C.byref(nd))
    if rc == EMISSING:
        return None
    _check(rc)
    return tuple(int(dims[i]) for i in range(min(nd.value, 2)))


def read_dataset(path, name, dtype=np.float64):
    shape = dataset_shape(path, name)
    if shape is None:
        raise H5Error(EMISSING, "no dataset %s in %s" % (name, path))
    code = {np.dtype(np.float64): F64, np.dtype(np.float32): F32, np.dtype(np.int32): I32}[np.dtype(dtype)]
    out = np.empty(shape, dtype=dtype)
    _check(lib().felbm_h5_read_dataset(_path(path), name.encode(), code, out.ctypes.data_as(_P), out.size))
    return out


def read_particles(path):
    """dict of the datasets present: position, velocity, id, Dm, plane_id,
    rho_hat, log_rho."""
    out = {}
    for name, dt in (("position", np.float64), ("velocity", np.float64), ("id", np.int32), ("Dm", np.float64),
                     ("plane_id", np.int32), ("rho_hat", np.float64), ("log_rho", np.float64)):
        if dataset_shape(path, name) is not None:
            out[name] = read_dataset(path, name, dt)
    if "position" not in out:
        raise H5Error(EMISSING, "%s has no position dataset" % path)
    return out


def _g(x):
    """A double as the driver's text streams print it (C++ default float format,
    6 significant digits: the same as printf %g)."""
    return "%g" % float(x)


class OutputSchedule:
    """The driver loop's file output (src/felbm_gpu_main.cu:596-770): before
    step t, a checkpoint when checkpoint_skip divides t (t past the restart
    step), the fields when file_skip divides t, the particles when
    particles_skip divides t; when log_skip divides t, a row of the flow
    monitor in timeseries_file (:620-626, :683-691), and when stretching_skip
    divides t the stretching rows drained since the last time in
    stretching_file (:635-671).  Paths follow the reference's names under
    output_dir; the two text files are opened (truncated) with the schedule,
    their column headers as the reference writes them."""

    TS_HEADER = "# iter   max_speed   mean_ux   mean_uy   mean_uz   mean_v2   total_order_parameter\n"
    ST_HEADER = ("# Lagrangian stretching of material line elements (Eq. 12), one row per step\n"
                 "# step   lambda   mean_log_rho   var_log_rho   n_active   lambda_w   lambda_nw   n_w   n_nw\n")

    def __init__(self, output_dir, output_name="output", file_skip=0, checkpoint_skip=0, particles_skip=0,
                 deflate=0, f32=False, restart_step=0, log_skip=0, timeseries_file=None, stretching_skip=0,
                 stretching_file=None, particles_format="h5"):
        self.dir = os.fspath(output_dir)
        self.name = output_name
        self.file_skip, self.checkpoint_skip, self.particles_skip = int(file_skip), int(checkpoint_skip), int(
            particles_skip)
        self.deflate, self.f32, self.restart_step = int(deflate), bool(f32), int(restart_step)
        # the periodic particle files; checkpoints always write HDF5 (they are restored from it)
        self.pext = ".csv" if str(particles_format).lower() == "csv" else ".h5"
        # log_skip defaults to file_skip, then 1; stretching_skip to log_skip (:620, :636)
        self.log_skip = int(log_skip) or self.file_skip or 1
        self.stretching_skip = int(stretching_skip) or self.log_skip
        self.ts = self.st = None
        if timeseries_file:
            self.ts = open(os.path.join(self.dir, timeseries_file), "w")
            self.ts.write(self.TS_HEADER)
        if stretching_file:
            self.st = open(os.path.join(self.dir, stretching_file), "w")
            self.st.write(self.ST_HEADER)
        self.written = []

    def path(self, kind, t):
        base = {"fields": self.name, "checkpoint": "checkpoint", "particles": "particles"}[kind]
        return os.path.join(self.dir, "%s_%d.h5" % (base, t))

    def before_step(self, gpu, t):
        if self.checkpoint_skip and t > self.restart_step and t % self.checkpoint_skip == 0:
            p = self.path("checkpoint", t)
            gpu.write_checkpoint(p, t)
            self.written.append(p)
            if gpu.has_tracers():
                p = self.path("particles", t)
                gpu.write_particles_h5(p)
                self.written.append(p)
        if self.ts is not None and t % self.log_skip == 0:
            self.ts.write(self.stats_row(t, gpu.flow_stats(), gpu.n))
            self.ts.flush()
        if self.file_skip and t % self.file_skip == 0:
            p = self.path("fields", t)
            gpu.write_fields_h5(p, self.deflate, self.f32)
            self.written.append(p)
        if self.particles_skip and gpu.has_tracers() and t % self.particles_skip == 0:
            p = self.path("particles", t)[:-3] + self.pext
            if p not in self.written:
                if self.pext == ".csv":
                    write_particles_csv(p, gpu.tracers_download()[0])
                else:
                    gpu.write_particles_h5(p)
                self.written.append(p)
        if self.st is not None and gpu.has_tracers() and t % self.stretching_skip == 0:
            rows, _ = gpu.tracers_stretch_rows()
            self.st.write("".join(self.stretch_row(r) for r in rows))
            self.st.flush()

    @staticmethod
    def stats_row(t, st, n):
        """iter, max |u|, <ux>, <uy>, <uz>, <u^2>, total c from flow_stats (:683-691)."""
        inv = 1.0 / float(n)
        return "%d  %s  %s  %s  %s  %s  %s\n" % (t, _g(math.sqrt(st[5])), _g(st[0] * inv), _g(st[1] * inv),
                                                 _g(st[2] * inv), _g(st[3] * inv), _g(st[4]))

    @staticmethod
    def stretch_row(r):
        """step, lambda, mean, var, n_active, lambda_w, lambda_nw, n_w, n_nw (:662-670)."""
        return "%d  %s  %s  %s  %d  %s  %s  %d  %d\n" % (int(r[0]), _g(r[1]), _g(r[2]), _g(r[3]), int(r[4]), _g(r[5]),
                                                       _g(r[6]), int(r[7]), int(r[8]))

    def finish(self, gpu):
        """End of a run (:781-783): drain the stretching rows not yet written."""
        if self.st is not None and gpu.has_tracers():
            rows, _ = gpu.tracers_stretch_rows()
            self.st.write("".join(self.stretch_row(r) for r in rows))
            self.st.flush()

    def close(self):
        for f in (self.ts, self.st):
            if f is not None:
                f.close()
        self.ts = self.st = None
