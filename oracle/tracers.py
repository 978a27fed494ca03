"""NumPy restatement of the tracer update (the oracle for the engine's
tracer kernel).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the oracle; the product
never loads it.

The reference advects its tracers with felbm_local's ParticleManager, which is
not available: its interpolation order, wall handling and random stream are
parity unpinned.  This restates the engine's documented algorithm
(felbm_gpu_amd/csrc/tracers.hpp) with the reference's driver semantics
(src/felbm_gpu_main.cu:752-777: the tracers of step t move on the u of step
t's field pass, refreshed every particles_velocity_skip steps; README.md:428-446:
Brownian kick sqrt(2 Dm dt) per axis, none when Dm = 0, no z kick in 2D):
  per substep dt = 1/substeps: u = trilinear interpolation over the 8 voxels
  around x (solid or out-of-box voxels carry 0); x' = x + u dt + sqrt(2 Dm dt) xi;
  periodic axes wrap; a move whose nearest voxel is solid or outside the box
  is rejected and counted.
The arithmetic order is the kernel's, so with Dm = 0 the positions agree bit
for bit; with Dm > 0 the kicks go through log/cos/sin, whose last bit may
differ between libraries.
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z):
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def normals3(seed, ids, draw):
    """Three N(0,1) values per tracer for (seed, tracer id, draw index)."""
    ids = np.asarray(ids, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = mix64(np.uint64(seed) ^ mix64(ids))
        base = key + np.uint64(4) * np.uint64(draw)
        r = [mix64(base + np.uint64(q)) for q in range(4)]
    u0 = ((r[0] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    v0 = (r[1] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u1 = ((r[2] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    v1 = (r[3] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    two_pi = 6.283185307179586476925286766559
    a0, a1 = np.sqrt(-2.0 * np.log(u0)), np.sqrt(-2.0 * np.log(u1))
    return a0 * np.cos(two_pi * v0), a0 * np.sin(two_pi * v0), a1 * np.cos(two_pi * v1)


class Tracers:
    """Tracers on a lattice given by its solid mask [nz][ny][nx].

    With rho_hat given, each tracer also carries a material line element
    (docs/HANDOFF_stretching.md:165-210, Eq. 12): per substep, with J the
    gradient of the trilinear interpolant at the substep's start position,
    lambda_s = rho_hat . J rho_hat; rho* = rho_hat + dt J rho_hat;
    log_rho += log|rho*|; rho_hat = rho*/|rho*|.  A tracer not blocked during
    the step contributes the mean of its lambda_s to the step's row and is
    classed wetting when c < 0.5 at its nearest node at the start of the step.
    """

    def __init__(self, solid, periodic, pos, dm, seed=0, substeps=1, velocity_skip=1, rho_hat=None, log_rho=None):
        self.solid = np.asarray(solid, dtype=np.uint8)
        nz, ny, nx = self.solid.shape
        self.N = (nx, ny, nz)
        self.per = tuple(int(p) for p in periodic)
        z, y, x = np.nonzero(self.solid == 0)
        self.idx = -np.ones(self.solid.shape, dtype=np.int64)
        self.idx[z, y, x] = np.arange(len(x))
        self.pos = np.array(pos, dtype=np.float64).reshape(-1, 3)
        self.dm = np.asarray(dm, dtype=np.float64).reshape(-1)
        self.seed, self.substeps, self.vskip = int(seed), int(substeps), int(velocity_skip)
        self.blocked = np.zeros(len(self.dm), dtype=np.int64)
        self.step_count = 0
        self.u = None
        self.c = None
        self.rho_hat = None if rho_hat is None else np.array(rho_hat, dtype=np.float64).reshape(-1, 3)
        self.log_rho = None
        if self.rho_hat is not None:
            self.log_rho = np.zeros(len(self.dm)) if log_rho is None else np.array(log_rho, dtype=np.float64)
        self.rows = []

    def _site(self, v):
        """site index of in-box voxel coordinates (x, y, z) arrays, -1 solid"""
        return self.idx[v[2], v[1], v[0]]

    def interp(self, p, u):
        fl = np.floor(p)
        o = fl.astype(np.int64)
        f = p - fl
        out = np.zeros_like(p)
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    c = (cx, cy, cz)
                    v = [o[:, a] + c[a] for a in range(3)]
                    ok = np.ones(len(p), dtype=bool)
                    for a in range(3):
                        if self.per[a]:
                            v[a] = np.mod(v[a], self.N[a])
                        else:
                            ok &= (v[a] >= 0) & (v[a] < self.N[a])
                            v[a] = np.clip(v[a], 0, self.N[a] - 1)
                    s = self._site(v)
                    ok &= s >= 0
                    w = ((f[:, 0] if cx else 1.0 - f[:, 0]) * (f[:, 1] if cy else 1.0 - f[:, 1])) * \
                        (f[:, 2] if cz else 1.0 - f[:, 2])
                    ss = np.where(ok, s, 0)
                    for a in range(3):
                        out[:, a] = np.where(ok, out[:, a] + w * u[a][ss], out[:, a])
        return out

    def interp_uJ(self, p, u):
        """u and J[a][b] = d u_a / d x_b of the trilinear interpolant."""
        fl = np.floor(p)
        o = fl.astype(np.int64)
        f = p - fl
        out = np.zeros_like(p)
        J = np.zeros((len(p), 3, 3))
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    c = (cx, cy, cz)
                    v = [o[:, a] + c[a] for a in range(3)]
                    ok = np.ones(len(p), dtype=bool)
                    for a in range(3):
                        if self.per[a]:
                            v[a] = np.mod(v[a], self.N[a])
                        else:
                            ok &= (v[a] >= 0) & (v[a] < self.N[a])
                            v[a] = np.clip(v[a], 0, self.N[a] - 1)
                    s = self._site(v)
                    ok &= s >= 0
                    wx = f[:, 0] if cx else 1.0 - f[:, 0]
                    wy = f[:, 1] if cy else 1.0 - f[:, 1]
                    wz = f[:, 2] if cz else 1.0 - f[:, 2]
                    w = (wx * wy) * wz
                    g = ((1.0 if cx else -1.0) * (wy * wz), (1.0 if cy else -1.0) * (wx * wz),
                         (1.0 if cz else -1.0) * (wx * wy))
                    ss = np.where(ok, s, 0)
                    for a in range(3):
                        va = u[a][ss]
                        out[:, a] = np.where(ok, out[:, a] + w * va, out[:, a])
                        for b in range(3):
                            J[:, a, b] = np.where(ok, J[:, a, b] + g[b] * va, J[:, a, b])
        return out, J

    def nearest_c(self, p):
        if self.c is None:
            return np.full(len(p), np.nan)
        v = []
        ok = np.ones(len(p), dtype=bool)
        for a in range(3):
            va = np.rint(p[:, a]).astype(np.int64)
            if self.per[a]:
                va = np.mod(va, self.N[a])
            else:
                ok &= (va >= 0) & (va < self.N[a])
                va = np.clip(va, 0, self.N[a] - 1)
            v.append(va)
        s = self._site(v)
        ok &= s >= 0
        return np.where(ok, self.c[np.where(ok, s, 0)], np.nan)

    def step(self, ux, uy, uz, c=None):
        """One LBM step of tracer motion on the step's velocity (site arrays);
        c (site array) feeds the phase split of the stretching rows."""
        if self.u is None or self.step_count % self.vskip == 0:
            self.u = (np.array(ux, dtype=np.float64), np.array(uy, dtype=np.float64), np.array(uz, dtype=np.float64))
            self.c = None if c is None else np.array(c, dtype=np.float64)
        dt = 1.0 / self.substeps
        amp = np.where(self.dm > 0.0, np.sqrt(2.0 * self.dm * dt), 0.0)
        ids = np.arange(len(self.dm), dtype=np.uint64)
        p = self.pos
        st = self.rho_hat is not None
        if st:
            c0 = self.nearest_c(p)
            rh, lr = self.rho_hat.copy(), self.log_rho.copy()
            lam = np.zeros(len(p))
            nblk = np.zeros(len(p), dtype=np.int64)
        for s in range(self.substeps):
            if st:
                u, J = self.interp_uJ(p, self.u)
                Jr = np.stack([(J[:, a, 0] * rh[:, 0] + J[:, a, 1] * rh[:, 1]) + J[:, a, 2] * rh[:, 2]
                               for a in range(3)], axis=1)
                lam = lam + ((rh[:, 0] * Jr[:, 0] + rh[:, 1] * Jr[:, 1]) + rh[:, 2] * Jr[:, 2])
                r = rh + dt * Jr
                g = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2])
                lr = lr + np.log(g)
                rh = r / g[:, None]
            else:
                u = self.interp(p, self.u)
            q = p + u * dt
            if np.any(self.dm > 0.0):
                xi = normals3(self.seed, ids, self.step_count * self.substeps + s)
                kick = self.dm > 0.0
                q[:, 0] = np.where(kick, q[:, 0] + amp * xi[0], q[:, 0])
                q[:, 1] = np.where(kick, q[:, 1] + amp * xi[1], q[:, 1])
                if self.N[2] > 1:
                    q[:, 2] = np.where(kick, q[:, 2] + amp * xi[2], q[:, 2])
            ok = np.ones(len(p), dtype=bool)
            v = []
            for a in range(3):
                L = float(self.N[a])
                if self.per[a]:
                    q[:, a] = q[:, a] - L * np.floor(q[:, a] / L)
                else:
                    ok &= (q[:, a] >= -0.5) & (q[:, a] < L - 0.5)
                va = np.rint(q[:, a]).astype(np.int64)
                if self.per[a]:
                    va = np.where(va >= self.N[a], va - self.N[a], va)
                v.append(np.clip(va, 0, self.N[a] - 1))
            ok &= self._site(v) >= 0
            p = np.where(ok[:, None], q, p)
            self.blocked += ~ok
            if st:
                nblk += ~ok
        self.pos = p
        if st:
            self.rho_hat, self.log_rho = rh, lr
            act = nblk == 0
            li = lam / self.substeps
            wet, dry = act & (c0 < 0.5), act & (c0 >= 0.5)
            n_act, n_w, n_nw = int(act.sum()), int(wet.sum()), int(dry.sum())
            mean = lr.sum() / len(lr)
            self.rows.append([self.step_count + 1, li[act].sum() / n_act if n_act else np.nan, mean,
                              (lr * lr).sum() / len(lr) - mean * mean, n_act,
                              li[wet].sum() / n_w if n_w else np.nan, li[dry].sum() / n_nw if n_nw else np.nan,
                              n_w, n_nw])
        self.step_count += 1
