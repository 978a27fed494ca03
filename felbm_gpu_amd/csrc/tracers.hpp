// tracers.hpp -- passive tracers advected on the device velocity field.
//
// The reference keeps its tracers on the host (felbm_local ParticleManager)
// and feeds them the GPU velocity by a device-to-host copy and a scatter into
// the global grid every step (src/felbm_gpu_main.cu:426-438, :471-491,
// :752-777).  Here the velocity snapshot never leaves HBM and one thread
// advances one tracer.  Per substep (dt = 1/substeps):
//   u  = trilinear interpolation of the snapshot over the 8 voxels around the
//        position; a solid or out-of-box voxel contributes u = 0, as the
//        reference's zero-filled global arrays do;
//   x' = x + u dt + sqrt(2 Dm dt) xi, xi ~ N(0,1) per axis from a counter-based
//        generator keyed by (seed, tracer, substep): no draw when Dm = 0 and no
//        z kick when Nz = 1 (README.md:428-446);
//   periodic axes wrap; a move whose nearest voxel is solid or outside the box
//   is rejected and counted (wall blocking).
// Voxel (x, y, z) sits at position (x, y, z).  The snapshot is the field
// pass's u* of the step, refreshed every tr_vskip steps (particles_velocity_skip).
//
// Lagrangian stretching (docs/HANDOFF_stretching.md:165-210, Eq. 12): each
// tracer carries a unit vector rho_hat and log|rho|; per substep, with J the
// gradient of the trilinear interpolant at the substep's start position,
//   lambda_s = rho_hat . J rho_hat,  rho* = rho_hat + dt J rho_hat,
//   log_rho += log|rho*|,  rho_hat = rho*/|rho*|   (Benettin renormalisation).
// A tracer not blocked during the step contributes lambda_i = mean_s lambda_s
// to the step's row and is classed wetting when c < 0.5 at its nearest node at
// the start of the step.  Row: step, lambda, <log rho>, Var(log rho), n_active,
// lambda_w, lambda_nw, n_w, n_nw (src/felbm_gpu_main.cu:643-671).
#pragma once
#include "kernels.hpp"

namespace fm {

// splitmix64 finaliser (Steele, Lea and Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014)
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// three standard normals for (seed, tracer id, draw index) by Box-Muller
__device__ inline void normals3(uint64_t seed, uint64_t id, uint64_t draw, double (&n)[3])
{
  #pragma clang fp contract(off)
  const uint64_t key = mix64(seed ^ mix64(id));
  const uint64_t r0 = mix64(key + 4*draw), r1 = mix64(key + 4*draw + 1);
  const uint64_t r2 = mix64(key + 4*draw + 2), r3 = mix64(key + 4*draw + 3);
  const double TWO_PI = 6.283185307179586476925286766559;
  const double u0 = (double)((r0 >> 11) + 1) * 0x1p-53, v0 = (double)(r1 >> 11) * 0x1p-53;
  const double u1 = (double)((r2 >> 11) + 1) * 0x1p-53, v1 = (double)(r3 >> 11) * 0x1p-53;
  const double a0 = sqrt(-2.0*log(u0)), a1 = sqrt(-2.0*log(u1));
  n[0] = a0*cos(TWO_PI*v0);
  n[1] = a0*sin(TWO_PI*v0);
  n[2] = a1*cos(TWO_PI*v1);
}

// site index of an in-box voxel, -1 when solid
template <typename real>
__device__ __forceinline__ int voxel_site(const Inst<real> &d, int x, int y, int z)
{
  const uint64_t w = ldg(d.words, (long)((z*d.Ny + y)*d.wpr + (x >> 5)));
  const unsigned lo = (unsigned)w, b = (unsigned)x & 31u;
  if (!((lo >> b) & 1u)) return -1;
  return (int)((unsigned)(w >> 32) + (unsigned)__popc(lo & ((1u << b) - 1u)));
}

// wrap a voxel coordinate onto the box; false when off a non-periodic axis
__device__ __forceinline__ bool wrap_voxel(int &v, int N, int per)
{
  if (per) { v %= N; if (v < 0) v += N; return true; }
  return v >= 0 && v < N;
}

template <typename real>
__device__ inline void interp_u(const Inst<real> &d, const double (&p)[3], double (&u)[3])
{
  #pragma clang fp contract(off)
  u[0] = 0.0; u[1] = 0.0; u[2] = 0.0;
  // a position off the lattice (non-finite, or beyond int range) reads no voxel
  for (int a = 0; a < 3; ++a) if (!(fabs(p[a]) < 1e9)) return;
  const double fl0 = floor(p[0]), fl1 = floor(p[1]), fl2 = floor(p[2]);
  const int x0 = (int)fl0, y0 = (int)fl1, z0 = (int)fl2;
  const double fx = p[0] - fl0, fy = p[1] - fl1, fz = p[2] - fl2;
  for (int cz = 0; cz < 2; ++cz)
    for (int cy = 0; cy < 2; ++cy)
      for (int cx = 0; cx < 2; ++cx) {
        int x = x0 + cx, y = y0 + cy, z = z0 + cz;
        if (!wrap_voxel(x, d.Nx, d.px) || !wrap_voxel(y, d.Ny, d.py) || !wrap_voxel(z, d.Nz, d.pz)) continue;
        const int s = voxel_site(d, x, y, z);
        if (s < 0) continue;
        const double w = ((cx ? fx : 1.0 - fx)*(cy ? fy : 1.0 - fy))*(cz ? fz : 1.0 - fz);
        u[0] += w*(double)ldg(d.tu, 4L*s);
        u[1] += w*(double)ldg(d.tu, 4L*s + 1);
        u[2] += w*(double)ldg(d.tu, 4L*s + 2);
      }
}

// u and its gradient J[a][b] = d u_a / d x_b of the trilinear interpolant
template <typename real>
__device__ inline void interp_uJ(const Inst<real> &d, const double (&p)[3], double (&u)[3], double (&J)[3][3])
{
  #pragma clang fp contract(off)
  for (int a = 0; a < 3; ++a) { u[a] = 0.0; J[a][0] = 0.0; J[a][1] = 0.0; J[a][2] = 0.0; }
  for (int a = 0; a < 3; ++a) if (!(fabs(p[a]) < 1e9)) return;
  const double fl0 = floor(p[0]), fl1 = floor(p[1]), fl2 = floor(p[2]);
  const int x0 = (int)fl0, y0 = (int)fl1, z0 = (int)fl2;
  const double fx = p[0] - fl0, fy = p[1] - fl1, fz = p[2] - fl2;
  for (int cz = 0; cz < 2; ++cz)
    for (int cy = 0; cy < 2; ++cy)
      for (int cx = 0; cx < 2; ++cx) {
        int x = x0 + cx, y = y0 + cy, z = z0 + cz;
        if (!wrap_voxel(x, d.Nx, d.px) || !wrap_voxel(y, d.Ny, d.py) || !wrap_voxel(z, d.Nz, d.pz)) continue;
        const int s = voxel_site(d, x, y, z);
        if (s < 0) continue;
        const double wx = cx ? fx : 1.0 - fx, wy = cy ? fy : 1.0 - fy, wz = cz ? fz : 1.0 - fz;
        const double w = (wx*wy)*wz;
        const double gx = (cx ? 1.0 : -1.0)*(wy*wz), gy = (cy ? 1.0 : -1.0)*(wx*wz), gz = (cz ? 1.0 : -1.0)*(wx*wy);
        const double v[3] = { (double)ldg(d.tu, 4L*s), (double)ldg(d.tu, 4L*s + 1), (double)ldg(d.tu, 4L*s + 2) };
        for (int a = 0; a < 3; ++a) {
          u[a] += w*v[a];
          J[a][0] += gx*v[a]; J[a][1] += gy*v[a]; J[a][2] += gz*v[a];
        }
      }
}

// c at the nearest node (the phase label of the stretching rows); NaN off the
// pore space or without a c snapshot
template <typename real>
__device__ inline double nearest_c(const Inst<real> &d, const double (&p)[3])
{
  if (!d.tr_phase) return __builtin_nan("");
  int v[3];
  const int N[3] = { d.Nx, d.Ny, d.Nz }, P[3] = { d.px, d.py, d.pz };
  for (int a = 0; a < 3; ++a) {
    if (!(fabs(p[a]) < 1e9)) return __builtin_nan("");
    v[a] = (int)rint(p[a]);
    if (!wrap_voxel(v[a], N[a], P[a])) return __builtin_nan("");
  }
  const int s = voxel_site(d, v[0], v[1], v[2]);
  return s < 0 ? __builtin_nan("") : (double)ldg(d.tu, 4L*s + 3);
}

// per-step stretching state of one slot's tracers
struct Stretch {
  double *rho_hat;     // [3 np]
  double *log_rho;     // [np]
  double *partials;    // [blocks][8]: sum lambda_i, n_active, sum log, sum log^2, sum lambda_w, n_w, sum lambda_nw, n_nw
  double *rows;        // [cap][9]
  int cap;
};

// fixed-order block sum of 8 doubles (4 waves): lane butterflies, then wave 0..3
__device__ inline void block_sum8(double (&v)[8], double *dst)
{
  __shared__ double sh[4][8];
  for (int q = 0; q < 8; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) for (int q = 0; q < 8; ++q) sh[wv][q] = v[q];
  __syncthreads();
  if (threadIdx.x < 8) {
    const int q = threadIdx.x;
    stg(dst, q, ((sh[0][q] + sh[1][q]) + sh[2][q]) + sh[3][q]);
  }
}

template <typename real>
__global__ void __launch_bounds__(256) k_tracers(const Inst<real> *__restrict__ D, int di, double *pos,
                                                 const double *__restrict__ dm, int *blocked, int np,
                                                 unsigned long long seed, int substeps, Stretch st)
{
  #pragma clang fp contract(off)
  const int j = blockIdx.x*blockDim.x + (int)threadIdx.x;
  double acc8[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
  if (j < np) {
  const Inst<real> &d = D[di];
  const unsigned long long step = ldg(d.acc, ACC_FLAGS + 2);
  double p[3] = { ldg(pos, 3L*j), ldg(pos, 3L*j + 1), ldg(pos, 3L*j + 2) };
  const double dt = 1.0/(double)substeps;
  const double Dm = ldg(dm, (long)j);
  const double amp = Dm > 0.0 ? sqrt(2.0*Dm*dt) : 0.0;
  const int N[3] = { d.Nx, d.Ny, d.Nz }, P[3] = { d.px, d.py, d.pz };
  int nblk = 0;
  double rh[3] = { 0, 0, 0 }, lr = 0.0, lam = 0.0;
  const double c0 = st.rows ? nearest_c(d, p) : 0.0;
  if (st.rows) { rh[0] = ldg(st.rho_hat, 3L*j); rh[1] = ldg(st.rho_hat, 3L*j + 1); rh[2] = ldg(st.rho_hat, 3L*j + 2); lr = ldg(st.log_rho, (long)j); }
  for (int s = 0; s < substeps; ++s) {
    double u[3];
    if (st.rows) {
      double J[3][3];
      interp_uJ(d, p, u, J);
      const double Jr[3] = { (J[0][0]*rh[0] + J[0][1]*rh[1]) + J[0][2]*rh[2],
                             (J[1][0]*rh[0] + J[1][1]*rh[1]) + J[1][2]*rh[2],
                             (J[2][0]*rh[0] + J[2][1]*rh[1]) + J[2][2]*rh[2] };
      lam += (rh[0]*Jr[0] + rh[1]*Jr[1]) + rh[2]*Jr[2];
      const double r[3] = { rh[0] + dt*Jr[0], rh[1] + dt*Jr[1], rh[2] + dt*Jr[2] };
      const double g = sqrt((r[0]*r[0] + r[1]*r[1]) + r[2]*r[2]);
      lr += log(g);
      rh[0] = r[0]/g; rh[1] = r[1]/g; rh[2] = r[2]/g;
    } else {
      interp_u(d, p, u);
    }
    double q[3] = { p[0] + u[0]*dt, p[1] + u[1]*dt, p[2] + u[2]*dt };
    if (Dm > 0.0) {
      double xi[3];
      normals3(seed, (uint64_t)j, step*(unsigned long long)substeps + (unsigned long long)s, xi);
      q[0] += amp*xi[0];
      q[1] += amp*xi[1];
      if (d.Nz > 1) q[2] += amp*xi[2];
    }
    bool ok = true;
    int v[3] = { 0, 0, 0 };
    for (int a = 0; a < 3; ++a) {
      if (!(fabs(q[a]) < 1e9)) { ok = false; continue; }   // non-finite or off the lattice
      if (P[a]) {
        const double L = (double)N[a];
        q[a] = q[a] - L*floor(q[a]/L);
      } else if (!(q[a] >= -0.5 && q[a] < (double)N[a] - 0.5)) {
        ok = false;
        continue;
      }
      v[a] = (int)rint(q[a]);
      if (P[a] && v[a] >= N[a]) v[a] -= N[a];
      if (v[a] < 0 || v[a] >= N[a]) ok = false;        // the lookup below stays in the table
    }
    if (ok && voxel_site(d, v[0], v[1], v[2]) < 0) ok = false;
    if (ok) { p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; }
    else ++nblk;
  }
  stg(pos, 3L*j, p[0]); stg(pos, 3L*j + 1, p[1]); stg(pos, 3L*j + 2, p[2]);
  if (nblk) stg(blocked, (long)j, ldg(blocked, (long)j) + nblk);
  if (st.rows) {
    stg(st.rho_hat, 3L*j, rh[0]); stg(st.rho_hat, 3L*j + 1, rh[1]); stg(st.rho_hat, 3L*j + 2, rh[2]);
    stg(st.log_rho, (long)j, lr);
    acc8[2] = lr; acc8[3] = lr*lr;
    if (nblk == 0) {
      const double li = lam/(double)substeps;
      acc8[0] = li; acc8[1] = 1.0;
      if (c0 < 0.5) { acc8[4] = li; acc8[5] = 1.0; }
      else if (c0 >= 0.5) { acc8[6] = li; acc8[7] = 1.0; }
    }
  }
  }
  if (st.rows) block_sum8(acc8, st.partials + 8L*blockIdx.x);
}

// the step's row from the block partials, in block order
template <typename real>
__global__ void k_tracer_row(const Inst<real> *__restrict__ D, int di, Stretch st, int nblocks, int np)
{
  #pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const Inst<real> &d = D[di];
  const unsigned long long step = ldg(d.acc, ACC_FLAGS + 2);
  double s[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
  for (int b = 0; b < nblocks; ++b)
    for (int q = 0; q < 8; ++q) s[q] += ldg(st.partials, 8L*b + q);
  const double nan = __builtin_nan("");
  const double mean = s[2]/(double)np;
  double *r = st.rows + 9L*(long)(step % (unsigned long long)st.cap);
  stg(r, 0, (double)(step + 1));
  stg(r, 1, s[1] > 0.0 ? s[0]/s[1] : nan);
  stg(r, 2, mean);
  stg(r, 3, s[3]/(double)np - mean*mean);
  stg(r, 4, s[1]);
  stg(r, 5, s[5] > 0.0 ? s[4]/s[5] : nan);
  stg(r, 6, s[7] > 0.0 ? s[6]/s[7] : nan);
  stg(r, 7, s[5]);
  stg(r, 8, s[7]);
}

// interpolated velocity at each tracer (the "velocity" the reference writes)
template <typename real>
__global__ void __launch_bounds__(256) k_tracer_velocity(const Inst<real> *__restrict__ D, int di,
                                                         const double *__restrict__ pos, double *uvw, int np)
{
  const int j = blockIdx.x*blockDim.x + (int)threadIdx.x;
  if (j >= np) return;
  const Inst<real> &d = D[di];
  const double p[3] = { ldg(pos, 3L*j), ldg(pos, 3L*j + 1), ldg(pos, 3L*j + 2) };
  double u[3];
  interp_u(d, p, u);
  stg(uvw, 3L*j, u[0]); stg(uvw, 3L*j + 1, u[1]); stg(uvw, 3L*j + 2, u[2]);
}

}  // namespace fm
