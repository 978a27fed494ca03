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
  const long ns = d.ns;
  for (int cz = 0; cz < 2; ++cz)
    for (int cy = 0; cy < 2; ++cy)
      for (int cx = 0; cx < 2; ++cx) {
        int x = x0 + cx, y = y0 + cy, z = z0 + cz;
        if (!wrap_voxel(x, d.Nx, d.px) || !wrap_voxel(y, d.Ny, d.py) || !wrap_voxel(z, d.Nz, d.pz)) continue;
        const int s = voxel_site(d, x, y, z);
        if (s < 0) continue;
        const double w = ((cx ? fx : 1.0 - fx)*(cy ? fy : 1.0 - fy))*(cz ? fz : 1.0 - fz);
        u[0] += w*(double)ldg(d.tu, s);
        u[1] += w*(double)ldg(d.tu, ns + s);
        u[2] += w*(double)ldg(d.tu, 2*ns + s);
      }
}

template <typename real>
__global__ void __launch_bounds__(256) k_tracers(const Inst<real> *__restrict__ D, int di, double *pos,
                                                 const double *__restrict__ dm, int *blocked, int np,
                                                 unsigned long long seed, int substeps)
{
  #pragma clang fp contract(off)
  const int j = blockIdx.x*blockDim.x + (int)threadIdx.x;
  if (j >= np) return;
  const Inst<real> &d = D[di];
  const unsigned long long step = ldg(d.acc, ACC_FLAGS + 2);
  double p[3] = { ldg(pos, 3L*j), ldg(pos, 3L*j + 1), ldg(pos, 3L*j + 2) };
  const double dt = 1.0/(double)substeps;
  const double Dm = ldg(dm, (long)j);
  const double amp = Dm > 0.0 ? sqrt(2.0*Dm*dt) : 0.0;
  const int N[3] = { d.Nx, d.Ny, d.Nz }, P[3] = { d.px, d.py, d.pz };
  int nblk = 0;
  for (int s = 0; s < substeps; ++s) {
    double u[3];
    interp_u(d, p, u);
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
