// kernels.hpp -- gfx950 kernels of the Lee-Liu D3Q19 step.
//
// One step is four launches over the flattened site list of every instance:
//   k_moments   own-site h,g -> c_raw, p0, j (+ exact FP64 mass partials)
//   k_cfield    c stencil (r=1) -> corrected c, mu, p*   (+ optional export)
//   k_mufield   mu stencil (r=1) -> lap(mu)
//   k_collide   equilibria + Guo force + BGK/MRT collision, push-streamed
//               into the next g buffer and the h buffer
// replacing the reference's 15-17 launches (multiphase_gpu.cuh:522-631).
//
// Layout: sites are the fluid voxels in x-fastest lexicographic order; every
// per-site array is SoA with stride ns (n rounded up to 64).  Neighbour sites
// are not stored: site_at() derives them from a per-row table of 32-voxel
// words (low 32 bits: fluid mask, high 32 bits: index of the first fluid voxel
// at or after the word), which stays resident in L2.  This replaces the
// reference's int32 stencil tables (cdm/cdp/bcase/bc1/bc2/avgnext/lcase/lc1/
// lc2, multiphase_gpu.cuh:93-193) and the streaming op program (:209-283).
//
// Streaming is push: the collide kernel writes post-collision population k of
// site x into slot k of x+e_k, or into slot opp(k) of x (halfway bounce-back)
// when x+e_k is a solid voxel, or zeroes slot opp(k) of x when x+e_k is
// outside a non-periodic box face.  The arrays therefore always hold the
// post-stream state, exactly what the reference's d_h/d_g hold between steps.
//
// FP64 arithmetic follows the reference's expression order term by term (the
// library is built with -ffp-contract=off for the FP64 instantiation), so the
// results are bit-identical to the CPU restatement of the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fm {

constexpr int Q = 19;
constexpr int NSLOT = 64;      // atomic spreading of the reduction partials

// D3Q19 in felbm order (d3q19.cuh:33-36)
constexpr int CX[19]  = { 0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1 };
constexpr int CY[19]  = { 0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 1,-1, 1,-1, 0, 0, 0, 0 };
constexpr int CZ[19]  = { 0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1 };
constexpr int OPP[19] = { 0, 2, 1, 4, 3, 6, 5, 8, 7,10, 9,12,11,14,13,16,15,18,17 };
constexpr double WD[19] = { 1.0/3.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0,
  1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0,
  1.0/36.0, 1.0/36.0, 1.0/36.0 };

// MRT moment basis (d'Humieres et al. 2002, D3Q19) in felbm direction order;
// rows rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz,
// pxz, mx, my, mz -- the order the relaxation vector of device_engine.cuh:625
// assumes.  Minv = M^T diag(1/|row|^2).  Compile-time, so zero entries fold
// away from the unrolled transform without changing any rounding.
struct MrtTab { double M[19][19]; double MI[19][19]; };
constexpr MrtTab make_mrt()
{
  MrtTab t{};
  for (int k = 0; k < 19; ++k) {
    int ex = CX[k], ey = CY[k], ez = CZ[k];
    int e2 = ex*ex + ey*ey + ez*ez;
    int r[19] = { 1, 19*e2 - 30, (21*e2*e2 - 53*e2 + 24)/2, ex, (5*e2 - 9)*ex, ey, (5*e2 - 9)*ey,
                  ez, (5*e2 - 9)*ez, 3*ex*ex - e2, (3*e2 - 5)*(3*ex*ex - e2), ey*ey - ez*ez,
                  (3*e2 - 5)*(ey*ey - ez*ez), ex*ey, ey*ez, ex*ez, (ey*ey - ez*ez)*ex,
                  (ez*ez - ex*ex)*ey, (ex*ex - ey*ey)*ez };
    for (int m = 0; m < 19; ++m) t.M[m][k] = (double)r[m];
  }
  for (int m = 0; m < 19; ++m) {
    double nrm = 0;
    for (int k = 0; k < 19; ++k) nrm += t.M[m][k]*t.M[m][k];
    for (int k = 0; k < 19; ++k) t.MI[k][m] = t.M[m][k] / nrm;
  }
  return t;
}
constexpr MrtTab MRTTAB = make_mrt();

// Per-instance descriptor (the batched DevParams, device_engine.cuh:31-59).
template <typename real>
struct Inst {
  int n, ns;
  int Nx, Ny, Nz, wpr;
  int sx, sxy;                 // coordinate packing: x | y<<sx | z<<sxy
  unsigned mx, my;
  int px, py, pz;
  int hbb, mrt, mass_on, par0;
  const uint64_t *words;
  const uint32_t *coords;
  real *h, *g0, *g1;
  real *craw, *p0, *jx, *jy, *jz;
  real *c, *mu, *ps, *lapmu;
  unsigned long long *acc;
  real rho0, rho1, drho, kappa, four_beta, bnd_coeff, tau0, tau1, mobility, mrt_lambda;
  real gx, gy, gz, fx, fy, fz, ff;
  double M0;
};

// accumulator layout (unsigned long long, NSLOT slots per quantity)
constexpr int ACC_MASS = 0;                 // [parity][5][NSLOT]: M.hi M.lo W.hi W.lo bad
constexpr int ACC_STATS = 10*NSLOT;         // [12][NSLOT]: 5 sums x (hi,lo), max bits, bad
constexpr int ACC_FLAGS = 22*NSLOT;         // [0] pending correction
constexpr int ACC_MISC = 22*NSLOT + 8;      // [5][NSLOT] single-instance reductions
constexpr int ACC_WORDS = 27*NSLOT + 8;

// ---- exact fixed-point sums --------------------------------------------------
// t (|t| < 8) rounded once to a multiple of 2^-60, split into 32-bit halves and
// summed as integers: order-independent, so every block count, batch layout
// and atomic schedule gives the same bits (the reference's double atomicAdd,
// device_engine.cuh:1124-1125, does not).
struct Fx { long long hi; unsigned long long lo; unsigned long long bad; };
__device__ __forceinline__ void fx_add(Fx &a, double t)
{
  if (!(fabs(t) < 8.0)) { a.bad += 1; return; }
  long long q = (long long)__builtin_rint(t * 0x1p60);
  a.hi += q >> 32;
  a.lo += (unsigned long long)(q & 0xffffffffLL);
}
__host__ __device__ inline double fx_value(long long hi, unsigned long long lo)
{
  long long h2 = hi + (long long)(lo >> 32);
  unsigned long long l2 = lo & 0xffffffffULL;
  return ((double)h2 * 4294967296.0 + (double)l2) * 0x1p-60;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
  #pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_max(unsigned long long v)
{
  #pragma unroll
  for (int o = 32; o > 0; o >>= 1) { unsigned long long u = __shfl_xor(v, o, 64); v = u > v ? u : v; }
  return v;
}

// block (256 threads = 4 waves) sum of NV values, one atomicAdd per value
template <int NV>
__device__ __forceinline__ void block_accumulate(unsigned long long (&v)[NV], unsigned long long *dst,
                                                 int stride, int slot, bool is_max_last = false)
{
  __shared__ unsigned long long sh[NV][4];
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  #pragma unroll
  for (int q = 0; q < NV; ++q) {
    bool mx = is_max_last && q == NV - 1;
    unsigned long long s = mx ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0) sh[q][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    int q = threadIdx.x;
    bool mx = is_max_last && q == NV - 1;
    unsigned long long s = sh[q][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = mx ? (sh[q][w] > s ? sh[q][w] : s) : s + sh[q][w];
    if (mx) atomicMax(dst + q*stride + slot, s);
    else if (s) atomicAdd(dst + q*stride + slot, s);
  }
}

// ---- block -> (instance, first site) with an XCD-aware remap -----------------
// Workgroups are dispatched round-robin over the 8 XCDs; the remap hands each
// XCD a contiguous run of the site list so its L2 holds the neighbour planes.
__device__ __forceinline__ int logical_block()
{
  unsigned b = blockIdx.x, nb = gridDim.x;
  return (int)((b & 7u) * (nb >> 3) + (b >> 3));      // nb is a multiple of 8
}

// ---- lattice lookup ---------------------------------------------------------------
// site at (x,y,z) (|offset| <= 2 from a valid voxel): >= 0 site index,
// -1 solid voxel, -2 outside a non-periodic face.
template <typename real>
__device__ __forceinline__ int site_at(const Inst<real> &d, int x, int y, int z)
{
  if (x < 0)        { if (!d.px) return -2; x += d.Nx; if (x < 0) x += d.Nx; }
  else if (x >= d.Nx) { if (!d.px) return -2; x -= d.Nx; if (x >= d.Nx) x -= d.Nx; }
  if (y < 0)        { if (!d.py) return -2; y += d.Ny; if (y < 0) y += d.Ny; }
  else if (y >= d.Ny) { if (!d.py) return -2; y -= d.Ny; if (y >= d.Ny) y -= d.Ny; }
  if (z < 0)        { if (!d.pz) return -2; z += d.Nz; if (z < 0) z += d.Nz; }
  else if (z >= d.Nz) { if (!d.pz) return -2; z -= d.Nz; if (z >= d.Nz) z -= d.Nz; }
  uint64_t w = d.words[((unsigned)z*(unsigned)d.Ny + (unsigned)y)*(unsigned)d.wpr + ((unsigned)x >> 5)];
  unsigned b = (unsigned)x & 31u, m = (unsigned)w;
  if (!((m >> b) & 1u)) return -1;
  return (int)(unsigned)(w >> 32) + __popc(m & ((1u << b) - 1u));
}

template <typename real>
__device__ __forceinline__ void decode(const Inst<real> &d, int i, int &x, int &y, int &z)
{
  uint32_t c = d.coords[i];
  x = (int)(c & d.mx); y = (int)((c >> d.sx) & d.my); z = (int)(c >> d.sxy);
}

// neighbour sites at x + e_m (r = 1) for m = 1..18; n1[0] = i
template <typename real>
__device__ __forceinline__ void neighbours1(const Inst<real> &d, int i, int x, int y, int z, int (&n1)[Q])
{
  n1[0] = i;
  #pragma unroll
  for (int m = 1; m < Q; ++m) n1[m] = site_at(d, x + CX[m], y + CY[m], z + CZ[m]);
}
template <typename real>
__device__ __forceinline__ void neighbours2(const Inst<real> &d, int x, int y, int z, int (&n2)[Q])
{
  n2[0] = -1;
  #pragma unroll
  for (int m = 1; m < Q; ++m) n2[m] = site_at(d, x + 2*CX[m], y + 2*CY[m], z + 2*CZ[m]);
}

template <typename real> __device__ __forceinline__ real W(int k) { return (real)WD[k]; }

// ---- mass-correction scalar from the partial sums of this step ------------------
// lambda = (M - M0)/W applied when W > 0 and M != M0 (multiphase_gpu.cuh:689-700)
template <typename real>
__device__ __forceinline__ double mass_lambda(const Inst<real> &d, unsigned cur, bool &apply)
{
  const unsigned long long *a = d.acc + ACC_MASS + cur*5*NSLOT;
  int lane = threadIdx.x & 63;
  unsigned long long mh = wave_sum(a[0*NSLOT + lane]), ml = wave_sum(a[1*NSLOT + lane]);
  unsigned long long wh = wave_sum(a[2*NSLOT + lane]), wl = wave_sum(a[3*NSLOT + lane]);
  double M = fx_value((long long)mh, ml), Wt = fx_value((long long)wh, wl);
  double drift = M - d.M0;
  apply = (Wt > 0.0 && drift != 0.0);
  return apply ? drift / Wt : 0.0;
}

// folded injection (device_engine.cuh:1128-1139): the reference adds
// (real)(w_k*dc) to each h_k; here the site sum takes dc once.
template <typename real>
__device__ __forceinline__ real corrected(real c, bool apply, double lambda)
{
  if (apply) {
    double cm = (double)c;
    double phi = cm*(1.0 - cm);
    if (phi > 0.0) return (real)(cm + (-lambda*phi));
  }
  return c;
}

// ==== K1: moments (k_moments, device_engine.cuh:116-135, pass 1 only) ===========
template <typename real>
__global__ void __launch_bounds__(256) k_moments(const Inst<real> *__restrict__ D,
                                                 const int2 *__restrict__ map, unsigned T)
{
  int2 bm = map[logical_block()];
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  const int i = bm.y + (int)threadIdx.x;
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  const real *g = cur ? d.g1 : d.g0;
  const int ns = d.ns;
  if (bm.y == 0) {                                  // clear the other parity's partials
    unsigned long long *o = d.acc + ACC_MASS + (cur ^ 1u)*5*NSLOT;
    for (int q = threadIdx.x; q < 5*NSLOT; q += blockDim.x) o[q] = 0ull;
  }
  Fx m = {0, 0, 0}, w = {0, 0, 0};
  if (i < d.n) {
    real c = 0, p = 0, ux = 0, uy = 0, uz = 0;
    #pragma unroll
    for (int k = 0; k < Q; ++k) {
      real hv = d.h[k*ns + i], gv = g[k*ns + i];
      c += hv; p += gv;
      if (CX[k] > 0) ux += gv; else if (CX[k] < 0) ux -= gv;
      if (CY[k] > 0) uy += gv; else if (CY[k] < 0) uy -= gv;
      if (CZ[k] > 0) uz += gv; else if (CZ[k] < 0) uz -= gv;
    }
    d.craw[i] = c; d.p0[i] = p; d.jx[i] = ux; d.jy[i] = uy; d.jz[i] = uz;
    if (d.mass_on) {              // k_mass_weight, device_engine.cuh:1109-1126
      double cm = (double)c;
      double phi = cm*(1.0 - cm); if (phi < 0.0) phi = 0.0;
      fx_add(m, cm); fx_add(w, phi);
    }
  }
  if (d.mass_on) {
    unsigned long long v[5] = { (unsigned long long)m.hi, m.lo, (unsigned long long)w.hi, w.lo, m.bad + w.bad };
    block_accumulate<5>(v, d.acc + ACC_MASS + cur*5*NSLOT, NSLOT, (int)(blockIdx.x & (NSLOT - 1)));
  }
}

// ---- shared stencil pieces -------------------------------------------------------
// central-difference vector gradient (k_grad_cd_mf, device_engine.cuh:897-913)
// with the near-wall cases of multiphase_gpu.cuh:113-121.
template <typename real, typename F>
__device__ __forceinline__ void grad_cd(const int (&n1)[Q], int hbb, real f0, F f,
                                        real &sx, real &sy, real &sz)
{
  sx = 0; sy = 0; sz = 0;
  #pragma unroll
  for (int m = 1; m < Q; ++m) {
    int in = n1[m], ip = n1[OPP[m]];
    real dd;
    if (in >= 0 && ip >= 0) dd = f(in) - f(ip);
    else if (hbb && ip >= 0) dd = f0 - f(ip);
    else if (hbb && in >= 0) dd = f(in) - f0;
    else continue;
    real coef = (real)3.0*(real)0.5*W<real>(m);
    real t = coef*dd;
    if (CX[m] > 0) sx += t; else if (CX[m] < 0) sx -= t;
    if (CY[m] > 0) sy += t; else if (CY[m] < 0) sy -= t;
    if (CZ[m] > 0) sz += t; else if (CZ[m] < 0) sz -= t;
  }
}

// biased per-direction value (bd_dir_val, device_engine.cuh:96-107) with the
// cases of multiphase_gpu.cuh:140-153
template <typename real, typename F>
__device__ __forceinline__ real bd_val(const int (&n1)[Q], const int (&n2)[Q], int hbb, int m, real f0, F f)
{
  int in = n1[m], ip = n1[OPP[m]];
  if (in >= 0) {
    if (n2[m] >= 0) { real f1 = f(in), f2 = f(n2[m]); return (real)(-1.5)*f0 + (real)2.0*f1 + (real)(-0.5)*f2; }
    real f1 = f(in);
    if (!hbb) return (real)(-2.0)*f0 + (real)2.0*f1;
    return (real)(-1.5)*f0 + (real)1.5*f1;
  }
  if (!hbb) {
    if (ip >= 0 && n2[OPP[m]] >= 0) { real f1 = f(ip), f2 = f(n2[OPP[m]]); return (real)(-1.5)*f0 + (real)2.0*f1 + (real)(-0.5)*f2; }
    return (real)0;
  }
  if (ip >= 0) { real f1 = f(ip); return (real)0.5*f0 + (real)(-0.5)*f1; }
  return (real)0;
}
template <typename real>
__device__ __forceinline__ bool bd_has(const int (&n1)[Q], int hbb, int m)
{
  int in = n1[m], ip = n1[OPP[m]];
  if (in >= 0) return true;
  if (!hbb) return false;     // non-hbb backward needs the 2-step check; handled by value 0
  return ip >= 0;
}

// biased vector gradient (k_grad_bd_mf, device_engine.cuh:924-945)
template <typename real, typename F>
__device__ __forceinline__ void grad_bd(const int (&n1)[Q], const int (&n2)[Q], int hbb, real f0, F f,
                                        real &sx, real &sy, real &sz)
{
  sx = 0; sy = 0; sz = 0;
  #pragma unroll
  for (int m = 1; m < Q; ++m) {
    int in = n1[m], ip = n1[OPP[m]];
    bool has;
    if (in >= 0) has = true;
    else if (!hbb) has = (ip >= 0 && n2[OPP[m]] >= 0);
    else has = (ip >= 0);
    if (!has) continue;
    real val = bd_val<real>(n1, n2, hbb, m, f0, f);
    real base = (real)3.0*W<real>(m);
    real t = base*val;
    if (CX[m] > 0) sx += t; else if (CX[m] < 0) sx -= t;
    if (CY[m] > 0) sy += t; else if (CY[m] < 0) sy -= t;
    if (CZ[m] > 0) sz += t; else if (CZ[m] < 0) sz -= t;
  }
}

// node-centred Laplacian (k_lap_mf, device_engine.cuh:996-1010) with the cases
// of multiphase_gpu.cuh:178-188; bnd = the second half of the 2n input
template <typename real, typename F>
__device__ __forceinline__ real laplacian(const int (&n1)[Q], int hbb, real f0, real bnd, F f)
{
  real s = 0;
  #pragma unroll
  for (int m = 1; m < Q; ++m) {
    int in = n1[m], ip = n1[OPP[m]];
    real w = (real)3.0*W<real>(m);
    if (in >= 0 && ip >= 0) { real f1 = f(ip), f2 = f(in); s += w*(f1 - (real)2*f0 + f2); }
    else if (ip >= 0 || in >= 0) {
      real f1 = f(ip >= 0 ? ip : in);
      if (hbb) s += w*(f1 - f0 - bnd);
      else     s += w*((real)2*f1 - (real)2*f0 - (real)2*bnd);
    }
  }
  return s;
}

// ==== K2: c stencil -> corrected c, mu, p* =========================================
// k_pack_lap_c + k_lap_mf + k_mu_axpy + k_grad_cd_mf(c) + k_vel_press_corr
// (device_engine.cuh:149-196).  exp != nullptr also writes rho, ux, uy, uz for
// a field download (multiphase_gpu.cuh:704-757).
template <typename real>
__global__ void __launch_bounds__(256) k_cfield(const Inst<real> *__restrict__ D,
                                                const int2 *__restrict__ map, unsigned T,
                                                real *exp, int exp_inst)
{
  int2 bm = map[logical_block()];
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  bool apply = false; double lambda = 0.0;
  if (d.mass_on && d.acc[ACC_FLAGS]) lambda = mass_lambda(d, cur, apply);
  const int i = bm.y + (int)threadIdx.x;
  if (i >= d.n) return;
  int x, y, z; decode(d, i, x, y, z);
  int n1[Q]; neighbours1(d, i, x, y, z, n1);
  const real *craw = d.craw;
  auto cf = [&](int j) -> real { return corrected<real>(craw[j], apply, lambda); };
  const real c = cf(i);
  // grad_c (central) and lap_c with the wetting boundary term
  real gx, gy, gz; grad_cd<real>(n1, d.hbb, c, cf, gx, gy, gz);
  real bnd = d.bnd_coeff*c*((real)1 - c);
  real lapc = laplacian<real>(n1, d.hbb, c, bnd, cf);
  real c_log = c*((real)1 - c);
  real rho = d.rho1 + d.drho*c;
  real inv = (real)1/((real)(1.0/3.0)*rho);
  real ux = d.jx[i]*inv, uy = d.jy[i]*inv, uz = d.jz[i]*inv;
  real mu = d.four_beta*c_log*((real)0.5 - c);
  mu += -d.kappa*lapc;
  real mor = mu/rho;
  ux = ux + (real)0.5*(d.gx + d.fx*d.ff/rho + mor*gx);
  uy = uy + (real)0.5*(d.gy + d.fy*d.ff/rho + mor*gy);
  uz = uz + (real)0.5*(d.gz + d.fz*d.ff/rho + mor*gz);
  real u_dot_drho = d.drho*(ux*gx + uy*gy + uz*gz);
  real p = d.p0[i];
  p += (real)0.5*u_dot_drho*(real)(1.0/3.0);
  d.c[i] = c; d.mu[i] = mu; d.ps[i] = p;
  if (exp && bm.x == exp_inst) { int ns = d.ns; exp[i] = rho; exp[ns + i] = ux; exp[2*ns + i] = uy; exp[3*ns + i] = uz; }
}

// ==== K3: lap(mu) (k_pack_lap_zero + k_lap_mf) ======================================
template <typename real>
__global__ void __launch_bounds__(256) k_mufield(const Inst<real> *__restrict__ D,
                                                 const int2 *__restrict__ map, unsigned T)
{
  int2 bm = map[logical_block()];
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  const int i = bm.y + (int)threadIdx.x;
  if (i >= d.n) return;
  int x, y, z; decode(d, i, x, y, z);
  int n1[Q]; neighbours1(d, i, x, y, z, n1);
  const real *mu = d.mu;
  auto f = [&](int j) -> real { return mu[j]; };
  d.lapmu[i] = laplacian<real>(n1, d.hbb, mu[i], (real)0, f);
}

// ==== K4: fused collision + push streaming ==========================================
// k_collide_fused_bgk (device_engine.cuh:459-545) and k_collide_fused_mrt
// (:547-651), with the field quantities of the reference's separate kernels
// (k_moments u, k_grad_cd/bd/md, k_vel_press_corr) recomputed in registers.
template <typename real, bool MRT>
__global__ void __launch_bounds__(256) k_collide(const Inst<real> *__restrict__ D,
                                                 const int2 *__restrict__ map, unsigned T, int stats)
{
  int2 bm = map[logical_block()];
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  const int i = bm.y + (int)threadIdx.x;
  const bool live = i < d.n;
  Fx s0 = {0,0,0}, s1 = {0,0,0}, s2 = {0,0,0}, s3 = {0,0,0}, s4 = {0,0,0};
  unsigned long long vmax = 0;
  if (threadIdx.x == 0 && bm.y == 0 && d.mass_on) d.acc[ACC_FLAGS] = 1ull;   // correction owed from now on
  if (live) {
    const int ns = d.ns;
    const real *gin = cur ? d.g1 : d.g0;
    real *gout = cur ? d.g0 : d.g1;
    int x, y, z; decode(d, i, x, y, z);
    int n1[Q], n2[Q];
    neighbours1(d, i, x, y, z, n1);
    neighbours2(d, x, y, z, n2);
    const real *cF = d.c, *pF = d.ps, *lF = d.lapmu;
    auto fc = [&](int j) -> real { return cF[j]; };
    auto fp = [&](int j) -> real { return pF[j]; };
    const real c = cF[i], p = pF[i], mu = d.mu[i], lmu = lF[i];
    const real drho = d.drho;
    const real CS2 = (real)(1.0/3.0), ALPHA = (real)3.0, GAMMA_C = (real)1.5, BETA_C = (real)4.5;
    const real rho = d.rho1 + drho*c;
    real cp = c < (real)0 ? (real)0 : (c > (real)1 ? (real)1 : c);
    real tau = (real)1/(cp/d.tau0 + ((real)1 - cp)/d.tau1);
    const real relax = (real)1/(tau + (real)0.5);
    // own-site post-stream g and its first moment (k_moments order)
    real gk[Q];
    real jx = 0, jy = 0, jz = 0;
    #pragma unroll
    for (int k = 0; k < Q; ++k) {
      real gv = gin[k*ns + i]; gk[k] = gv;
      if (CX[k] > 0) jx += gv; else if (CX[k] < 0) jx -= gv;
      if (CY[k] > 0) jy += gv; else if (CY[k] < 0) jy -= gv;
      if (CZ[k] > 0) jz += gv; else if (CZ[k] < 0) jz -= gv;
    }
    // gradients of c and p* (cd, bd, md)
    real e_dcx, e_dcy, e_dcz, e_dpx, e_dpy, e_dpz, b_cx, b_cy, b_cz, b_px, b_py, b_pz;
    grad_cd<real>(n1, d.hbb, c, fc, e_dcx, e_dcy, e_dcz);
    grad_bd<real>(n1, n2, d.hbb, c, fc, b_cx, b_cy, b_cz);
    grad_cd<real>(n1, d.hbb, p, fp, e_dpx, e_dpy, e_dpz);
    grad_bd<real>(n1, n2, d.hbb, p, fp, b_px, b_py, b_pz);
    const real f_dcx = (real)0.5*(e_dcx + b_cx), f_dcy = (real)0.5*(e_dcy + b_cy), f_dcz = (real)0.5*(e_dcz + b_cz);
    const real f_dpx = (real)0.5*(e_dpx + b_px), f_dpy = (real)0.5*(e_dpy + b_py), f_dpz = (real)0.5*(e_dpz + b_pz);
    // u* = u + 1/2 (g + f ff/rho + (mu/rho) grad_cd c)   (k_vel_press_corr)
    real inv = (real)1/(CS2*rho);
    real ux = jx*inv, uy = jy*inv, uz = jz*inv;
    {
      real mor = mu/rho;
      ux = ux + (real)0.5*(d.gx + d.fx*d.ff/rho + mor*e_dcx);
      uy = uy + (real)0.5*(d.gy + d.fy*d.ff/rho + mor*e_dcy);
      uz = uz + (real)0.5*(d.gz + d.fz*d.ff/rho + mor*e_dcz);
    }
    if (stats) {                  // k_flow_stats, device_engine.cuh:660-683
      double vx = (double)ux, vy = (double)uy, vz = (double)uz;
      double v2 = vx*vx + vy*vy + vz*vz;
      fx_add(s0, vx); fx_add(s1, vy); fx_add(s2, vz); fx_add(s3, v2); fx_add(s4, (double)c);
      vmax = (unsigned long long)__double_as_longlong(v2);
    }
    const real bfx = rho*d.gx + d.fx*d.ff, bfy = rho*d.gy + d.fy*d.ff, bfz = rho*d.gz + d.fz*d.ff;
    real *hO = d.h;
    if constexpr (!MRT) {
      const real drcx = drho*e_dcx*CS2, drcy = drho*e_dcy*CS2, drcz = drho*e_dcz*CS2;
      const real e_fgx = mu*e_dcx + bfx, e_fgy = mu*e_dcy + bfy, e_fgz = mu*e_dcz + bfz;
      const real a_over_rho = ALPHA/rho;
      const real e_fhx = e_dcx - c*a_over_rho*(e_dpx - e_fgx), e_fhy = e_dcy - c*a_over_rho*(e_dpy - e_fgy), e_fhz = e_dcz - c*a_over_rho*(e_dpz - e_fgz);
      const real u_sq = ux*ux + uy*uy + uz*uz;
      const real u_fh_e = ux*e_fhx + uy*e_fhy + uz*e_fhz;
      const real u_drhocd = ux*drcx + uy*drcy + uz*drcz;
      const real u_fg_e = ux*e_fgx + uy*e_fgy + uz*e_fgz;
      const real drho_cs2 = drho*CS2;
      const real drmx = drho_cs2*f_dcx, drmy = drho_cs2*f_dcy, drmz = drho_cs2*f_dcz;
      const real c_over = c/CS2/rho;
      const real f_fgx = mu*f_dcx + bfx, f_fgy = mu*f_dcy + bfy, f_fgz = mu*f_dcz + bfz;
      const real f_fhx = f_dcx - c_over*(f_dpx - f_fgx), f_fhy = f_dcy - c_over*(f_dpy - f_fgy), f_fhz = f_dcz - c_over*(f_dpz - f_fgz);
      const real u_fh_f = ux*f_fhx + uy*f_fhy + uz*f_fhz;
      const real u_drm = ux*drmx + uy*drmy + uz*drmz;
      const real u_fg_f = ux*f_fgx + uy*f_fgy + uz*f_fgz;
      #pragma unroll
      for (int k = 0; k < Q; ++k) {
        const real w = W<real>(k);
        real uk = 0;
        if (CX[k] > 0) uk += ux; else if (CX[k] < 0) uk -= ux;
        if (CY[k] > 0) uk += uy; else if (CY[k] < 0) uk -= uy;
        if (CZ[k] > 0) uk += uz; else if (CZ[k] < 0) uk -= uz;
        real u_term = uk*(ALPHA + BETA_C*uk) - GAMMA_C*u_sq;
        real G = w*((real)1 + u_term);
        real f_proj = 0;
        if (CX[k] > 0) f_proj += bfx; else if (CX[k] < 0) f_proj -= bfx;
        if (CY[k] > 0) f_proj += bfy; else if (CY[k] < 0) f_proj -= bfy;
        if (CZ[k] > 0) f_proj += bfz; else if (CZ[k] < 0) f_proj -= bfz;
        real cd_c = 0, cd_p = 0, bd_c = 0, bd_p = 0, avgk;
        if (k > 0) {
          int in = n1[k], ip = n1[OPP[k]];
          if (in >= 0 && ip >= 0) { cd_c = (real)0.5*(cF[in] - cF[ip]); cd_p = (real)0.5*(pF[in] - pF[ip]); }
          else if (d.hbb && ip >= 0) { cd_c = (real)0.5*(c - cF[ip]); cd_p = (real)0.5*(p - pF[ip]); }
          else if (d.hbb && in >= 0) { cd_c = (real)0.5*(cF[in] - c); cd_p = (real)0.5*(pF[in] - p); }
          bd_c = bd_val<real>(n1, n2, d.hbb, k, c, fc); bd_p = bd_val<real>(n1, n2, d.hbb, k, p, fp);
          avgk = in >= 0 ? (real)0.5*(lmu + lF[in]) : lmu;
        } else avgk = (real)0.5*(lmu + lmu);
        real drho_dk = drho*cd_c*CS2;
        real eh = c*G;
        eh -= (real)0.5*G*(cd_c - c*a_over_rho*(cd_p - mu*cd_c - f_proj));
        eh += (real)0.5*G*u_fh_e;
        real eg = w*(p + rho*CS2*u_term);
        eg -= (real)0.5*(drho_dk*(G - w) + (mu*cd_c + f_proj)*G);
        eg += (real)0.5*((G - w)*u_drhocd + G*u_fg_e);
        real dGk = G - w;
        real dc_dk = (real)0.5*(cd_c + bd_c), dp_dk = (real)0.5*(cd_p + bd_p);
        real drho_dk2 = drho_cs2*dc_dk;
        real ffk = mu*dc_dk + f_proj;
        real fh = G*(dc_dk - c_over*(dp_dk - ffk) - u_fh_f + d.mobility*avgk);
        real fg = (drho_dk2 - u_drm)*dGk + (ffk - u_fg_f)*G;
        real gnew = gk[k] + relax*(eg - gk[k]) + fg;
        real hnew = eh + fh;
        int dst = n1[k];
        if (dst >= 0) { hO[k*ns + dst] = hnew; gout[k*ns + dst] = gnew; }
        else if (dst == -1 && d.hbb) { hO[OPP[k]*ns + i] = hnew; gout[OPP[k]*ns + i] = gnew; }
        else { hO[OPP[k]*ns + i] = (real)0; gout[OPP[k]*ns + i] = (real)0; }
      }
    } else {
      const real drho_cs2 = drho*CS2, rho_cs2 = rho*CS2, c_over = c/rho_cs2;
      const real drcx = drho_cs2*e_dcx, drcy = drho_cs2*e_dcy, drcz = drho_cs2*e_dcz;
      const real e_fgx = mu*e_dcx + bfx, e_fgy = mu*e_dcy + bfy, e_fgz = mu*e_dcz + bfz;
      const real e_fhx = e_dcx - c_over*(e_dpx - e_fgx), e_fhy = e_dcy - c_over*(e_dpy - e_fgy), e_fhz = e_dcz - c_over*(e_dpz - e_fgz);
      const real gamma_u_sq = GAMMA_C*(ux*ux + uy*uy + uz*uz);
      const real u_term_h = ux*e_fhx + uy*e_fhy + uz*e_fhz;
      const real u_force_g = ux*e_fgx + uy*e_fgy + uz*e_fgz;
      const real u_drho = ux*drcx + uy*drcy + uz*drcz;
      const real drmx = drho_cs2*f_dcx, drmy = drho_cs2*f_dcy, drmz = drho_cs2*f_dcz;
      const real f_fgx = mu*f_dcx + bfx, f_fgy = mu*f_dcy + bfy, f_fgz = mu*f_dcz + bfz;
      const real f_fhx = f_dcx - c_over*(f_dpx - f_fgx), f_fhy = f_dcy - c_over*(f_dpy - f_fgy), f_fhz = f_dcz - c_over*(f_dpz - f_fgz);
      const real u_fh_f = ux*f_fhx + uy*f_fhy + uz*f_fhz;
      const real u_drm = ux*drmx + uy*drmy + uz*drmz;
      const real u_fg_f = ux*f_fgx + uy*f_fgy + uz*f_fgz;
      real cg[Q], fgs[Q];
      #pragma unroll
      for (int k = 0; k < Q; ++k) {
        const real w = W<real>(k);
        real uk = 0;
        if (CX[k] > 0) uk += ux; else if (CX[k] < 0) uk -= ux;
        if (CY[k] > 0) uk += uy; else if (CY[k] < 0) uk -= uy;
        if (CZ[k] > 0) uk += uz; else if (CZ[k] < 0) uk -= uz;
        real u_term = uk*(ALPHA + BETA_C*uk) - gamma_u_sq;
        real G = w*((real)1 + u_term);
        real w_u_term = G - w;
        real f_proj = 0;
        if (CX[k] > 0) f_proj += bfx; else if (CX[k] < 0) f_proj -= bfx;
        if (CY[k] > 0) f_proj += bfy; else if (CY[k] < 0) f_proj -= bfy;
        if (CZ[k] > 0) f_proj += bfz; else if (CZ[k] < 0) f_proj -= bfz;
        real cd_c = 0, cd_p = 0, bd_c = 0, bd_p = 0, avgk;
        if (k > 0) {
          int in = n1[k], ip = n1[OPP[k]];
          if (in >= 0 && ip >= 0) { cd_c = (real)0.5*(cF[in] - cF[ip]); cd_p = (real)0.5*(pF[in] - pF[ip]); }
          else if (d.hbb && ip >= 0) { cd_c = (real)0.5*(c - cF[ip]); cd_p = (real)0.5*(p - pF[ip]); }
          else if (d.hbb && in >= 0) { cd_c = (real)0.5*(cF[in] - c); cd_p = (real)0.5*(pF[in] - p); }
          bd_c = bd_val<real>(n1, n2, d.hbb, k, c, fc); bd_p = bd_val<real>(n1, n2, d.hbb, k, p, fp);
          avgk = in >= 0 ? (real)0.5*(lmu + lF[in]) : lmu;
        } else avgk = (real)0.5*(lmu + lmu);
        real drho_dk = drho_cs2*cd_c;
        real eh = c*G + (real)0.5*G*(u_term_h - cd_c + c_over*(cd_p - mu*cd_c - f_proj));
        real eg = w*p + rho_cs2*w_u_term + (real)0.5*((u_drho - drho_dk)*w_u_term + (u_force_g + mu*cd_c + f_proj)*G);
        real dGk = G - w;
        real dc_dk = (real)0.5*(cd_c + bd_c), dp_dk = (real)0.5*(cd_p + bd_p);
        real drho_dk2 = drho_cs2*dc_dk;
        real ffk = mu*dc_dk + f_proj;
        real fh = G*(dc_dk - c_over*(dp_dk - ffk) - u_fh_f + d.mobility*avgk);
        real fg = (drho_dk2 - u_drm)*dGk + (ffk - u_fg_f)*G;
        real hnew = eh + fh;
        int dst = n1[k];
        if (dst >= 0) hO[k*ns + dst] = hnew;
        else if (dst == -1 && d.hbb) hO[OPP[k]*ns + i] = hnew;
        else hO[OPP[k]*ns + i] = (real)0;
        cg[k] = eg - gk[k]; fgs[k] = fg;
      }
      // cg <- M^-1 S M cg (device_engine.cuh:617-648); S from relax and mrt_lambda
      const real s0r = relax; const real taur = (real)1/s0r - (real)0.5;
      const real s1r = (real)2*taur/((real)2*d.mrt_lambda + taur);
      real xx[Q];
      #pragma unroll
      for (int m = 0; m < Q; ++m) xx[m] = 0;
      #pragma unroll
      for (int nn = 0; nn < Q; ++nn) {
        real yv = cg[nn];
        #pragma unroll
        for (int m = 1; m < Q; ++m) {
          if (MRTTAB.M[m][nn] != 0.0) {
            const bool odd = (m == 4 || m == 6 || m == 8 || m >= 16);
            const bool cons = (m == 3 || m == 5 || m == 7);
            if (!cons) {
              real sm = odd ? s1r : s0r;
              xx[m] += sm*(real)MRTTAB.M[m][nn]*yv;
            }
          }
        }
      }
      #pragma unroll
      for (int m = 0; m < Q; ++m) {
        real acc = 0;
        #pragma unroll
        for (int nn = 0; nn < Q; ++nn) {
          const bool cons = (nn == 0 || nn == 3 || nn == 5 || nn == 7);
          if (MRTTAB.MI[m][nn] != 0.0 && !cons) acc += (real)MRTTAB.MI[m][nn]*xx[nn];
        }
        real gnew = gk[m] + acc + fgs[m];
        int dst = n1[m];
        if (dst >= 0) gout[m*ns + dst] = gnew;
        else if (dst == -1 && d.hbb) gout[OPP[m]*ns + i] = gnew;
        else gout[OPP[m]*ns + i] = (real)0;
      }
    }
  }
  if (stats) {
    unsigned long long v[12] = { (unsigned long long)s0.hi, s0.lo, (unsigned long long)s1.hi, s1.lo,
                                 (unsigned long long)s2.hi, s2.lo, (unsigned long long)s3.hi, s3.lo,
                                 (unsigned long long)s4.hi, s4.lo, s0.bad + s3.bad + s4.bad, vmax };
    block_accumulate<12>(v, d.acc + ACC_STATS, NSLOT, (int)(blockIdx.x & (NSLOT - 1)), true);
  }
}

}  // namespace fm

namespace fm {

// ==== single-instance helpers (not on the step path) ================================

// mass sums of the current state (k_mass_weight, device_engine.cuh:1109-1126);
// with apply, the sum of the folded-corrected c instead (the effective mass).
template <typename real>
__global__ void __launch_bounds__(256) k_mass_single(Inst<real> d, unsigned cur, int apply, double lambda)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  const real *g = cur ? d.g1 : d.g0; (void)g;
  Fx m = {0, 0, 0}, w = {0, 0, 0};
  if (i < d.n) {
    real c = 0;
    #pragma unroll
    for (int k = 0; k < Q; ++k) c += d.h[k*d.ns + i];
    if (apply) c = corrected<real>(c, true, lambda);
    double cm = (double)c;
    double phi = cm*(1.0 - cm); if (phi < 0.0) phi = 0.0;
    fx_add(m, cm); fx_add(w, phi);
  }
  unsigned long long v[5] = { (unsigned long long)m.hi, m.lo, (unsigned long long)w.hi, w.lo, m.bad + w.bad };
  block_accumulate<5>(v, d.acc + ACC_MISC, NSLOT, (int)(blockIdx.x & (NSLOT - 1)));
}

// h with the owed correction injected per population, as the reference's
// k_inject_mass (device_engine.cuh:1128-1139) leaves it; written to out.
template <typename real>
__global__ void __launch_bounds__(256) k_inject_copy(Inst<real> d, int apply, double lambda, real *out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const int ns = d.ns;
  real c = 0;
  #pragma unroll
  for (int k = 0; k < Q; ++k) c += d.h[k*ns + i];
  double cm = (double)c, phi = cm*(1.0 - cm);
  bool inj = apply && phi > 0.0;
  double delta_c = -lambda*phi;
  #pragma unroll
  for (int k = 0; k < Q; ++k) {
    real v = d.h[k*ns + i];
    if (inj) v += (real)(WD[k]*delta_c);
    out[k*ns + i] = v;
  }
}

// reference-format stencil/streaming tables (multiphase_gpu.cuh:93-208)
// regenerated from the engine's neighbour logic, for bitwise comparison
template <typename real>
__global__ void __launch_bounds__(256) k_export_tables(Inst<real> d, int *cdm, int *cdp, int *bcase,
                                                       int *bc1, int *bc2, int *avgnext, int *lcase,
                                                       int *lc1, int *lc2, int *src)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const int n = d.n;
  int x, y, z; decode(d, i, x, y, z);
  int n1[Q], n2[Q]; neighbours1(d, i, x, y, z, n1); neighbours2(d, x, y, z, n2);
  for (int m = 1; m < Q; ++m) {
    int in = n1[m], ip = n1[OPP[m]];
    size_t idx = (size_t)(m - 1)*n + i;
    int a = -1, b = -1;
    if (in >= 0 && ip >= 0) { a = ip; b = in; }
    else if (d.hbb && ip >= 0) { a = ip; b = i; }
    else if (d.hbb && in >= 0) { a = i; b = in; }
    cdm[idx] = a; cdp[idx] = b;
    int cs = 0, c1 = -1, c2 = -1;
    if (in >= 0) {
      if (n2[m] >= 0) { cs = 1; c1 = in; c2 = n2[m]; }
      else { cs = d.hbb ? 3 : 2; c1 = in; }
    } else if (!d.hbb) {
      if (ip >= 0 && n2[OPP[m]] >= 0) { cs = 1; c1 = ip; c2 = n2[OPP[m]]; }
    } else if (ip >= 0) { cs = 4; c1 = ip; }
    bcase[idx] = cs; bc1[idx] = c1; bc2[idx] = c2;
    cs = 0; c1 = -1; c2 = -1;
    if (in >= 0 && ip >= 0) { cs = 1; c1 = ip; c2 = in; }
    else if (ip >= 0) { cs = d.hbb ? 3 : 2; c1 = ip; }
    else if (in >= 0) { cs = d.hbb ? 3 : 2; c1 = in; }
    lcase[idx] = cs; lc1[idx] = c1; lc2[idx] = c2;
  }
  for (int m = 0; m < Q; ++m) avgnext[(size_t)m*n + i] = n1[m] >= 0 ? n1[m] : -1;
  // src: population k at i after streaming comes from (k, x - e_k), or
  // bounce-back slot opp(k) of i, or nothing (-2)
  src[i] = i;
  for (int k = 1; k < Q; ++k) {
    int s = n1[OPP[k]];
    int v = s >= 0 ? k*n + s : ((s == -1 && d.hbb) ? OPP[k]*n + i : -2);
    src[(size_t)k*n + i] = v;
  }
}

template <typename real>
__global__ void k_zero_words(unsigned long long *p, int count)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

}  // namespace fm
