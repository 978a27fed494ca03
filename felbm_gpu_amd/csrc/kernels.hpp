// kernels.hpp -- gfx950 kernels of the Lee-Liu D3Q19 step.
//
// One step is four launches over the flattened site list of every instance:
//   k_moments   gathered (streamed) h,g -> c_raw, p0, j (+ exact FP64 mass partials)
//   k_cfield    c stencil (r=1) -> corrected c, mu, p*   (+ optional export)
//   k_mufield   mu stencil (r=1) -> lap(mu)
//   k_collide   equilibria + Guo force + BGK/MRT collision; post-collision
//               populations stored at their own site (the next g buffer, h)
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
// Streaming is a gather (pull): the collide kernel stores post-collision
// population k of site x in its own slot k (contiguous stores), as zero when it
// would leave a non-periodic box face or hit a solid without halfway
// bounce-back; the next moments and collide passes read population k of x from
// slot k of x - e_k, or -- x - e_k not a site -- from slot opp(k) of x (the
// bounced-back population).  State uploads and downloads convert to and from
// the post-stream layout the reference's d_h/d_g hold between steps.
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
struct MrtTab { double M[19][19]; double MI[19][19]; double N[19]; };
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
    t.N[m] = nrm;
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
  unsigned nwords;
  real *h, *g0, *g1;
  real *craw, *p0, *jx, *jy, *jz;
  real *c, *mu, *ps, *clo;      // clo: FP32 remainder Sigma h - c_raw of the moments pass
  unsigned long long *acc;
  real rho0, rho1, drho, kappa, four_beta, bnd_coeff, tau0, tau1, mobility, mrt_lambda;
  real gx, gy, gz, fx, fy, fz, ff;
  double M0;
  double lam; int lam_on;     // this step's mass-correction lambda, written by k_mass_lambda
  // open boundaries (the DevParams open-BC block, device_engine.cuh:45-58)
  int obc;                     // inlet/outlet lists present
  const int *ent;              // [n]: -1, or (entry << 4) | ENT_* bits
  real *ov;                    // [entries][8]: c, rho, p, ux, uy, uz, own raw c
  int use_in_vel, use_in_prs, use_in_fluid, use_out_prs, use_out_fluid, inlet_mode;
  double u_in[3], p_in, p_out, c_out_fixed, rho_out_fixed;
  double inlet_c_a, inlet_period, inlet_duty, inlet_ramp, split_pos;
  // tracers: the velocity snapshot they are advected on
  int trk;                     // tracers present
  int tr_vskip;                // refresh the snapshot every tr_vskip steps
  int tr_phase;                // the snapshot also holds c (phase split of the stretching rows)
  real *tu;                    // [ns][4] snapshot of (u*, c) of the field pass, one record per site
};

// site classes of the open boundary
constexpr int ENT_BCPROC = 1;     // fields imposed: no velocity/pressure correction
constexpr int ENT_EQ = 2;         // g replaced by the equilibrium of the site's fields
constexpr int ENT_NOSTREAM = 4;   // buffer node: no collision, stores its equilibrium

// one open-boundary node: `site` takes the imposed fields of `ref` (itself for
// inlet/outlet nodes, the mirrored vertex for a buffer node); rkind is the
// ref's class (0 inlet, 1 outlet, 3 neither), rcoord its split-axis coordinate
struct ObEntry { int desc, site, ref, rkind, eid, pad; double rcoord; };

// accumulator layout (unsigned long long, NSLOT slots per quantity)
constexpr int ACC_MASS = 0;                 // [parity][5][NSLOT]: M.hi M.lo W.hi W.lo bad
constexpr int ACC_STATS = 10*NSLOT;         // [12][NSLOT]: 5 sums x (hi,lo), max bits, bad
constexpr int ACC_FLAGS = 22*NSLOT;         // [0] pending correction, [1] open-boundary step, [2] tracer step
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
// A pass launched over the blocks of one instance gets its descriptor index as
// the argument sdi (first site = logical block x sites per workgroup) and reads
// no block map; sdi < 0 selects the map of a batch (one dependent load less
// per wave: c and mu stencil passes 0.163/0.138 -> 0.160/0.136 ms at 300^3).
// Workgroups are dispatched round-robin over the 8 XCDs; the remap hands each
// XCD a contiguous run of the site list so its L2 holds the neighbour planes.
__device__ __forceinline__ int logical_block()
{
  unsigned b = blockIdx.x, nb = gridDim.x;
  return (int)((b & 7u) * (nb >> 3) + (b >> 3));      // nb is a multiple of 8
}

// A block-map entry (descriptor, first site) of an instance with non-periodic
// faces or open-boundary nodes carries MAP_INTERIOR when none of its sites is
// within 2 voxels of a non-periodic face or an open-boundary node (nor site
// 0, whose block advances the step counters): the kernels of the open
// variant run such a workgroup through the periodic code, the same
// arithmetic without the in-box tests and boundary-node lookups.
constexpr int MAP_INTERIOR = 1 << 30;
__device__ __forceinline__ int2 map_entry(const int2 *__restrict__ map, int sdi)
{
  return sdi >= 0 ? make_int2(sdi, logical_block()*(int)blockDim.x) : map[logical_block()];
}

// ---- memory access ------------------------------------------------------------------
// Generic pointers from the descriptor are cast to address space 1 for plain
// global accesses; the SoA arrays go through raw buffer loads/stores so the
// per-direction offset k*ns lives in a scalar register (no VALU address math).
template <typename T>
__device__ __forceinline__ T ldg(const T *p, long i) { return ((const __attribute__((address_space(1))) T *)p)[i]; }
template <typename T>
__device__ __forceinline__ void stg(T *p, long i, T v) { ((__attribute__((address_space(1))) T *)p)[i] = v; }

using rsrc_t = __amdgpu_buffer_rsrc_t;
// raw buffer resource: base, stride 0, num_records in bytes, gfx9 data format word
__device__ __forceinline__ rsrc_t make_rsrc(const void *p, unsigned bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T bld(rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ float bld<float>(rsrc_t r, unsigned vo, unsigned so)
{ return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0)); }
template <> __device__ __forceinline__ double bld<double>(rsrc_t r, unsigned vo, unsigned so)
{ return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0)); }
template <> __device__ __forceinline__ uint64_t bld<uint64_t>(rsrc_t r, unsigned vo, unsigned so)
{ return __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0)); }
// loads of data read once per step: non-temporal, so they do not displace
// the stencil fields re-read from L2 (not for gathers whose lines adjacent
// waves share through L2)
template <typename T> __device__ __forceinline__ T bld_nt(rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ float bld_nt<float>(rsrc_t r, unsigned vo, unsigned so)
{ return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 2)); }
template <> __device__ __forceinline__ double bld_nt<double>(rsrc_t r, unsigned vo, unsigned so)
{ return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 2)); }
__device__ __forceinline__ void bst(float v, rsrc_t r, unsigned vo, unsigned so)
{ __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo, (int)so, 0); }
__device__ __forceinline__ void bst(double v, rsrc_t r, unsigned vo, unsigned so)
{ __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, (int)vo, (int)so, 0); }

// field block layout (ns values each): c_raw; [p0, jx, jy, jz] interleaved per
// site (4 ns); [c, p*] interleaved per site (2 ns); mu; the FP32 remainder
// Sigma h - c_raw of the moments pass (unused in FP64); then 4 x ns
// holding either what the mu-stencil pass writes for the collision -- FP32:
// 12-byte per-site records [c, p*, lapmu] (one vector load per neighbour; the
// collision reads mu only at its own site, from its field, so the records the
// +-2 planes keep in L2 are 12 rather than 16 bytes a site: collision 0.725 ->
// 0.699 ms, mu pass 0.135 -> 0.129 ms at 300^3); FP64: lapmu alone (F_LMU),
// the collision reading the [c, p*] pairs, whose 16 bytes a record would
// double -- or, between steps, the exported rho, ux, uy, uz of a field
// download.  A pass reads its own site's inputs with one vector access per
// group.
enum { F_CRAW = 0, F_PJ = 1, F_CP = 5, F_MU = 7, F_CLO = 8, F_REC = 9, F_LMU = 9, F_COUNT = 13 };

// record loads/stores of one site, vo = site byte offset in a scalar field
// (i * sizeof(real)).  The loaded vector is bit-cast as a whole: a bit-cast of one element of the
// builtin's result (bit_cast<float>(v[1])) compiles to a read of element 0 with
// this toolchain, which narrows the load to a single dword.
template <typename real> struct Rec2 { real c, p; };
template <typename real> struct Rec4 { real c, p, l, m; };

// [p0, jx, jy, jz] at byte offset 4*io and [c, p*] at 2*io of their blocks
// (AUX: the buffer instruction's cache policy, 2 = non-temporal)
template <int AUX = 0>
__device__ __forceinline__ void bst_pj(float p, float x, float y, float z, rsrc_t r, unsigned io, unsigned so)
{
  typedef __attribute__((ext_vector_type(4))) unsigned u4;
  __builtin_amdgcn_raw_buffer_store_b128(u4{__builtin_bit_cast(unsigned, p), __builtin_bit_cast(unsigned, x),
                                            __builtin_bit_cast(unsigned, y), __builtin_bit_cast(unsigned, z)},
                                         r, (int)(io*4u), (int)so, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ void bst_pj(double p, double x, double y, double z, rsrc_t r, unsigned io, unsigned so)
{
  typedef __attribute__((ext_vector_type(2))) unsigned u2;
  typedef __attribute__((ext_vector_type(4))) unsigned u4;
  const u2 a = __builtin_bit_cast(u2, p), b = __builtin_bit_cast(u2, x), c = __builtin_bit_cast(u2, y), e = __builtin_bit_cast(u2, z);
  __builtin_amdgcn_raw_buffer_store_b128(u4{a[0], a[1], b[0], b[1]}, r, (int)(io*4u), (int)so, AUX);
  __builtin_amdgcn_raw_buffer_store_b128(u4{c[0], c[1], e[0], e[1]}, r, (int)(io*4u + 16u), (int)so, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ Rec4<float> bld_pj(const float *, rsrc_t r, unsigned io, unsigned so)
{
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const f32x4 f = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(io*4u), (int)so, AUX));
  return { f[0], f[1], f[2], f[3] };
}
template <int AUX = 0>
__device__ __forceinline__ Rec4<double> bld_pj(const double *, rsrc_t r, unsigned io, unsigned so)
{
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  const f64x2 a = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(io*4u), (int)so, AUX));
  const f64x2 b = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(io*4u + 16u), (int)so, AUX));
  return { a[0], a[1], b[0], b[1] };
}
template <int AUX = 0>
__device__ __forceinline__ void bst_cp(float c, float p, rsrc_t r, unsigned io, unsigned so)
{
  typedef __attribute__((ext_vector_type(2))) unsigned u2;
  __builtin_amdgcn_raw_buffer_store_b64(u2{__builtin_bit_cast(unsigned, c), __builtin_bit_cast(unsigned, p)},
                                        r, (int)(io*2u), (int)so, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ void bst_cp(double c, double p, rsrc_t r, unsigned io, unsigned so)
{
  typedef __attribute__((ext_vector_type(2))) unsigned u2;
  typedef __attribute__((ext_vector_type(4))) unsigned u4;
  const u2 a = __builtin_bit_cast(u2, c), b = __builtin_bit_cast(u2, p);
  __builtin_amdgcn_raw_buffer_store_b128(u4{a[0], a[1], b[0], b[1]}, r, (int)(io*2u), (int)so, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ Rec2<float> bld_cp(const float *, rsrc_t r, unsigned io, unsigned so)
{
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  const f32x2 f = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(io*2u), (int)so, AUX));
  return { f[0], f[1] };
}
template <int AUX = 0>
__device__ __forceinline__ Rec2<double> bld_cp(const double *, rsrc_t r, unsigned io, unsigned so)
{
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  const f64x2 f = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(io*2u), (int)so, AUX));
  return { f[0], f[1] };
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// 12-byte records [c, p*, lapmu] at 12*i (vo = 4*i)
__device__ __forceinline__ Rec4<float> bld_r3(rsrc_t r, unsigned vo, unsigned so)
{
  typedef float f32x3 __attribute__((ext_vector_type(3)));
  const f32x3 f = __builtin_bit_cast(f32x3, __builtin_amdgcn_raw_buffer_load_b96(r, (int)(vo*3u), (int)so, 0));
  return { f[0], f[1], f[2], 0.0f };
}
__device__ __forceinline__ Rec2<float> bld_r2(rsrc_t r, unsigned vo, unsigned so)
{
  const f32x2 f = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(vo*3u), (int)so, 0));
  return { f[0], f[1] };
}
__device__ __forceinline__ void bst_r3(float c, float p, float l, rsrc_t r, unsigned vo, unsigned so)
{
  typedef __attribute__((ext_vector_type(3))) unsigned u3;
  u3 v = { __builtin_bit_cast(unsigned, c), __builtin_bit_cast(unsigned, p), __builtin_bit_cast(unsigned, l) };
  __builtin_amdgcn_raw_buffer_store_b96(v, r, (int)(vo*3u), (int)so, 0);
}

// ---- lattice lookup ---------------------------------------------------------------
// The wrapped coordinates x-2..x+2 (and y, z) of one site, computed once;
// bit a*5+j of ok says whether offset j-2 on axis a exists (always on a
// periodic axis).  Out-of-box coordinates are still wrapped, so every word
// load stays inside the table.
// yr and zr hold the wrapped y and z already scaled to row-word offsets.
// The rings read each row once, as a 16-byte window of the two words xb and
// xb+1 that hold x-2..x+2; anyw[S-1] says whether some lane of the wave is
// within S of an x face, where x+dx (|dx| <= S) may wrap outside that window.
struct Nbhd { int xw[5], yr[5], zr[5]; unsigned ok, xb; bool anyw[2]; };

// offsets are within +-2, so one correction wraps any axis of length >= 2;
// an axis of length 1 maps everything to 0
__device__ __forceinline__ int wrap_axis(int v, int N, int per, bool &ok)
{
  ok = per || (v >= 0 && v < N);
  v = v < 0 ? v + N : v;
  v = v >= N ? v - N : v;
  return N == 1 ? 0 : v;
}

template <typename real>
__device__ __forceinline__ void decode(const Inst<real> &d, int i, int &x, int &y, int &z)
{
  uint32_t c = ldg(d.coords, i);
  x = (int)(c & d.mx); y = (int)((c >> d.sx) & d.my); z = (int)(c >> d.sxy);
}

template <typename real>
__device__ __forceinline__ void make_nbhd(const Inst<real> &d, int x, int y, int z, Nbhd &nb)
{
  nb.ok = 0;
  // row-word offsets stay below 2^32 (checked when the instance is created) and
  // each factor below 2^24, so the full-rate 24-bit multiply is exact
  const unsigned wpr = (unsigned)d.wpr, plane = (unsigned)d.Ny*wpr;
  #pragma unroll
  for (int j = 0; j < 5; ++j) {
    bool a, b, c;
    nb.xw[j] = wrap_axis(x + j - 2, d.Nx, d.px, a);
    nb.yr[j] = (int)__umul24((unsigned)wrap_axis(y + j - 2, d.Ny, d.py, b), wpr);
    nb.zr[j] = (int)__umul24((unsigned)wrap_axis(z + j - 2, d.Nz, d.pz, c), plane);
    nb.ok |= (a ? 1u << j : 0u) | (b ? 1u << (5 + j) : 0u) | (c ? 1u << (10 + j) : 0u);
  }
  nb.xb = (unsigned)(x >= 2 ? x - 2 : 0) >> 5;
  nb.anyw[0] = __ballot(x < 1 || x + 1 >= d.Nx) != 0ull;
  nb.anyw[1] = __ballot(x < 2 || x + 2 >= d.Nx) != 0ull;
}

// the words wb and wb+1 of the row starting at word `row` (a word past the
// table reads as 0 and is never selected)
struct RowWin { uint64_t a, b; };
__device__ __forceinline__ RowWin row_win(rsrc_t rW, unsigned row, unsigned wb)
{
  typedef __attribute__((ext_vector_type(2))) unsigned long long u64x2;
  const u64x2 v = __builtin_bit_cast(u64x2, __builtin_amdgcn_raw_buffer_load_b128(rW, (int)((row + wb)*8u), 0, 0));
  return { v.x, v.y };
}

// The voxel at offset (dx,dy,dz), |d| <= 2.  Returns the byte offset (times
// RB) of its site, or of the centre site itself when it is not a site, and sets
// fluid / solid-inside-the-box flags.  Branch-free.
template <typename real, bool OPEN>
__device__ __forceinline__ unsigned nb_lookup(const Inst<real> & /*d*/, rsrc_t rW, const Nbhd &nb,
                                              int dx, int dy, int dz, unsigned self_off,
                                              bool &fluid, bool &solid)
{
  const int xx = nb.xw[dx + 2];
  const unsigned widx = (unsigned)nb.zr[dz + 2] + (unsigned)nb.yr[dy + 2] + ((unsigned)xx >> 5);
  const uint64_t w = bld<uint64_t>(rW, widx*8u, 0u);
  const unsigned lo = (unsigned)w, b = (unsigned)xx & 31u;
  const unsigned idx = (unsigned)(w >> 32) + (unsigned)__popc(lo & ((1u << b) - 1u));
  bool f = (lo >> b) & 1u;
  bool inbox = true;
  if constexpr (OPEN) {
    const unsigned bits = (1u << (dx + 2)) | (1u << (7 + dy)) | (1u << (12 + dz));
    inbox = (nb.ok & bits) == bits;
  }
  fluid = f && inbox;
  solid = !f && inbox;
  return fluid ? idx*(unsigned)sizeof(real) : self_off;
}

// The words of the 9 neighbour rows of ring S: a two-word window for the rows
// entered off x, plus the word of the one position that can wrap round x;
// the word of x itself for the diagonal rows.
struct RingWords { RowWin L[3][3]; uint64_t F[3][3]; };
template <typename real, int S>
__device__ __forceinline__ void ring_words_global(const Inst<real> &d, rsrc_t rW, const Nbhd &nb, RingWords &rw)
{
  (void)d;
  const int x2 = nb.xw[2];
  const unsigned wf = (unsigned)(x2 < S ? nb.xw[2 - S] : nb.xw[2 + S]) >> 5, w0 = (unsigned)x2 >> 5;
  #pragma unroll
  for (int a = 0; a < 3; ++a)
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      const unsigned row = (unsigned)nb.zr[2 + S*(c - 1)] + (unsigned)nb.yr[2 + S*(a - 1)];
      if (a == 1 || c == 1) {
        rw.L[a][c] = row_win(rW, row, nb.xb);
        rw.F[a][c] = nb.anyw[S - 1] ? bld<uint64_t>(rW, (row + wf)*8u, 0u) : 0ull;
      } else {
        rw.L[a][c] = RowWin{0ull, 0ull};
        rw.F[a][c] = bld<uint64_t>(rW, (row + w0)*8u, 0u);
      }
    }
}

// ---- workgroup row cache --------------------------------------------------------
// A workgroup's sites are consecutive in the x-fastest site list.  When they
// lie in one z plane, in rows y0..y1, every ring-S neighbour row of each of
// them is one of (y0-S .. y1+S) x (z0-S, z0, z0+S): the workgroup reads those
// rows whole into LDS, a few coalesced loads per wave, and each lane takes its
// neighbour words from there instead of reading a window per row from the
// table (9 to 14 vector memory instructions per ring and wave: the passes are
// bound by their vector memory instructions).  A workgroup that spans two
// planes, or whose rows do not fit, reads the table.  The moments and c
// passes use it (0.414 -> 0.396 and 0.148 -> 0.145 ms at 300^3); in the mu
// pass it gained nothing, and in the collision, at 3 workgroups per CU, the
// fill and barrier in front of every wave cost more than the windows
// (0.662 -> 0.666 ms with ring 1 cached, 0.708 with both rings).
constexpr int RC_WORDS = 512;                  // per ring: 4 KiB of LDS
struct RowCache { int y0, nr, z0, on1, on2; };

// the plan, from the first and last site of the workgroup (uniform loads)
template <typename real>
__device__ __forceinline__ RowCache rc_plan(const Inst<real> &d, int first)
{
  const int last = min(first + (int)blockDim.x - 1, d.n - 1);
  int x0, y0, z0, x1, y1, z1;
  decode(d, first, x0, y0, z0);
  decode(d, last, x1, y1, z1);
  RowCache rc;
  rc.y0 = y0; rc.z0 = z0; rc.nr = y1 - y0 + 1;
  rc.on1 = z0 == z1 && 3*(rc.nr + 2)*d.wpr + 1 <= RC_WORDS;
  rc.on2 = z0 == z1 && 3*(rc.nr + 4)*d.wpr + 1 <= RC_WORDS;
  return rc;
}

// rows (y0-S+j, z0+S(c-1)), j < nr+2S, c < 3, each wpr words, in that order;
// the caller synchronises the workgroup before reading them
template <typename real, int S>
__device__ __forceinline__ void rc_fill(const Inst<real> &d, const RowCache &rc, uint64_t *lds)
{
  const int wpr = d.wpr, nrs = rc.nr + 2*S, total = 3*nrs*wpr;
  const float inv = 1.0f/(float)wpr;
  for (int e = (int)threadIdx.x; e < total; e += (int)blockDim.x) {
    int q = (int)((float)e*inv), r = e - q*wpr;           // e = q*wpr + r, q = row slot
    if (r < 0) { --q; r += wpr; } else if (r >= wpr) { ++q; r -= wpr; }
    const int c = q >= 2*nrs ? 2 : (q >= nrs ? 1 : 0), j = q - c*nrs;
    bool ok;
    const int y = wrap_axis(rc.y0 - S + j, d.Ny, 1, ok), z = wrap_axis(rc.z0 + S*(c - 1), d.Nz, 1, ok);
    lds[e] = ldg(d.words, ((long)z*d.Ny + y)*wpr + r);
  }
}

// The neighbour words of ring S from the workgroup's row cache (rc_fill):
// the same words ring_words_global reads from the table.  ly = the lane's row
// y minus the cache's first own row.
template <int S>
__device__ __forceinline__ void ring_words_lds(const uint64_t *lds, const RowCache &rc, int wpr, int ly,
                                               const Nbhd &nb, RingWords &rw)
{
  const int x2 = nb.xw[2];
  const unsigned wf = (unsigned)(x2 < S ? nb.xw[2 - S] : nb.xw[2 + S]) >> 5, w0 = (unsigned)x2 >> 5;
  const int nrs = rc.nr + 2*S;
  #pragma unroll
  for (int a = 0; a < 3; ++a)
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      const unsigned base = (unsigned)((c*nrs + ly + S*a)*wpr);
      if (a == 1 || c == 1) {
        rw.L[a][c] = RowWin{lds[base + nb.xb], lds[base + nb.xb + 1u]};
        rw.F[a][c] = nb.anyw[S - 1] ? lds[base + wf] : 0ull;
      } else {
        rw.L[a][c] = RowWin{0ull, 0ull};
        rw.F[a][c] = lds[base + w0];
      }
    }
}

// byte offsets, fluid and solid bits of ring S from its words
template <typename real, bool OPEN, int S, bool SELF>
__device__ __forceinline__ void ring_offsets(const Nbhd &nb, const RingWords &rw, unsigned self_off,
                                             unsigned (&vo)[Q], unsigned &fl, unsigned &so)
{
  const auto &L = rw.L;
  const auto &F = rw.F;
  vo[0] = self_off; fl = S == 1 ? 1u : 0u; so = 0u;
  #pragma unroll
  for (int m = 1; m < Q; ++m) {
    const int dx = S*CX[m], dy = S*CY[m], dz = S*CZ[m];
    const RowWin &Lw = L[CY[m] + 1][CZ[m] + 1];
    const uint64_t Fw = F[CY[m] + 1][CZ[m] + 1];
    const int xx = nb.xw[dx + 2];
    const unsigned wi = (unsigned)xx >> 5;
    const uint64_t w = (CY[m] != 0 && CZ[m] != 0) ? Fw : (wi == nb.xb ? Lw.a : (wi == nb.xb + 1u ? Lw.b : Fw));
    const unsigned lo = (unsigned)w, b = (unsigned)xx & 31u;
    const unsigned idx = (unsigned)(w >> 32) + (unsigned)__popc(lo & ((1u << b) - 1u));
    // the flag words are built with integer bit operations (a bool per
    // direction became a lane mask held in scalar registers, 36 of them
    // across the two rings of the collision pass)
    const unsigned fb = (lo >> b) & 1u;
    unsigned ib = 1u;
    if constexpr (OPEN) {
      const unsigned bits = (1u << (dx + 2)) | (1u << (7 + dy)) | (1u << (12 + dz));
      ib = (nb.ok & bits) == bits ? 1u : 0u;
    }
    const unsigned fluid = fb & ib;
    if constexpr (SELF) vo[m] = fluid ? idx*(unsigned)sizeof(real) : self_off;
    else vo[m] = idx*(unsigned)sizeof(real);
    fl |= fluid << m; so |= ((fb ^ 1u) & ib) << m;
  }
}

// One ring (S = 1: x + e_m, S = 2: x + 2 e_m).  The rows entered off x
// (dy == 0 or dz == 0) are read as a window of the two words that hold
// x-2..x+2; only one position of the ring can wrap round the x axis (x - S
// when x < S, x + S when x + S >= Nx), so those rows add the one word of that
// position, read in the same round trip in waves that have a lane within S
// of an x face.  The diagonal rows (dy, dz both nonzero) hold only x itself:
// one word.  (Two-word windows for the wrapped part and for the diagonal
// rows held 18 more registers.)  Reading the wrapped words only after the
// local ones came back (a second round trip in half the waves) cost more
// than it saved.
// SELF = false: a neighbour that is not a site gets the offset of the next
// site along its row instead of the own offset, so a wave's load for one
// direction stays within the lines of that row (the caller substitutes the
// own value); the own offset would add the lines of the own row to every
// such load.
template <typename real, bool OPEN, int S, bool SELF = true>
__device__ __forceinline__ void ring_f(const Inst<real> &d, rsrc_t rW, const Nbhd &nb, unsigned self_off,
                                       unsigned (&vo)[Q], unsigned &fl, unsigned &so)
{
  RingWords rw;
  ring_words_global<real, S>(d, rW, nb, rw);
  ring_offsets<real, OPEN, S, SELF>(nb, rw, self_off, vo, fl, so);
}

// r = 1 ring: byte offsets vo[m] of x + e_m (own offset when not a site),
// fluid bits fl and solid-in-box bits so (bit m)
template <typename real, bool OPEN>
__device__ __forceinline__ void ring1(const Inst<real> &d, rsrc_t rW, const Nbhd &nb, unsigned self_off,
                                      unsigned (&vo)[Q], unsigned &fl, unsigned &so)
{
  ring_f<real, OPEN, 1>(d, rW, nb, self_off, vo, fl, so);
}

// ---- pull streaming ---------------------------------------------------------------
// The arrays hold post-collision populations at their own site (slot k = the
// population leaving along e_k; the collision writes it as 0 when it would
// leave the box, or hit a solid without halfway bounce-back).  Streaming is a
// gather: the post-stream population k of x is slot k of x - e_k = x + e_opp(k)
// when that is a site, else slot opp(k) of x itself -- the bounced (or zero)
// population.  Opposite slots are adjacent (felbm order pairs 1-2, 3-4, ...),
// so the scalar offset is min(k, opp k)*ns and the lane offset carries the rest.
__device__ __forceinline__ unsigned pull_off(int k, const unsigned (&vo)[Q], unsigned fl, unsigned io, unsigned nsb)
{
  const int o = OPP[k];
  const unsigned bulk = (k > o) ? vo[o] + nsb : vo[o];
  const unsigned bb = (o > k) ? io + nsb : io;
  return ((fl >> o) & 1u) ? bulk : bb;
}
__device__ __forceinline__ unsigned pull_soff(int k, unsigned nsb)
{
  const int o = OPP[k];
  return (unsigned)(k < o ? k : o)*nsb;
}

// site codes for table export: >= 0 site, -1 solid voxel, -2 outside the box
template <typename real>
__device__ __forceinline__ int nb_site(const Inst<real> &d, const Nbhd &nb, int dx, int dy, int dz)
{
  rsrc_t rW = make_rsrc(d.words, d.nwords*8u);
  bool f, s;
  unsigned off = nb_lookup<real, true>(d, rW, nb, dx, dy, dz, 0xffffffffu, f, s);
  return f ? (int)(off / (unsigned)sizeof(real)) : (s ? -1 : -2);
}
template <typename real>
__device__ __forceinline__ void neighbours1(const Inst<real> &d, const Nbhd &nb, int i, int (&n1)[Q])
{
  n1[0] = i;
  #pragma unroll
  for (int m = 1; m < Q; ++m) n1[m] = nb_site(d, nb, CX[m], CY[m], CZ[m]);
}
template <typename real>
__device__ __forceinline__ void neighbours2(const Inst<real> &d, const Nbhd &nb, int (&n2)[Q])
{
  n2[0] = -1;
  #pragma unroll
  for (int m = 1; m < Q; ++m) n2[m] = nb_site(d, nb, 2*CX[m], 2*CY[m], 2*CZ[m]);
}

template <typename real> __device__ __forceinline__ real W(int k) { return (real)WD[k]; }
// FP32 weights: 1/3, 1/18, 1/36 rounded to nearest sum to 1 + 2^-26, and the
// equilibria built on them then add that fraction of c (and of p) at every site
// every step -- a steady mass gain that the mass correction moves onto the
// interface.  These sum to exactly 1: 1/18 one ulp below its nearest float,
// 1/36 its half, and the rest weight the exact remainder 1 - 12 w1.
constexpr float WF0 = 0x1.555558p-2f, WF1 = 0x1.c71c70p-5f, WF2 = 0x1.c71c70p-6f;
template <> __device__ __forceinline__ float W<float>(int k) { return k == 0 ? WF0 : (k < 7 ? WF1 : WF2); }
__host__ __device__ constexpr float wf32(int k) { return k == 0 ? WF0 : (k < 7 ? WF1 : WF2); }

// Reference pressure of the stored state.  The FP32 state carries the
// pressure as a deviation from the rest pressure 1: the p fields hold p - 1 and
// g slot k holds g_k - w_k (FP32 weights; the zero population is stored as
// -w_k).  An absolute FP32 pressure near 1 rounds at 6e-8, and that rounding,
// fed through the pressure gradients of the h forcing, was the largest FP32
// error in c; the deviations round 10^2-10^3 times finer.  Every equilibrium
// and forcing term is affine in (g, p) with the offset w_k p_ref cancelling, so
// the passes compute the deviations directly.  The FP64 state is absolute.
template <typename real> __host__ __device__ constexpr double p_ref() { return sizeof(real) == 4 ? 1.0 : 0.0; }
// FP32 g deviations below 2^-31 are stored as 0 (below the resolution of an
// absolute FP32 g), so the exchanged double state (w_k + s, exact in double)
// converts back to the same FP32 bits
constexpr float G_FLUSH = 0x1p-31f;
__host__ __device__ inline float g_flush(float s) { return (s < G_FLUSH && s > -G_FLUSH) ? 0.0f : s; }

// ---- mass-correction scalar from the partial sums of this step ------------------
// lambda = (M - M0)/W applied when W > 0 and M != M0 (multiphase_gpu.cuh:689-700)
template <typename real>
__device__ __forceinline__ double mass_lambda(const Inst<real> &d, unsigned cur, bool &apply)
{
  const unsigned long long *a = d.acc + ACC_MASS + cur*5*NSLOT;
  int lane = threadIdx.x & 63;
  unsigned long long mh = wave_sum(ldg(a, 0*NSLOT + lane)), ml = wave_sum(ldg(a, 1*NSLOT + lane));
  unsigned long long wh = wave_sum(ldg(a, 2*NSLOT + lane)), wl = wave_sum(ldg(a, 3*NSLOT + lane));
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
  if constexpr (sizeof(real) == 4) {
    // FP32 engine: the same correction in single precision
    const float phi = c*(1.0f - c);
    return (apply && phi > 0.0f) ? c - (float)lambda*phi : c;
  } else {
    double cm = (double)c;
    double phi = cm*(1.0 - cm);
    return (apply && phi > 0.0) ? (real)(cm + (-lambda*phi)) : c;
  }
}

// lambda of the pending correction, once per instance and step (one wave per
// descriptor in [first, first + count)), written into the descriptor where the
// c stencil and collision passes read it with their other scalar parameters
template <typename real>
__global__ void __launch_bounds__(64) k_mass_lambda(Inst<real> *D, int first, int count, unsigned T)
{
  const int q = first + (int)blockIdx.x;
  if ((int)blockIdx.x >= count) return;
  const Inst<real> &d = D[q];
  if (!d.mass_on) return;
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  bool apply = false;
  double lambda = 0.0;
  if (ldg(d.acc, ACC_FLAGS)) lambda = mass_lambda(d, cur, apply);
  if (threadIdx.x == 0) {
    stg(&D[q].lam, 0, lambda);
    stg(&D[q].lam_on, 0, apply ? 1 : 0);
  }
}

template <typename real>
__device__ __forceinline__ void acc_vec(int m, real t, real &sx, real &sy, real &sz)
{
  if (CX[m] > 0) sx += t; else if (CX[m] < 0) sx -= t;
  if (CY[m] > 0) sy += t; else if (CY[m] < 0) sy -= t;
  if (CZ[m] > 0) sz += t; else if (CZ[m] < 0) sz -= t;
}

// ==== K1: moments (k_moments, device_engine.cuh:116-135, pass 1 only) ===========
template <typename real, bool OPEN>
__device__ __forceinline__ void moments_body(const Inst<real> *__restrict__ D, const int2 bm, unsigned T)
{
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  const int i = bm.y + (int)threadIdx.x;
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  constexpr unsigned RB = sizeof(real);
  const unsigned nsb = (unsigned)d.ns*RB;
  const rsrc_t rH = make_rsrc(d.h, 19u*nsb), rG = make_rsrc(cur ? d.g1 : d.g0, 19u*nsb);
  const rsrc_t rF = make_rsrc(d.craw, (unsigned)F_COUNT*nsb);
  if (bm.y == 0) {                                  // clear the other parity's partials
    unsigned long long *o = d.acc + ACC_MASS + (cur ^ 1u)*5*NSLOT;
    for (int q = threadIdx.x; q < 5*NSLOT; q += blockDim.x) stg(o, q, 0ull);
  }
  if (bm.y >= d.n) return;                          // a padding workgroup: no sites
  __shared__ uint64_t rcw[RC_WORDS];
  const RowCache rc = rc_plan(d, bm.y);
  if (rc.on1) { rc_fill<real, 1>(d, rc, rcw); __syncthreads(); }
  Fx m = {0, 0, 0}, w = {0, 0, 0};
  if (i < d.n) {
    const unsigned io = (unsigned)i*RB;
    const rsrc_t rW = make_rsrc(d.words, d.nwords*8u);
    int x, y, z; decode(d, i, x, y, z);
    Nbhd nb; make_nbhd(d, x, y, z, nb);
    unsigned vo[Q], fl, so;
    {
      RingWords rw;
      if (rc.on1) ring_words_lds<1>(rcw, rc, d.wpr, y - rc.y0, nb, rw);
      else ring_words_global<real, 1>(d, rW, nb, rw);
      ring_offsets<real, OPEN, 1, true>(nb, rw, io, vo, fl, so);
    }
    (void)so;
    // c accumulates in FP64 (the FP64 state sums alike).  The FP32 field c
    // keeps Sigma h to half an ulp (6e-8 near c = 1), below which the slow
    // changes of a bulk phase fall; the exact remainder goes to the collision
    // pass, which adds it to the populations it writes, so Sigma h carries on
    // from step to step without the rounding.
    double cs = 0;
    real p = 0, ux = 0, uy = 0, uz = 0;
    #pragma unroll
    for (int k = 0; k < Q; ++k) {           // streaming (gather) and moments
      const unsigned off = pull_off(k, vo, fl, io, nsb), soff = pull_soff(k, nsb);
      real hv = bld<real>(rH, off, soff), gv = bld<real>(rG, off, soff);
      cs += (double)hv; p += gv;
      acc_vec<real>(k, gv, ux, uy, uz);
    }
    const real c = (real)cs;
    if constexpr (sizeof(real) == 4)        // read once, by this site's own collision
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)(cs - (double)c)), rF, (int)io,
                                            (int)(F_CLO*nsb), 2);
    // [p0, j] is read once, by this site's own c pass: written non-temporal
    bst(c, rF, io, F_CRAW*nsb); bst_pj<2>(p, ux, uy, uz, rF, io, F_PJ*nsb);
    if (d.mass_on) {              // k_mass_weight, device_engine.cuh:1109-1126
      double cm = cs;
      double phi = cm*(1.0 - cm); if (phi < 0.0) phi = 0.0;
      fx_add(m, cm); fx_add(w, phi);
    }
  }
  if (d.mass_on) {
    unsigned long long v[5] = { (unsigned long long)m.hi, m.lo, (unsigned long long)w.hi, w.lo, m.bad + w.bad };
    block_accumulate<5>(v, d.acc + ACC_MASS + cur*5*NSLOT, NSLOT, (int)(blockIdx.x & (NSLOT - 1)));
  }
}

template <typename real, bool OPEN>
__global__ void __launch_bounds__(256) k_moments(const Inst<real> *__restrict__ D,
                                                 const int2 *__restrict__ map, int sdi, unsigned T)
{
  const int2 bm = map_entry(map, sdi);
  if constexpr (OPEN) {
    if (bm.x >= 0 && (bm.x & MAP_INTERIOR)) { moments_body<real, false>(D, make_int2(bm.x & ~MAP_INTERIOR, bm.y), T); return; }
  }
  moments_body<real, OPEN>(D, bm, T);
}

// ==== open boundaries: imposed fields (k_open_bnd_fields + k_open_bnd_buf_fields,
// device_engine.cuh:738-759, :807-815) ================================================

// injection schedule (d_inlet_gate, device_engine.cuh:699-736), evaluated in FP64
template <typename real>
__device__ inline double inlet_gate(const Inst<real> &d, unsigned step, double coord)
{
  const double w = d.inlet_ramp;
  if (d.inlet_mode == 1) {
    const double Per = d.inlet_period;
    if (Per <= 0.0) return 1.0;
    // the reference floors the period count in single precision
    const double tau = (double)step - Per*(double)floorf((float)((double)step/Per));
    const double lo = 0.5*Per - 0.5*d.inlet_duty*Per, hi = 0.5*Per + 0.5*d.inlet_duty*Per;
    if (w <= 0.0) return (tau >= lo && tau < hi) ? 1.0 : 0.0;
    return 0.5*(tanh(2.0*(tau - lo)/w) - tanh(2.0*(tau - hi)/w));
  }
  if (d.inlet_mode == 2) {
    const double x = coord - d.split_pos;
    if (w <= 0.0) return (x < 0.0) ? 1.0 : 0.0;
    return 0.5*(1.0 - tanh(2.0*x/w));
  }
  if (d.inlet_mode == 3) {
    const double Per = d.inlet_period;
    if (Per <= 0.0) return 1.0;
    double t = coord - d.split_pos;
    t = t - Per*floor(t/Per);
    const double lo = 0.5*Per - 0.5*d.inlet_duty*Per, hi = 0.5*Per + 0.5*d.inlet_duty*Per;
    if (w <= 0.0) return (t >= lo && t < hi) ? 1.0 : 0.0;
    return 0.5*(tanh(2.0*(t - lo)/w) - tanh(2.0*(t - hi)/w));
  }
  return 1.0;
}

// One thread per open-boundary node, after the moments pass.  The fields the
// reference imposes (c, rho, p, u) are computed from the mirrored vertex's
// moments and stored in ov; k_open_commit then writes c into the c_raw field,
// so every stencil of the step reads the imposed value, while the node's own
// raw c (kept in ov) sets its bulk chemical potential and relaxation rate, which
// the reference leaves as computed from the distributions.
template <typename real>
__global__ void __launch_bounds__(256) k_open_fields(const Inst<real> *__restrict__ D,
                                                     const ObEntry *__restrict__ E, int ne, unsigned T)
{
  const int j = blockIdx.x*blockDim.x + (int)threadIdx.x;
  if (j >= ne) return;
  const ObEntry e = E[j];
  const Inst<real> &d = D[e.desc];
  (void)T;
  const long ns = d.ns;
  // the moments pass's sums at the mirrored vertex (k_moments, device_engine.cuh:116-135)
  real c = ldg(d.craw, (long)e.ref), p = ldg(d.craw, F_PJ*ns + 4*(long)e.ref);
  real ux = ldg(d.craw, F_PJ*ns + 4*(long)e.ref + 1), uy = ldg(d.craw, F_PJ*ns + 4*(long)e.ref + 2);
  real uz = ldg(d.craw, F_PJ*ns + 4*(long)e.ref + 3);
  const real cr = c;
  real rho = d.rho1 + d.drho*c;
  const real inv = (real)1/((real)(1.0/3.0)*rho);
  ux *= inv; uy *= inv; uz *= inv;
  const unsigned step = (unsigned)ldg(d.acc, ACC_FLAGS + 1);
  if (e.rkind == 0) {
    if (d.use_in_vel) { ux = (real)d.u_in[0]; uy = (real)d.u_in[1]; uz = (real)d.u_in[2]; }
    if (d.use_in_prs) p = (real)d.p_in;
    if (d.use_in_fluid) {
      const double gate = inlet_gate(d, step, e.rcoord);
      const double cv = (1.0 - d.inlet_c_a) + (2.0*d.inlet_c_a - 1.0)*gate;
      c = (real)cv;
      rho = (real)(cv*(double)d.rho0 + (1.0 - cv)*(double)d.rho1);
    }
  } else if (e.rkind == 1) {
    if (d.use_out_prs) p = (real)d.p_out;
    if (d.use_out_fluid) { c = (real)d.c_out_fixed; rho = (real)d.rho_out_fixed; }
  }
  const real craw_own = (e.site != e.ref) ? ldg(d.craw, (long)e.site) : cr;
  real *o = d.ov + 8*(long)e.eid;
  stg(o, 0, c); stg(o, 1, rho); stg(o, 2, p); stg(o, 3, ux); stg(o, 4, uy); stg(o, 5, uz); stg(o, 6, craw_own);
}

// the imposed c into c_raw, once every node has read its vertex's raw moments
template <typename real>
__global__ void __launch_bounds__(256) k_open_commit(const Inst<real> *__restrict__ D,
                                                     const ObEntry *__restrict__ E, int ne)
{
  const int j = blockIdx.x*blockDim.x + (int)threadIdx.x;
  if (j >= ne) return;
  const ObEntry e = E[j];
  const Inst<real> &d = D[e.desc];
  stg(d.craw, (long)e.site, ldg(d.ov, 8*(long)e.eid));
}

// equilibrium the open boundary writes (k_open_bnd / k_open_bnd_buf,
// device_engine.cuh:772-779, :838-845)
template <typename real>
__device__ __forceinline__ void plain_eq(int k, real ux, real uy, real uz, real pp, real rho, real c,
                                         real &h, real &g)
{
  const real w = W<real>(k);
  const real u_sq = ux*ux + uy*uy + uz*uz;
  const real uk = ux*(real)CX[k] + uy*(real)CY[k] + uz*(real)CZ[k];
  const real u_term = uk*((real)3.0 + (real)4.5*uk) - (real)1.5*u_sq;
  h = c*w*((real)1 + u_term);
  g = w*(pp + rho*(real)(1.0/3.0)*u_term);
}

// ---- per-direction stencil values ---------------------------------------------------
// Neighbour values are fetched at the own site when the neighbour is not a site,
// which makes the halfway-bounce-back central difference f(x+e)-f(x-e) already
// equal to the reference's backward / forward / none cases (multiphase_gpu.cuh:
// 113-121): f0-f(x-e), f(x+e)-f0 and f0-f0 = 0.
template <typename real, bool HBB>
__device__ __forceinline__ real cd_num(bool hin, bool hip, real fin, real fip, bool &use)
{
  if constexpr (HBB) { use = true; return fin - fip; }
  use = hin && hip;
  return fin - fip;
}
// biased value (bd_dir_val, device_engine.cuh:96-107; cases multiphase_gpu.cuh:140-153)
template <typename real, bool HBB>
__device__ __forceinline__ real bd_num(bool hin, bool hin2, bool hip, bool hip2,
                                       real fin, real fin2, real fip, real fip2, real f0)
{
  if constexpr (HBB) {
    real v3 = (real)(-1.5)*f0 + (real)2.0*fin + (real)(-0.5)*fin2;
    real vf = (real)(-1.5)*f0 + (real)1.5*fin;
    real vb = (real)0.5*f0 + (real)(-0.5)*fip;
    return hin ? (hin2 ? v3 : vf) : (hip ? vb : (real)0);
  } else {
    real f1 = hin ? fin : fip, f2 = hin ? fin2 : fip2;
    real v3 = (real)(-1.5)*f0 + (real)2.0*f1 + (real)(-0.5)*f2;
    real vf = (real)(-2.0)*f0 + (real)2.0*fin;
    return hin ? (hin2 ? v3 : vf) : ((hip && hip2) ? v3 : (real)0);
  }
}
// Laplacian contribution (k_lap_mf, device_engine.cuh:996-1010; cases :178-188)
template <typename real, bool HBB>
__device__ __forceinline__ real lap_term(bool hin, bool hip, real fin, real fip, real f0, real bnd, real w)
{
  real tc = w*(fip - (real)2*f0 + fin);
  real f1 = hip ? fip : fin;
  real t1;
  if constexpr (HBB) t1 = w*(f1 - f0 - bnd);
  else t1 = w*((real)2*f1 - (real)2*f0 - (real)2*bnd);
  return (hin && hip) ? tc : ((hin || hip) ? t1 : (real)0);
}

// ==== K2: c stencil -> corrected c, mu, p* =========================================
// k_pack_lap_c + k_lap_mf + k_mu_axpy + k_grad_cd_mf(c) + k_vel_press_corr
// (device_engine.cuh:149-196).  exp != nullptr also writes rho, ux, uy, uz for
// a field download (multiphase_gpu.cuh:704-757).
template <typename real, bool HBB, bool OPEN>
__device__ __forceinline__ void cfield_body(const Inst<real> *__restrict__ D, const int2 bm, unsigned T, real *exp, int exp_inst)
{
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  (void)cur;
  bool apply = false; double lambda = 0.0;
  if (d.mass_on) {                 // set by k_mass_lambda for this step
    apply = d.lam_on != 0;
    lambda = d.lam;
  }
  if (bm.y >= d.n) return;                          // a padding workgroup: no sites
  __shared__ uint64_t rcw[RC_WORDS];
  const RowCache rc = rc_plan(d, bm.y);
  if (rc.on1) { rc_fill<real, 1>(d, rc, rcw); __syncthreads(); }
  const int i = bm.y + (int)threadIdx.x;
  if (i >= d.n) return;
  constexpr unsigned RB = sizeof(real);
  const unsigned nsb = (unsigned)d.ns*RB, io = (unsigned)i*RB;
  const rsrc_t rF = make_rsrc(d.craw, (unsigned)F_COUNT*nsb), rW = make_rsrc(d.words, d.nwords*8u);
  int x, y, z; decode(d, i, x, y, z);
  Nbhd nb; make_nbhd(d, x, y, z, nb);
  unsigned vo[Q], fl, so;
  {
    RingWords rw;
    if (rc.on1) ring_words_lds<1>(rcw, rc, d.wpr, y - rc.y0, nb, rw);
    else ring_words_global<real, 1>(d, rW, nb, rw);
    ring_offsets<real, OPEN, 1, false>(nb, rw, io, vo, fl, so);
  }
  const real craw = bld<real>(rF, io, F_CRAW*nsb);
  const real c = corrected<real>(craw, apply, lambda);
  real cn[Q];
  #pragma unroll
  for (int m = 1; m < Q; ++m) {
    const real v = bld<real>(rF, vo[m], F_CRAW*nsb);
    cn[m] = corrected<real>(((fl >> m) & 1u) ? v : craw, apply, lambda);
  }
  real gx = 0, gy = 0, gz = 0, lapc = 0;
  const real bnd = d.bnd_coeff*c*((real)1 - c);
  if constexpr (sizeof(real) == 4) {
    // FP32: a direction pair (m, opp m) contributes its central difference and
    // its Laplacian term twice, so each pair is evaluated once
    #pragma unroll
    for (int m = 1; m < Q; m += 2) {
      const int o = m + 1;
      const bool hin = (fl >> m) & 1u, hip = (fl >> o) & 1u;
      bool use; const float dd = cd_num<float, HBB>(hin, hip, cn[m], cn[o], use);
      acc_vec<float>(m, use ? (3.0f*W<float>(m))*dd : 0.0f, gx, gy, gz);
      lapc += lap_term<float, HBB>(hin, hip, cn[m], cn[o], c, bnd, 6.0f*W<float>(m));
    }
  } else {
    #pragma unroll
    for (int m = 1; m < Q; ++m) {
      const int o = OPP[m];
      const bool hin = (fl >> m) & 1u, hip = (fl >> o) & 1u;
      bool use; real dd = cd_num<real, HBB>(hin, hip, cn[m], cn[o], use);
      real t = ((real)3.0*(real)0.5*W<real>(m))*dd;
      acc_vec<real>(m, use ? t : (real)0, gx, gy, gz);
      lapc += lap_term<real, HBB>(hin, hip, cn[m], cn[o], c, bnd, (real)3.0*W<real>(m));
    }
  }
  // open-boundary node: imposed rho (and u, p when the fields are imposed);
  // mu_bulk from the node's own raw c, as the reference's moments pass has it
  int eb = -1;
  const real *ovp = nullptr;
  if constexpr (OPEN) {
    if (d.obc) { eb = ldg(d.ent, i); if (eb >= 0) ovp = d.ov + 8*(long)(eb >> 4); }
  }
  const real cm = ovp ? ldg(ovp, 6) : c;
  real c_log = cm*((real)1 - cm);
  real rho = ovp ? ldg(ovp, 1) : d.rho1 + d.drho*c;
  real irho, inv;                  // 1/rho and 1/(rho cs^2)
  if constexpr (sizeof(real) == 4) { irho = __builtin_amdgcn_rcpf(rho); inv = 3.0f*irho; }
  else { irho = (real)0; inv = (real)1/((real)(1.0/3.0)*rho); }
  // p0, jx, jy, jz: read once, non-temporal (keeps L2 for the c_raw stencil)
  const Rec4<real> pj = bld_pj<2>((const real *)nullptr, rF, io, F_PJ*nsb);
  real ux = pj.p*inv, uy = pj.l*inv, uz = pj.m*inv;
  real mu = d.four_beta*c_log*((real)0.5 - cm);
  mu += -d.kappa*lapc;
  real p = pj.c;
  if (eb >= 0 && (eb & ENT_BCPROC)) {
    // k_vel_press_corr skips nodes whose fields the boundary imposed
    ux = ldg(ovp, 3); uy = ldg(ovp, 4); uz = ldg(ovp, 5); p = ldg(ovp, 2);
  } else {
    if constexpr (sizeof(real) == 4) {
      const float mor = mu*irho, ffr = d.ff*irho;
      ux = ux + 0.5f*(d.gx + d.fx*ffr + mor*gx);
      uy = uy + 0.5f*(d.gy + d.fy*ffr + mor*gy);
      uz = uz + 0.5f*(d.gz + d.fz*ffr + mor*gz);
    } else {
      real mor = mu/rho;
      ux = ux + (real)0.5*(d.gx + d.fx*d.ff/rho + mor*gx);
      uy = uy + (real)0.5*(d.gy + d.fy*d.ff/rho + mor*gy);
      uz = uz + (real)0.5*(d.gz + d.fz*d.ff/rho + mor*gz);
    }
    real u_dot_drho = d.drho*(ux*gx + uy*gy + uz*gz);
    p += (real)0.5*u_dot_drho*(real)(1.0/3.0);
  }
  bst_cp(c, p, rF, io, F_CP*nsb); bst(mu, rF, io, F_MU*nsb);
  if (exp && bm.x == exp_inst) {
    const long ns = d.ns;
    stg(exp, i, rho); stg(exp, ns + i, ux); stg(exp, 2*ns + i, uy); stg(exp, 3*ns + i, uz);
  }
  if (!exp && d.trk) {          // tracers read this step's u (felbm_gpu_main.cu:752-763)
    const unsigned long long ts = ldg(d.acc, ACC_FLAGS + 2);
    if (ts % (unsigned long long)d.tr_vskip == 0) {
      // one record per site, [ux, uy, uz, c], in one vector store (c is read
      // only for the phase split of the stretching rows).  Non-temporal: the
      // tracer kernel reads the records of a few cells per tracer, and the
      // full snapshot (16 B per site, FP32) would otherwise displace the mu
      // stencil data of the next pass from L2.
      typedef real v4 __attribute__((ext_vector_type(4)));
      __builtin_nontemporal_store(v4{ux, uy, uz, c}, (__attribute__((address_space(1))) v4 *)d.tu + i);
    }
  }
}

template <typename real, bool HBB, bool OPEN>
__global__ void __launch_bounds__(256) k_cfield(const Inst<real> *__restrict__ D,
                                                const int2 *__restrict__ map, int sdi, unsigned T,
                                                real *exp, int exp_inst)
{
  const int2 bm = map_entry(map, sdi);
  if constexpr (OPEN) {
    if (bm.x >= 0 && (bm.x & MAP_INTERIOR)) { cfield_body<real, HBB, false>(D, make_int2(bm.x & ~MAP_INTERIOR, bm.y), T, exp, exp_inst); return; }
  }
  cfield_body<real, HBB, OPEN>(D, bm, T, exp, exp_inst);
}

// ==== K3: lap(mu) (k_pack_lap_zero + k_lap_mf) ======================================
template <typename real, bool HBB, bool OPEN>
__device__ __forceinline__ void mufield_body(const Inst<real> *__restrict__ D, const int2 bm, unsigned /*T*/)
{
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  if (bm.y >= d.n) return;                          // a padding workgroup: no sites
  const int i = bm.y + (int)threadIdx.x;
  if (i >= d.n) return;
  constexpr unsigned RB = sizeof(real);
  const unsigned nsb = (unsigned)d.ns*RB, io = (unsigned)i*RB;
  const rsrc_t rF = make_rsrc(d.craw, (unsigned)F_COUNT*nsb), rW = make_rsrc(d.words, d.nwords*8u);
  int x, y, z; decode(d, i, x, y, z);
  Nbhd nb; make_nbhd(d, x, y, z, nb);
  unsigned vo[Q], fl, so;
  ring_f<real, OPEN, 1, false>(d, rW, nb, io, vo, fl, so);
  const real m0 = bld<real>(rF, io, F_MU*nsb);
  real mn[Q];
  #pragma unroll
  for (int m = 1; m < Q; ++m) {
    const real v = bld<real>(rF, vo[m], F_MU*nsb);
    mn[m] = ((fl >> m) & 1u) ? v : m0;
  }
  real s = 0;
  if constexpr (sizeof(real) == 4) {
    // FP32: both directions of a pair give the same term
    #pragma unroll
    for (int m = 1; m < Q; m += 2)
      s += lap_term<float, HBB>((fl >> m) & 1u, (fl >> (m + 1)) & 1u, mn[m], mn[m + 1], m0, 0.0f, 6.0f*W<float>(m));
  } else {
    #pragma unroll
    for (int m = 1; m < Q; ++m)
      s += lap_term<real, HBB>((fl >> m) & 1u, (fl >> OPP[m]) & 1u, mn[m], mn[OPP[m]], m0, (real)0, (real)3.0*W<real>(m));
  }
  if constexpr (sizeof(real) == 4) {
    // own [c, p*], read once: non-temporal, like the other single-use streams,
    // so L2 keeps the stencil data (mu here, the records in the collision pass)
    const Rec2<real> cp = bld_cp<2>((const real *)nullptr, rF, io, F_CP*nsb);
    bst_r3(cp.c, cp.p, s, rF, io, F_REC*nsb);
  } else {
    // FP64: lapmu alone; the collision takes c and p* from the [c, p*] pairs
    // and mu from its field (one 16-byte and one 8-byte load per ring-1
    // neighbour instead of two 16-byte record loads: mu pass 0.244 -> 0.181 ms,
    // collision 2.23 -> 2.15 ms at 300^3)
    bst(s, rF, io, F_LMU*nsb);
  }
}

template <typename real, bool HBB, bool OPEN>
__global__ void __launch_bounds__(256) k_mufield(const Inst<real> *__restrict__ D,
                                                 const int2 *__restrict__ map, int sdi, unsigned T)
{
  const int2 bm = map_entry(map, sdi);
  if constexpr (OPEN) {
    if (bm.x >= 0 && (bm.x & MAP_INTERIOR)) { mufield_body<real, HBB, false>(D, make_int2(bm.x & ~MAP_INTERIOR, bm.y), T); return; }
  }
  mufield_body<real, HBB, OPEN>(D, bm, T);
}

// Store one post-collision population k of the site at byte offset io into
// its own slot k (contiguous across the wave); the next gather reads it from
// x + e_k, or -- neighbour not a site -- from slot opp(k) of x as the
// bounced-back population (halfway bounce-back off a solid) or zero (outside
// the box, or no bounce-back; `zero` is the stored form of a zero population).
// vo is unused: the gather side resolves it.
// The FP32 store carries the non-temporal hint: nothing reads the slot before
// the next step, and the line need not displace the stencil data in L2.
template <typename real, bool HBB, bool OPEN>
__device__ __forceinline__ void store_post(real v, rsrc_t r, int k, unsigned vo, bool fluid, bool solid,
                                     unsigned io, unsigned nsb, real zero = (real)0)
{
  (void)vo;
  real val = v;
  if constexpr (OPEN || !HBB) val = (fluid || (HBB && solid)) ? v : zero;
  (void)solid;
  if constexpr (sizeof(real) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, val), r, (int)io, (int)((unsigned)k*nsb), 2);
  else bst(val, r, io, (unsigned)k*nsb);
}

// ==== K4: fused collision, post-collision store (FP64, reference order) ==============
// k_collide_fused_bgk (device_engine.cuh:459-545) and k_collide_fused_mrt
// (:547-651), with the field quantities of the reference's separate kernels
// (k_moments u, k_grad_cd/bd/md, k_vel_press_corr) recomputed in registers.
template <typename real, bool MRT, bool HBB, bool OPEN>
__device__ __forceinline__ void collide_body(const Inst<real> *__restrict__ D, const int2 bm, unsigned T, int stats)
{
  if (bm.x < 0) return;
  const Inst<real> &d = D[bm.x];
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  const int i = bm.y + (int)threadIdx.x;
  const bool live = i < d.n;
  Fx s0 = {0,0,0}, s1 = {0,0,0}, s2 = {0,0,0}, s3 = {0,0,0}, s4 = {0,0,0};
  unsigned long long vmax = 0;
  if (threadIdx.x == 0 && bm.y == 0) {
    if (d.mass_on) stg(d.acc, ACC_FLAGS, 1ull);                   // correction owed from now on
    if (OPEN && d.obc) stg(d.acc, ACC_FLAGS + 1, ldg(d.acc, ACC_FLAGS + 1) + 1ull);   // ++open_step
    if (d.trk) stg(d.acc, ACC_FLAGS + 2, ldg(d.acc, ACC_FLAGS + 2) + 1ull);            // ++tracer step
  }
  if (live) {
    constexpr unsigned RB = sizeof(real);
    const unsigned nsb = (unsigned)d.ns*RB, io = (unsigned)i*RB;
    const rsrc_t rF = make_rsrc(d.craw, (unsigned)F_COUNT*nsb), rW = make_rsrc(d.words, d.nwords*8u);
    const rsrc_t rGi = make_rsrc(cur ? d.g1 : d.g0, 19u*nsb), rGo = make_rsrc(cur ? d.g0 : d.g1, 19u*nsb);
    const rsrc_t rH = make_rsrc(d.h, 19u*nsb);
    int x, y, z; decode(d, i, x, y, z);
    Nbhd nb; make_nbhd(d, x, y, z, nb);
    unsigned vo[Q], fl, so, fl2;
    ring_f<real, OPEN, 1, false>(d, rW, nb, io, vo, fl, so);
    // open-boundary node (ENT_* bits); a buffer node does not collide and
    // stores the equilibrium the reference writes into it (k_open_bnd_buf)
    int eb = -1;
    const real *ovp = nullptr;
    if constexpr (OPEN) {
      if (d.obc) { eb = ldg(d.ent, i); if (eb >= 0) ovp = d.ov + 8*(long)(eb >> 4); }
    }
    const bool nostream = eb >= 0 && (eb & ENT_NOSTREAM);
    if (nostream) {
      const real oc = ldg(ovp, 0), orho = ldg(ovp, 1), op = ldg(ovp, 2);
      const real oux = ldg(ovp, 3), ouy = ldg(ovp, 4), ouz = ldg(ovp, 5);
      #pragma unroll
      for (int k = 0; k < Q; ++k) {
        real hv, gv;
        plain_eq<real>(k, oux, ouy, ouz, op, orho, oc, hv, gv);
        store_post<real, HBB, OPEN>(hv, rH, k, vo[k], (fl >> k) & 1u, (so >> k) & 1u, io, nsb);
        store_post<real, HBB, OPEN>(gv, rGo, k, vo[k], (fl >> k) & 1u, (so >> k) & 1u, io, nsb);
      }
      if (stats) {
        double vx = (double)oux, vy = (double)ouy, vz = (double)ouz;
        double v2 = vx*vx + vy*vy + vz*vz;
        fx_add(s0, vx); fx_add(s1, vy); fx_add(s2, vz); fx_add(s3, v2); fx_add(s4, (double)oc);
        vmax = (unsigned long long)__double_as_longlong(v2);
      }
    } else {
      real cN[Q], pN[Q], c2[Q], p2[Q], lN[Q];
      const Rec2<real> r0 = bld_cp((const real *)nullptr, rF, io, F_CP*nsb);
      const real lmu = bld<real>(rF, io, F_LMU*nsb), mu = bld<real>(rF, io, F_MU*nsb);
      {
        // a neighbour that is not a site takes the own values
        unsigned vo2[Q];
        { unsigned so2; ring_f<real, OPEN, 2, false>(d, rW, nb, io, vo2, fl2, so2); }
        #pragma unroll
        for (int m = 1; m < Q; ++m) {
          const Rec2<real> r1 = bld_cp((const real *)nullptr, rF, vo[m], F_CP*nsb);
          const real l1 = bld<real>(rF, vo[m], F_LMU*nsb);
          const Rec2<real> r2 = bld_cp((const real *)nullptr, rF, vo2[m], F_CP*nsb);
          const bool s1 = (fl >> m) & 1u, s2 = (fl2 >> m) & 1u;
          cN[m] = s1 ? r1.c : r0.c; pN[m] = s1 ? r1.p : r0.p; lN[m] = s1 ? l1 : lmu;
          c2[m] = s2 ? r2.c : r0.c; p2[m] = s2 ? r2.p : r0.p;
        }
      }
      const real c = r0.c, p = r0.p;
      const real drho = d.drho;
      const real CS2 = (real)(1.0/3.0), ALPHA = (real)3.0, GAMMA_C = (real)1.5, BETA_C = (real)4.5;
      const real rho = ovp ? ldg(ovp, 1) : d.rho1 + drho*c;
      const real crl = ovp ? ldg(ovp, 6) : c;     // relax from the raw c (k_moments)
      const real cp = crl < (real)0 ? (real)0 : (crl > (real)1 ? (real)1 : crl);
      const real tau = (real)1/(cp/d.tau0 + ((real)1 - cp)/d.tau1);
      const real relax = (real)1/(tau + (real)0.5);
      real gk[Q];
      real jx = 0, jy = 0, jz = 0;
      #pragma unroll
      for (int k = 0; k < Q; ++k) { gk[k] = bld<real>(rGi, pull_off(k, vo, fl, io, nsb), pull_soff(k, nsb)); acc_vec<real>(k, gk[k], jx, jy, jz); }
      // per-direction central and biased values, and the vector gradients in
      // the reference's direction order
      real cdc[Q], cdp[Q], bdc[Q], bdp[Q];
      real e_dcx = 0, e_dcy = 0, e_dcz = 0, e_dpx = 0, e_dpy = 0, e_dpz = 0;
      real b_cx = 0, b_cy = 0, b_cz = 0, b_px = 0, b_py = 0, b_pz = 0;
      cdc[0] = cdp[0] = bdc[0] = bdp[0] = 0;
      #pragma unroll
      for (int m = 1; m < Q; ++m) {
        const int o = OPP[m];
        const bool hin = (fl >> m) & 1u, hip = (fl >> o) & 1u, h2 = (fl2 >> m) & 1u, hp2 = (fl2 >> o) & 1u;
        bool use;
        real nc = cd_num<real, HBB>(hin, hip, cN[m], cN[o], use);
        real np = cd_num<real, HBB>(hin, hip, pN[m], pN[o], use);
        const real coef = (real)3.0*(real)0.5*W<real>(m), base = (real)3.0*W<real>(m);
        acc_vec<real>(m, use ? coef*nc : (real)0, e_dcx, e_dcy, e_dcz);
        acc_vec<real>(m, use ? coef*np : (real)0, e_dpx, e_dpy, e_dpz);
        cdc[m] = use ? (real)0.5*nc : (real)0;
        cdp[m] = use ? (real)0.5*np : (real)0;
        real vc = bd_num<real, HBB>(hin, h2, hip, hp2, cN[m], c2[m], cN[o], c2[o], c);
        real vp = bd_num<real, HBB>(hin, h2, hip, hp2, pN[m], p2[m], pN[o], p2[o], p);
        acc_vec<real>(m, base*vc, b_cx, b_cy, b_cz);
        acc_vec<real>(m, base*vp, b_px, b_py, b_pz);
        bdc[m] = vc; bdp[m] = vp;
      }
      const real f_dcx = (real)0.5*(e_dcx + b_cx), f_dcy = (real)0.5*(e_dcy + b_cy), f_dcz = (real)0.5*(e_dcz + b_cz);
      const real f_dpx = (real)0.5*(e_dpx + b_px), f_dpy = (real)0.5*(e_dpy + b_py), f_dpz = (real)0.5*(e_dpz + b_pz);
      // u* = u + 1/2 (g + f ff/rho + (mu/rho) grad_cd c)   (k_vel_press_corr)
      const real inv = (real)1/(CS2*rho);
      real ux = jx*inv, uy = jy*inv, uz = jz*inv;
      if (eb >= 0 && (eb & ENT_BCPROC)) {
        ux = ldg(ovp, 3); uy = ldg(ovp, 4); uz = ldg(ovp, 5);
      } else {
        real mor = mu/rho;
        ux = ux + (real)0.5*(d.gx + d.fx*d.ff/rho + mor*e_dcx);
        uy = uy + (real)0.5*(d.gy + d.fy*d.ff/rho + mor*e_dcy);
        uz = uz + (real)0.5*(d.gz + d.fz*d.ff/rho + mor*e_dcz);
      }
      if (eb >= 0 && (eb & ENT_EQ)) {
        // k_open_bnd overwrote g with the equilibrium of the node's fields
        #pragma unroll
        for (int k = 0; k < Q; ++k) { real hv; plain_eq<real>(k, ux, uy, uz, p, rho, c, hv, gk[k]); }
      }
      if (stats) {                  // k_flow_stats, device_engine.cuh:660-683
        double vx = (double)ux, vy = (double)uy, vz = (double)uz;
        double v2 = vx*vx + vy*vy + vz*vz;
        fx_add(s0, vx); fx_add(s1, vy); fx_add(s2, vz); fx_add(s3, v2); fx_add(s4, (double)c);
        vmax = (unsigned long long)__double_as_longlong(v2);
      }
      const real bfx = rho*d.gx + d.fx*d.ff, bfy = rho*d.gy + d.fy*d.ff, bfz = rho*d.gz + d.fz*d.ff;
      const real mob = d.mobility;
      // avg_dir: 1/2 (lmu + lapmu(x+e_k)); a non-site neighbour fetched the own
      // value, and 1/2 (lmu + lmu) == lmu exactly, the reference's self case
      auto avg_of = [&](int k) -> real { return (real)0.5*(lmu + (k ? lN[k] : lmu)); };
      auto dir_dot = [&](int k, real ax, real ay, real az) -> real {
        real s = 0;
        if (CX[k] > 0) s += ax; else if (CX[k] < 0) s -= ax;
        if (CY[k] > 0) s += ay; else if (CY[k] < 0) s -= ay;
        if (CZ[k] > 0) s += az; else if (CZ[k] < 0) s -= az;
        return s;
      };
      auto storek = [&](real v, rsrc_t r, int k) {
        store_post<real, HBB, OPEN>(v, r, k, vo[k], (fl >> k) & 1u, (so >> k) & 1u, io, nsb);
      };
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
          real uk = dir_dot(k, ux, uy, uz);
          real u_term = uk*(ALPHA + BETA_C*uk) - GAMMA_C*u_sq;
          real G = w*((real)1 + u_term);
          real f_proj = dir_dot(k, bfx, bfy, bfz);
          const real cd_c = cdc[k], cd_p = cdp[k];
          real drho_dk = drho*cd_c*CS2;
          real eh = c*G;
          eh -= (real)0.5*G*(cd_c - c*a_over_rho*(cd_p - mu*cd_c - f_proj));
          eh += (real)0.5*G*u_fh_e;
          real eg = w*(p + rho*CS2*u_term);
          eg -= (real)0.5*(drho_dk*(G - w) + (mu*cd_c + f_proj)*G);
          eg += (real)0.5*((G - w)*u_drhocd + G*u_fg_e);
          real dGk = G - w;
          real dc_dk = (real)0.5*(cd_c + bdc[k]), dp_dk = (real)0.5*(cd_p + bdp[k]);
          real drho_dk2 = drho_cs2*dc_dk;
          real ffk = mu*dc_dk + f_proj;
          real fh = G*(dc_dk - c_over*(dp_dk - ffk) - u_fh_f + mob*avg_of(k));
          real fg = (drho_dk2 - u_drm)*dGk + (ffk - u_fg_f)*G;
          real gnew = gk[k] + relax*(eg - gk[k]) + fg;
          storek(eh + fh, rH, k);
          storek(gnew, rGo, k);
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
        // relaxation rates (device_engine.cuh:622-625, :640-643)
        const real s0r = relax; const real taur = (real)1/s0r - (real)0.5;
        const real s1r = (real)2*taur/((real)2*d.mrt_lambda + taur);
        // x = S M (eq_g - g), accumulated column by column in the reference's
        // order as each direction is produced, keeping its (s*M)*y rounding
        real xx[Q];
        #pragma unroll
        for (int m = 0; m < Q; ++m) xx[m] = 0;
        real fgs[Q];
        #pragma unroll
        for (int k = 0; k < Q; ++k) {
          const real w = W<real>(k);
          real uk = dir_dot(k, ux, uy, uz);
          real u_term = uk*(ALPHA + BETA_C*uk) - gamma_u_sq;
          real G = w*((real)1 + u_term);
          real w_u_term = G - w;
          real f_proj = dir_dot(k, bfx, bfy, bfz);
          const real cd_c = cdc[k], cd_p = cdp[k];
          real drho_dk = drho_cs2*cd_c;
          real eh = c*G + (real)0.5*G*(u_term_h - cd_c + c_over*(cd_p - mu*cd_c - f_proj));
          real eg = w*p + rho_cs2*w_u_term + (real)0.5*((u_drho - drho_dk)*w_u_term + (u_force_g + mu*cd_c + f_proj)*G);
          real dGk = G - w;
          real dc_dk = (real)0.5*(cd_c + bdc[k]), dp_dk = (real)0.5*(cd_p + bdp[k]);
          real drho_dk2 = drho_cs2*dc_dk;
          real ffk = mu*dc_dk + f_proj;
          real fh = G*(dc_dk - c_over*(dp_dk - ffk) - u_fh_f + mob*avg_of(k));
          fgs[k] = (drho_dk2 - u_drm)*dGk + (ffk - u_fg_f)*G;
          storek(eh + fh, rH, k);
          const real yv = eg - gk[k];
          #pragma unroll
          for (int m = 1; m < Q; ++m) {
            if (MRTTAB.M[m][k] != 0.0 && m != 3 && m != 5 && m != 7) {
              const bool odd = (m == 4 || m == 6 || m == 8 || m >= 16);
              const real sm = odd ? s1r : s0r;
              xx[m] += sm*(real)MRTTAB.M[m][k]*yv;
            }
          }
        }
        #pragma unroll
        for (int k = 0; k < Q; ++k) {
          real acc = 0;
          #pragma unroll
          for (int m = 0; m < Q; ++m) {
            const bool cons = (m == 0 || m == 3 || m == 5 || m == 7);
            if (MRTTAB.MI[k][m] != 0.0 && !cons) acc += (real)MRTTAB.MI[k][m]*xx[m];
          }
          storek(gk[k] + acc + fgs[k], rGo, k);
        }
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

template <typename real, bool MRT, bool HBB, bool OPEN>
__global__ void __launch_bounds__(256) k_collide(const Inst<real> *__restrict__ D,
                                                 const int2 *__restrict__ map, int sdi, unsigned T, int stats)
{
  const int2 bm = map_entry(map, sdi);
  if constexpr (OPEN) {
    if (bm.x >= 0 && (bm.x & MAP_INTERIOR)) { collide_body<real, MRT, HBB, false>(D, make_int2(bm.x & ~MAP_INTERIOR, bm.y), T, stats); return; }
  }
  collide_body<real, MRT, HBB, OPEN>(D, bm, T, stats);
}

}  // namespace fm


namespace fm {

// ==== single-instance helpers (not on the step path) ================================

// mass sums of the current state (k_mass_weight, device_engine.cuh:1109-1126);
// with apply, the sum of the folded-corrected c instead (the effective mass).
// post-stream populations of site i gathered from the post-collision array a
template <typename real>
__device__ inline void gather_site(const Inst<real> &d, const real *a, int i, real (&v)[Q])
{
  constexpr unsigned RB = sizeof(real);
  const unsigned nsb = (unsigned)d.ns*RB, io = (unsigned)i*RB;
  const rsrc_t rW = make_rsrc(d.words, d.nwords*8u), rA = make_rsrc(a, 19u*nsb);
  int x, y, z; decode(d, i, x, y, z);
  Nbhd nb; make_nbhd(d, x, y, z, nb);
  unsigned vo[Q], fl, so;
  ring1<real, true>(d, rW, nb, io, vo, fl, so);
  #pragma unroll
  for (int k = 0; k < Q; ++k) v[k] = bld<real>(rA, pull_off(k, vo, fl, io, nsb), pull_soff(k, nsb));
}

template <typename real>
__global__ void __launch_bounds__(256) k_mass_single(Inst<real> d, unsigned cur, int apply, double lambda)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  const real *g = cur ? d.g1 : d.g0; (void)g;
  Fx m = {0, 0, 0}, w = {0, 0, 0};
  if (i < d.n) {
    real hv[Q];
    gather_site(d, d.h, i, hv);
    double cs = 0;                  // summed as the moments pass sums it
    #pragma unroll
    for (int k = 0; k < Q; ++k) cs += (double)hv[k];
    const real c0 = (real)cs;
    const real c = apply ? corrected<real>(c0, true, lambda) : c0;
    double cm = (double)c;
    if constexpr (sizeof(real) == 4) {
      // the populations carry the exact sum and the exact correction (the
      // moments remainder and the correction residual, k_collide_f32)
      const float phi = c0*(1.0f - c0);                 // as corrected() forms it
      cm = cs - ((apply && phi > 0.0f) ? (double)(float)lambda*(double)phi : 0.0);
    }
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
  real hv[Q];
  gather_site(d, d.h, i, hv);
  double cs = 0;                    // summed as the moments pass sums it
  #pragma unroll
  for (int k = 0; k < Q; ++k) cs += (double)hv[k];
  const real c = (real)cs;
  double cm = (double)c, phi = cm*(1.0 - cm);
  bool inj = apply && phi > 0.0;
  double delta_c = -lambda*phi;
  #pragma unroll
  for (int k = 0; k < Q; ++k) {
    real v = hv[k];
    if (inj) v += (real)(WD[k]*delta_c);
    out[k*ns + i] = v;
  }
}

// post-stream copy of a post-collision array (state download)
template <typename real>
__global__ void __launch_bounds__(256) k_gather_copy(Inst<real> d, const real *a, real *out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  real v[Q];
  gather_site(d, a, i, v);
  #pragma unroll
  for (int k = 0; k < Q; ++k) out[(long)k*d.ns + i] = v[k];
}

// inverse of the gather (state upload): slot k of x holds the post-stream
// population k of x + e_k when that is a site, else post-stream opp(k) of x
template <typename real>
__global__ void __launch_bounds__(256) k_unstream(Inst<real> d, const real *in, real *out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  constexpr unsigned RB = sizeof(real);
  const unsigned nsb = (unsigned)d.ns*RB, io = (unsigned)i*RB;
  const rsrc_t rW = make_rsrc(d.words, d.nwords*8u), rI = make_rsrc(in, 19u*nsb), rO = make_rsrc(out, 19u*nsb);
  int x, y, z; decode(d, i, x, y, z);
  Nbhd nb; make_nbhd(d, x, y, z, nb);
  unsigned vo[Q], fl, so;
  ring1<real, true>(d, rW, nb, io, vo, fl, so);
  #pragma unroll
  for (int k = 0; k < Q; ++k) {
    const int o = OPP[k];
    const bool f = (fl >> k) & 1u;
    // slot k of x + e_k, or slot o of x; the scalar part min(k, o)*ns is uniform
    const unsigned off = f ? vo[k] + (k > o ? nsb : 0u) : io + (o > k ? nsb : 0u);
    bst(bld<real>(rI, off, pull_soff(k, nsb)), rO, io, (unsigned)k*nsb);
  }
}

// The rest state of felbm_init_rest in the post-stream layout (population k
// of site i at k*ns + i; k_unstream then stores it for the next gather):
// h_k = c_i w_k (which = 0) or g_k = w_k p0 (which = 1), formed in double and
// rounded once to the state's type (FP32 g as its deviation from w_k, flushed
// below 2^-31) -- the operations and roundings of a host-built state passed
// through upload_state, so the two give the same bits.  c is in engine order.
template <typename real>
__global__ void __launch_bounds__(256) k_rest_stage(Inst<real> d, const double *c, double p0, int which, real *out)
{
  const int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const size_t ns = (size_t)d.ns;
  const double ci = c[i];
  #pragma unroll
  for (int k = 0; k < Q; ++k) {
    real v;
    if (which == 0) v = (real)__dmul_rn(ci, WD[k]);
    else if constexpr (sizeof(real) == 4) v = g_flush((float)__dsub_rn(__dmul_rn(WD[k], p0), (double)wf32(k)*p_ref<real>()));
    else v = (real)__dmul_rn(WD[k], p0);
    out[(size_t)k*ns + (size_t)i] = v;
  }
}

// reference-format stencil/streaming tables (multiphase_gpu.cuh:93-208)
// regenerated from the engine's neighbour logic, for bitwise comparison
// iperm (may be null): the caller's number of each engine site; tables are
// then laid out and filled in the caller's numbering
template <typename real>
__global__ void __launch_bounds__(256) k_export_tables(Inst<real> d, int *cdm, int *cdp, int *bcase,
                                                       int *bc1, int *bc2, int *avgnext, int *lcase,
                                                       int *lc1, int *lc2, int *src, const int *iperm)
{
  const int e = blockIdx.x*blockDim.x + threadIdx.x;
  if (e >= d.n) return;
  const int n = d.n;
  int x, y, z; decode(d, e, x, y, z);
  Nbhd nb; make_nbhd(d, x, y, z, nb);
  int n1[Q], n2[Q]; neighbours1(d, nb, e, n1); neighbours2(d, nb, n2);
  if (iperm) {
    #pragma unroll
    for (int m = 0; m < Q; ++m) { if (n1[m] >= 0) n1[m] = iperm[n1[m]]; if (n2[m] >= 0) n2[m] = iperm[n2[m]]; }
  }
  const int i = iperm ? iperm[e] : e;
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

// set_forcing_factor as a stream-ordered write into one descriptor, so a
// per-step forcing schedule needs no host synchronisation
template <typename real>
__global__ void k_set_ff(Inst<real> *D, int di, real ff)
{
  if (threadIdx.x == 0) stg(&D[di].ff, 0, ff);
}

template <typename real>
__global__ void k_zero_words(unsigned long long *p, int count)
{
  int i = blockIdx.x*blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

}  // namespace fm
