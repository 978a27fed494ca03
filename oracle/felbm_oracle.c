/*
 * felbm_oracle.c -- FP64 CPU restatement of the felbm Lee-Liu D3Q19 step.
 *
 * TEST INFRASTRUCTURE: the parity oracle and CPU baseline for the HIP
 * engine; never linked into the product.  Built with -ffp-contract=off so
 * every operation rounds exactly as written (the HIP FP64 build uses the same
 * flag and the same expression order, which makes the two bit-identical).
 *
 * Reference files followed (paths relative to the reference root):
 *   include/felbm_gpu/d3q19.cuh, device_engine.cuh, multiphase_gpu.cuh,
 *   src/felbm_gpu_main.cu (loop order, forcing factor, mass correction
 *   placement).
 *
 * Parity pins (tests/test_oracle*.py):
 *   D3Q19 tables          closed form + isotropy identities
 *   MRT matrices          M*Minv = I, conserved rows, orthogonality,
 *                         single-relaxation reduction to BGK
 *   stencil tables        all-fluid closed forms; hand-built wall cases
 *   gradients/Laplacian   exact on linear / quadratic fields
 *   in-place program      the reference's own LCG verifier
 *                         (multiphase_gpu.cuh:247-265), permutation invariant
 *   fused collision       equals the unfused equilibria/force/apply chain
 *   mass correction       closed form: total restored, momentum unchanged,
 *                         bulk untouched
 *   solver                rest state fixed point, Poiseuille profile,
 *                         single-phase inertness
 *   multiphase terms      (tests/test_oracle_physics.py) the flat tanh
 *                         interface is stationary with mu at discretisation
 *                         level; Laplace law dp = (d-1) sigma / R in 2D and 3D
 *                         and at density ratio 5; hydrostatic dp/dy = -rho g in
 *                         each phase; a flat interface in uniform flow moves
 *                         with it (Galilean check); the static contact angle
 *                         from phi = 6 sigma cos(theta) at 60, 90, 120 degrees
 *   open boundaries       injection schedules' closed forms and pressure-
 *                         driven Poiseuille; the reference's node lists
 *                         (OpenBoundaryOperator, absent tree) are unpinned
 */
#include "felbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define Q ORC_Q

/* d3q19.cuh:33-44: felbm ordering (rest; +-x; +-y; +-z; xy, yz, zx diagonals) */
static const int CX[19]  = { 0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1 };
static const int CY[19]  = { 0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 1,-1, 1,-1, 0, 0, 0, 0 };
static const int CZ[19]  = { 0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1 };
static const int OPP[19] = { 0, 2, 1, 4, 3, 6, 5, 8, 7,10, 9,12,11,14,13,16,15,18,17 };
static double WK[19];
static double MM[19][19], MI[19][19];
static int consts_ready = 0;

/* d3q19.cuh:47-50 */
static const double CS2 = 1.0/3.0, ALPHA = 3.0, GAMMA_C = 1.5, BETA_C = 4.5;

/* MRT basis: d'Humieres et al. (2002) D3Q19 moments in felbm direction order,
 * rows rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz,
 * pxz, mx, my, mz.  The row order is fixed by the relaxation vector of
 * device_engine.cuh:625,643.  The CPU statics (EquilibriumDistributionMRT)
 * are in the absent felbm_local tree; the invariants in the tests pin the
 * structure, the exact bits of M and Minv are parity unpinned against them.
 * Minv = M^T diag(1/|row|^2) (rows are mutually orthogonal). */
static void build_consts(void)
{
  if (consts_ready) return;
  for (int k = 0; k < 19; ++k) WK[k] = (k == 0) ? 1.0/3.0 : (k < 7 ? 1.0/18.0 : 1.0/36.0);
  for (int k = 0; k < 19; ++k) {
    int ex = CX[k], ey = CY[k], ez = CZ[k];
    int e2 = ex*ex + ey*ey + ez*ez;
    double row[19];
    row[0]  = 1;
    row[1]  = 19*e2 - 30;
    row[2]  = (21*e2*e2 - 53*e2 + 24)/2;
    row[3]  = ex;             row[4]  = (5*e2 - 9)*ex;
    row[5]  = ey;             row[6]  = (5*e2 - 9)*ey;
    row[7]  = ez;             row[8]  = (5*e2 - 9)*ez;
    row[9]  = 3*ex*ex - e2;   row[10] = (3*e2 - 5)*(3*ex*ex - e2);
    row[11] = ey*ey - ez*ez;  row[12] = (3*e2 - 5)*(ey*ey - ez*ez);
    row[13] = ex*ey;          row[14] = ey*ez;           row[15] = ex*ez;
    row[16] = (ey*ey - ez*ez)*ex;
    row[17] = (ez*ez - ex*ex)*ey;
    row[18] = (ex*ex - ey*ey)*ez;
    for (int m = 0; m < 19; ++m) MM[m][k] = row[m];
  }
  for (int m = 0; m < 19; ++m) {
    double nrm = 0;
    for (int k = 0; k < 19; ++k) nrm += MM[m][k]*MM[m][k];
    for (int k = 0; k < 19; ++k) MI[k][m] = MM[m][k] / nrm;
  }
  consts_ready = 1;
}

void orc_d3q19(int *cx, int *cy, int *cz, int *opp, double *w)
{
  build_consts();
  for (int k = 0; k < 19; ++k) { cx[k] = CX[k]; cy[k] = CY[k]; cz[k] = CZ[k]; opp[k] = OPP[k]; w[k] = WK[k]; }
}

void orc_mrt_matrices(double *M, double *Minv)
{
  build_consts();
  memcpy(M, MM, sizeof(MM)); memcpy(Minv, MI, sizeof(MI));
}

/* ---- exact fixed-point sums ------------------------------------------------
 * Each term t (|t| < 8) is rounded once to a multiple of 2^-60 and summed as a
 * split 32/32-bit integer, so the total is independent of summation order.
 * The HIP engine uses the identical definition; it replaces the reference's
 * order-nondeterministic double atomicAdd (device_engine.cuh:1124-1125). */
typedef struct { int64_t hi; uint64_t lo; uint64_t bad; } fxacc;

static inline void fx_add(fxacc *a, double t)
{
  if (!(fabs(t) < 8.0)) { a->bad++; return; }
  int64_t q = llrint(t * 0x1p60);
  a->hi += q >> 32;
  a->lo += (uint64_t)(q & 0xffffffffLL);
}
static inline void fx_merge(fxacc *a, const fxacc *b) { a->hi += b->hi; a->lo += b->lo; a->bad += b->bad; }
static inline double fx_value(const fxacc *a)
{
  int64_t hi = a->hi + (int64_t)(a->lo >> 32);
  uint64_t lo = a->lo & 0xffffffffULL;
  if (a->bad) return NAN;
  return ((double)hi * 4294967296.0 + (double)lo) * 0x1p-60;
}

double orc_fixed_sum(const double *v, int64_t n)
{
  fxacc a = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) fx_add(&a, v[i]);
  return fx_value(&a);
}

/* ---- lattice ------------------------------------------------------------ */
int orc_lattice_init(orc_lattice *L, int Nx, int Ny, int Nz, const uint8_t *solid,
                     int hbb, int px, int py, int pz)
{
  build_consts();
  memset(L, 0, sizeof(*L));
  L->Nx = Nx; L->Ny = Ny; L->Nz = Nz; L->hbb = hbb;
  L->periodic[0] = px; L->periodic[1] = py; L->periodic[2] = pz;
  int64_t V = (int64_t)Nx*Ny*Nz;
  L->vox2idx = (int32_t*)malloc(sizeof(int32_t)*V);
  int n = 0;
  for (int64_t v = 0; v < V; ++v) L->vox2idx[v] = solid[v] ? -1 : n++;
  L->n = n;
  L->cx = (int32_t*)malloc(sizeof(int32_t)*(n ? n : 1));
  L->cy = (int32_t*)malloc(sizeof(int32_t)*(n ? n : 1));
  L->cz = (int32_t*)malloc(sizeof(int32_t)*(n ? n : 1));
  L->is_streamed = (uint8_t*)malloc(n ? n : 1);
  for (int64_t v = 0; v < V; ++v) {
    int i = L->vox2idx[v];
    if (i < 0) continue;
    L->cx[i] = (int)(v % Nx); L->cy[i] = (int)((v / Nx) % Ny); L->cz[i] = (int)(v / ((int64_t)Nx*Ny));
    L->is_streamed[i] = 1;
  }
  return 0;
}

void orc_lattice_free(orc_lattice *L)
{
  free(L->vox2idx); free(L->cx); free(L->cy); free(L->cz); free(L->is_streamed);
  memset(L, 0, sizeof(*L));
}

/* coords_to_idx(x+d) with periodic wrap on periodic axes; -1 = solid or
 * outside a non-periodic axis. */
static inline int nbr(const orc_lattice *L, int i, int dx, int dy, int dz)
{
  int x = L->cx[i] + dx, y = L->cy[i] + dy, z = L->cz[i] + dz;
  if (x < 0 || x >= L->Nx) { if (!L->periodic[0]) return -1; x = ((x % L->Nx) + L->Nx) % L->Nx; }
  if (y < 0 || y >= L->Ny) { if (!L->periodic[1]) return -1; y = ((y % L->Ny) + L->Ny) % L->Ny; }
  if (z < 0 || z >= L->Nz) { if (!L->periodic[2]) return -1; z = ((z % L->Nz) + L->Nz) % L->Nz; }
  return L->vox2idx[(int64_t)x + (int64_t)L->Nx*((int64_t)y + (int64_t)L->Ny*z)];
}
static inline int is_known(const orc_lattice *L, int i, int m) { return nbr(L, i, -CX[m], -CY[m], -CZ[m]) >= 0; }
static inline int voxel_inside(const orc_lattice *L, int i, int dx, int dy, int dz)
{
  int x = L->cx[i] + dx, y = L->cy[i] + dy, z = L->cz[i] + dz;
  if (!L->periodic[0] && (x < 0 || x >= L->Nx)) return 0;
  if (!L->periodic[1] && (y < 0 || y >= L->Ny)) return 0;
  if (!L->periodic[2] && (z < 0 || z >= L->Nz)) return 0;
  return 1;
}

/* ---- table builders: multiphase_gpu.cuh:93-193 ---------------------------- */
static void build_tables(orc_engine *E)
{
  const orc_lattice *L = &E->L;
  int n = E->n; int hbb = L->hbb;
  size_t Q1n = (size_t)(Q-1)*n, Vn = (size_t)Q*n;
  E->cdm = malloc(sizeof(int32_t)*(Q1n+1)); E->cdp = malloc(sizeof(int32_t)*(Q1n+1));
  E->bcase = malloc(sizeof(int32_t)*(Q1n+1)); E->bc1 = malloc(sizeof(int32_t)*(Q1n+1)); E->bc2 = malloc(sizeof(int32_t)*(Q1n+1));
  E->avgnext = malloc(sizeof(int32_t)*(Vn+1));
  E->lcase = malloc(sizeof(int32_t)*(Q1n+1)); E->lc1 = malloc(sizeof(int32_t)*(Q1n+1)); E->lc2 = malloc(sizeof(int32_t)*(Q1n+1));
  E->src = malloc(sizeof(int32_t)*(Vn+1));
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    for (int m = 1; m < Q; ++m) {
      int ex = CX[m], ey = CY[m], ez = CZ[m], mo = OPP[m];
      int inx = nbr(L, i, ex, ey, ez), ipx = nbr(L, i, -ex, -ey, -ez);
      int in2 = nbr(L, i, 2*ex, 2*ey, 2*ez), ip2 = nbr(L, i, -2*ex, -2*ey, -2*ez);
      int Km = is_known(L, i, m), Ko = is_known(L, i, mo);
      int Nok = inx >= 0, Pok = ipx >= 0;
      size_t idx = (size_t)(m-1)*n + i;
      /* grad_cd column pair: multiphase_gpu.cuh:113-121 */
      int a = -1, b = -1;
      if (Km && Ko && Pok && Nok) { a = ipx; b = inx; }
      else if (hbb && Km && Pok)  { a = ipx; b = i; }
      else if (hbb && Ko && Nok)  { a = i;   b = inx; }
      E->cdm[idx] = a; E->cdp[idx] = b;
      /* grad_bd cases: multiphase_gpu.cuh:140-153 */
      int cs = 0, c1 = -1, c2 = -1;
      if (Ko && inx >= 0) {
        if (is_known(L, inx, mo) && in2 >= 0) { cs = 1; c1 = inx; c2 = in2; }
        else if (!hbb) { cs = 2; c1 = inx; }
        else           { cs = 3; c1 = inx; }
      } else {
        if (!hbb) {
          if (Km && ipx >= 0 && is_known(L, ipx, m) && ip2 >= 0) { cs = 1; c1 = ipx; c2 = ip2; }
        } else {
          if (Km && ipx >= 0) { cs = 4; c1 = ipx; }
        }
      }
      E->bcase[idx] = cs; E->bc1[idx] = c1; E->bc2[idx] = c2;
      /* lap cases: multiphase_gpu.cuh:178-188 */
      cs = 0; c1 = -1; c2 = -1;
      if (Km && Ko && Pok && Nok) { cs = 1; c1 = ipx; c2 = inx; }
      else if (Km && Pok)         { cs = hbb ? 3 : 2; c1 = ipx; }
      else if (Ko && Nok)         { cs = hbb ? 3 : 2; c1 = inx; }
      E->lcase[idx] = cs; E->lc1[idx] = c1; E->lc2[idx] = c2;
    }
    /* avg_dir: multiphase_gpu.cuh:173-177 */
    for (int m = 0; m < Q; ++m) {
      int nx = nbr(L, i, CX[m], CY[m], CZ[m]);
      E->avgnext[(size_t)m*n + i] = (is_known(L, i, OPP[m]) && nx >= 0) ? nx : -1;
    }
    /* streaming source (the StreamingOperator rows reduced as in
     * multiphase_gpu.cuh:200-208).  The operator lives in the absent tree;
     * this engine defines it: rest = self; bulk pull from x - e_k; halfway
     * bounce-back from slot opp(k) of the same site when x - e_k is a solid
     * voxel inside the box; empty (-2) when x - e_k is outside a
     * non-periodic box face. */
    E->src[i] = i;
    for (int k = 1; k < Q; ++k) {
      int s = nbr(L, i, -CX[k], -CY[k], -CZ[k]);
      int v;
      if (s >= 0) v = k*n + s;
      else if (voxel_inside(L, i, -CX[k], -CY[k], -CZ[k]) && hbb) v = OPP[k]*n + i;
      else v = -2;
      E->src[(size_t)k*n + i] = v;
    }
  }
}

/* in-place program builder + host verification: multiphase_gpu.cuh:213-283 */
static int build_inplace(orc_engine *E)
{
  int n = E->n; int64_t Vn = (int64_t)Q*n;
  int32_t *oa = malloc(sizeof(int32_t)*(Vn+1)), *ob = malloc(sizeof(int32_t)*(Vn+1));
  uint8_t *ot = malloc(Vn+1); char *dn = calloc(Vn+1, 1);
  int64_t no = 0; int ok = 1;
  for (int64_t j = 0; j < Vn && ok; ++j) {
    if (dn[j]) continue;
    int k = (int)(j / n), i = (int)(j - (int64_t)k*n), kb = OPP[k];
    int64_t jb = (int64_t)kb*n + i; int64_t sc = E->src[j];
    if (sc == jb) { dn[j] = 1; }
    else if (sc >= 0) {
      int sk = (int)(sc / n), x = (int)(sc - (int64_t)sk*n); int64_t q = (int64_t)kb*n + x;
      if (sk != k || E->src[q] != (int64_t)kb*n + i || dn[q]) { ok = 0; break; }
      oa[no] = (int32_t)j; ob[no] = (int32_t)q; ot[no] = 0; ++no; dn[j] = dn[q] = 1;
    } else if (sc == -1) {
      if (E->src[jb] != -1 || dn[jb]) { ok = 0; break; }
      oa[no] = (int32_t)j; ob[no] = (int32_t)jb; ot[no] = 1; ++no; dn[j] = dn[jb] = 1;
    } else {
      oa[no] = (int32_t)j; ob[no] = (int32_t)j; ot[no] = 2; ++no; dn[j] = 1;
    }
  }
  free(dn);
  E->oa = oa; E->ob = ob; E->ot = ot; E->n_ops = no;
  if (ok) ok = orc_verify_inplace(E);
  E->inplace_ok = ok;
  return ok;
}

/* the reference's host verifier (multiphase_gpu.cuh:247-265): the op program
 * applied to the reversed layout of an LCG vector must equal the gather. */
int orc_verify_inplace(const orc_engine *E)
{
  int n = E->n; int64_t Vn = (int64_t)Q*n;
  double *f = malloc(sizeof(double)*(Vn+1)), *ref = malloc(sizeof(double)*(Vn+1)), *r = malloc(sizeof(double)*(Vn+1));
  unsigned long long seed = 0x243F6A8885A308D3ULL;
  for (int64_t j = 0; j < Vn; ++j) {
    seed = seed*6364136223846793005ULL + 1442695040888963407ULL;
    f[j] = (double)(seed >> 11) * (1.0/9007199254740992.0) - 0.5;
  }
  for (int64_t j = 0; j < Vn; ++j) {
    int k = (int)(j / n), i = (int)(j - (int64_t)k*n); int64_t jb = (int64_t)OPP[k]*n + i; int64_t sc = E->src[j];
    ref[j] = sc >= 0 ? f[sc] : (sc == -1 ? 0.5*(f[j] + f[jb]) : 0.0);
    r[j] = f[jb];
  }
  for (int64_t q = 0; q < E->n_ops; ++q) {
    int a = E->oa[q], b = E->ob[q];
    if (E->ot[q] == 0) { double t = r[a]; r[a] = r[b]; r[b] = t; }
    else if (E->ot[q] == 1) { double m = 0.5*(r[a] + r[b]); r[a] = m; r[b] = m; }
    else r[a] = 0.0;
  }
  int ok = 1;
  for (int64_t j = 0; j < Vn; ++j) if (r[j] != ref[j]) { ok = 0; break; }
  free(f); free(ref); free(r);
  return ok;
}

int64_t orc_num_ops(const orc_engine *E) { return E->n_ops; }

/* ---- engine lifecycle ------------------------------------------------------ */
static double *dz(size_t m) { return (double*)calloc(m ? m : 1, sizeof(double)); }

int orc_engine_init(orc_engine *E, int Nx, int Ny, int Nz, const uint8_t *solid,
                    int hbb, int px, int py, int pz,
                    const orc_params *P, const orc_switches *S)
{
  memset(E, 0, sizeof(*E));
  orc_lattice_init(&E->L, Nx, Ny, Nz, solid, hbb, px, py, pz);
  E->P = *P; E->S = *S;
  int n = E->L.n; E->n = n; E->Vn = Q*n;
  size_t Vn = (size_t)Q*n;
  E->h = dz(Vn); E->g = dz(Vn); E->h2 = dz(Vn); E->g2 = dz(Vn);
  E->c = dz(n); E->p = dz(n); E->rho = dz(n); E->mu = dz(n); E->ux = dz(n); E->uy = dz(n); E->uz = dz(n);
  E->relax = dz(n); E->lapc = dz(n); E->lapmu = dz(n); E->xtnd = dz(2*(size_t)n);
  for (int d = 0; d < 3; ++d) { E->gcc[d] = dz(n); E->gcb[d] = dz(n); E->gcm[d] = dz(n); E->gpc[d] = dz(n); E->gpb[d] = dz(n); E->gpm[d] = dz(n); }
  if (!S->fuse_collision) {
    E->gc_cd = dz(Vn); E->gp_cd = dz(Vn); E->gc_bd = dz(Vn); E->gp_bd = dz(Vn); E->avg = dz(Vn);
    E->eqh = dz(Vn); E->eqg = dz(Vn); E->collh = dz(Vn); E->collg = dz(Vn); E->fh = dz(Vn); E->fg = dz(Vn);
  }
  E->bcproc = calloc(n ? n : 1, 1);
  build_tables(E);
  build_inplace(E);
  if (E->S.stream_inplace && (!E->inplace_ok || !E->S.fuse_collision)) E->S.stream_inplace = 0;
  return 0;
}

void orc_engine_free(orc_engine *E)
{
  void *ps[] = { E->h, E->g, E->h2, E->g2, E->c, E->p, E->rho, E->mu, E->ux, E->uy, E->uz, E->relax,
                 E->lapc, E->lapmu, E->xtnd, E->gc_cd, E->gp_cd, E->gc_bd, E->gp_bd, E->avg, E->eqh,
                 E->eqg, E->collh, E->collg, E->fh, E->fg, E->cdm, E->cdp, E->bcase, E->bc1, E->bc2,
                 E->avgnext, E->lcase, E->lc1, E->lc2, E->src, E->oa, E->ob, E->ot, E->inlet,
                 E->outlet, E->buf_verts, E->buf_refs, E->inlet_coord, E->bcproc };
  for (size_t q = 0; q < sizeof(ps)/sizeof(ps[0]); ++q) free(ps[q]);
  for (int d = 0; d < 3; ++d) { free(E->gcc[d]); free(E->gcb[d]); free(E->gcm[d]); free(E->gpc[d]); free(E->gpb[d]); free(E->gpm[d]); }
  orc_lattice_free(&E->L);
  memset(E, 0, sizeof(*E));
}

void orc_get_tables(const orc_engine *E, int32_t *cdm, int32_t *cdp, int32_t *bcase,
                    int32_t *bc1, int32_t *bc2, int32_t *avgnext, int32_t *lcase,
                    int32_t *lc1, int32_t *lc2, int32_t *src)
{
  size_t Q1n = (size_t)(Q-1)*E->n, Vn = (size_t)Q*E->n;
  if (cdm) memcpy(cdm, E->cdm, 4*Q1n);
  if (cdp) memcpy(cdp, E->cdp, 4*Q1n);
  if (bcase) memcpy(bcase, E->bcase, 4*Q1n);
  if (bc1) memcpy(bc1, E->bc1, 4*Q1n);
  if (bc2) memcpy(bc2, E->bc2, 4*Q1n);
  if (avgnext) memcpy(avgnext, E->avgnext, 4*Vn);
  if (lcase) memcpy(lcase, E->lcase, 4*Q1n);
  if (lc1) memcpy(lc1, E->lc1, 4*Q1n);
  if (lc2) memcpy(lc2, E->lc2, 4*Q1n);
  if (src) memcpy(src, E->src, 4*Vn);
}

double *orc_field(orc_engine *E, const char *nm)
{
  struct { const char *k; double *v; } t[] = {
    {"c", E->c}, {"p", E->p}, {"rho", E->rho}, {"mu", E->mu}, {"ux", E->ux}, {"uy", E->uy},
    {"uz", E->uz}, {"relax", E->relax}, {"lapc", E->lapc}, {"lapmu", E->lapmu}, {"h", E->h},
    {"g", E->g}, {"gcc_x", E->gcc[0]}, {"gcc_y", E->gcc[1]}, {"gcc_z", E->gcc[2]},
    {"gcb_x", E->gcb[0]}, {"gcb_y", E->gcb[1]}, {"gcb_z", E->gcb[2]},
    {"gpc_x", E->gpc[0]}, {"gpc_y", E->gpc[1]}, {"gpc_z", E->gpc[2]},
    {"gpb_x", E->gpb[0]}, {"gpb_y", E->gpb[1]}, {"gpb_z", E->gpb[2]} };
  for (size_t q = 0; q < sizeof(t)/sizeof(t[0]); ++q) if (!strcmp(nm, t[q].k)) return t[q].v;
  return NULL;
}

/* ---- state ------------------------------------------------------------------ */
void orc_upload_state(orc_engine *E, const double *h, const double *g)
{
  memcpy(E->h, h, sizeof(double)*(size_t)E->Vn);
  memcpy(E->g, g, sizeof(double)*(size_t)E->Vn);
  E->pending = 0;
}

/* Rest equilibrium h_k = w_k c, g_k = w_k p0 (the initializers live in the
 * absent tree; this is the u = 0 equilibrium of device_engine.cuh:776-777). */
void orc_init_rest(orc_engine *E, const double *c_site, double p0)
{
  int n = E->n;
  for (int k = 0; k < Q; ++k)
    for (int i = 0; i < n; ++i) {
      E->h[(size_t)k*n + i] = c_site[i]*WK[k];
      E->g[(size_t)k*n + i] = WK[k]*p0;
    }
  E->pending = 0;
}

/* ---- field pass kernels --------------------------------------------------- */

/* mass correction folded into the moments: the injection
 * h_k += (real)(w_k*dc) of device_engine.cuh:1128-1139 changes the site sum by
 * dc; folded, c = c_raw + dc rounded once.  Only used in mass_mode 2. */
static double fold_lambda(orc_engine *E, int *apply)
{
  double M, W;
  orc_mass_weight(E, &M, &W);
  double drift = M - E->target_mass;
  E->last_drift = drift;
  *apply = (W > 0.0 && drift != 0.0);
  return *apply ? drift / W : 0.0;
}

/* k_moments: device_engine.cuh:116-135 */
static void k_moments(orc_engine *E)
{
  int n = E->n; const orc_params *P = &E->P;
  double drho = P->rho0 - P->rho1, four_beta = 4.0*P->beta;
  int apply = 0; double lambda = 0.0;
  if (E->S.mass_mode == 2 && E->pending) lambda = fold_lambda(E, &apply);
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double c = 0, p = 0, ux = 0, uy = 0, uz = 0;
    for (int k = 0; k < Q; ++k) {
      double hv = E->h[(size_t)k*n + i], gv = E->g[(size_t)k*n + i];
      c += hv; p += gv; ux += (double)CX[k]*gv; uy += (double)CY[k]*gv; uz += (double)CZ[k]*gv;
    }
    if (apply && E->L.is_streamed[i]) {
      double phi = c*(1.0 - c);
      if (phi > 0.0) c = c + (-lambda*phi);
    }
    double c_log = c*(1.0 - c);
    double rho = P->rho1 + drho*c;
    double inv = 1.0/(CS2*rho);
    ux *= inv; uy *= inv; uz *= inv;
    double mu = four_beta*c_log*(0.5 - c);
    double cp = c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
    double tau = 1.0/(cp/P->tau0 + (1.0 - cp)/P->tau1);
    double rp = 1.0/(tau + 0.5);
    E->c[i] = c; E->p[i] = p; E->rho[i] = rho; E->mu[i] = mu;
    E->ux[i] = ux; E->uy[i] = uy; E->uz[i] = uz; E->relax[i] = rp;
  }
  if (E->S.mass_mode == 2) E->pending = 0;
}

/* k_grad_cd_mf: device_engine.cuh:897-913 */
static void grad_cd(orc_engine *E, const double *f, double *gx, double *gy, double *gz)
{
  int n = E->n;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double sx = 0, sy = 0, sz = 0;
    for (int m = 1; m < Q; ++m) {
      size_t idx = (size_t)(m-1)*n + i;
      int cm = E->cdm[idx];
      if (cm < 0) continue;
      int cp = E->cdp[idx];
      double d = f[cp] - f[cm];
      double coef = ALPHA*0.5*WK[m];
      sx += coef*(double)CX[m]*d; sy += coef*(double)CY[m]*d; sz += coef*(double)CZ[m]*d;
    }
    gx[i] = sx; gy[i] = sy; gz[i] = sz;
  }
}

/* k_grad_bd_mf: device_engine.cuh:924-945 */
static void grad_bd(orc_engine *E, const double *f, double *gx, double *gy, double *gz)
{
  int n = E->n;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double sx = 0, sy = 0, sz = 0, f0 = f[i];
    for (int m = 1; m < Q; ++m) {
      size_t idx = (size_t)(m-1)*n + i;
      int cs = E->bcase[idx];
      if (cs == 0) continue;
      double f1 = f[E->bc1[idx]], val;
      if (cs == 1) { double f2 = f[E->bc2[idx]]; val = -1.5*f0 + 2.0*f1 + -0.5*f2; }
      else if (cs == 2) val = -2.0*f0 + 2.0*f1;
      else if (cs == 3) val = -1.5*f0 + 1.5*f1;
      else              val = 0.5*f0 + -0.5*f1;
      double base = ALPHA*WK[m];
      sx += base*(double)CX[m]*val; sy += base*(double)CY[m]*val; sz += base*(double)CZ[m]*val;
    }
    gx[i] = sx; gy[i] = sy; gz[i] = sz;
  }
}

/* k_grad_md: device_engine.cuh:138-146 */
static void grad_md(int n, double **cd, double **bd, double **md)
{
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) md[d][i] = 0.5*(cd[d][i] + bd[d][i]);
}

/* k_lap_mf: device_engine.cuh:996-1010 over the 2n input of k_pack_lap_*
 * (device_engine.cuh:159-173) */
static void lap(orc_engine *E, const double *xtnd, double *out)
{
  int n = E->n;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double f0 = xtnd[i], bnd = xtnd[n + i], s = 0;
    for (int m = 1; m < Q; ++m) {
      size_t idx = (size_t)(m-1)*n + i;
      int cs = E->lcase[idx];
      if (cs == 0) continue;
      double w = ALPHA*WK[m];
      if (cs == 1) { double f1 = xtnd[E->lc1[idx]], f2 = xtnd[E->lc2[idx]]; s += w*(f1 - 2.0*f0 + f2); }
      else if (cs == 2) { double f1 = xtnd[E->lc1[idx]]; s += w*(2.0*f1 - 2.0*f0 - 2.0*bnd); }
      else { double f1 = xtnd[E->lc1[idx]]; s += w*(f1 - f0 - bnd); }
    }
    out[i] = s;
  }
}

/* k_vel_press_corr: device_engine.cuh:177-196 */
static void vel_press_corr(orc_engine *E)
{
  int n = E->n; const orc_params *P = &E->P; double drho = P->rho0 - P->rho1;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if (E->P.open_bnd && E->bcproc[i]) continue;
    double mu = E->mu[i], rho = E->rho[i];
    double gx = E->gcc[0][i], gy = E->gcc[1][i], gz = E->gcc[2][i];
    double mor = mu/rho;
    double ux = E->ux[i] + 0.5*(P->gx + P->fx*P->forcing_factor/rho + mor*gx);
    double uy = E->uy[i] + 0.5*(P->gy + P->fy*P->forcing_factor/rho + mor*gy);
    double uz = E->uz[i] + 0.5*(P->gz + P->fz*P->forcing_factor/rho + mor*gz);
    E->ux[i] = ux; E->uy[i] = uy; E->uz[i] = uz;
    double u_dot_drho = drho*(ux*gx + uy*gy + uz*gz);
    E->p[i] += 0.5*u_dot_drho*CS2;
  }
}

/* ---- open boundaries: device_engine.cuh:699-846 ----------------------------- */
static double inlet_gate(const orc_params *P, unsigned step, double coord)
{
  double w = P->inlet_ramp;
  if (P->inlet_mode == 1) {
    double Per = P->inlet_period;
    if (Per <= 0.0) return 1.0;
    /* the reference floors in single precision here (floorf) */
    double tau = (double)step - Per*(double)floorf((float)((double)step/Per));
    double lo = 0.5*Per - 0.5*P->inlet_duty*Per, hi = 0.5*Per + 0.5*P->inlet_duty*Per;
    if (w <= 0.0) return (tau >= lo && tau < hi) ? 1.0 : 0.0;
    return 0.5*(tanh(2.0*(tau - lo)/w) - tanh(2.0*(tau - hi)/w));
  }
  if (P->inlet_mode == 2) {
    double x = coord - P->split_pos;
    if (w <= 0.0) return (x < 0.0) ? 1.0 : 0.0;
    return 0.5*(1.0 - tanh(2.0*x/w));
  }
  if (P->inlet_mode == 3) {
    double Per = P->inlet_period;
    if (Per <= 0.0) return 1.0;
    double t = coord - P->split_pos;
    t = t - Per*floor(t/Per);
    double lo = 0.5*Per - 0.5*P->inlet_duty*Per, hi = 0.5*Per + 0.5*P->inlet_duty*Per;
    if (w <= 0.0) return (t >= lo && t < hi) ? 1.0 : 0.0;
    return 0.5*(tanh(2.0*(t - lo)/w) - tanh(2.0*(t - hi)/w));
  }
  return 1.0;
}

int orc_set_open_bnd(orc_engine *E, const int32_t *inlet, int n_in,
                     const int32_t *outlet, int n_out, const double *icoord,
                     const int32_t *buf_verts, const int32_t *buf_refs, int n_buf,
                     int copy_to_buffers)
{
  orc_params *P = &E->P; int n = E->n;
  P->n_inlet = n_in; P->n_outlet = n_out;
  memset(E->bcproc, 0, n ? n : 1);
  if (P->use_in_vel || P->use_in_prs || P->use_in_fluid) for (int q = 0; q < n_in; ++q) E->bcproc[inlet[q]] = 1;
  if (P->use_out_prs || P->use_out_fluid) for (int q = 0; q < n_out; ++q) E->bcproc[outlet[q]] = 1;
  if (copy_to_buffers) for (int q = 0; q < n_buf; ++q) E->bcproc[buf_verts[q]] = 1;
  E->inlet = malloc(sizeof(int32_t)*(n_in + 1)); memcpy(E->inlet, inlet, sizeof(int32_t)*n_in);
  E->outlet = malloc(sizeof(int32_t)*(n_out + 1)); memcpy(E->outlet, outlet, sizeof(int32_t)*n_out);
  E->inlet_coord = NULL;
  if (icoord) { E->inlet_coord = malloc(sizeof(double)*(n_in + 1)); memcpy(E->inlet_coord, icoord, sizeof(double)*n_in); }
  P->n_buf = n_buf; P->copy_to_buf = copy_to_buffers ? 1 : 0;
  E->buf_verts = malloc(sizeof(int32_t)*(n_buf + 1)); E->buf_refs = malloc(sizeof(int32_t)*(n_buf + 1));
  if (n_buf) { memcpy(E->buf_verts, buf_verts, sizeof(int32_t)*n_buf); memcpy(E->buf_refs, buf_refs, sizeof(int32_t)*n_buf); }
  /* buffer (ghost) nodes are not streamed (collision skipped) */
  for (int q = 0; q < n_buf; ++q) E->L.is_streamed[buf_verts[q]] = 0;
  return 0;
}

/* enforce_open_bnd_fields: multiphase_gpu.cuh:404-418; kernels :738-759, :807-815 */
static void enforce_open_bnd_fields(orc_engine *E)
{
  const orc_params *P = &E->P;
  if (!P->open_bnd) return;
  for (int j = 0; j < P->n_inlet; ++j) {
    int i = E->inlet[j];
    if (P->use_in_vel) { E->ux[i] = P->u_in_x; E->uy[i] = P->u_in_y; E->uz[i] = P->u_in_z; }
    if (P->use_in_prs) E->p[i] = P->p_in;
    if (P->use_in_fluid) {
      double gate = inlet_gate(P, E->open_step, E->inlet_coord ? E->inlet_coord[j] : 0.0);
      double c = (1.0 - P->inlet_c_a) + (2.0*P->inlet_c_a - 1.0)*gate;
      E->c[i] = c; E->rho[i] = c*P->rho0 + (1.0 - c)*P->rho1;
    }
  }
  for (int j = 0; j < P->n_outlet; ++j) {
    int i = E->outlet[j];
    if (P->use_out_prs) E->p[i] = P->p_out;
    if (P->use_out_fluid) { E->c[i] = P->c_out_fixed; E->rho[i] = P->rho_out_fixed; }
  }
  if (P->copy_to_buf)
    for (int j = 0; j < P->n_buf; ++j) {
      int b = E->buf_verts[j], r = E->buf_refs[j];
      E->c[b] = E->c[r]; E->rho[b] = E->rho[r]; E->p[b] = E->p[r];
      E->ux[b] = E->ux[r]; E->uy[b] = E->uy[r]; E->uz[b] = E->uz[r];
    }
}

static void write_plain_eq(orc_engine *E, int i, double ux, double uy, double uz, double pp, double rho, double c)
{
  int n = E->n;
  double u_sq = ux*ux + uy*uy + uz*uz;
  for (int k = 0; k < Q; ++k) {
    double w = WK[k];
    double uk = ux*(double)CX[k] + uy*(double)CY[k] + uz*(double)CZ[k];
    double u_term = uk*(ALPHA + BETA_C*uk) - GAMMA_C*u_sq;
    E->h[(size_t)k*n + i] = c*w*(1.0 + u_term);
    E->g[(size_t)k*n + i] = w*(pp + rho*CS2*u_term);
  }
}

/* apply_open_bnd: multiphase_gpu.cuh:427-446; kernels :761-798, :826-846 */
static void apply_open_bnd(orc_engine *E)
{
  const orc_params *P = &E->P;
  if (!P->open_bnd) return;
  if (P->copy_to_buf) {
    for (int j = 0; j < P->n_buf; ++j) {
      int b = E->buf_verts[j], r = E->buf_refs[j];
      write_plain_eq(E, b, E->ux[r], E->uy[r], E->uz[r], E->p[r], E->rho[r], E->c[r]);
    }
  } else {
    for (int j = 0; j < P->n_inlet; ++j) {
      int i = E->inlet[j];
      double ux, uy, uz, pp, rho, c;
      if (P->use_in_vel) { ux = P->u_in_x; uy = P->u_in_y; uz = P->u_in_z; }
      else { ux = E->ux[i]; uy = E->uy[i]; uz = E->uz[i]; }
      pp = P->use_in_prs ? P->p_in : E->p[i];
      if (P->use_in_fluid) {
        double gate = inlet_gate(P, E->open_step, E->inlet_coord ? E->inlet_coord[j] : 0.0);
        c = (1.0 - P->inlet_c_a) + (2.0*P->inlet_c_a - 1.0)*gate;
        rho = c*P->rho0 + (1.0 - c)*P->rho1;
      } else { rho = E->rho[i]; c = E->c[i]; }
      write_plain_eq(E, i, ux, uy, uz, pp, rho, c);
    }
    for (int j = 0; j < P->n_outlet; ++j) {
      int i = E->outlet[j];
      double pp = P->use_out_prs ? P->p_out : E->p[i], rho, c;
      if (P->use_out_fluid) { rho = P->rho_out_fixed; c = P->c_out_fixed; }
      else { rho = E->rho[i]; c = E->c[i]; }
      write_plain_eq(E, i, E->ux[i], E->uy[i], E->uz[i], pp, rho, c);
    }
  }
  E->open_step++;
}

/* ---- per-direction helpers: device_engine.cuh:89-113 ----------------------- */
static inline double cd_dir_val(const orc_engine *E, const double *f, int i, int m)
{
  if (m == 0) return 0.0;
  size_t idx = (size_t)(m-1)*E->n + i; int cm = E->cdm[idx];
  return (cm < 0) ? 0.0 : 0.5*(f[E->cdp[idx]] - f[cm]);
}
static inline double bd_dir_val(const orc_engine *E, const double *f, double f0, int i, int m)
{
  if (m == 0) return 0.0;
  size_t idx = (size_t)(m-1)*E->n + i; int cs = E->bcase[idx];
  if (cs == 0) return 0.0;
  double f1 = f[E->bc1[idx]];
  if (cs == 1) { double f2 = f[E->bc2[idx]]; return -1.5*f0 + 2.0*f1 + -0.5*f2; }
  else if (cs == 2) return -2.0*f0 + 2.0*f1;
  else if (cs == 3) return -1.5*f0 + 1.5*f1;
  else return 0.5*f0 + -0.5*f1;
}
static inline double avg_dir_val(const orc_engine *E, const double *x, double xi, int i, int m)
{
  int nx = E->avgnext[(size_t)m*E->n + i];
  return (nx < 0) ? xi : 0.5*(xi + x[nx]);
}

/* ---- fused collision: BGK device_engine.cuh:459-545 ------------------------- */
static void collide_fused_bgk(orc_engine *E)
{
  int n = E->n; const orc_params *P = &E->P;
  double drho = P->rho0 - P->rho1;
  int inplace = E->S.stream_inplace;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if (!E->L.is_streamed[i]) continue;
    double c = E->c[i], rho = E->rho[i], p = E->p[i], mu = E->mu[i];
    double ux = E->ux[i], uy = E->uy[i], uz = E->uz[i], relax = E->relax[i], lmu = E->lapmu[i];
    double e_dpx = E->gpc[0][i], e_dpy = E->gpc[1][i], e_dpz = E->gpc[2][i];
    double e_dcx = E->gcc[0][i], e_dcy = E->gcc[1][i], e_dcz = E->gcc[2][i];
    double drcx = drho*e_dcx*CS2, drcy = drho*e_dcy*CS2, drcz = drho*e_dcz*CS2;
    double bfx = rho*P->gx + P->fx*P->forcing_factor, bfy = rho*P->gy + P->fy*P->forcing_factor, bfz = rho*P->gz + P->fz*P->forcing_factor;
    double e_fgx = mu*e_dcx + bfx, e_fgy = mu*e_dcy + bfy, e_fgz = mu*e_dcz + bfz;
    double a_over_rho = ALPHA/rho;
    double e_fhx = e_dcx - c*a_over_rho*(e_dpx - e_fgx), e_fhy = e_dcy - c*a_over_rho*(e_dpy - e_fgy), e_fhz = e_dcz - c*a_over_rho*(e_dpz - e_fgz);
    double u_sq = ux*ux + uy*uy + uz*uz;
    double u_fh_e = ux*e_fhx + uy*e_fhy + uz*e_fhz;
    double u_drhocd = ux*drcx + uy*drcy + uz*drcz;
    double u_fg_e = ux*e_fgx + uy*e_fgy + uz*e_fgz;
    double f_dpx = E->gpm[0][i], f_dpy = E->gpm[1][i], f_dpz = E->gpm[2][i];
    double f_dcx = E->gcm[0][i], f_dcy = E->gcm[1][i], f_dcz = E->gcm[2][i];
    double drho_cs2 = drho*CS2;
    double drmx = drho_cs2*f_dcx, drmy = drho_cs2*f_dcy, drmz = drho_cs2*f_dcz;
    double c_over = c/CS2/rho;
    double f_fgx = mu*f_dcx + bfx, f_fgy = mu*f_dcy + bfy, f_fgz = mu*f_dcz + bfz;
    double f_fhx = f_dcx - c_over*(f_dpx - f_fgx), f_fhy = f_dcy - c_over*(f_dpy - f_fgy), f_fhz = f_dcz - c_over*(f_dpz - f_fgz);
    double u_fh_f = ux*f_fhx + uy*f_fhy + uz*f_fhz;
    double u_drm = ux*drmx + uy*drmy + uz*drmz;
    double u_fg_f = ux*f_fgx + uy*f_fgy + uz*f_fgz;
    double gout[19];
    for (int k = 0; k < Q; ++k) {
      double w = WK[k];
      double uk = ux*(double)CX[k] + uy*(double)CY[k] + uz*(double)CZ[k];
      double u_term = uk*(ALPHA + BETA_C*uk) - GAMMA_C*u_sq;
      double G = w*(1.0 + u_term);
      double f_proj = bfx*(double)CX[k] + bfy*(double)CY[k] + bfz*(double)CZ[k];
      double cd_c = cd_dir_val(E, E->c, i, k), cd_p = cd_dir_val(E, E->p, i, k);
      double drho_dk = drho*cd_c*CS2;
      double eh = c*G;
      eh -= 0.5*G*(cd_c - c*a_over_rho*(cd_p - mu*cd_c - f_proj));
      eh += 0.5*G*u_fh_e;
      double eg = w*(p + rho*CS2*u_term);
      eg -= 0.5*(drho_dk*(G - w) + (mu*cd_c + f_proj)*G);
      eg += 0.5*((G - w)*u_drhocd + G*u_fg_e);
      double dGk = G - w;
      double bd_c = bd_dir_val(E, E->c, c, i, k), bd_p = bd_dir_val(E, E->p, p, i, k);
      double dc_dk = 0.5*(cd_c + bd_c), dp_dk = 0.5*(cd_p + bd_p);
      double drho_dk2 = drho_cs2*dc_dk;
      double ff = mu*dc_dk + f_proj;
      double avgk = avg_dir_val(E, E->lapmu, lmu, i, k);
      double fh = G*(dc_dk - c_over*(dp_dk - ff) - u_fh_f + P->mobility*avgk);
      double fg = (drho_dk2 - u_drm)*dGk + (ff - u_fg_f)*G;
      double gk = E->g[(size_t)k*n + i];
      double gnew = gk + relax*(eg - gk) + fg;
      if (inplace) { E->h[(size_t)OPP[k]*n + i] = eh + fh; gout[k] = gnew; }
      else { E->h[(size_t)k*n + i] = eh + fh; E->g[(size_t)k*n + i] = gnew; }
    }
    if (inplace) for (int k = 0; k < Q; ++k) E->g[(size_t)OPP[k]*n + i] = gout[k];
  }
}

/* MRT relaxation of one site's raw collision term: device_engine.cuh:622-648.
 * In FP64 the real_t (mrt_fast) and double forms are the same arithmetic. */
static void mrt_relax(double *cg, double relax, double lambda_mrt)
{
  double s0 = relax; double tau = 1.0/s0 - 0.5;
  double s1 = 2.0*tau/(2.0*lambda_mrt + tau);
  double s[19] = { 0.0, s0, s0, 0.0, s1, 0.0, s1, 0.0, s1, s0, s0, s0, s0, s0, s0, s0, s1, s1, s1 };
  double xx[19];
  for (int k = 0; k < 19; ++k) xx[k] = 0.0;
  for (int nn = 0; nn < 19; ++nn) { double y = cg[nn]; for (int m = 1; m < 19; ++m) xx[m] += s[m]*MM[m][nn]*y; }
  for (int m = 0; m < 19; ++m) { double acc = 0.0; for (int nn = 0; nn < 19; ++nn) acc += MI[m][nn]*xx[nn]; cg[m] = acc; }
}

/* ---- fused collision: MRT device_engine.cuh:547-651.  Non-streamed sites are
 * skipped (the CPU semantics the unfused apply documents at :306-312; the
 * reference's fused MRT kernel instead zeroes h there). */
static void collide_fused_mrt(orc_engine *E)
{
  int n = E->n; const orc_params *P = &E->P;
  double drho = P->rho0 - P->rho1;
  int inplace = E->S.stream_inplace;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if (!E->L.is_streamed[i]) continue;
    double c = E->c[i], rho = E->rho[i], p = E->p[i], mu = E->mu[i];
    double ux = E->ux[i], uy = E->uy[i], uz = E->uz[i], lmu = E->lapmu[i];
    double drho_cs2 = drho*CS2, rho_cs2 = rho*CS2, c_over = c/rho_cs2;
    double bfx = rho*P->gx + P->fx*P->forcing_factor, bfy = rho*P->gy + P->fy*P->forcing_factor, bfz = rho*P->gz + P->fz*P->forcing_factor;
    double e_dpx = E->gpc[0][i], e_dpy = E->gpc[1][i], e_dpz = E->gpc[2][i];
    double e_dcx = E->gcc[0][i], e_dcy = E->gcc[1][i], e_dcz = E->gcc[2][i];
    double drcx = drho_cs2*e_dcx, drcy = drho_cs2*e_dcy, drcz = drho_cs2*e_dcz;
    double e_fgx = mu*e_dcx + bfx, e_fgy = mu*e_dcy + bfy, e_fgz = mu*e_dcz + bfz;
    double e_fhx = e_dcx - c_over*(e_dpx - e_fgx), e_fhy = e_dcy - c_over*(e_dpy - e_fgy), e_fhz = e_dcz - c_over*(e_dpz - e_fgz);
    double gamma_u_sq = GAMMA_C*(ux*ux + uy*uy + uz*uz);
    double u_term_h = ux*e_fhx + uy*e_fhy + uz*e_fhz;
    double u_force_g = ux*e_fgx + uy*e_fgy + uz*e_fgz;
    double u_drho = ux*drcx + uy*drcy + uz*drcz;
    double f_dpx = E->gpm[0][i], f_dpy = E->gpm[1][i], f_dpz = E->gpm[2][i];
    double f_dcx = E->gcm[0][i], f_dcy = E->gcm[1][i], f_dcz = E->gcm[2][i];
    double drmx = drho_cs2*f_dcx, drmy = drho_cs2*f_dcy, drmz = drho_cs2*f_dcz;
    double f_fgx = mu*f_dcx + bfx, f_fgy = mu*f_dcy + bfy, f_fgz = mu*f_dcz + bfz;
    double f_fhx = f_dcx - c_over*(f_dpx - f_fgx), f_fhy = f_dcy - c_over*(f_dpy - f_fgy), f_fhz = f_dcz - c_over*(f_dpz - f_fgz);
    double u_fh_f = ux*f_fhx + uy*f_fhy + uz*f_fhz;
    double u_drm = ux*drmx + uy*drmy + uz*drmz;
    double u_fg_f = ux*f_fgx + uy*f_fgy + uz*f_fgz;
    double cg[19], fgs[19], go[19];
    for (int k = 0; k < Q; ++k) {
      double w = WK[k];
      double uk = ux*(double)CX[k] + uy*(double)CY[k] + uz*(double)CZ[k];
      double u_term = uk*(ALPHA + BETA_C*uk) - gamma_u_sq;
      double G = w*(1.0 + u_term);
      double w_u_term = G - w;
      double f_proj = bfx*(double)CX[k] + bfy*(double)CY[k] + bfz*(double)CZ[k];
      double cd_c = cd_dir_val(E, E->c, i, k), cd_p = cd_dir_val(E, E->p, i, k);
      double drho_dk = drho_cs2*cd_c;
      double eh = c*G + 0.5*G*(u_term_h - cd_c + c_over*(cd_p - mu*cd_c - f_proj));
      double eg = w*p + rho_cs2*w_u_term + 0.5*((u_drho - drho_dk)*w_u_term + (u_force_g + mu*cd_c + f_proj)*G);
      double dGk = G - w;
      double bd_c = bd_dir_val(E, E->c, c, i, k), bd_p = bd_dir_val(E, E->p, p, i, k);
      double dc_dk = 0.5*(cd_c + bd_c), dp_dk = 0.5*(cd_p + bd_p);
      double drho_dk2 = drho_cs2*dc_dk;
      double ff = mu*dc_dk + f_proj;
      double avgk = avg_dir_val(E, E->lapmu, lmu, i, k);
      double fh = G*(dc_dk - c_over*(dp_dk - ff) - u_fh_f + P->mobility*avgk);
      double fg = (drho_dk2 - u_drm)*dGk + (ff - u_fg_f)*G;
      double gk = E->g[(size_t)k*n + i];
      int wk = inplace ? OPP[k] : k;
      E->h[(size_t)wk*n + i] = eh + fh;
      cg[k] = eg - gk; fgs[k] = fg; go[k] = gk;
    }
    mrt_relax(cg, E->relax[i], P->mrt_lambda);
    if (inplace) for (int k = 0; k < Q; ++k) E->g[(size_t)OPP[k]*n + i] = go[k] + cg[k] + fgs[k];
    else         for (int k = 0; k < Q; ++k) E->g[(size_t)k*n + i] = go[k] + cg[k] + fgs[k];
  }
}

/* ---- unfused chain (multiphase_gpu.cuh:206-253): dir buffers, force,
 * equilibria, collision term, apply.  Cross-checks the fused kernels. ------- */
static void dir_buffers(orc_engine *E)
{
  int n = E->n;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    for (int m = 0; m < Q; ++m) {      /* k_cd_dir_mf / k_bd_dir_mf / k_avg_dir_mf: :950-989 */
      size_t j = (size_t)m*n + i;
      E->gc_cd[j] = cd_dir_val(E, E->c, i, m); E->gp_cd[j] = cd_dir_val(E, E->p, i, m);
      E->gc_bd[j] = bd_dir_val(E, E->c, E->c[i], i, m); E->gp_bd[j] = bd_dir_val(E, E->p, E->p[i], i, m);
      E->avg[j] = avg_dir_val(E, E->lapmu, E->lapmu[i], i, m);
    }
  }
}

/* k_force_term: device_engine.cuh:255-296 */
static void force_term(orc_engine *E)
{
  int n = E->n; const orc_params *P = &E->P; double drho = P->rho0 - P->rho1;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if (!E->L.is_streamed[i]) { for (int k = 0; k < Q; ++k) { E->fh[(size_t)k*n + i] = 0; E->fg[(size_t)k*n + i] = 0; } continue; }
    double c = E->c[i], rho = E->rho[i], mu = E->mu[i];
    double ux = E->ux[i], uy = E->uy[i], uz = E->uz[i];
    double dpx = E->gpm[0][i], dpy = E->gpm[1][i], dpz = E->gpm[2][i];
    double dcx = E->gcm[0][i], dcy = E->gcm[1][i], dcz = E->gcm[2][i];
    double drho_cs2 = drho*CS2;
    double drmx = drho_cs2*dcx, drmy = drho_cs2*dcy, drmz = drho_cs2*dcz;
    double bfx = rho*P->gx + P->fx*P->forcing_factor, bfy = rho*P->gy + P->fy*P->forcing_factor, bfz = rho*P->gz + P->fz*P->forcing_factor;
    double c_over = c/CS2/rho;
    double fgx = mu*dcx + bfx, fgy = mu*dcy + bfy, fgz = mu*dcz + bfz;
    double fhx = dcx - c_over*(dpx - fgx), fhy = dcy - c_over*(dpy - fgy), fhz = dcz - c_over*(dpz - fgz);
    double u_fh = ux*fhx + uy*fhy + uz*fhz, u_drm = ux*drmx + uy*drmy + uz*drmz, u_fg = ux*fgx + uy*fgy + uz*fgz;
    double u_sq = ux*ux + uy*uy + uz*uz;
    for (int k = 0; k < Q; ++k) {
      size_t j = (size_t)k*n + i;
      double w = WK[k];
      double uk = ux*(double)CX[k] + uy*(double)CY[k] + uz*(double)CZ[k];
      double u_term = uk*(ALPHA + BETA_C*uk) - GAMMA_C*u_sq;
      double Gk = w*(1.0 + u_term), dGk = Gk - w;
      double dc_dk = 0.5*(E->gc_cd[j] + E->gc_bd[j]), dp_dk = 0.5*(E->gp_cd[j] + E->gp_bd[j]);
      double drho_dk = drho_cs2*dc_dk;
      double fproj = bfx*(double)CX[k] + bfy*(double)CY[k] + bfz*(double)CZ[k];
      double ff = mu*dc_dk + fproj;
      E->fh[j] = Gk*(dc_dk - c_over*(dp_dk - ff) - u_fh + P->mobility*E->avg[j]);
      E->fg[j] = (drho_dk - u_drm)*dGk + (ff - u_fg)*Gk;
    }
  }
}

/* k_equilibria (BGK, :199-240) and k_equilibria_mrt (:1016-1054) */
static void equilibria(orc_engine *E, int mrt)
{
  int n = E->n; const orc_params *P = &E->P; double drho = P->rho0 - P->rho1;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if (!E->L.is_streamed[i]) { for (int k = 0; k < Q; ++k) { E->eqh[(size_t)k*n + i] = 0; E->eqg[(size_t)k*n + i] = 0; } continue; }
    double c = E->c[i], rho = E->rho[i], p = E->p[i], mu = E->mu[i];
    double ux = E->ux[i], uy = E->uy[i], uz = E->uz[i];
    double dpx = E->gpc[0][i], dpy = E->gpc[1][i], dpz = E->gpc[2][i];
    double dcx = E->gcc[0][i], dcy = E->gcc[1][i], dcz = E->gcc[2][i];
    double bfx = rho*P->gx + P->fx*P->forcing_factor, bfy = rho*P->gy + P->fy*P->forcing_factor, bfz = rho*P->gz + P->fz*P->forcing_factor;
    double fgx = mu*dcx + bfx, fgy = mu*dcy + bfy, fgz = mu*dcz + bfz;
    if (!mrt) {
      double drcx = drho*dcx*CS2, drcy = drho*dcy*CS2, drcz = drho*dcz*CS2;
      double a_over_rho = ALPHA/rho;
      double fhx = dcx - c*a_over_rho*(dpx - fgx), fhy = dcy - c*a_over_rho*(dpy - fgy), fhz = dcz - c*a_over_rho*(dpz - fgz);
      double u_sq = ux*ux + uy*uy + uz*uz;
      for (int k = 0; k < Q; ++k) {
        size_t j = (size_t)k*n + i;
        double w = WK[k];
        double uk = ux*(double)CX[k] + uy*(double)CY[k] + uz*(double)CZ[k];
        double u_term = uk*(ALPHA + BETA_C*uk) - GAMMA_C*u_sq;
        double G = w*(1.0 + u_term);
        double dc_dk = E->gc_cd[j], dp_dk = E->gp_cd[j];
        double drho_dk = drho*dc_dk*CS2;
        double f_proj = bfx*(double)CX[k] + bfy*(double)CY[k] + bfz*(double)CZ[k];
        double u_fh = ux*fhx + uy*fhy + uz*fhz, u_drhocd = ux*drcx + uy*drcy + uz*drcz, u_fg = ux*fgx + uy*fgy + uz*fgz;
        double eh = c*G;
        eh -= 0.5*G*(dc_dk - c*a_over_rho*(dp_dk - mu*dc_dk - f_proj));
        eh += 0.5*G*u_fh;
        E->eqh[j] = eh;
        double eg = w*(p + rho*CS2*u_term);
        eg -= 0.5*(drho_dk*(G - w) + (mu*dc_dk + f_proj)*G);
        eg += 0.5*((G - w)*u_drhocd + G*u_fg);
        E->eqg[j] = eg;
      }
    } else {
      double drho_cs2 = drho*CS2;
      double drcx = drho_cs2*dcx, drcy = drho_cs2*dcy, drcz = drho_cs2*dcz;
      double rho_cs2 = rho*CS2, c_over = c/rho_cs2;
      double fhx = dcx - c_over*(dpx - fgx), fhy = dcy - c_over*(dpy - fgy), fhz = dcz - c_over*(dpz - fgz);
      double gamma_u_sq = GAMMA_C*(ux*ux + uy*uy + uz*uz);
      double u_term_h = ux*fhx + uy*fhy + uz*fhz, u_force_g = ux*fgx + uy*fgy + uz*fgz, u_drho = ux*drcx + uy*drcy + uz*drcz;
      for (int k = 0; k < Q; ++k) {
        size_t j = (size_t)k*n + i;
        double w = WK[k];
        double uk = ux*(double)CX[k] + uy*(double)CY[k] + uz*(double)CZ[k];
        double w_u_term = w*(uk*(ALPHA + BETA_C*uk) - gamma_u_sq);
        double G = w + w_u_term;
        double dc_dk = E->gc_cd[j], dp_dk = E->gp_cd[j];
        double drho_dk = drho_cs2*dc_dk;
        double f_proj = bfx*(double)CX[k] + bfy*(double)CY[k] + bfz*(double)CZ[k];
        E->eqh[j] = c*G + 0.5*G*(u_term_h - dc_dk + c_over*(dp_dk - mu*dc_dk - f_proj));
        E->eqg[j] = w*p + rho_cs2*w_u_term + 0.5*((u_drho - drho_dk)*w_u_term + (u_force_g + mu*dc_dk + f_proj)*G);
      }
    }
  }
}

/* k_collision_term (:245-252), k_collision_term_raw + k_mrt_relax_g
 * (:1057-1101), k_collide_apply (:299-315) */
static void collide_unfused(orc_engine *E)
{
  int n = E->n; int mrt = E->S.use_mrt;
  dir_buffers(E);
  force_term(E);
  equilibria(E, mrt);
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double cg[19];
    for (int k = 0; k < Q; ++k) {
      size_t j = (size_t)k*n + i;
      E->collh[j] = E->eqh[j] - E->h[j];
      cg[k] = mrt ? (E->eqg[j] - E->g[j]) : E->relax[i]*(E->eqg[j] - E->g[j]);
    }
    if (mrt) mrt_relax(cg, E->relax[i], E->P.mrt_lambda);
    for (int k = 0; k < Q; ++k) E->collg[(size_t)k*n + i] = cg[k];
    if (!E->L.is_streamed[i]) continue;
    for (int k = 0; k < Q; ++k) {
      size_t j = (size_t)k*n + i;
      E->h[j] += E->collh[j] + E->fh[j];
      E->g[j] += E->collg[j] + E->fg[j];
    }
  }
}

/* k_stream_gather (:882-889) and k_stream_inplace (:859-873) */
static void stream(orc_engine *E)
{
  int n = E->n; int64_t Vn = (int64_t)Q*n;
  if (E->S.stream_inplace) {
    for (int64_t q = 0; q < E->n_ops; ++q) {
      int a = E->oa[q], b = E->ob[q];
      if (E->ot[q] == 0) { double t = E->h[a]; E->h[a] = E->h[b]; E->h[b] = t; t = E->g[a]; E->g[a] = E->g[b]; E->g[b] = t; }
      else if (E->ot[q] == 1) { double m = 0.5*(E->h[a] + E->h[b]); E->h[a] = m; E->h[b] = m; m = 0.5*(E->g[a] + E->g[b]); E->g[a] = m; E->g[b] = m; }
      else { E->h[a] = 0; E->g[a] = 0; }
    }
    return;
  }
  #pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < Vn; ++j) {
    int s = E->src[j];
    if (s >= 0) { E->h2[j] = E->h[s]; E->g2[j] = E->g[s]; }
    else if (s == -1) {
      int k = (int)(j / n), i = (int)(j - (int64_t)k*n); int64_t o = (int64_t)OPP[k]*n + i;
      E->h2[j] = 0.5*(E->h[j] + E->h[o]); E->g2[j] = 0.5*(E->g[j] + E->g[o]);
    } else { E->h2[j] = 0; E->g2[j] = 0; }
  }
  double *t = E->h; E->h = E->h2; E->h2 = t; t = E->g; E->g = E->g2; E->g2 = t;
}

/* the field pass shared by step and download (multiphase_gpu.cuh:525-541) */
static void field_pass(orc_engine *E, int full)
{
  int n = E->n; double kappa = E->P.kappa;
  k_moments(E);
  enforce_open_bnd_fields(E);
  grad_cd(E, E->c, E->gcc[0], E->gcc[1], E->gcc[2]);
  if (full) { grad_bd(E, E->c, E->gcb[0], E->gcb[1], E->gcb[2]); grad_md(n, E->gcc, E->gcb, E->gcm); }
  /* k_pack_lap_c + lap + k_mu_axpy */
  for (int i = 0; i < n; ++i) { double cv = E->c[i]; E->xtnd[i] = cv; E->xtnd[n + i] = E->P.bnd_coeff*cv*(1.0 - cv); }
  lap(E, E->xtnd, E->lapc);
  for (int i = 0; i < n; ++i) E->mu[i] += -kappa*E->lapc[i];
  if (full) {
    for (int i = 0; i < n; ++i) { E->xtnd[i] = E->mu[i]; E->xtnd[n + i] = 0.0; }
    lap(E, E->xtnd, E->lapmu);
  }
  vel_press_corr(E);
  if (full) {
    grad_cd(E, E->p, E->gpc[0], E->gpc[1], E->gpc[2]);
    grad_bd(E, E->p, E->gpb[0], E->gpb[1], E->gpb[2]);
    grad_md(n, E->gpc, E->gpb, E->gpm);
  }
}

/* MultiPhaseGPU::step, multiphase_gpu.cuh:522-631 */
void orc_step(orc_engine *E)
{
  field_pass(E, 1);
  apply_open_bnd(E);
  if (E->S.fuse_collision) { if (E->S.use_mrt) collide_fused_mrt(E); else collide_fused_bgk(E); }
  else collide_unfused(E);
  stream(E);
  E->steps_done++;
}

/* ---- mass correction: multiphase_gpu.cuh:634-700, device_engine.cuh:1109-1139 */
void orc_mass_weight(orc_engine *E, double *M, double *W)
{
  int n = E->n;
  fxacc am = {0, 0, 0}, aw = {0, 0, 0};
  #pragma omp parallel
  {
    fxacc lm = {0, 0, 0}, lw = {0, 0, 0};
    #pragma omp for schedule(static)
    for (int i = 0; i < n; ++i) {
      if (!E->L.is_streamed[i]) continue;
      double c = 0.0;
      for (int k = 0; k < Q; ++k) c += E->h[(size_t)k*n + i];
      double phi = c*(1.0 - c); if (phi < 0.0) phi = 0.0;
      fx_add(&lm, c); fx_add(&lw, phi);
    }
    #pragma omp critical
    { fx_merge(&am, &lm); fx_merge(&aw, &lw); }
  }
  *M = fx_value(&am); *W = fx_value(&aw);
}

/* current total mass of the effective state (a pending folded correction
 * counts as applied) */
double orc_current_mass(orc_engine *E)
{
  double M, W;
  orc_mass_weight(E, &M, &W);
  if (E->S.mass_mode == 2 && E->pending) {
    double drift = M - E->target_mass;
    if (W > 0.0 && drift != 0.0) {
      double lambda = drift / W; int n = E->n;
      fxacc am = {0, 0, 0};
      for (int i = 0; i < n; ++i) {
        if (!E->L.is_streamed[i]) continue;
        double c = 0.0;
        for (int k = 0; k < Q; ++k) c += E->h[(size_t)k*n + i];
        double phi = c*(1.0 - c);
        if (phi > 0.0) c = c + (-lambda*phi);
        fx_add(&am, c);
      }
      M = fx_value(&am);
    }
  }
  return M;
}

void orc_record_target_mass(orc_engine *E) { E->target_mass = orc_current_mass(E); }

static void inject_mass(orc_engine *E, double lambda)
{
  int n = E->n;
  #pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    if (!E->L.is_streamed[i]) continue;
    double c = 0.0; for (int k = 0; k < Q; ++k) c += E->h[(size_t)k*n + i];
    double phi = c*(1.0 - c);
    if (phi <= 0.0) continue;
    double delta_c = -lambda*phi;
    for (int k = 0; k < Q; ++k) E->h[(size_t)k*n + i] += WK[k]*delta_c;
  }
}

/* apply_mass_correction (multiphase_gpu.cuh:689-700).  mass_mode 1 injects now
 * as the reference does; mode 2 defers the same correction into the next
 * moments (and into download_state). */
double orc_apply_mass_correction(orc_engine *E)
{
  if (E->S.mass_mode == 0) return 0.0;
  if (E->S.mass_mode == 2) { E->pending = 1; return 0.0; }
  double M, W; orc_mass_weight(E, &M, &W);
  double drift = M - E->target_mass;
  if (W > 0.0 && drift != 0.0) inject_mass(E, drift / W);
  E->last_drift = drift;
  return drift;
}

void orc_download_state(orc_engine *E, double *h, double *g)
{
  if (E->S.mass_mode == 2 && E->pending) {
    double M, W; orc_mass_weight(E, &M, &W);
    double drift = M - E->target_mass;
    size_t Vn = (size_t)E->Vn;
    double *hs = malloc(sizeof(double)*Vn); memcpy(hs, E->h, sizeof(double)*Vn);
    if (W > 0.0 && drift != 0.0) inject_mass(E, drift / W);
    memcpy(h, E->h, sizeof(double)*Vn);
    memcpy(E->h, hs, sizeof(double)*Vn); free(hs);
  } else memcpy(h, E->h, sizeof(double)*(size_t)E->Vn);
  memcpy(g, E->g, sizeof(double)*(size_t)E->Vn);
}

/* driver loop body: felbm_gpu_main.cu:742-745 */
void orc_run(orc_engine *E, int nsteps)
{
  for (int t = 0; t < nsteps; ++t) { orc_step(E); orc_apply_mass_correction(E); }
}

/* k_flow_stats (device_engine.cuh:660-683) over the last step's fields:
 * sum ux, uy, uz, u^2, c and max u^2; exact fixed-point sums. */
void orc_flow_stats(orc_engine *E, double out[6])
{
  int n = E->n;
  fxacc a[5]; memset(a, 0, sizeof(a)); double mx = 0.0;
  for (int i = 0; i < n; ++i) {
    double vx = E->ux[i], vy = E->uy[i], vz = E->uz[i];
    double v2 = vx*vx + vy*vy + vz*vz;
    fx_add(&a[0], vx); fx_add(&a[1], vy); fx_add(&a[2], vz); fx_add(&a[3], v2); fx_add(&a[4], E->c[i]);
    if (v2 > mx) mx = v2;
  }
  for (int q = 0; q < 5; ++q) out[q] = fx_value(&a[q]);
  out[5] = mx;
}

/* download_raw (multiphase_gpu.cuh:730-757): recompute the field pass on the
 * current state; out = rho, c, ux, uy, uz, p. */
void orc_download_fields(orc_engine *E, double *out)
{
  int n = E->n;
  int pend = E->pending;
  field_pass(E, 0);
  E->pending = pend;   /* a download does not consume the owed correction */
  memcpy(out + 0*(size_t)n, E->rho, sizeof(double)*n);
  memcpy(out + 1*(size_t)n, E->c, sizeof(double)*n);
  memcpy(out + 2*(size_t)n, E->ux, sizeof(double)*n);
  memcpy(out + 3*(size_t)n, E->uy, sizeof(double)*n);
  memcpy(out + 4*(size_t)n, E->uz, sizeof(double)*n);
  memcpy(out + 5*(size_t)n, E->p, sizeof(double)*n);
}

/* heap handles for foreign-function callers */
orc_engine *orc_new(int Nx, int Ny, int Nz, const uint8_t *solid, int hbb, int px, int py, int pz,
                    const orc_params *P, const orc_switches *S)
{
  orc_engine *E = (orc_engine*)calloc(1, sizeof(orc_engine));
  orc_engine_init(E, Nx, Ny, Nz, solid, hbb, px, py, pz, P, S);
  return E;
}
void orc_delete(orc_engine *E) { if (E) { orc_engine_free(E); free(E); } }
int orc_n(const orc_engine *E) { return E->n; }
int orc_inplace_active(const orc_engine *E) { return E->S.stream_inplace; }
void orc_set_forcing_factor(orc_engine *E, double ff) { E->P.forcing_factor = ff; }

/* forcing_factor_at of the driver loop, src/felbm_gpu_main.cu:587-594.
 * Pinned by closed forms (quarter-period values, signs, the constant default). */
double orc_forcing_factor_at(const char *timedep, double period, long long t)
{
  const double PI = 3.14159265358979323846;
  if (timedep && strstr(timedep, "sinusoidal")) return sin(2.0*PI*(double)t/period);
  if (timedep && strstr(timedep, "square")) {
    double x = sin(2.0*PI*(double)t/period);
    return (x < 0) ? -1.0 : ((x > 0) ? 1.0 : 0.0);
  }
  return 1.0;
}
double orc_get_target_mass(const orc_engine *E) { return E->target_mass; }
void orc_set_target_mass(orc_engine *E, double m) { E->target_mass = m; }
void orc_get_coords(const orc_engine *E, int32_t *xyz)
{
  for (int i = 0; i < E->n; ++i) { xyz[3*i] = E->L.cx[i]; xyz[3*i+1] = E->L.cy[i]; xyz[3*i+2] = E->L.cz[i]; }
}
