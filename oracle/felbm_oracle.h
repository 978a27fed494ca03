/*
 * felbm_oracle.h -- FP64 CPU restatement of the felbm Lee-Liu D3Q19 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the HIP
 * engine and the CPU baseline timed beside it.  Only tests/, the smoke() entry
 * point and bench.py's cpu_baseline leg may load it; the product path never
 * does.
 *
 * It restates, in the reference's own "k*n+i" SoA layout and with the same
 * arithmetic expression order:
 *   include/felbm_gpu/device_engine.cuh   (all kernels of the step)
 *   include/felbm_gpu/multiphase_gpu.cuh  (table builders, step order,
 *                                          mass correction, stats, download)
 *   include/felbm_gpu/d3q19.cuh           (velocity set, constants)
 * Every function cites the reference lines it follows.
 *
 * What the reference takes from the absent felbm_local tree is defined here
 * explicitly (see orc_lattice below): compressed fluid indexing, is_known,
 * is_streamed, the streaming source table and the MRT moment basis.
 *
 * Parity status per function is recorded in the comment above it; functions
 * with no independent check say "parity unpinned".
 */
#ifndef FELBM_ORACLE_H
#define FELBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_Q 19

/* ---- lattice: the SubDomain semantics this engine defines ---------------
 * Sites are the fluid voxels of an Nx*Ny*Nz box, numbered in x-fastest
 * lexicographic order (voxel v = x + Nx*(y + Ny*z)).  Solid voxels are not
 * sites; coords_to_idx of a solid voxel is the sentinel -1 and is_solid(-1)
 * is true.  is_known(i,m) <=> the voxel at x - e_m (periodic wrap on periodic
 * axes) is a fluid site.  Off a non-periodic axis a voxel does not exist
 * (treated like solid).  is_streamed(i) is 1 for every site unless an open
 * boundary marks it otherwise.                                               */
typedef struct {
  int Nx, Ny, Nz;
  int periodic[3];
  int hbb;            /* use_halfway_bb */
  int n;              /* number of sites */
  int32_t *vox2idx;   /* [Nx*Ny*Nz] site index or -1 */
  int32_t *cx, *cy, *cz; /* [n] coordinates */
  uint8_t *is_streamed;  /* [n] */
} orc_lattice;

/* Per-run scalars; mirrors DevParams (device_engine.cuh:31-59). */
typedef struct {
  double rho0, rho1, kappa, beta, bnd_coeff, tau0, tau1, mobility;
  double forcing_factor;
  double gx, gy, gz, fx, fy, fz;
  double mrt_lambda;
  /* open boundary (device_engine.cuh:45-58) */
  int open_bnd, n_inlet, n_outlet, n_buf, copy_to_buf;
  int use_in_vel, use_in_prs, use_in_fluid, use_out_prs, use_out_fluid;
  double u_in_x, u_in_y, u_in_z, p_in, p_out, c_out_fixed, rho_out_fixed;
  int inlet_mode, split_axis;
  double inlet_c_a, inlet_period, inlet_duty, inlet_ramp, split_pos;
} orc_params;

/* switches (felbm_gpu_main.cu:210-222) */
typedef struct {
  int use_mrt;
  int fuse_collision;   /* 1: k_collide_fused_*; 0: unfused equilibria/force/apply chain */
  int stream_inplace;   /* 1: reversed writes + disjoint op program; 0: gather ping-pong */
  int mass_mode;        /* 0 off, 1 reference injection, 2 folded into the next moments */
} orc_switches;

typedef struct {
  orc_lattice L;
  orc_params P;
  orc_switches S;
  int n, Vn;
  /* state (k*n+i) */
  double *h, *g, *h2, *g2;
  /* fields */
  double *c, *p, *rho, *mu, *ux, *uy, *uz, *relax, *lapc, *lapmu, *xtnd;
  double *gcc[3], *gcb[3], *gcm[3], *gpc[3], *gpb[3], *gpm[3];
  /* unfused temporaries (Q*n) */
  double *gc_cd, *gp_cd, *gc_bd, *gp_bd, *avg, *eqh, *eqg, *collh, *collg, *fh, *fg;
  /* matrix-free tables in the reference format */
  int32_t *cdm, *cdp, *bcase, *bc1, *bc2, *avgnext, *lcase, *lc1, *lc2, *src;
  /* in-place op program */
  int64_t n_ops; int32_t *oa, *ob; uint8_t *ot;
  int inplace_ok;
  /* open boundary lists */
  int32_t *inlet, *outlet, *buf_verts, *buf_refs; double *inlet_coord;
  uint8_t *bcproc;
  unsigned open_step;
  /* mass correction */
  double target_mass;
  int pending;          /* folded mode: a correction is owed by the next moments */
  double last_drift;
  int64_t steps_done;
} orc_engine;

/* D3Q19 + MRT constants */
void orc_d3q19(int *cx, int *cy, int *cz, int *opp, double *w);
void orc_mrt_matrices(double *M, double *Minv);

/* lattice + engine lifecycle */
int  orc_lattice_init(orc_lattice *L, int Nx, int Ny, int Nz, const uint8_t *solid,
                      int hbb, int px, int py, int pz);
void orc_lattice_free(orc_lattice *L);
int  orc_engine_init(orc_engine *E, int Nx, int Ny, int Nz, const uint8_t *solid,
                     int hbb, int px, int py, int pz,
                     const orc_params *P, const orc_switches *S);
void orc_engine_free(orc_engine *E);

/* state I/O (multiphase_gpu.cuh:354-359, 449-455) */
void orc_upload_state(orc_engine *E, const double *h, const double *g);
void orc_download_state(orc_engine *E, double *h, double *g);
void orc_init_rest(orc_engine *E, const double *c_site, double p0);

/* one reference step (multiphase_gpu.cuh:522-631) and the mass corrector */
void   orc_step(orc_engine *E);
void   orc_run(orc_engine *E, int nsteps);  /* step + apply_mass_correction, as the driver loop */
void   orc_record_target_mass(orc_engine *E);
double orc_current_mass(orc_engine *E);
double orc_apply_mass_correction(orc_engine *E);
void   orc_mass_weight(orc_engine *E, double *M, double *W);

/* monitoring (device_engine.cuh:660-683) over the fields of the last step */
void orc_flow_stats(orc_engine *E, double out[6]);
/* download_raw (multiphase_gpu.cuh:730-757): out = rho,c,ux,uy,uz,p (6*n) */
void orc_download_fields(orc_engine *E, double *out);

/* open boundary setup (multiphase_gpu.cuh:369-400) */
int orc_set_open_bnd(orc_engine *E, const int32_t *inlet, int n_in,
                     const int32_t *outlet, int n_out, const double *icoord,
                     const int32_t *buf_verts, const int32_t *buf_refs, int n_buf,
                     int copy_to_buffers);

/* tables in the reference layout, for bitwise comparison */
void orc_get_tables(const orc_engine *E, int32_t *cdm, int32_t *cdp, int32_t *bcase,
                    int32_t *bc1, int32_t *bc2, int32_t *avgnext, int32_t *lcase,
                    int32_t *lc1, int32_t *lc2, int32_t *src);
int64_t orc_num_ops(const orc_engine *E);
int  orc_verify_inplace(const orc_engine *E);

/* field pointers, for tests */
double *orc_field(orc_engine *E, const char *name);

/* exact, order-independent fixed-point sum (shared with the engine's definition) */
double orc_fixed_sum(const double *v, int64_t n);

#ifdef __cplusplus
}
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif
orc_engine *orc_new(int Nx, int Ny, int Nz, const uint8_t *solid, int hbb, int px, int py, int pz,
                    const orc_params *P, const orc_switches *S);
void orc_delete(orc_engine *E);
int orc_n(const orc_engine *E);
int orc_inplace_active(const orc_engine *E);
void orc_set_forcing_factor(orc_engine *E, double ff);
double orc_forcing_factor_at(const char *timedep, double period, long long t);
double orc_get_target_mass(const orc_engine *E);
void orc_set_target_mass(orc_engine *E, double m);
void orc_get_coords(const orc_engine *E, int32_t *xyz);
#ifdef __cplusplus
}
#endif
