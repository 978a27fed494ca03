/*
 * felbm_mi355x.h -- C ABI of the MI355X engine for the felbm Lee-Liu
 * free-energy multiphase D3Q19 step.
 *
 * The engine is the drop-in replacement for felbm_gpu::MultiPhaseGPU
 * (include/felbm_gpu/multiphase_gpu.cuh:23-783 of the reference), the GPU
 * peer of the CPU lbm::EngineMultiPhase.  Each declaration below names the
 * reference member it replaces.  All calls return FELBM_OK (0) or a negative
 * error code; felbm_last_error() describes the last failure of the calling
 * thread.  Nothing here exits the process (the reference's GPU_CHECK calls
 * exit(EXIT_FAILURE), gpu_common.cuh:14-22).
 *
 * One handle drives one GPU and holds one or more independent lattice
 * instances ("slots").  felbm_step advances every slot in the same launches,
 * which is both the plain engine (one slot) and the batched form (many).
 * A handle is not thread-safe: drive it from one host thread.
 *
 * Layouts at the boundary are the reference's: distributions are double,
 * population k of site i at [k*n + i] (Vn = 19*n); sites are the fluid voxels,
 * numbered as the caller's site list says (felbm_geometry.site_xyz) or else
 * in x-fastest lexicographic order; masks are uint8 [nz][ny][nx] with
 * 1 = solid.
 */
#ifndef FELBM_MI355X_H
#define FELBM_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FELBM_OK           0
#define FELBM_EINVAL      (-1)
#define FELBM_ENOMEM      (-2)
#define FELBM_EHIP        (-3)
#define FELBM_ESTATE      (-4)
#define FELBM_EUNSUPPORTED (-5)
#define FELBM_ENONFINITE  (-6)

typedef struct felbm_engine felbm_engine;

/* Geometry of one instance (replaces lbm::SubDomain, multiphase_gpu.cuh:75-79).
 * site_xyz (optional) is the caller's numbering of the sites: site i is the
 * voxel (site_xyz[3i], site_xyz[3i+1], site_xyz[3i+2]), as SubDomain's
 * idx_to_coords gives it (src/felbm_gpu_main.cu:124-143); n_sites entries,
 * each a fluid voxel of `solid`, each listed once.  Every per-site array that
 * crosses the boundary (state, fields, rest c, open-boundary node lists,
 * exported tables) is then in that numbering.  NULL: the fluid voxels in
 * x-fastest lexicographic order. */
typedef struct {
  int nx, ny, nz;
  int periodic[3];           /* 1 = periodic axis */
  int use_halfway_bb;        /* Settings::use_halfway_bb */
  const uint8_t *solid;      /* [nz][ny][nx], 1 = solid voxel */
  const int32_t *site_xyz;   /* [n_sites][3] (x, y, z) or NULL */
  int64_t n_sites;
} felbm_geometry;

/* Physics (ParametersMultiPhase + Settings fields read at
 * multiphase_gpu.cuh:331-343; derived values as compare_cpu_gpu.cu:111-118). */
typedef struct {
  double rho0, rho1;                 /* phase_density(0), phase_density(1) */
  double kappa, beta;
  double boundary_coefficient;
  double tau0, tau1;                 /* relaxation_time(0), relaxation_time(1) */
  double mobility;
  double forcing_factor;
  double acceleration[3];            /* Settings::acceleration */
  double forcing[3];                 /* Settings::forcing */
  double mrt_lambda;                 /* Settings::mrt_lambda */
} felbm_params;

/* Operator switches (felbm_gpu_main.cu:210-222).  The six fast-path switches
 * select algebraically equivalent paths in the reference; this engine always
 * runs its fused form and accepts them for configuration compatibility. */
typedef struct {
  int use_mrt;
  int correct_op_mass;
  int stream_matrix_free, grad_matrix_free, fused, fuse_collision;
  int mrt_fast_transform, stream_inplace;
} felbm_switches;

/* ---- lifecycle -------------------------------------------------------------
 * precision: 8 = FP64 distributions (reference double build),
 *            4 = FP32 distributions with FP64 reductions (-DFELBM_SINGLE). */
/* One engine per GPU holding any number of lattice instances (slots); the
 * reference constructs one MultiPhaseGPU per lattice (felbm_gpu_main.cu:314). */
int  felbm_create(felbm_engine **out, int device, int precision);
int  felbm_destroy(felbm_engine *e);                         /* MultiPhaseGPU::free, :759 */
/* MultiPhaseGPU::init (multiphase_gpu.cuh:75-351); returns the slot id.
 * One instance holds fewer than 2^32 bytes per distribution array
 * (19 * sites * sizeof(real)): at most about 56.5M FP32 or 28M FP64 sites;
 * a larger one returns FELBM_EUNSUPPORTED. */
int  felbm_add_instance(felbm_engine *e, const felbm_geometry *geom,
                        const felbm_params *prm, const felbm_switches *sw, int *slot);
/* retire one instance of the engine (no counterpart: the reference frees a
 * whole MultiPhaseGPU, free() :759) */
int  felbm_remove_instance(felbm_engine *e, int slot);
int  felbm_num_sites(felbm_engine *e, int slot, int64_t *n);  /* SubDomain::size_sites */
int  felbm_state_size(felbm_engine *e, int slot, int64_t *vn);/* state_size, :457 */
int  felbm_num_instances(felbm_engine *e, int *count);       /* no counterpart */

/* ---- state ---------------------------------------------------------------- */
int  felbm_upload_state(felbm_engine *e, int slot, const double *h, const double *g);   /* :354 */
int  felbm_download_state(felbm_engine *e, int slot, double *h, double *g);            /* :449 */
/* rest equilibrium h_k = w_k c, g_k = w_k p0 from a per-site c (n values),
   built on the device.  No MultiPhaseGPU member: the reference driver builds
   the initial populations on the host and passes them to upload_state
   (felbm_gpu_main.cu:322); for a rest initial condition this call gives the
   same state bit for bit with 8 bytes per site over the bus. */
int  felbm_init_rest(felbm_engine *e, int slot, const double *c_site, double p0);

/* ---- stepping -------------------------------------------------------------
 * step() (multiphase_gpu.cuh:522-631) followed by apply_mass_correction()
 * (:689-700) when correct_op_mass is on, nsteps times, for every slot.
 * stats_last != 0 accumulates the monitoring sums during the final step. */
int  felbm_step(felbm_engine *e, int nsteps, int stats_last);
/* P.forcing_factor (felbm_gpu_main.cu:743); stream-ordered, no host sync */
int  felbm_set_forcing_factor(felbm_engine *e, int slot, double ff);
/* forcing_factor_at(t) of the driver (felbm_gpu_main.cu:587-594): timedep
 * containing "sinusoidal" -> sin(2 pi t / period), "square" -> its sign,
 * anything else -> 1 */
int  felbm_forcing_factor_at(const char *timedep, double period, int64_t t, double *ff);
/* wait for the engine's stream (no counterpart: the reference's step()
 * blocks on its per-step mass-correction copy, multiphase_gpu.cuh:636-639) */
int  felbm_synchronize(felbm_engine *e);
/* capture `nsteps` (even) steps in a HIP graph and replay it `replays` times
 * (no counterpart: that per-step host copy keeps the reference's step out of
 * graph capture) */
int  felbm_step_graph(felbm_engine *e, int nsteps, int replays);

/* ---- mass correction (record_target_mass / get/set / current_mass) -------- */
int  felbm_record_target_mass(felbm_engine *e, int slot);               /* :685 */
int  felbm_get_target_mass(felbm_engine *e, int slot, double *m);       /* :458 */
int  felbm_set_target_mass(felbm_engine *e, int slot, double m);        /* :459 */
int  felbm_current_mass(felbm_engine *e, int slot, double *m);          /* :687 */

/* ---- monitoring: flow_stats (:645-683) ------------------------------------
 * out = { sum ux, sum uy, sum uz, sum u^2, sum c, max u^2 } over the fields
 * of the last step run with stats_last != 0; if none, over a fresh field
 * pass of the current state. */
int  felbm_flow_stats(felbm_engine *e, int slot, double out[6]);
int  felbm_flow_stats_launch(felbm_engine *e);                 /* async copy-back */
int  felbm_flow_stats_ready(felbm_engine *e, int *ready);
int  felbm_flow_stats_collect(felbm_engine *e, int slot, double out[6]);

/* ---- fields: download / download_raw (:704-757) ---------------------------
 * out = rho, c, ux, uy, uz, p, each n doubles, recomputed from the current
 * state exactly as the reference's download does. */
int  felbm_download_fields(felbm_engine *e, int slot, double *out);

/* ---- open boundaries: set_open_bnd (:369-400) + the DevParams open-BC
 * block (device_engine.cuh:45-58, filled at felbm_gpu_main.cu:340-364). */
typedef struct {
  int use_in_vel, use_in_prs, use_in_fluid, use_out_prs, use_out_fluid;
  double u_in[3];
  double p_in, p_out;
  double c_out_fixed, rho_out_fixed;
  double inlet_c_a;
  int inlet_mode;          /* 0 single, 1 alternate, 2 split, 3 stripes */
  int split_axis;
  double inlet_period, inlet_duty, inlet_ramp, split_pos;
} felbm_open_bnd;
int  felbm_set_open_bnd(felbm_engine *e, int slot, const felbm_open_bnd *ob,
                        const int32_t *inlet, int n_inlet,
                        const int32_t *outlet, int n_outlet,
                        const double *inlet_coord,
                        const int32_t *buf_verts, const int32_t *buf_refs, int n_buf,
                        int copy_to_buffers);

/* ---- tracers: ParticleManager construction, update() and output_state()
 * (felbm_gpu_main.cu:426-438, :752-777; README.md:421-464).  np tracers at
 * positions xyz[3*np] (voxel (x,y,z) at position (x,y,z)) with molecular
 * diffusivities dm[np]; each step advances them by `substeps` Euler substeps
 * on the step's velocity, refreshed every velocity_skip steps
 * (particles_velocity_skip), with a Brownian kick sqrt(2 Dm dt) per axis from
 * the counter-based stream `seed`.  np = 0 removes them. */
int  felbm_tracers_init(felbm_engine *e, int slot, int64_t np, const double *xyz, const double *dm,
                        uint64_t seed, int substeps, int velocity_skip);
/* positions [3*np], interpolated velocities [3*np] and per-tracer counts of
 * rejected (wall-blocked) substeps [np]; any pointer may be NULL */
int  felbm_tracers_download(felbm_engine *e, int slot, double *xyz, double *uvw, int32_t *blocked);
int  felbm_tracers_count(felbm_engine *e, int slot, int64_t *np);
/* Sets the slot's tracer step counter (0 after felbm_tracers_init), which
 * numbers the Brownian draws, phases the particles_velocity_skip refresh and
 * fills the step column of the stretching rows.  A restart from
 * particles_<step>.h5 (felbm_gpu_main.cu:566-575) sets it to the checkpoint
 * step, so the Dm > 0 tracers continue their random stream; the reference's
 * ParticleManager re-seeds instead (docs/HANDOFF_stretching.md:543-549). */
int  felbm_tracers_set_step(felbm_engine *e, int slot, uint64_t step);
/* Lagrangian stretching of the slot's tracers (particles_stretching,
 * docs/HANDOFF_stretching.md:165-210; rows as src/felbm_gpu_main.cu:643-671).
 * enable = 0 turns it off.  rho_hat [3*np] / log_rho [np] restore a saved state
 * (particles_<step>.h5); NULL gives isotropic unit vectors from orient_seed
 * (particles_stretching_seed) and log_rho = 0.  phase_split also snapshots c
 * for the wetting / non-wetting split.  One row per step is kept on the device
 * in a ring of row_capacity rows. */
int  felbm_tracers_stretching(felbm_engine *e, int slot, int enable, const double *rho_hat,
                              const double *log_rho, uint64_t orient_seed, int phase_split,
                              int row_capacity);
/* take_stretch_rows: the rows of the steps since the last call, oldest first,
 * 9 doubles each (step, lambda, mean_log_rho, var_log_rho, n_active, lambda_w,
 * lambda_nw, n_w, n_nw); at most max_rows; *lost counts rows the ring dropped */
int  felbm_tracers_stretch_rows(felbm_engine *e, int slot, double *rows, int max_rows, int *count,
                                int64_t *lost);
int  felbm_tracers_stretch_state(felbm_engine *e, int slot, double *rho_hat, double *log_rho);

/* ---- verification helpers -------------------------------------------------
 * Reference-format tables (multiphase_gpu.cuh:93-208) regenerated on the GPU
 * from the engine's own neighbour logic, for bitwise comparison. Any pointer
 * may be NULL. (Q-1)*n entries except avgnext and src (Q*n). */
int  felbm_export_tables(felbm_engine *e, int slot, int32_t *cdm, int32_t *cdp,
                         int32_t *bcase, int32_t *bc1, int32_t *bc2, int32_t *avgnext,
                         int32_t *lcase, int32_t *lc1, int32_t *lc2, int32_t *src);
/* felbm_export_tables without any device allocation (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc.  The tables
 * are staged in a static 32 MiB device workspace that every GPU the library
 * runs on holds until the process ends, so instances of at most 46,091 sites
 * (182 int32 per site) are accepted; larger ones return FELBM_EINVAL with a
 * message.  The workspace is shared by every engine on a GPU: calls that
 * overlap on one GPU (engines driven from two host threads) race. */
int  felbm_export_tables_noalloc(felbm_engine *e, int slot, int32_t *cdm, int32_t *cdp,
                                 int32_t *bcase, int32_t *bc1, int32_t *bc2, int32_t *avgnext,
                                 int32_t *lcase, int32_t *lc1, int32_t *lc2, int32_t *src);
/* the 19x19 MRT moment matrix and its inverse used by the engine (the
 * reference's constant tables, device_engine.cuh:65-81) */
int  felbm_mrt_matrices(double *M, double *Minv);
/* kernel timings of one step, in ms, averaged over nsteps: moments,
 * c-stencil, mu-stencil, collide (measurement aid, no counterpart) */
int  felbm_kernel_times(felbm_engine *e, int nsteps, float *ms4);

/* Tile-shape autotuner for the batched engine (no reference counterpart: the
 * reference launches fixed shapes).  Times each pass at 64, 128 and 256 sites
 * per workgroup over nsteps steps (after two warm-up steps per shape) on the
 * instances present, keeps the fastest shape per pass, and restores every
 * instance's state bit for bit, so the tuning steps leave no trace.  Refused
 * (FELBM_ESTATE) while tracers are present.  bs4: the chosen sites per
 * workgroup of moments, c-stencil, mu-stencil, collide; ms12: pass times in ms
 * as [shape 64, 128, 256][pass]; either may be NULL.  Results do not depend
 * on the shapes. */
int  felbm_autotune(felbm_engine *e, int nsteps, int *bs4, float *ms12);
/* set the sites per workgroup of the four passes (64, 128 or 256 each) */
int  felbm_set_block_sizes(felbm_engine *e, const int *bs4);

/* streaming-bandwidth probe: `reps` device-to-device copies of `bytes`;
 * gbs = (read + write) bytes / s.  Also the calibration kernel for the HBM
 * byte counters.  No reference counterpart (measurement aid). */
int  felbm_copy_probe(int device, int64_t bytes, int reps, double *gbs);

/* hand-written streaming probe: mode 0 reads `bytes`, 1 writes them, 2 copies
 * them (16 bytes per lane, non-temporal, grid-stride), `reps` timed launches;
 * gbs = bytes moved / s.  The achievable HBM ceiling the kernels are compared
 * with.  No reference counterpart (measurement aid). */
int  felbm_stream_probe(int device, int64_t bytes, int reps, int mode, double *gbs);

/* message of the last failing call on this thread (the reference exits the
 * process on any CUDA error, gpu_common.cuh:14-22) */
const char *felbm_last_error(void);
const char *felbm_version(void);                             /* no counterpart */

/* ---- geometry generators (DomainInitializer_* stand-ins) ------------------ */
int     felbm_geom_cylinders(int nx, int ny, int nz, int count, double rmin, double rmax,
                             double gap, uint64_t seed, int max_tries, uint8_t *solid);
int     felbm_geom_spheres(int nx, int ny, int nz, double rmin, double rmax, double porosity,
                           uint64_t seed, uint8_t *solid);
int64_t felbm_keep_largest_cluster(int nx, int ny, int nz, int px, int py, int pz,
                                   uint8_t *solid);

#ifdef __cplusplus
}
#endif
#endif
