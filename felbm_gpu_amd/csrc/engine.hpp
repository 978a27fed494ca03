// engine.hpp -- precision-independent engine interface behind the C ABI.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include "felbm_mi355x.h"

namespace fm {

void set_error(const std::string &msg);

// Static device workspace (capi.cpp) of the allocation-free table export:
// 32 MiB per GPU; scratch_init looks its address up once per device (the
// device current), scratch_ptr returns it (nullptr before set-up).
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxScratchDevices = 64;
int scratch_init(int device);
void *scratch_ptr(int device);

struct EngineBase {
  virtual ~EngineBase() {}
  virtual int add_instance(const felbm_geometry *g, const felbm_params *p, const felbm_switches *s, int *slot) = 0;
  virtual int remove_instance(int slot) = 0;
  virtual int num_sites(int slot, int64_t *n) = 0;
  virtual int num_instances(int *count) = 0;
  virtual int upload_state(int slot, const double *h, const double *g) = 0;
  virtual int download_state(int slot, double *h, double *g) = 0;
  virtual int init_rest(int slot, const double *c, double p0) = 0;
  virtual int step(int nsteps, int stats_last) = 0;
  virtual int step_graph(int nsteps, int replays) = 0;
  virtual int set_forcing_factor(int slot, double ff) = 0;
  virtual int synchronize() = 0;
  virtual int record_target_mass(int slot) = 0;
  virtual int get_target_mass(int slot, double *m) = 0;
  virtual int set_target_mass(int slot, double m) = 0;
  virtual int current_mass(int slot, double *m) = 0;
  virtual int flow_stats(int slot, double *out) = 0;
  virtual int flow_stats_launch() = 0;
  virtual int flow_stats_ready(int *ready) = 0;
  virtual int flow_stats_collect(int slot, double *out) = 0;
  virtual int download_fields(int slot, double *out) = 0;
  virtual int set_open_bnd(int slot, const felbm_open_bnd *ob, const int32_t *inlet, int n_inlet,
                           const int32_t *outlet, int n_outlet, const double *icoord,
                           const int32_t *bv, const int32_t *br, int n_buf, int copy_to_buffers) = 0;
  virtual int export_tables(int slot, int32_t **tabs) = 0;
  virtual int export_tables_noalloc(int slot, int32_t **tabs) = 0;
  virtual int kernel_times(int nsteps, float *ms4) = 0;
  virtual int autotune(int nsteps, int *bs4, float *ms12) = 0;
  virtual int set_block_sizes(const int *bs4) = 0;
  virtual int tracers_init(int slot, int64_t np, const double *xyz, const double *dm, uint64_t seed, int substeps,
                           int vskip) = 0;
  virtual int tracers_download(int slot, double *xyz, double *uvw, int32_t *blocked) = 0;
  virtual int tracers_count(int slot, int64_t *np) = 0;
  virtual int tracers_set_step(int slot, uint64_t step) = 0;
  virtual int tracers_stretching(int slot, int enable, const double *rho_hat, const double *log_rho,
                                 uint64_t orient_seed, int phase_split, int row_capacity) = 0;
  virtual int tracers_stretch_rows(int slot, double *rows, int max_rows, int *count, int64_t *lost) = 0;
  virtual int tracers_stretch_state(int slot, double *rho_hat, double *log_rho) = 0;
};

EngineBase *make_engine_f64(int device, int *err);
EngineBase *make_engine_f32(int device, int *err);

}  // namespace fm
