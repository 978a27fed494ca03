// capi.cpp -- the extern "C" boundary (include/felbm_mi355x.h).
#include <cmath>
#include <hip/hip_runtime.h>
#include <string>
#include <mutex>
#include <new>
#include "engine.hpp"
#include "kernels.hpp"

namespace fm {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// Static device workspace of felbm_export_tables_noalloc: the staging area its
// kernel writes the ten tables into before they are copied to the caller's
// host arrays.  Every GPU the library runs on holds its own copy, at its own
// address, until the process ends.  In host code `scratch` names a host-side
// stand-in, never the device memory: its address on each device comes from
// hipGetSymbolAddress, once per device, at engine creation (scratch_init).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxScratchDevices] = {};

int scratch_init(int device)
{
  if (device < 0 || device >= kMaxScratchDevices) { set_error("device index beyond the workspace table"); return FELBM_EUNSUPPORTED; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return FELBM_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) { set_error("workspace set-up: device not current"); return FELBM_EHIP; }
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("workspace set-up: hipGetSymbolAddress failed: ") + hipGetErrorName(e));
    return FELBM_EHIP;
  }
  g_scratch_ptr[device] = p;
  return FELBM_OK;
}

void *scratch_ptr(int device)
{
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}
}

struct felbm_engine { fm::EngineBase *impl; int precision; };

#define ENG(e) do { if (!(e) || !(e)->impl) { fm::set_error("null engine handle"); return FELBM_EINVAL; } } while (0)

extern "C" {

const char *felbm_last_error(void) { return fm::g_err.c_str(); }
const char *felbm_version(void) { return "felbm-mi355x 0.1 (gfx950)"; }

int felbm_create(felbm_engine **out, int device, int precision)
{
  if (!out) { fm::set_error("null output pointer"); return FELBM_EINVAL; }
  *out = nullptr;
  if (precision != 8 && precision != 4) { fm::set_error("precision must be 8 (FP64) or 4 (FP32)"); return FELBM_EINVAL; }
  int err = FELBM_OK;
  fm::EngineBase *impl = precision == 8 ? fm::make_engine_f64(device, &err) : fm::make_engine_f32(device, &err);
  if (!impl) return err ? err : FELBM_EHIP;
  felbm_engine *e = new (std::nothrow) felbm_engine{impl, precision};
  if (!e) { delete impl; fm::set_error("out of host memory"); return FELBM_ENOMEM; }
  *out = e;
  return FELBM_OK;
}

int felbm_destroy(felbm_engine *e)
{
  if (!e) return FELBM_OK;
  delete e->impl; delete e;
  return FELBM_OK;
}

int felbm_add_instance(felbm_engine *e, const felbm_geometry *g, const felbm_params *p, const felbm_switches *s, int *slot)
{ ENG(e); return e->impl->add_instance(g, p, s, slot); }
int felbm_remove_instance(felbm_engine *e, int slot) { ENG(e); return e->impl->remove_instance(slot); }
int felbm_num_sites(felbm_engine *e, int slot, int64_t *n) { ENG(e); if (!n) return FELBM_EINVAL; return e->impl->num_sites(slot, n); }
int felbm_state_size(felbm_engine *e, int slot, int64_t *vn)
{ ENG(e); if (!vn) return FELBM_EINVAL; int64_t n; int r = e->impl->num_sites(slot, &n); if (!r) *vn = 19 * n; return r; }
int felbm_num_instances(felbm_engine *e, int *count) { ENG(e); if (!count) return FELBM_EINVAL; return e->impl->num_instances(count); }
int felbm_upload_state(felbm_engine *e, int slot, const double *h, const double *g) { ENG(e); return e->impl->upload_state(slot, h, g); }
int felbm_download_state(felbm_engine *e, int slot, double *h, double *g)
{ ENG(e); if (!h || !g) { fm::set_error("null output"); return FELBM_EINVAL; } return e->impl->download_state(slot, h, g); }
int felbm_init_rest(felbm_engine *e, int slot, const double *c, double p0)
{ ENG(e); if (!c) { fm::set_error("null c"); return FELBM_EINVAL; } return e->impl->init_rest(slot, c, p0); }
int felbm_step(felbm_engine *e, int nsteps, int stats_last) { ENG(e); return e->impl->step(nsteps, stats_last); }
int felbm_step_graph(felbm_engine *e, int nsteps, int replays) { ENG(e); return e->impl->step_graph(nsteps, replays); }
int felbm_set_forcing_factor(felbm_engine *e, int slot, double ff) { ENG(e); return e->impl->set_forcing_factor(slot, ff); }

int felbm_tracers_init(felbm_engine *e, int slot, int64_t np, const double *xyz, const double *dm, uint64_t seed,
                       int substeps, int velocity_skip)
{ ENG(e); return e->impl->tracers_init(slot, np, xyz, dm, seed, substeps, velocity_skip); }
int felbm_tracers_download(felbm_engine *e, int slot, double *xyz, double *uvw, int32_t *blocked)
{ ENG(e); return e->impl->tracers_download(slot, xyz, uvw, blocked); }
int felbm_tracers_stretching(felbm_engine *e, int slot, int enable, const double *rho_hat, const double *log_rho,
                             uint64_t orient_seed, int phase_split, int row_capacity)
{ ENG(e); return e->impl->tracers_stretching(slot, enable, rho_hat, log_rho, orient_seed, phase_split, row_capacity); }
int felbm_tracers_stretch_rows(felbm_engine *e, int slot, double *rows, int max_rows, int *count, int64_t *lost)
{ ENG(e); return e->impl->tracers_stretch_rows(slot, rows, max_rows, count, lost); }
int felbm_tracers_stretch_state(felbm_engine *e, int slot, double *rho_hat, double *log_rho)
{ ENG(e); return e->impl->tracers_stretch_state(slot, rho_hat, log_rho); }
int felbm_tracers_count(felbm_engine *e, int slot, int64_t *np)
{ ENG(e); if (!np) { fm::set_error("null output pointer"); return FELBM_EINVAL; } return e->impl->tracers_count(slot, np); }
int felbm_tracers_set_step(felbm_engine *e, int slot, uint64_t step)
{ ENG(e); return e->impl->tracers_set_step(slot, step); }

// forcing_factor_at of the reference driver (src/felbm_gpu_main.cu:587-594):
// the time-dependence name is matched as a substring, anything else is constant
int felbm_forcing_factor_at(const char *timedep, double period, int64_t t, double *ff)
{
  if (!ff) { fm::set_error("null output pointer"); return FELBM_EINVAL; }
  const std::string s = timedep ? timedep : "";
  const double two_pi_t = 2.0 * 3.14159265358979323846 * (double)t;
  if (s.find("sinusoidal") != std::string::npos) { *ff = std::sin(two_pi_t / period); return FELBM_OK; }
  if (s.find("square") != std::string::npos) {
    const double x = std::sin(two_pi_t / period);
    *ff = (x < 0) ? -1.0 : ((x > 0) ? 1.0 : 0.0);
    return FELBM_OK;
  }
  *ff = 1.0;
  return FELBM_OK;
}
int felbm_synchronize(felbm_engine *e) { ENG(e); return e->impl->synchronize(); }
int felbm_record_target_mass(felbm_engine *e, int slot) { ENG(e); return e->impl->record_target_mass(slot); }
int felbm_get_target_mass(felbm_engine *e, int slot, double *m) { ENG(e); if (!m) return FELBM_EINVAL; return e->impl->get_target_mass(slot, m); }
int felbm_set_target_mass(felbm_engine *e, int slot, double m) { ENG(e); return e->impl->set_target_mass(slot, m); }
int felbm_current_mass(felbm_engine *e, int slot, double *m) { ENG(e); if (!m) return FELBM_EINVAL; return e->impl->current_mass(slot, m); }
int felbm_flow_stats(felbm_engine *e, int slot, double out[6]) { ENG(e); if (!out) return FELBM_EINVAL; return e->impl->flow_stats(slot, out); }
int felbm_flow_stats_launch(felbm_engine *e) { ENG(e); return e->impl->flow_stats_launch(); }
int felbm_flow_stats_ready(felbm_engine *e, int *ready) { ENG(e); if (!ready) return FELBM_EINVAL; return e->impl->flow_stats_ready(ready); }
int felbm_flow_stats_collect(felbm_engine *e, int slot, double out[6]) { ENG(e); if (!out) return FELBM_EINVAL; return e->impl->flow_stats_collect(slot, out); }
int felbm_download_fields(felbm_engine *e, int slot, double *out) { ENG(e); if (!out) return FELBM_EINVAL; return e->impl->download_fields(slot, out); }
int felbm_set_open_bnd(felbm_engine *e, int slot, const felbm_open_bnd *ob, const int32_t *inlet, int n_inlet,
                       const int32_t *outlet, int n_outlet, const double *icoord,
                       const int32_t *bv, const int32_t *br, int n_buf, int copy_to_buffers)
{ ENG(e); return e->impl->set_open_bnd(slot, ob, inlet, n_inlet, outlet, n_outlet, icoord, bv, br, n_buf, copy_to_buffers); }
int felbm_export_tables(felbm_engine *e, int slot, int32_t *cdm, int32_t *cdp, int32_t *bcase, int32_t *bc1,
                        int32_t *bc2, int32_t *avgnext, int32_t *lcase, int32_t *lc1, int32_t *lc2, int32_t *src)
{
  ENG(e);
  int32_t *t[10] = { cdm, cdp, bcase, bc1, bc2, avgnext, lcase, lc1, lc2, src };
  return e->impl->export_tables(slot, t);
}
int felbm_export_tables_noalloc(felbm_engine *e, int slot, int32_t *cdm, int32_t *cdp, int32_t *bcase, int32_t *bc1,
                                int32_t *bc2, int32_t *avgnext, int32_t *lcase, int32_t *lc1, int32_t *lc2, int32_t *src)
{
  ENG(e);
  int32_t *t[10] = { cdm, cdp, bcase, bc1, bc2, avgnext, lcase, lc1, lc2, src };
  return e->impl->export_tables_noalloc(slot, t);
}
int felbm_kernel_times(felbm_engine *e, int nsteps, float *ms4) { ENG(e); if (!ms4) return FELBM_EINVAL; return e->impl->kernel_times(nsteps, ms4); }
int felbm_autotune(felbm_engine *e, int nsteps, int *bs4, float *ms12) { ENG(e); return e->impl->autotune(nsteps, bs4, ms12); }
int felbm_set_block_sizes(felbm_engine *e, const int *bs4) { ENG(e); if (!bs4) return FELBM_EINVAL; return e->impl->set_block_sizes(bs4); }

int felbm_mrt_matrices(double *M, double *Minv)
{
  if (!M || !Minv) return FELBM_EINVAL;
  for (int a = 0; a < 19; ++a) for (int b = 0; b < 19; ++b) { M[a*19 + b] = fm::MRTTAB.M[a][b]; Minv[a*19 + b] = fm::MRTTAB.MI[a][b]; }
  return FELBM_OK;
}

}  // extern "C"

namespace fm {
// Streaming kernels of felbm_stream_probe: 16 bytes per lane, four independent
// accesses in flight per lane per iteration, grid-stride over n4 f32x4 values,
// non-temporal loads and stores (every byte is touched once per launch).
__global__ void __launch_bounds__(256) k_stream_read(const f32x4 *__restrict__ a, size_t n4, float *__restrict__ sink)
{
  const size_t step = (size_t)gridDim.x * 256u;
  size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
  float s = 0.0f;
  for (; i + 3 * step < n4; i += 4 * step) {
    const f32x4 v0 = __builtin_nontemporal_load(a + i), v1 = __builtin_nontemporal_load(a + i + step);
    const f32x4 v2 = __builtin_nontemporal_load(a + i + 2 * step), v3 = __builtin_nontemporal_load(a + i + 3 * step);
    s += (v0[0] + v1[1]) + (v2[2] + v3[3]);
  }
  for (; i < n4; i += step) { const f32x4 v = __builtin_nontemporal_load(a + i); s += v[0]; }
  if (s == 1.0f) sink[blockIdx.x] = s;      // the buffer holds zeros: never stored, keeps the loads
}
__global__ void __launch_bounds__(256) k_stream_write(f32x4 *__restrict__ b, size_t n4)
{
  const size_t step = (size_t)gridDim.x * 256u;
  const f32x4 z = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n4; i += step) __builtin_nontemporal_store(z, b + i);
}
__global__ void __launch_bounds__(256) k_stream_copy(const f32x4 *__restrict__ a, f32x4 *__restrict__ b, size_t n4)
{
  const size_t step = (size_t)gridDim.x * 256u;
  size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
  for (; i + step < n4; i += 2 * step) {
    const f32x4 v0 = __builtin_nontemporal_load(a + i), v1 = __builtin_nontemporal_load(a + i + step);
    __builtin_nontemporal_store(v0, b + i); __builtin_nontemporal_store(v1, b + i + step);
  }
  for (; i < n4; i += step) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
}

// hand-written streaming bandwidth probe: mode 0 reads `bytes`, 1 writes them,
// 2 copies them (read + write counted); `reps` timed launches after one
// warm-up launch; gbs = bytes moved / s.  No reference counterpart
// (measurement aid: the achievable ceiling beside the 8 TB/s peak).
extern "C" int felbm_stream_probe(int device, int64_t bytes, int reps, int mode, double *gbs)
{
  if (!gbs || bytes < 16 || reps <= 0 || mode < 0 || mode > 2) { fm::set_error("invalid stream probe arguments"); return FELBM_EINVAL; }
  if (hipSetDevice(device) != hipSuccess) { fm::set_error("hipSetDevice failed"); return FELBM_EHIP; }
  const size_t n4 = (size_t)bytes / 16;
  void *a = nullptr, *b = nullptr;
  float *sink = nullptr;
  int nsm = 256;
  (void)hipDeviceGetAttribute(&nsm, hipDeviceAttributeMultiprocessorCount, device);
  const int grid = nsm * 16;
  if (hipMalloc(&a, n4 * 16) != hipSuccess) { fm::set_error("allocation failed"); return FELBM_ENOMEM; }
  if (hipMalloc(&b, n4 * 16) != hipSuccess) { (void)hipFree(a); fm::set_error("allocation failed"); return FELBM_ENOMEM; }
  if (hipMalloc((void **)&sink, (size_t)grid * 4) != hipSuccess) { (void)hipFree(a); (void)hipFree(b); fm::set_error("allocation failed"); return FELBM_ENOMEM; }
  hipStream_t s; hipEvent_t e0, e1;
  int rc = FELBM_OK;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) rc = FELBM_EHIP;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = FELBM_EHIP;
  if (!rc) {
    (void)hipMemsetAsync(a, 0, n4 * 16, s);
    auto launch = [&]() {
      if (mode == 0) hipLaunchKernelGGL(fm::k_stream_read, dim3(grid), dim3(256), 0, s, (const fm::f32x4 *)a, n4, sink);
      else if (mode == 1) hipLaunchKernelGGL(fm::k_stream_write, dim3(grid), dim3(256), 0, s, (fm::f32x4 *)b, n4);
      else hipLaunchKernelGGL(fm::k_stream_copy, dim3(grid), dim3(256), 0, s, (const fm::f32x4 *)a, (fm::f32x4 *)b, n4);
    };
    launch();
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, s);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) rc = FELBM_EHIP;
    float ms = 0; (void)hipEventElapsedTime(&ms, e0, e1);
    const double moved = (double)n4 * 16.0 * (mode == 2 ? 2.0 : 1.0) * reps;
    *gbs = ms > 0 ? moved / (ms * 1e-3) / 1e9 : 0.0;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipStreamDestroy(s);
  }
  (void)hipFree(a); (void)hipFree(b); (void)hipFree(sink);
  if (rc) fm::set_error("stream probe failed");
  return rc;
}

extern "C" int felbm_copy_probe(int device, int64_t bytes, int reps, double *gbs)
{
  // device-to-device copy of `bytes`, `reps` times: the achievable streaming
  // bandwidth (reads + writes) and the calibration kernel for HBM counters
  if (!gbs || bytes <= 0 || reps <= 0) { fm::set_error("invalid copy probe arguments"); return FELBM_EINVAL; }
  if (hipSetDevice(device) != hipSuccess) { fm::set_error("hipSetDevice failed"); return FELBM_EHIP; }
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, (size_t)bytes) != hipSuccess) { fm::set_error("allocation failed"); return FELBM_ENOMEM; }
  if (hipMalloc(&b, (size_t)bytes) != hipSuccess) { (void)hipFree(a); fm::set_error("allocation failed"); return FELBM_ENOMEM; }
  hipStream_t s; hipEvent_t e0, e1;
  int rc = FELBM_OK;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) rc = FELBM_EHIP;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = FELBM_EHIP;
  if (!rc) {
    (void)hipMemsetAsync(a, 1, (size_t)bytes, s);
    (void)hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, s);
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) (void)hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, s);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) rc = FELBM_EHIP;
    float ms = 0; (void)hipEventElapsedTime(&ms, e0, e1);
    *gbs = ms > 0 ? 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9 : 0.0;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipStreamDestroy(s);
  }
  (void)hipFree(a); (void)hipFree(b);
  if (rc) fm::set_error("copy probe failed");
  return rc;
}
