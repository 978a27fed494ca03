// engine_impl.hpp -- host side of the engine, templated on the distribution
// precision.  Included by engine_f64.hip and engine_f32.hip, which are built
// with different floating-point contraction settings.
//
// Replaces felbm_gpu::MultiPhaseGPU (multiphase_gpu.cuh:23-783): owns the
// device state of every instance, runs the step, the mass corrector, the
// monitoring reductions and the state/field transfers.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include "engine.hpp"
#include "kernels.hpp"
#include "collide_f32.hpp"
#include "tracers.hpp"

namespace fm {

#define FM_HIP(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) {                 \
    set_error(std::string("HIP error ") + hipGetErrorName(e_) + " at " + __FILE__ + ":" +   \
              std::to_string(__LINE__) + " (" #call ")"); return FELBM_EHIP; } } while (0)
#define FM_LAUNCH() FM_HIP(hipGetLastError())

inline int ceil_log2(int v) { int b = 0; while ((1 << b) < v) ++b; return b; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

template <typename real>
struct Slot {
  bool used = false;
  int n = 0, ns = 0, nblocks = 0;
  int Nx = 0, Ny = 0, Nz = 0, wpr = 0;
  int per[3] = {1, 1, 1};
  int hbb = 1, mrt = 0, mass_on = 0, open = 0;
  // per group of 64 sites: 1 when a site of the group lies within 2 voxels of
  // a non-periodic face (its ring-2 stencils leave the box)
  std::vector<uint8_t> near_face;
  int par0 = 0;
  double M0 = 0.0;
  double params[32];
  felbm_params prm{};
  char *base = nullptr;           // single device allocation
  size_t bytes = 0;
  Inst<real> d{};                 // descriptor (device pointers filled)
  int2 *dmap = nullptr;           // single-instance block map
  int dmap_len = 0;
  int stats_valid = 0;
  // the caller's site numbering (felbm_geometry.site_xyz): caller site i is
  // engine site perm[i], engine site e is caller site iperm[e] (device copy
  // d_iperm for the table export); empty when it is the engine's own order
  std::vector<int> perm, iperm;
  int *d_iperm = nullptr;
  // open boundaries
  int obc = 0;
  char *ob_base = nullptr;        // ent [n] + ov [entries*8]
  std::vector<ObEntry> entries;   // desc filled at rebuild
  int e0 = 0;                     // first entry in the engine-wide list
  // tracers
  int tr_np = 0, tr_substeps = 1;
  unsigned long long tr_seed = 0;
  char *tr_base = nullptr;        // pos [3np] f64, dm [np] f64, uvw [3np] f64, blocked [np] i32, tu [ns][4] real
  double *tr_pos = nullptr, *tr_dm = nullptr, *tr_uvw = nullptr;
  int *tr_blocked = nullptr;
  // Lagrangian stretching (rho_hat, log_rho, block partials, row ring)
  char *tr_sbase = nullptr;
  Stretch st{};
  unsigned long long tr_drained = 0;   // tracer steps already returned as rows
};

template <typename real>
class Engine : public EngineBase {
public:
  explicit Engine(int device) : dev_(device) {}
  ~Engine() override { release(); }

  int init(int *err) {
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) { set_error("no HIP device available"); *err = FELBM_EHIP; return *err; }
    if (dev_ < 0 || dev_ >= nd) { set_error("device index out of range"); *err = FELBM_EINVAL; return *err; }
    if (hipSetDevice(dev_) != hipSuccess) { set_error("hipSetDevice failed"); *err = FELBM_EHIP; return *err; }
    // the device address of the static workspace, once per device
    if (int e = scratch_init(dev_)) { *err = e; return e; }
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) { set_error("stream create failed"); *err = FELBM_EHIP; return *err; }
    for (int q = 0; q < 10; ++q) if (hipEventCreate(&ev_[q]) != hipSuccess) { set_error("event create failed"); *err = FELBM_EHIP; return *err; }
    if (hipHostMalloc((void **)&pinned_stats_, sizeof(unsigned long long)*12*NSLOT*kMaxAsyncSlots, hipHostMallocDefault) != hipSuccess) { *err = FELBM_ENOMEM; return *err; }
    *err = FELBM_OK;
    return FELBM_OK;
  }

  // ---- lifecycle ---------------------------------------------------------------
  int add_instance(const felbm_geometry *g, const felbm_params *p, const felbm_switches *s, int *slot) override {
    if (!g || !p || !s || !slot || !g->solid) { set_error("null argument"); return FELBM_EINVAL; }
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) { set_error("box dimensions must be positive"); return FELBM_EINVAL; }
    int bx = ceil_log2(g->nx), by = ceil_log2(g->ny), bz = ceil_log2(g->nz);
    if (bx + by + bz > 32) { set_error("box too large for 32-bit packed coordinates"); return FELBM_EUNSUPPORTED; }
    FM_HIP(hipSetDevice(dev_));
    Slot<real> S;
    S.Nx = g->nx; S.Ny = g->ny; S.Nz = g->nz;
    for (int a = 0; a < 3; ++a) S.per[a] = g->periodic[a] ? 1 : 0;
    S.hbb = g->use_halfway_bb ? 1 : 0;
    S.mrt = s->use_mrt ? 1 : 0;
    S.mass_on = s->correct_op_mass ? 1 : 0;
    S.prm = *p;
    // lattice words + packed coordinates
    const int wpr = (g->nx + 31) / 32;
    const int64_t rows = (int64_t)g->ny * g->nz;
    std::vector<uint64_t> words((size_t)(rows * wpr));
    std::vector<uint32_t> coords;
    int64_t cnt = 0;
    for (int z = 0; z < g->nz; ++z)
      for (int y = 0; y < g->ny; ++y) {
        const uint8_t *row = g->solid + ((int64_t)z * g->ny + y) * g->nx;
        for (int w = 0; w < wpr; ++w) {
          uint32_t mask = 0; uint64_t start = (uint64_t)cnt;
          for (int b = 0; b < 32; ++b) {
            int x = w * 32 + b;
            if (x >= g->nx) break;
            if (!row[x]) {
              mask |= 1u << b;
              coords.push_back((uint32_t)x | ((uint32_t)y << bx) | ((uint32_t)z << (bx + by)));
              ++cnt;
            }
          }
          words[(size_t)((((int64_t)z * g->ny) + y) * wpr + w)] = (start << 32) | mask;
        }
      }
    // one zero word past the last row: a 16-byte row window read at the table's
    // last word stays inside the buffer
    words.push_back(0);
    if (cnt == 0) { set_error("geometry has no fluid sites"); return FELBM_EINVAL; }
    if (cnt >= (int64_t)1 << 31) { set_error("more than 2^31 sites"); return FELBM_EUNSUPPORTED; }
    if (g->site_xyz) {
      // the caller's numbering: each listed voxel must be fluid and listed once
      if (g->n_sites != cnt) {
        set_error("site list has " + std::to_string(g->n_sites) + " entries, the mask has " + std::to_string(cnt) +
                  " fluid voxels");
        return FELBM_EINVAL;
      }
      S.perm.assign((size_t)cnt, -1); S.iperm.assign((size_t)cnt, -1);
      bool ident = true;
      for (int64_t i = 0; i < cnt; ++i) {
        const int x = g->site_xyz[3 * i], y = g->site_xyz[3 * i + 1], z = g->site_xyz[3 * i + 2];
        const std::string at = " (site " + std::to_string(i) + " at " + std::to_string(x) + ", " + std::to_string(y) +
                               ", " + std::to_string(z) + ")";
        if (x < 0 || x >= g->nx || y < 0 || y >= g->ny || z < 0 || z >= g->nz) {
          set_error("site outside the box" + at); return FELBM_EINVAL;
        }
        const uint64_t w = words[(size_t)((((int64_t)z * g->ny) + y) * wpr + (x >> 5))];
        const unsigned b = (unsigned)x & 31u;
        if (!((w >> b) & 1u)) { set_error("site on a solid voxel" + at); return FELBM_EINVAL; }
        const int e = (int)(w >> 32) + __builtin_popcount((unsigned)w & ((1u << b) - 1u));
        if (S.iperm[(size_t)e] >= 0) { set_error("voxel listed twice" + at); return FELBM_EINVAL; }
        S.perm[(size_t)i] = e; S.iperm[(size_t)e] = (int)i;
        ident = ident && e == (int)i;
      }
      if (ident) { S.perm.clear(); S.iperm.clear(); }
    }
    S.n = (int)cnt; S.ns = (int)align_up((size_t)cnt, 64); S.wpr = wpr;
    if ((int64_t)g->ny * wpr >= ((int64_t)1 << 24)) {     // row offsets use 24-bit multiplies
      set_error("box too large: ny * ceil(nx/32) must stay below 2^24"); return FELBM_EUNSUPPORTED;
    }
    if ((uint64_t)19 * (uint64_t)S.ns * sizeof(real) >= ((uint64_t)1 << 32) || words.size() * 8 >= ((size_t)1 << 32)) {
      set_error("instance too large: 19*sites*sizeof(real) must stay below 4 GiB");
      return FELBM_EUNSUPPORTED;
    }
    S.nblocks = (S.n + 255) / 256;
    if (!(S.per[0] && S.per[1] && S.per[2])) {
      const int dims[3] = { g->nx, g->ny, g->nz };
      S.near_face.assign((size_t)((S.n + 63) / 64), 0);
      for (int64_t i = 0; i < cnt; ++i) {
        const uint32_t c = coords[(size_t)i];
        const int v[3] = { (int)(c & ((1u << bx) - 1u)), (int)((c >> bx) & ((1u << by) - 1u)), (int)(c >> (bx + by)) };
        for (int a = 0; a < 3; ++a)
          if (!S.per[a] && (v[a] < 2 || v[a] > dims[a] - 3)) { S.near_face[(size_t)(i / 64)] = 1; break; }
      }
    }
    // one allocation: h, g0, g1 (19 ns), 9 fields + 4 export (ns), words, coords, acc
    const size_t RB = sizeof(real), ns = (size_t)S.ns;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off = align_up(off + b, 256); return o; };
    size_t o_h = take(19 * ns * RB), o_g0 = take(19 * ns * RB), o_g1 = take(19 * ns * RB);
    size_t o_f = take(13 * ns * RB);
    size_t o_w = take(words.size() * 8), o_c = take(coords.size() * 4);
    size_t o_a = take(sizeof(unsigned long long) * ACC_WORDS);
    size_t o_m = take(sizeof(int2) * (size_t)(align_up((size_t)S.nblocks, 8)));
    const size_t o_p = S.iperm.empty() ? 0 : take(S.iperm.size() * sizeof(int));
    S.bytes = off;
    if (hipMalloc((void **)&S.base, off) != hipSuccess) { set_error("device allocation failed"); return FELBM_ENOMEM; }
    {
      hipError_t he = hipMemsetAsync(S.base, 0, off, stream_);
      if (he == hipSuccess) he = hipMemcpyAsync(S.base + o_w, words.data(), words.size() * 8, hipMemcpyHostToDevice, stream_);
      if (he == hipSuccess) he = hipMemcpyAsync(S.base + o_c, coords.data(), coords.size() * 4, hipMemcpyHostToDevice, stream_);
      if (he == hipSuccess && !S.iperm.empty())
        he = hipMemcpyAsync(S.base + o_p, S.iperm.data(), S.iperm.size() * sizeof(int), hipMemcpyHostToDevice, stream_);
      if (he == hipSuccess) he = hipStreamSynchronize(stream_);
      if (he != hipSuccess) {
        (void)hipStreamSynchronize(stream_);
        (void)hipFree(S.base);
        set_error(std::string("add_instance: ") + hipGetErrorName(he));
        return FELBM_EHIP;
      }
    }
    Inst<real> &d = S.d;
    d.n = S.n; d.ns = S.ns; d.Nx = S.Nx; d.Ny = S.Ny; d.Nz = S.Nz; d.wpr = wpr;
    d.sx = bx; d.sxy = bx + by; d.mx = (1u << bx) - 1u; d.my = (1u << by) - 1u;
    d.px = S.per[0]; d.py = S.per[1]; d.pz = S.per[2];
    d.hbb = S.hbb; d.mrt = S.mrt; d.mass_on = S.mass_on;
    d.words = (const uint64_t *)(S.base + o_w); d.coords = (const uint32_t *)(S.base + o_c);
    d.nwords = (unsigned)words.size();
    S.open = !(S.per[0] && S.per[1] && S.per[2]);
    d.h = (real *)(S.base + o_h); d.g0 = (real *)(S.base + o_g0); d.g1 = (real *)(S.base + o_g1);
    real *f = (real *)(S.base + o_f);
    d.craw = f; d.p0 = f + F_PJ * ns; d.jx = d.jy = d.jz = d.p0;      // [p0, jx, jy, jz] per site
    d.c = d.ps = f + F_CP * ns;                                        // [c, p*] per site
    d.mu = f + F_MU * ns; d.clo = f + F_CLO * ns;
    export_off_ = 9;   // exports at f + 9*ns .. 12*ns
    d.acc = (unsigned long long *)(S.base + o_a);
    S.d_iperm = S.iperm.empty() ? nullptr : (int *)(S.base + o_p);
    set_params(S);
    // admission parity: g0 holds the state at the current step counter
    S.par0 = (int)(T_ & 1u); d.par0 = S.par0;
    // single-instance block map
    S.dmap = (int2 *)(S.base + o_m);
    S.dmap_len = (int)align_up((size_t)S.nblocks, 8);
    int sid = -1;
    for (size_t q = 0; q < slots_.size(); ++q) if (!slots_[q].used) { sid = (int)q; break; }
    if (sid < 0) { sid = (int)slots_.size(); slots_.push_back(Slot<real>()); }
    S.used = true;
    slots_[sid] = S;
    dirty_ = true;
    if (int e = rebuild()) {
      // roll the slot back: the engine keeps the instances it had
      (void)hipStreamSynchronize(stream_);
      (void)hipFree(slots_[sid].base);
      slots_[sid] = Slot<real>();
      dirty_ = true;
      (void)rebuild();
      return e;
    }
    *slot = sid;
    return FELBM_OK;
  }

  int remove_instance(int slot) override {
    if (int e = check(slot)) return e;
    FM_HIP(hipStreamSynchronize(stream_));
    drop_graph();
    FM_HIP(hipFree(slots_[slot].base));
    if (slots_[slot].ob_base) FM_HIP(hipFree(slots_[slot].ob_base));
    if (slots_[slot].tr_base) FM_HIP(hipFree(slots_[slot].tr_base));
    if (slots_[slot].tr_sbase) FM_HIP(hipFree(slots_[slot].tr_sbase));
    slots_[slot] = Slot<real>();
    dirty_ = true;
    return rebuild();
  }

  int num_sites(int slot, int64_t *n) override { if (int e = check(slot)) return e; *n = slots_[slot].n; return FELBM_OK; }
  int num_instances(int *count) override {
    int c = 0; for (auto &s : slots_) c += s.used; *count = c; return FELBM_OK;
  }

  // ---- state -------------------------------------------------------------------
  int upload_state(int slot, const double *h, const double *g) override {
    if (int e = check(slot)) return e;
    if (!h || !g) { set_error("null state"); return FELBM_EINVAL; }
    Slot<real> &S = slots_[slot];
    const size_t n = (size_t)S.n, ns = (size_t)S.ns;
    std::vector<real> t(19 * ns, (real)0);
    unsigned cur = cur_of(S);
    real *stage = cur ? S.d.g0 : S.d.g1;      // the dead g buffer between steps
    const int nb = (S.n + 255) / 256;
    // the arrays hold post-collision populations: store the state so the
    // next step's gather reproduces it (k_unstream inverts the gather)
    for (int pass = 0; pass < 2; ++pass) {
      const double *src = pass ? g : h;
      if (pass && p_ref<real>() != 0.0) {        // FP32 g slots hold g_k - w_k (p_ref, kernels.hpp)
        for (int k = 0; k < 19; ++k) {
          const double off = (double)wf32(k) * p_ref<real>();
          for (size_t i = 0; i < n; ++i) t[k * ns + eng(S, i)] = (real)g_flush((float)(src[k * n + i] - off));
        }
      } else {
        for (int k = 0; k < 19; ++k) for (size_t i = 0; i < n; ++i) t[k * ns + eng(S, i)] = (real)src[k * n + i];
      }
      real *dst = pass ? (cur ? S.d.g1 : S.d.g0) : S.d.h;
      FM_HIP(hipMemcpyAsync(stage, t.data(), t.size() * sizeof(real), hipMemcpyHostToDevice, stream_));
      hipLaunchKernelGGL(k_unstream<real>, dim3(nb), dim3(256), 0, stream_, S.d, (const real *)stage, dst);
      FM_LAUNCH();
      FM_HIP(hipStreamSynchronize(stream_));
    }
    FM_HIP(hipMemsetAsync(S.d.acc + ACC_FLAGS, 0, sizeof(unsigned long long), stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    S.stats_valid = 0;
    return FELBM_OK;
  }

  // The rest state h_k = c w_k, g_k = w_k p0 (the reference's initial
  // condition), built on the device: only c crosses the bus (8 B per site
  // instead of the 304 B of a double state), staged in the field block, which
  // holds nothing between steps that the next step reads.  Same bits as
  // upload_state of the host-built state (k_rest_stage).
  int init_rest(int slot, const double *c, double p0) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    const size_t n = (size_t)S.n;
    std::vector<double> ce;
    const double *src = c;
    if (!S.perm.empty()) {
      ce.resize(n);
      for (size_t i = 0; i < n; ++i) ce[eng(S, i)] = c[i];
      src = ce.data();
    }
    double *cd = (double *)S.d.craw;            // 13 ns reals >= n doubles
    const unsigned cur = cur_of(S);
    real *stage = cur ? S.d.g0 : S.d.g1;        // the dead g buffer between steps
    const int nb = (S.n + 255) / 256;
    FM_HIP(hipMemcpyAsync(cd, src, n * sizeof(double), hipMemcpyHostToDevice, stream_));
    for (int pass = 0; pass < 2; ++pass) {
      hipLaunchKernelGGL(k_rest_stage<real>, dim3(nb), dim3(256), 0, stream_, S.d, (const double *)cd, p0, pass, stage);
      FM_LAUNCH();
      real *dst = pass ? (cur ? S.d.g1 : S.d.g0) : S.d.h;
      hipLaunchKernelGGL(k_unstream<real>, dim3(nb), dim3(256), 0, stream_, S.d, (const real *)stage, dst);
      FM_LAUNCH();
    }
    FM_HIP(hipMemsetAsync(S.d.acc + ACC_FLAGS, 0, sizeof(unsigned long long), stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    S.stats_valid = 0;
    return FELBM_OK;
  }

  int download_state(int slot, double *h, double *g) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (int e = flush()) return e;
    const size_t n = (size_t)S.n, ns = (size_t)S.ns;
    unsigned cur = cur_of(S);
    real *stage = cur ? S.d.g0 : S.d.g1;      // the dead g buffer between steps
    bool apply = false; double lambda = 0.0;
    if (int e = owed_lambda(S, apply, lambda)) return e;
    int nb = (S.n + 255) / 256;
    hipLaunchKernelGGL(k_inject_copy<real>, dim3(nb), dim3(256), 0, stream_, S.d, apply ? 1 : 0, lambda, stage);
    FM_LAUNCH();
    std::vector<real> t(19 * ns);
    FM_HIP(hipMemcpyAsync(t.data(), stage, t.size() * sizeof(real), hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    for (int k = 0; k < 19; ++k) for (size_t i = 0; i < n; ++i) h[k * n + i] = (double)t[k * ns + eng(S, i)];
    hipLaunchKernelGGL(k_gather_copy<real>, dim3(nb), dim3(256), 0, stream_, S.d, (const real *)(cur ? S.d.g1 : S.d.g0), stage);
    FM_LAUNCH();
    FM_HIP(hipMemcpyAsync(t.data(), stage, t.size() * sizeof(real), hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    for (int k = 0; k < 19; ++k) {
      const double off = p_ref<real>() != 0.0 ? (double)wf32(k) * p_ref<real>() : 0.0;
      for (size_t i = 0; i < n; ++i) g[k * n + i] = (double)t[k * ns + eng(S, i)] + off;
    }
    return FELBM_OK;
  }

  // ---- stepping ------------------------------------------------------------------
  int step(int nsteps, int stats_last) override {
    if (nsteps < 0) { set_error("negative step count"); return FELBM_EINVAL; }
    if (int e = flush()) return e;
    if (nmap_ == 0) return FELBM_OK;
    for (int t = 0; t < nsteps; ++t) {
      bool st = stats_last && t == nsteps - 1;
      if (int e = launch_step(st)) return e;
    }
    if (nsteps > 0) for (auto &s : slots_) if (s.used) s.stats_valid = stats_last ? 1 : 0;
    return FELBM_OK;
  }

  int step_graph(int nsteps, int replays) override {
    if (nsteps <= 0 || (nsteps & 1) || replays < 0) { set_error("graph steps must be a positive even count"); return FELBM_EINVAL; }
    if (int e = flush()) return e;
    if (nmap_ == 0) return FELBM_OK;
    if (!graph_exec_ || graph_steps_ != nsteps || graph_ver_ != map_ver_ || (graph_T_ & 1u) != (T_ & 1u)) {
      if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
      unsigned T0 = T_;
      FM_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      for (int t = 0; t < nsteps; ++t) {
        int e = launch_step(false);
        if (e) { hipGraph_t gg; (void)hipStreamEndCapture(stream_, &gg); T_ = T0; return e; }
      }
      hipGraph_t graph;
      FM_HIP(hipStreamEndCapture(stream_, &graph));
      FM_HIP(hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
      T_ = T0;
      graph_steps_ = nsteps; graph_ver_ = map_ver_; graph_T_ = T0;
    }
    for (int r = 0; r < replays; ++r) { FM_HIP(hipGraphLaunch(graph_exec_, stream_)); T_ += (unsigned)nsteps; }
    for (auto &s : slots_) if (s.used) s.stats_valid = 0;
    return FELBM_OK;
  }

  int set_forcing_factor(int slot, double ff) override {
    if (int e = check(slot)) return e;
    slots_[slot].prm.forcing_factor = ff;
    slots_[slot].d.ff = (real)ff;
    // descriptors on the device are current: update this one in stream order
    if (!dirty_ && !dirty_desc_ && d_desc_ && slot_desc_.size() > (size_t)slot && slot_desc_[slot] >= 0) {
      hipLaunchKernelGGL(k_set_ff<real>, dim3(1), dim3(64), 0, stream_, d_desc_, slot_desc_[slot], (real)ff);
      FM_LAUNCH();
    } else {
      dirty_desc_ = true;
    }
    return FELBM_OK;
  }
  int synchronize() override { FM_HIP(hipStreamSynchronize(stream_)); return FELBM_OK; }

  // ---- mass correction ------------------------------------------------------------
  int record_target_mass(int slot) override {
    double m; if (int e = current_mass(slot, &m)) return e;
    slots_[slot].M0 = m; slots_[slot].d.M0 = m; dirty_desc_ = true;
    return FELBM_OK;
  }
  int get_target_mass(int slot, double *m) override { if (int e = check(slot)) return e; *m = slots_[slot].M0; return FELBM_OK; }
  int set_target_mass(int slot, double m) override {
    if (int e = check(slot)) return e;
    slots_[slot].M0 = m; slots_[slot].d.M0 = m; dirty_desc_ = true; return FELBM_OK;
  }
  int current_mass(int slot, double *m) override {
    if (int e = check(slot)) return e;
    if (int e = flush()) return e;
    Slot<real> &S = slots_[slot];
    bool apply = false; double lambda = 0.0;
    if (int e = owed_lambda(S, apply, lambda)) return e;
    double M, W; uint64_t bad;
    if (int e = mass_sums(S, apply, lambda, M, W, bad)) return e;
    if (bad) { set_error("non-finite or out-of-range order parameter in the state"); return FELBM_ENONFINITE; }
    *m = M;
    return FELBM_OK;
  }

  // ---- monitoring -------------------------------------------------------------------
  int flow_stats(int slot, double *out) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (!S.stats_valid) return stats_from_fields(S, out);
    std::vector<unsigned long long> a(12 * NSLOT);
    FM_HIP(hipMemcpyAsync(a.data(), S.d.acc + ACC_STATS, a.size() * 8, hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    return reduce_stats(a.data(), out);
  }
  int flow_stats_launch() override {
    int q = 0;
    for (auto &s : slots_) {
      if (s.used) FM_HIP(hipMemcpyAsync(pinned_stats_ + (size_t)q * 12 * NSLOT, s.d.acc + ACC_STATS,
                                        12 * NSLOT * 8, hipMemcpyDeviceToHost, stream_));
      if (++q >= kMaxAsyncSlots) break;
    }
    FM_HIP(hipEventRecord(ev_[9], stream_));
    stats_launched_ = true;
    return FELBM_OK;
  }
  int flow_stats_ready(int *ready) override {
    if (!stats_launched_) { *ready = 0; return FELBM_OK; }
    hipError_t e = hipEventQuery(ev_[9]);
    *ready = (e == hipSuccess) ? 1 : 0;
    if (e != hipSuccess && e != hipErrorNotReady) { set_error("event query failed"); return FELBM_EHIP; }
    return FELBM_OK;
  }
  int flow_stats_collect(int slot, double *out) override {
    if (int e = check(slot)) return e;
    if (!stats_launched_) { set_error("flow_stats_launch was not called"); return FELBM_ESTATE; }
    if (slot >= kMaxAsyncSlots) { set_error("async stats cover the first 256 slots"); return FELBM_EUNSUPPORTED; }
    FM_HIP(hipEventSynchronize(ev_[9]));
    return reduce_stats(pinned_stats_ + (size_t)slot * 12 * NSLOT, out);
  }

  // ---- fields ----------------------------------------------------------------------
  int download_fields(int slot, double *out) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (int e = field_pass(S)) return e;
    const size_t n = (size_t)S.n, ns = (size_t)S.ns;
    real *exp = S.d.craw + export_off_ * ns;
    std::vector<real> t(4 * ns), cp(2 * ns);
    FM_HIP(hipMemcpyAsync(t.data(), exp, t.size() * sizeof(real), hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipMemcpyAsync(cp.data(), S.d.c, cp.size() * sizeof(real), hipMemcpyDeviceToHost, stream_));   // [c, p*] pairs
    FM_HIP(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n; ++i) {
      const size_t e = eng(S, i);
      out[i] = (double)t[e]; out[n + i] = (double)cp[2 * e];
      out[2 * n + i] = (double)t[ns + e]; out[3 * n + i] = (double)t[2 * ns + e]; out[4 * n + i] = (double)t[3 * ns + e];
      out[5 * n + i] = (double)cp[2 * e + 1] + p_ref<real>();    // p* is stored as p - p_ref
    }
    return FELBM_OK;
  }

  // MultiPhaseGPU::set_open_bnd (multiphase_gpu.cuh:369-400) with the DevParams
  // open-BC block (felbm_gpu_main.cu:340-364)
  int set_open_bnd(int slot, const felbm_open_bnd *ob, const int32_t *inlet, int n_in, const int32_t *outlet,
                   int n_out, const double *icoord, const int32_t *bv, const int32_t *br, int n_buf,
                   int copy) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (!ob || n_in < 0 || n_out < 0 || n_buf < 0 || (n_in && !inlet) || (n_out && !outlet) ||
        (n_buf && (!bv || !br))) { set_error("invalid open-boundary lists"); return FELBM_EINVAL; }
    if (S.mass_on) {
      set_error("correct_op_mass must be false with open boundaries (mass enters and leaves)");
      return FELBM_EINVAL;
    }
    if (n_buf && !copy) { set_error("buffer nodes are only used with copy_to_buffers"); return FELBM_EINVAL; }
    if ((ob->inlet_mode == 2 || ob->inlet_mode == 3) && n_in && !icoord) {
      set_error("split and stripes inlet modes need the inlet coordinates"); return FELBM_EINVAL;
    }
    const int n = S.n;
    // node lists in the caller's numbering -> engine sites (out-of-range
    // indices are passed on and refused below)
    std::vector<int32_t> in_e, out_e, bv_e, br_e;
    if (!S.perm.empty()) {
      auto map_list = [&](const int32_t *a, int cnt, std::vector<int32_t> &o) {
        o.resize((size_t)cnt);
        for (int q = 0; q < cnt; ++q) o[(size_t)q] = (a[q] >= 0 && a[q] < n) ? S.perm[(size_t)a[q]] : -1;
      };
      map_list(inlet, n_in, in_e); map_list(outlet, n_out, out_e); map_list(bv, n_buf, bv_e); map_list(br, n_buf, br_e);
      inlet = in_e.data(); outlet = out_e.data(); bv = bv_e.data(); br = br_e.data();
    }
    std::vector<int> ent((size_t)n, -1), kind((size_t)n, 3), pos((size_t)n, -1);
    std::vector<ObEntry> E;
    const bool in_proc = ob->use_in_vel || ob->use_in_prs || ob->use_in_fluid;
    const bool out_proc = ob->use_out_prs || ob->use_out_fluid;
    auto add = [&](int site, int ref, int rkind, double rc, int bits) -> bool {
      if (site < 0 || site >= n || ref < 0 || ref >= n || ent[site] >= 0) return false;
      ObEntry e{}; e.desc = -1; e.site = site; e.ref = ref; e.rkind = rkind; e.eid = (int)E.size(); e.rcoord = rc;
      ent[site] = (e.eid << 4) | bits;
      E.push_back(e);
      return true;
    };
    for (int q = 0; q < n_in; ++q) {
      if (!add(inlet[q], inlet[q], 0, icoord ? icoord[q] : 0.0, (in_proc ? ENT_BCPROC : 0) | (copy ? 0 : ENT_EQ))) {
        set_error("inlet node out of range or listed twice"); return FELBM_EINVAL;
      }
      kind[inlet[q]] = 0; pos[inlet[q]] = q;
    }
    for (int q = 0; q < n_out; ++q) {
      if (!add(outlet[q], outlet[q], 1, 0.0, (out_proc ? ENT_BCPROC : 0) | (copy ? 0 : ENT_EQ))) {
        set_error("outlet node out of range or listed twice"); return FELBM_EINVAL;
      }
      kind[outlet[q]] = 1;
    }
    for (int q = 0; q < n_buf; ++q) {
      const int r = br[q];
      if (r < 0 || r >= n) { set_error("buffer reference out of range"); return FELBM_EINVAL; }
      const double rc = (kind[r] == 0 && icoord) ? icoord[pos[r]] : 0.0;
      if (!add(bv[q], r, kind[r], rc, ENT_BCPROC | ENT_NOSTREAM)) {
        set_error("buffer node out of range or listed twice"); return FELBM_EINVAL;
      }
    }
    FM_HIP(hipStreamSynchronize(stream_));
    drop_graph();                     // a captured step graph holds the boundary lists
    if (S.ob_base) { FM_HIP(hipFree(S.ob_base)); S.ob_base = nullptr; }
    const size_t ent_bytes = align_up((size_t)n * sizeof(int), 256);
    const size_t bytes = ent_bytes + std::max<size_t>(1, E.size()) * 8 * sizeof(real);
    if (hipMalloc((void **)&S.ob_base, bytes) != hipSuccess) { set_error("device allocation failed"); return FELBM_ENOMEM; }
    FM_HIP(hipMemcpyAsync(S.ob_base, ent.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, stream_));
    FM_HIP(hipMemsetAsync(S.ob_base + ent_bytes, 0, bytes - ent_bytes, stream_));
    Inst<real> &d = S.d;
    d.ent = (const int *)S.ob_base; d.ov = (real *)(S.ob_base + ent_bytes);
    d.obc = 1;
    d.use_in_vel = ob->use_in_vel; d.use_in_prs = ob->use_in_prs; d.use_in_fluid = ob->use_in_fluid;
    d.use_out_prs = ob->use_out_prs; d.use_out_fluid = ob->use_out_fluid; d.inlet_mode = ob->inlet_mode;
    for (int a = 0; a < 3; ++a) d.u_in[a] = ob->u_in[a];
    d.p_in = ob->p_in - p_ref<real>(); d.p_out = ob->p_out - p_ref<real>();   // as the state stores pressure
    d.c_out_fixed = ob->c_out_fixed; d.rho_out_fixed = ob->rho_out_fixed;
    d.inlet_c_a = ob->inlet_c_a; d.inlet_period = ob->inlet_period; d.inlet_duty = ob->inlet_duty;
    d.inlet_ramp = ob->inlet_ramp; d.split_pos = ob->split_pos;
    // the injection schedule starts at step 0 (open_step, multiphase_gpu.cuh:445)
    static const unsigned long long zero = 0;
    FM_HIP(hipMemcpyAsync(d.acc + ACC_FLAGS + 1, &zero, sizeof(zero), hipMemcpyHostToDevice, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    S.obc = 1;
    S.entries = std::move(E);
    dirty_ = true;
    return FELBM_OK;
  }

  int export_tables(int slot, int32_t **tabs) override {
    if (int e = check(slot)) return e;
    if (int e = flush()) return e;
    Slot<real> &S = slots_[slot];
    const size_t n = (size_t)S.n, Q1n = 18 * n, Vn = 19 * n;
    size_t sz[10] = { Q1n, Q1n, Q1n, Q1n, Q1n, Vn, Q1n, Q1n, Q1n, Vn };
    size_t tot = 0; for (int q = 0; q < 10; ++q) tot += sz[q];
    int *buf = nullptr;
    FM_HIP(hipMalloc((void **)&buf, tot * 4));
    int *p[10]; size_t o = 0;
    for (int q = 0; q < 10; ++q) { p[q] = buf + o; o += sz[q]; }
    hipLaunchKernelGGL(k_export_tables<real>, dim3((S.n + 255) / 256), dim3(256), 0, stream_, S.d,
                       p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], (const int *)S.d_iperm);
    hipError_t le = hipGetLastError();
    std::vector<int> h(tot);
    if (le == hipSuccess) le = hipMemcpyAsync(h.data(), buf, tot * 4, hipMemcpyDeviceToHost, stream_);
    if (le == hipSuccess) le = hipStreamSynchronize(stream_);
    (void)hipFree(buf);
    if (le != hipSuccess) { set_error(std::string("export_tables: ") + hipGetErrorName(le)); return FELBM_EHIP; }
    o = 0;
    for (int q = 0; q < 10; ++q) { if (tabs[q]) std::memcpy(tabs[q], h.data() + o, sz[q] * 4); o += sz[q]; }
    return FELBM_OK;
  }

  // export_tables with the staging area in the library's static workspace:
  // no device allocation or release.  The kernel writes every entry of the ten
  // tables, so nothing in the workspace needs clearing; calls that overlap on
  // one GPU (two engines on two host threads) share the workspace and race.
  int export_tables_noalloc(int slot, int32_t **tabs) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    const size_t n = (size_t)S.n, Q1n = 18 * n, Vn = 19 * n;
    const size_t sz[10] = { Q1n, Q1n, Q1n, Q1n, Q1n, Vn, Q1n, Q1n, Q1n, Vn };
    size_t tot = 0; for (int q = 0; q < 10; ++q) tot += sz[q];
    if (tot * sizeof(int) > kScratchBytes) {
      set_error("export_tables_noalloc: " + std::to_string(S.n) + " sites need " + std::to_string(tot * sizeof(int)) +
                " bytes of workspace; the static workspace holds " + std::to_string(kScratchBytes) + " (at most " +
                std::to_string(kScratchBytes / (182 * sizeof(int))) + " sites)");
      return FELBM_EINVAL;
    }
    int *buf = (int *)scratch_ptr(dev_);
    if (!buf) { set_error("export_tables_noalloc: workspace not set up on this device"); return FELBM_ESTATE; }
    int *p[10]; size_t o = 0;
    for (int q = 0; q < 10; ++q) { p[q] = buf + o; o += sz[q]; }
    hipLaunchKernelGGL(k_export_tables<real>, dim3((S.n + 255) / 256), dim3(256), 0, stream_, S.d,
                       p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], (const int *)S.d_iperm);
    FM_LAUNCH();
    for (int q = 0; q < 10; ++q)
      if (tabs[q]) FM_HIP(hipMemcpyAsync(tabs[q], p[q], sz[q] * sizeof(int), hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    return FELBM_OK;
  }

  // ---- tracers (ParticleManager, felbm_gpu_main.cu:426-438, :752-777) ------------
  int tracers_init(int slot, int64_t np, const double *xyz, const double *dm, uint64_t seed, int substeps,
                   int vskip) override {
    if (int e = check(slot)) return e;
    if (np < 0 || np > (1 << 30) || (np && (!xyz || !dm)) || substeps < 1 || vskip < 1) {
      set_error("invalid tracer arguments"); return FELBM_EINVAL;
    }
    Slot<real> &S = slots_[slot];
    FM_HIP(hipStreamSynchronize(stream_));
    drop_graph();                     // a captured step graph holds the tracer buffers
    if (S.tr_base) { FM_HIP(hipFree(S.tr_base)); S.tr_base = nullptr; }
    if (S.tr_sbase) { FM_HIP(hipFree(S.tr_sbase)); S.tr_sbase = nullptr; }
    S.tr_np = 0; S.d.trk = 0; S.d.tu = nullptr; S.d.tr_phase = 0; S.st = Stretch{}; S.tr_drained = 0;
    dirty_desc_ = true;
    if (np == 0) return FELBM_OK;
    const size_t n = (size_t)np, ns = (size_t)S.ns;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off = align_up(off + b, 256); return o; };
    const size_t o_p = take(3 * n * 8), o_d = take(n * 8), o_u = take(3 * n * 8), o_b = take(n * 4);
    const size_t o_t = take(4 * ns * sizeof(real));   // per site: u (3) and c for the phase split
    if (hipMalloc((void **)&S.tr_base, off) != hipSuccess) { set_error("device allocation failed"); return FELBM_ENOMEM; }
    FM_HIP(hipMemsetAsync(S.tr_base, 0, off, stream_));
    S.tr_pos = (double *)(S.tr_base + o_p); S.tr_dm = (double *)(S.tr_base + o_d);
    S.tr_uvw = (double *)(S.tr_base + o_u); S.tr_blocked = (int *)(S.tr_base + o_b);
    FM_HIP(hipMemcpyAsync(S.tr_pos, xyz, 3 * n * 8, hipMemcpyHostToDevice, stream_));
    FM_HIP(hipMemcpyAsync(S.tr_dm, dm, n * 8, hipMemcpyHostToDevice, stream_));
    static const unsigned long long zero = 0;
    FM_HIP(hipMemcpyAsync(S.d.acc + ACC_FLAGS + 2, &zero, sizeof(zero), hipMemcpyHostToDevice, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    S.tr_np = (int)np; S.tr_seed = seed; S.tr_substeps = substeps;
    S.d.trk = 1; S.d.tr_vskip = vskip; S.d.tu = (real *)(S.tr_base + o_t);
    return FELBM_OK;
  }

  int tracers_download(int slot, double *xyz, double *uvw, int32_t *blocked) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (!S.tr_np) { set_error("no tracers on this slot"); return FELBM_ESTATE; }
    if (int e = flush()) return e;
    const size_t n = (size_t)S.tr_np;
    if (uvw) {
      hipLaunchKernelGGL(k_tracer_velocity<real>, dim3((S.tr_np + 255) / 256), dim3(256), 0, stream_, d_desc_,
                         slot_desc_[slot], (const double *)S.tr_pos, S.tr_uvw, S.tr_np);
      FM_LAUNCH();
      FM_HIP(hipMemcpyAsync(uvw, S.tr_uvw, 3 * n * 8, hipMemcpyDeviceToHost, stream_));
    }
    if (xyz) FM_HIP(hipMemcpyAsync(xyz, S.tr_pos, 3 * n * 8, hipMemcpyDeviceToHost, stream_));
    if (blocked) FM_HIP(hipMemcpyAsync(blocked, S.tr_blocked, n * 4, hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    return FELBM_OK;
  }

  // stretching on/off for the slot's tracers; rho_hat/log_rho restore a saved
  // state (NULL: isotropic unit vectors from orient_seed, log_rho = 0)
  int tracers_stretching(int slot, int enable, const double *rho_hat, const double *log_rho, uint64_t orient_seed,
                         int phase_split, int row_capacity) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (!S.tr_np) { set_error("no tracers on this slot"); return FELBM_ESTATE; }
    FM_HIP(hipStreamSynchronize(stream_));
    drop_graph();                     // a captured step graph holds the stretching buffers
    if (S.tr_sbase) { FM_HIP(hipFree(S.tr_sbase)); S.tr_sbase = nullptr; }
    S.st = Stretch{}; S.d.tr_phase = 0; dirty_desc_ = true;
    if (!enable) return FELBM_OK;
    if (row_capacity < 1) { set_error("row capacity must be positive"); return FELBM_EINVAL; }
    const size_t n = (size_t)S.tr_np, nb = (n + 255) / 256;
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off = align_up(off + b, 256); return o; };
    const size_t o_r = take(3 * n * 8), o_l = take(n * 8), o_p = take(nb * 8 * 8), o_w = take((size_t)row_capacity * 9 * 8);
    if (hipMalloc((void **)&S.tr_sbase, off) != hipSuccess) { set_error("device allocation failed"); return FELBM_ENOMEM; }
    std::vector<double> rh(3 * n), lr(n, 0.0);
    if (rho_hat) {
      std::copy(rho_hat, rho_hat + 3 * n, rh.begin());
    } else {
      // isotropic orientations: normalised triples of standard normals
      for (size_t t = 0; t < n; ++t) {
        double v[3], nrm = 0.0;
        for (int a = 0; a < 3; ++a) {
          const uint64_t r0 = mix64(orient_seed ^ mix64(6 * t + 2 * a)), r1 = mix64(orient_seed ^ mix64(6 * t + 2 * a + 1));
          const double u0 = (double)((r0 >> 11) + 1) * 0x1p-53, v0 = (double)(r1 >> 11) * 0x1p-53;
          v[a] = std::sqrt(-2.0 * std::log(u0)) * std::cos(6.283185307179586 * v0);
          nrm += v[a] * v[a];
        }
        nrm = std::sqrt(nrm);
        for (int a = 0; a < 3; ++a) rh[3 * t + a] = v[a] / nrm;
      }
    }
    if (log_rho) std::copy(log_rho, log_rho + n, lr.begin());
    FM_HIP(hipMemcpyAsync(S.tr_sbase + o_r, rh.data(), 3 * n * 8, hipMemcpyHostToDevice, stream_));
    FM_HIP(hipMemcpyAsync(S.tr_sbase + o_l, lr.data(), n * 8, hipMemcpyHostToDevice, stream_));
    FM_HIP(hipMemsetAsync(S.tr_sbase + o_p, 0, off - o_p, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    S.st.rho_hat = (double *)(S.tr_sbase + o_r); S.st.log_rho = (double *)(S.tr_sbase + o_l);
    S.st.partials = (double *)(S.tr_sbase + o_p); S.st.rows = (double *)(S.tr_sbase + o_w); S.st.cap = row_capacity;
    S.d.tr_phase = phase_split ? 1 : 0;
    unsigned long long ts = 0;
    FM_HIP(hipMemcpy(&ts, S.d.acc + ACC_FLAGS + 2, 8, hipMemcpyDeviceToHost));
    S.tr_drained = ts;
    return FELBM_OK;
  }

  // rows of the steps since the last call, oldest first (at most max_rows);
  // rows older than the ring capacity are counted in *lost
  int tracers_stretch_rows(int slot, double *rows, int max_rows, int *count, int64_t *lost) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (!S.st.rows) { set_error("stretching is not enabled on this slot"); return FELBM_ESTATE; }
    if (!count || max_rows < 0 || (max_rows && !rows)) { set_error("invalid row buffer"); return FELBM_EINVAL; }
    unsigned long long ts = 0;
    FM_HIP(hipMemcpyAsync(&ts, S.d.acc + ACC_FLAGS + 2, 8, hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    unsigned long long first = S.tr_drained, lostn = 0;
    if (ts - first > (unsigned long long)S.st.cap) { lostn = ts - first - S.st.cap; first = ts - S.st.cap; }
    const unsigned long long take_n = std::min<unsigned long long>(ts - first, (unsigned long long)max_rows);
    std::vector<double> ring((size_t)S.st.cap * 9);
    FM_HIP(hipMemcpy(ring.data(), S.st.rows, ring.size() * 8, hipMemcpyDeviceToHost));
    for (unsigned long long r = 0; r < take_n; ++r) {
      const size_t src = (size_t)((first + r) % (unsigned long long)S.st.cap);
      std::copy(ring.begin() + 9 * src, ring.begin() + 9 * src + 9, rows + 9 * r);
    }
    S.tr_drained = first + take_n;
    *count = (int)take_n;
    if (lost) *lost = (int64_t)lostn;
    return FELBM_OK;
  }

  int tracers_stretch_state(int slot, double *rho_hat, double *log_rho) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (!S.st.rows) { set_error("stretching is not enabled on this slot"); return FELBM_ESTATE; }
    const size_t n = (size_t)S.tr_np;
    if (rho_hat) FM_HIP(hipMemcpyAsync(rho_hat, S.st.rho_hat, 3 * n * 8, hipMemcpyDeviceToHost, stream_));
    if (log_rho) FM_HIP(hipMemcpyAsync(log_rho, S.st.log_rho, n * 8, hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    return FELBM_OK;
  }

  int tracers_count(int slot, int64_t *np) override {
    if (int e = check(slot)) return e;
    *np = slots_[slot].tr_np;
    return FELBM_OK;
  }

  // tracer step counter: the Brownian draw index, the velocity-refresh phase
  // and the step column of the stretching rows all follow it
  int tracers_set_step(int slot, uint64_t step) override {
    if (int e = check(slot)) return e;
    Slot<real> &S = slots_[slot];
    if (!S.tr_np) { set_error("no tracers on this slot"); return FELBM_ESTATE; }
    const unsigned long long v = step;
    FM_HIP(hipMemcpyAsync(S.d.acc + ACC_FLAGS + 2, &v, sizeof(v), hipMemcpyHostToDevice, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    S.tr_drained = v;
    return FELBM_OK;
  }

  int kernel_times(int nsteps, float *ms4) override {
    if (int e = flush()) return e;
    if (nmap_ == 0 || nsteps <= 0) { for (int q = 0; q < 4; ++q) ms4[q] = 0; return FELBM_OK; }
    double acc[4] = {0, 0, 0, 0};
    for (int t = 0; t < nsteps; ++t) {
      if (int e = launch_step(false, true)) return e;
      FM_HIP(hipEventSynchronize(ev_[4]));
      for (int q = 0; q < 4; ++q) { float ms = 0; FM_HIP(hipEventElapsedTime(&ms, ev_[q], ev_[q + 1])); acc[q] += ms; }
    }
    for (int q = 0; q < 4; ++q) ms4[q] = (float)(acc[q] / nsteps);
    for (auto &s : slots_) if (s.used) s.stats_valid = 0;
    return FELBM_OK;
  }

  // Tile-shape autotuner: every pass timed at each workgroup shape (HIP
  // events around its launches, after two warm-up steps per shape) on the
  // instances present; the fastest shape per pass is kept.  The state of
  // every instance is copied aside first and restored afterwards, so the
  // tuning steps leave no trace.  Not with tracers (their positions live
  // outside the instance allocation).
  int autotune(int nsteps, int *bs4, float *ms12) override {
    if (nsteps <= 0) { set_error("autotune needs at least one timed step"); return FELBM_EINVAL; }
    if (int e = flush()) return e;
    for (auto &s : slots_)
      if (s.used && s.tr_np) { set_error("autotune before tracers are added (the tuning steps would move them)"); return FELBM_ESTATE; }
    float ms[3][4] = {};
    if (nmap_ > 0) {
      std::vector<std::pair<Slot<real> *, char *>> saved;
      auto drop = [&]() { for (auto &x : saved) (void)hipFree(x.second); saved.clear(); };
      for (auto &s : slots_) {
        if (!s.used) continue;
        char *p = nullptr;
        if (hipMalloc((void **)&p, s.bytes) != hipSuccess) { drop(); set_error("device allocation failed (autotune state copy)"); return FELBM_ENOMEM; }
        saved.push_back({&s, p});
        if (hipMemcpyAsync(p, s.base, s.bytes, hipMemcpyDeviceToDevice, stream_) != hipSuccess) {
          drop(); set_error("state copy failed (autotune)"); return FELBM_EHIP;
        }
      }
      const unsigned T0 = T_;
      int keep[5]; std::copy(bsi_, bsi_ + 5, keep);
      int rc = FELBM_OK;
      for (int bi = 0; bi < 3 && rc == FELBM_OK; ++bi) {
        for (int q = 1; q <= 4; ++q) bsi_[q] = bi;
        for (int w = 0; w < 2 && rc == FELBM_OK; ++w) rc = launch_step(false);
        if (rc == FELBM_OK) rc = kernel_times(nsteps, ms[bi]);
      }
      // the state goes back whatever happened while timing
      for (auto &x : saved)
        if (hipMemcpyAsync(x.first->base, x.second, x.first->bytes, hipMemcpyDeviceToDevice, stream_) != hipSuccess &&
            rc == FELBM_OK) {
          set_error("state restore failed (autotune)"); rc = FELBM_EHIP;
        }
      if (hipStreamSynchronize(stream_) != hipSuccess && rc == FELBM_OK) { set_error("autotune synchronisation failed"); rc = FELBM_EHIP; }
      drop();
      T_ = T0;
      for (auto &s : slots_) if (s.used) s.stats_valid = 0;
      ++map_ver_;                     // a captured step graph has the old shapes
      if (rc != FELBM_OK) { std::copy(keep, keep + 5, bsi_); return rc; }
      for (int q = 1; q <= 4; ++q) {
        int best = 2;
        for (int bi = 0; bi < 3; ++bi) if (ms[bi][q - 1] < ms[best][q - 1]) best = bi;
        bsi_[q] = best;
      }
    }
    if (bs4) for (int q = 0; q < 4; ++q) bs4[q] = kBS[bsi_[q + 1]];
    if (ms12) for (int bi = 0; bi < 3; ++bi) for (int q = 0; q < 4; ++q) ms12[bi * 4 + q] = ms[bi][q];
    return FELBM_OK;
  }

  int set_block_sizes(const int *bs4) override {
    int idx[4];
    for (int q = 0; q < 4; ++q) {
      idx[q] = -1;
      for (int bi = 0; bi < 3; ++bi) if (bs4[q] == kBS[bi]) idx[q] = bi;
      if (idx[q] < 0) { set_error("sites per workgroup must be 64, 128 or 256"); return FELBM_EINVAL; }
    }
    for (int q = 0; q < 4; ++q) bsi_[q + 1] = idx[q];
    ++map_ver_;
    return FELBM_OK;
  }

private:
  static constexpr int kMaxAsyncSlots = 256;
  int dev_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[10] = {};
  std::vector<Slot<real>> slots_;
  std::vector<int> slot_desc_;      // slot -> descriptor index
  Inst<real> *d_desc_ = nullptr; int desc_cap_ = 0, ndesc_ = 0;
  int2 *d_map_ = nullptr;
  // block maps at each workgroup shape (sites per workgroup), per variant
  static constexpr int kBS[3] = { 64, 128, 256 };
  int2 *d_var_[3][8] = {};
  int nvar_[3][8] = {};
  int vsdi_[3][8] = {};             // the one descriptor of a variant map, or -1 (several instances)
  int bsi_[5] = { 2, 2, 2, 2, 2 };  // shape index of passes 1-4 (moments, c, mu, collide)
  int map_cap_ = 0;
  int nmap_ = 0;
  static int variant(const Slot<real> &s) { return (s.mrt ? 1 : 0) | (s.hbb ? 2 : 0) | ((s.open || s.obc) ? 4 : 0); }
  ObEntry *d_ent_all_ = nullptr;    // open-boundary nodes of every slot
  int n_ent_all_ = 0, ent_cap_ = 0;
  bool dirty_ = false, dirty_desc_ = false;
  unsigned T_ = 0;
  unsigned map_ver_ = 0;
  int export_off_ = 9;
  unsigned long long *pinned_stats_ = nullptr;
  bool stats_launched_ = false;
  hipGraphExec_t graph_exec_ = nullptr;
  int graph_steps_ = 0; unsigned graph_ver_ = ~0u, graph_T_ = 0;

  // A captured step graph holds its kernels' arguments by value: the tracer
  // buffers and counts, the block maps and shapes.  Dropped before any of them
  // is freed or changed; the next step_graph captures again.
  void drop_graph() {
    if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
    ++map_ver_;
  }

  int check(int slot) {
    if (slot < 0 || slot >= (int)slots_.size() || !slots_[slot].used) { set_error("invalid slot"); return FELBM_EINVAL; }
    return FELBM_OK;
  }
  unsigned cur_of(const Slot<real> &S) const { return (T_ ^ (unsigned)S.par0) & 1u; }
  // engine site of the caller's site i
  static size_t eng(const Slot<real> &S, size_t i) { return S.perm.empty() ? i : (size_t)S.perm[i]; }

  void set_params(Slot<real> &S) {
    const felbm_params &p = S.prm;
    Inst<real> &d = S.d;
    d.rho0 = (real)p.rho0; d.rho1 = (real)p.rho1; d.drho = d.rho0 - d.rho1;
    d.kappa = (real)p.kappa; d.four_beta = (real)(4.0 * p.beta); d.bnd_coeff = (real)p.boundary_coefficient;
    d.tau0 = (real)p.tau0; d.tau1 = (real)p.tau1; d.mobility = (real)p.mobility; d.mrt_lambda = (real)p.mrt_lambda;
    d.gx = (real)p.acceleration[0]; d.gy = (real)p.acceleration[1]; d.gz = (real)p.acceleration[2];
    d.fx = (real)p.forcing[0]; d.fy = (real)p.forcing[1]; d.fz = (real)p.forcing[2];
    d.ff = (real)p.forcing_factor;
    d.M0 = S.M0;
  }

  void release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    for (auto &s : slots_) if (s.used && s.base) (void)hipFree(s.base);
    for (auto &s : slots_) if (s.used && s.ob_base) (void)hipFree(s.ob_base);
    for (auto &s : slots_) if (s.used && s.tr_base) (void)hipFree(s.tr_base);
    for (auto &s : slots_) if (s.used && s.tr_sbase) (void)hipFree(s.tr_sbase);
    if (d_ent_all_) (void)hipFree(d_ent_all_);
    slots_.clear();
    if (d_desc_) (void)hipFree(d_desc_);
    if (d_map_) (void)hipFree(d_map_);
    if (pinned_stats_) (void)hipHostFree(pinned_stats_);
    for (auto &e : ev_) if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
  }

  // rebuild descriptor array and block maps after membership changes
  int rebuild() {
    slot_desc_.assign(slots_.size(), -1);
    std::vector<Inst<real>> descs;
    std::vector<int2> all, var[3][8];
    for (size_t q = 0; q < slots_.size(); ++q) {
      if (!slots_[q].used) continue;
      int di = (int)descs.size();
      slot_desc_[q] = di;
      descs.push_back(slots_[q].d);
      const int key = variant(slots_[q]);
      for (int b = 0; b < slots_[q].nblocks; ++b) all.push_back(make_int2(di, b * 256));
      // open variant: a 256-site group with no site within 2 voxels of a
      // non-periodic face, no open-boundary node and not site 0 is interior,
      // and every block of it, at any workgroup shape, is marked MAP_INTERIOR
      // (its workgroup runs the periodic code).  Deciding per 256-site group
      // keeps each site on the same code at every shape: in FP32 the two
      // bodies may round a term differently where FMA contraction differs.
      std::vector<uint8_t> edge, inner;
      if (key & 4) {
        const Slot<real> &S = slots_[q];
        edge = S.near_face;
        edge.resize((size_t)((S.n + 63) / 64), 0);
        for (const ObEntry &e : S.entries) edge[(size_t)(e.site / 64)] = 1;
        edge[0] = 1;
        inner.assign((size_t)((S.n + 255) / 256), 1);
        for (size_t gq = 0; gq < edge.size(); ++gq) if (edge[gq]) inner[gq / 4] = 0;
      }
      for (int bi = 0; bi < 3; ++bi) {
        for (int b = 0; b * kBS[bi] < slots_[q].n; ++b) {
          int x = di;
          if ((key & 4) && inner[(size_t)(b * kBS[bi] / 256)]) x |= MAP_INTERIOR;
          var[bi][key].push_back(make_int2(x, b * kBS[bi]));
        }
      }
    }
    auto pad = [](std::vector<int2> &v) { while (v.size() % 8) v.push_back(make_int2(-1, 0)); };
    for (int bi = 0; bi < 3; ++bi)
      for (int k = 0; k < 8; ++k) {
        // one instance, no interior marks: the kernel takes the descriptor
        // index as an argument and reads no map
        int one = var[bi][k].empty() ? -1 : var[bi][k][0].x;
        for (const int2 &e : var[bi][k]) if (e.x != one) { one = -1; break; }
        vsdi_[bi][k] = one;
      }
    pad(all); for (auto &vb : var) for (auto &v : vb) pad(v);
    nmap_ = (int)all.size();
    int need = nmap_;
    for (int bi = 0; bi < 3; ++bi)
      for (int k = 0; k < 8; ++k) { nvar_[bi][k] = (int)var[bi][k].size(); need += nvar_[bi][k]; }
    if ((int)descs.size() > desc_cap_) {
      if (d_desc_) FM_HIP(hipFree(d_desc_));
      desc_cap_ = std::max((int)descs.size(), 2 * desc_cap_);
      FM_HIP(hipMalloc((void **)&d_desc_, sizeof(Inst<real>) * desc_cap_));
    }
    if (need > map_cap_) {
      if (d_map_) FM_HIP(hipFree(d_map_));
      map_cap_ = std::max(need, 2 * map_cap_);
      FM_HIP(hipMalloc((void **)&d_map_, sizeof(int2) * map_cap_));
    }
    if (!descs.empty()) FM_HIP(hipMemcpyAsync(d_desc_, descs.data(), sizeof(Inst<real>) * descs.size(), hipMemcpyHostToDevice, stream_));
    ndesc_ = (int)descs.size();
    if (nmap_) FM_HIP(hipMemcpyAsync(d_map_, all.data(), sizeof(int2) * nmap_, hipMemcpyHostToDevice, stream_));
    {
      int2 *p = d_map_ + nmap_;
      for (int bi = 0; bi < 3; ++bi)
        for (int k = 0; k < 8; ++k) {
          d_var_[bi][k] = p;
          if (nvar_[bi][k])
            FM_HIP(hipMemcpyAsync(p, var[bi][k].data(), sizeof(int2) * nvar_[bi][k], hipMemcpyHostToDevice, stream_));
          p += nvar_[bi][k];
        }
    }
    // open-boundary nodes of every slot, one list for one launch per step
    {
      std::vector<ObEntry> all_e;
      for (size_t q = 0; q < slots_.size(); ++q) {
        if (!slots_[q].used || !slots_[q].obc) continue;
        slots_[q].e0 = (int)all_e.size();
        for (ObEntry e : slots_[q].entries) { e.desc = slot_desc_[q]; all_e.push_back(e); }
      }
      n_ent_all_ = (int)all_e.size();
      if (n_ent_all_ > ent_cap_) {
        if (d_ent_all_) FM_HIP(hipFree(d_ent_all_));
        ent_cap_ = std::max(n_ent_all_, 2 * ent_cap_);
        FM_HIP(hipMalloc((void **)&d_ent_all_, sizeof(ObEntry) * ent_cap_));
      }
      if (n_ent_all_) FM_HIP(hipMemcpyAsync(d_ent_all_, all_e.data(), sizeof(ObEntry) * n_ent_all_, hipMemcpyHostToDevice, stream_));
    }
    // single-instance maps (field downloads) carry the new descriptor indices
    std::vector<std::vector<int2>> single(slots_.size());
    for (size_t q = 0; q < slots_.size(); ++q) {
      if (!slots_[q].used) continue;
      single[q].assign((size_t)slots_[q].dmap_len, make_int2(-1, 0));
      for (int b = 0; b < slots_[q].nblocks; ++b) single[q][b] = make_int2(slot_desc_[q], b * 256);
      FM_HIP(hipMemcpyAsync(slots_[q].dmap, single[q].data(), single[q].size() * sizeof(int2), hipMemcpyHostToDevice, stream_));
    }
    FM_HIP(hipStreamSynchronize(stream_));
    dirty_ = false; dirty_desc_ = false;
    ++map_ver_;
    return FELBM_OK;
  }

  int flush() {
    if (dirty_) return rebuild();
    if (dirty_desc_) {
      std::vector<Inst<real>> descs;
      for (auto &s : slots_) if (s.used) descs.push_back(s.d);
      if (!descs.empty()) FM_HIP(hipMemcpyAsync(d_desc_, descs.data(), sizeof(Inst<real>) * descs.size(), hipMemcpyHostToDevice, stream_));
      FM_HIP(hipStreamSynchronize(stream_));
      dirty_desc_ = false;
    }
    return FELBM_OK;
  }

  int zero_stats() {
    for (auto &s : slots_) if (s.used) {
      hipLaunchKernelGGL(k_zero_words<real>, dim3((12 * NSLOT + 255) / 256), dim3(256), 0, stream_, s.d.acc + ACC_STATS, 12 * NSLOT);
      FM_LAUNCH();
    }
    return FELBM_OK;
  }

  template <bool MRT, bool HBB, bool OPEN>
  void launch_collide(dim3 grid, const int2 *map, int sdi, bool stats, int bs) {
    if constexpr (sizeof(real) == 4)
      hipLaunchKernelGGL((k_collide_f32<MRT, HBB, OPEN>), grid, dim3(bs), 0, stream_, (const Inst<float> *)d_desc_, map, sdi, T_, stats ? 1 : 0);
    else
      hipLaunchKernelGGL((k_collide<real, MRT, HBB, OPEN>), grid, dim3(bs), 0, stream_, d_desc_, map, sdi, T_, stats ? 1 : 0);
  }
  template <bool MRT, bool HBB, bool OPEN>
  void launch_variant(int stage, dim3 grid, const int2 *map, int sdi, bool stats, real *exp, int exp_inst, int bs) {
    if (stage == 2) hipLaunchKernelGGL((k_cfield<real, HBB, OPEN>), grid, dim3(bs), 0, stream_, d_desc_, map, sdi, T_, exp, exp_inst);
    else if (stage == 3) hipLaunchKernelGGL((k_mufield<real, HBB, OPEN>), grid, dim3(bs), 0, stream_, d_desc_, map, sdi, T_);
    else launch_collide<MRT, HBB, OPEN>(grid, map, sdi, stats, bs);
  }
  void launch_key(int key, int stage, dim3 grid, const int2 *map, int sdi, bool stats, real *exp = nullptr, int exp_inst = -1,
                  int bs = 256) {
    switch (key) {
      case 0: launch_variant<false, false, false>(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
      case 1: launch_variant<true,  false, false>(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
      case 2: launch_variant<false, true,  false>(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
      case 3: launch_variant<true,  true,  false>(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
      case 4: launch_variant<false, false, true >(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
      case 5: launch_variant<true,  false, true >(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
      case 6: launch_variant<false, true,  true >(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
      default: launch_variant<true,  true,  true >(stage, grid, map, sdi, stats, exp, exp_inst, bs); break;
    }
  }

  // the moments pass gathers through the neighbour rings: open (non-periodic)
  // variants need the in-box checks
  void launch_moments(int key, dim3 grid, const int2 *map, int sdi, int bs = 256) {
    if (key & 4) hipLaunchKernelGGL((k_moments<real, true>), grid, dim3(bs), 0, stream_, d_desc_, map, sdi, T_);
    else hipLaunchKernelGGL((k_moments<real, false>), grid, dim3(bs), 0, stream_, d_desc_, map, sdi, T_);
  }

  int launch_step(bool stats, bool timed = false) {
    if (stats) if (int e = zero_stats()) return e;
    dim3 blk(256);
    if (timed) FM_HIP(hipEventRecord(ev_[0], stream_));
    int bi = bsi_[1];
    for (int k = 0; k < 8; ++k) if (nvar_[bi][k]) { launch_moments(k, dim3(nvar_[bi][k]), d_var_[bi][k], vsdi_[bi][k], kBS[bi]); FM_LAUNCH(); }
    if (n_ent_all_) {
      hipLaunchKernelGGL(k_open_fields<real>, dim3((n_ent_all_ + 255) / 256), blk, 0, stream_, d_desc_, d_ent_all_, n_ent_all_, T_);
      FM_LAUNCH();
      hipLaunchKernelGGL(k_open_commit<real>, dim3((n_ent_all_ + 255) / 256), blk, 0, stream_, d_desc_, d_ent_all_, n_ent_all_);
      FM_LAUNCH();
    }
    if (ndesc_) {
      hipLaunchKernelGGL(k_mass_lambda<real>, dim3(ndesc_), dim3(64), 0, stream_, d_desc_, 0, ndesc_, T_);
      FM_LAUNCH();
    }
    if (timed) FM_HIP(hipEventRecord(ev_[1], stream_));
    bi = bsi_[2];
    for (int k = 0; k < 8; ++k) if (nvar_[bi][k]) { launch_key(k, 2, dim3(nvar_[bi][k]), d_var_[bi][k], vsdi_[bi][k], stats, nullptr, -1, kBS[bi]); FM_LAUNCH(); }
    if (timed) FM_HIP(hipEventRecord(ev_[2], stream_));
    bi = bsi_[3];
    for (int k = 0; k < 8; ++k) if (nvar_[bi][k]) { launch_key(k, 3, dim3(nvar_[bi][k]), d_var_[bi][k], vsdi_[bi][k], stats, nullptr, -1, kBS[bi]); FM_LAUNCH(); }
    // tracers move on this step's u snapshot before the collision pass
    // advances the step counter they key their draws with
    for (size_t q = 0; q < slots_.size(); ++q) {
      Slot<real> &S = slots_[q];
      if (!S.used || !S.tr_np) continue;
      const int nb = (S.tr_np + 255) / 256;
      hipLaunchKernelGGL(k_tracers<real>, dim3(nb), blk, 0, stream_, d_desc_, slot_desc_[q],
                         S.tr_pos, (const double *)S.tr_dm, S.tr_blocked, S.tr_np, S.tr_seed, S.tr_substeps, S.st);
      FM_LAUNCH();
      if (S.st.rows) {
        hipLaunchKernelGGL(k_tracer_row<real>, dim3(1), dim3(64), 0, stream_, d_desc_, slot_desc_[q], S.st, nb, S.tr_np);
        FM_LAUNCH();
      }
    }
    if (timed) FM_HIP(hipEventRecord(ev_[3], stream_));
    bi = bsi_[4];
    for (int k = 0; k < 8; ++k) if (nvar_[bi][k]) { launch_key(k, 4, dim3(nvar_[bi][k]), d_var_[bi][k], vsdi_[bi][k], stats, nullptr, -1, kBS[bi]); FM_LAUNCH(); }
    if (timed) FM_HIP(hipEventRecord(ev_[4], stream_));
    ++T_;
    return FELBM_OK;
  }

  // lambda owed by the state (a correction is pending after any step)
  int owed_lambda(Slot<real> &S, bool &apply, double &lambda) {
    apply = false; lambda = 0.0;
    if (!S.mass_on) return FELBM_OK;
    unsigned long long pend = 0;
    FM_HIP(hipMemcpyAsync(&pend, S.d.acc + ACC_FLAGS, 8, hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    if (!pend) return FELBM_OK;
    double M, W; uint64_t bad;
    if (int e = mass_sums(S, false, 0.0, M, W, bad)) return e;
    double drift = M - S.M0;
    apply = (W > 0.0 && drift != 0.0);
    lambda = apply ? drift / W : 0.0;
    return FELBM_OK;
  }

  int mass_sums(Slot<real> &S, bool apply, double lambda, double &M, double &W, uint64_t &bad) {
    FM_HIP(hipMemsetAsync(S.d.acc + ACC_MISC, 0, 5 * NSLOT * 8, stream_));
    hipLaunchKernelGGL(k_mass_single<real>, dim3((S.n + 255) / 256), dim3(256), 0, stream_, S.d, cur_of(S), apply ? 1 : 0, lambda);
    FM_LAUNCH();
    std::vector<unsigned long long> a(5 * NSLOT);
    FM_HIP(hipMemcpyAsync(a.data(), S.d.acc + ACC_MISC, a.size() * 8, hipMemcpyDeviceToHost, stream_));
    FM_HIP(hipStreamSynchronize(stream_));
    unsigned long long s[5] = {0, 0, 0, 0, 0};
    for (int q = 0; q < 5; ++q) for (int l = 0; l < NSLOT; ++l) s[q] += a[q * NSLOT + l];
    M = fx_value((long long)s[0], s[1]); W = fx_value((long long)s[2], s[3]); bad = s[4];
    return FELBM_OK;
  }

  int reduce_stats(const unsigned long long *a, double *out) {
    unsigned long long s[12] = {0};
    for (int q = 0; q < 11; ++q) for (int l = 0; l < NSLOT; ++l) s[q] += a[q * NSLOT + l];
    for (int l = 0; l < NSLOT; ++l) s[11] = std::max(s[11], a[11 * NSLOT + l]);
    for (int q = 0; q < 5; ++q) out[q] = fx_value((long long)s[2 * q], s[2 * q + 1]);
    double mx; std::memcpy(&mx, &s[11], 8); out[5] = mx;
    if (s[10]) { set_error("non-finite or out-of-range value in the monitored fields"); return FELBM_ENONFINITE; }
    return FELBM_OK;
  }

  // field pass on one instance: moments + c stencil with export (download_raw)
  int field_pass(Slot<real> &S) {
    if (int e = flush()) return e;
    const int di = slot_desc_[&S - slots_.data()];
    const size_t ns = (size_t)S.ns;
    dim3 blk(256), grid(S.dmap_len);
    launch_moments(variant(S), grid, S.dmap, di); FM_LAUNCH();
    if (S.obc && !S.entries.empty()) {
      const int ne = (int)S.entries.size();
      hipLaunchKernelGGL(k_open_fields<real>, dim3((ne + 255) / 256), blk, 0, stream_, d_desc_, d_ent_all_ + S.e0, ne, T_);
      FM_LAUNCH();
      hipLaunchKernelGGL(k_open_commit<real>, dim3((ne + 255) / 256), blk, 0, stream_, d_desc_, d_ent_all_ + S.e0, ne);
      FM_LAUNCH();
    }
    hipLaunchKernelGGL(k_mass_lambda<real>, dim3(1), dim3(64), 0, stream_, d_desc_, di, 1, T_);
    FM_LAUNCH();
    launch_key(variant(S), 2, grid, S.dmap, di, false, S.d.craw + export_off_ * ns, di); FM_LAUNCH();
    // the step's own moments pass must find this parity's partials at zero
    unsigned cur = cur_of(S);
    FM_HIP(hipMemsetAsync(S.d.acc + ACC_MASS + cur * 5 * NSLOT, 0, 5 * NSLOT * 8, stream_));
    return FELBM_OK;
  }

  int stats_from_fields(Slot<real> &S, double *out) {
    std::vector<double> f((size_t)6 * S.n);
    if (int e = download_fields((int)(&S - slots_.data()), f.data())) return e;
    // exact fixed-point sums on the host, same definition as the kernels
    long long hi[5] = {0}; unsigned long long lo[5] = {0}, bad = 0; double mx = 0;
    const size_t n = (size_t)S.n;
    for (size_t i = 0; i < n; ++i) {
      double vx = f[2 * n + i], vy = f[3 * n + i], vz = f[4 * n + i];
      double v2 = vx * vx + vy * vy + vz * vz;
      double t[5] = { vx, vy, vz, v2, f[n + i] };
      for (int q = 0; q < 5; ++q) {
        if (!(std::fabs(t[q]) < 8.0)) { ++bad; continue; }
        long long qq = (long long)std::rint(t[q] * 0x1p60);
        hi[q] += qq >> 32; lo[q] += (unsigned long long)(qq & 0xffffffffLL);
      }
      if (v2 > mx) mx = v2;
    }
    for (int q = 0; q < 5; ++q) out[q] = fx_value(hi[q], lo[q]);
    out[5] = mx;
    if (bad) { set_error("non-finite or out-of-range value in the monitored fields"); return FELBM_ENONFINITE; }
    return FELBM_OK;
  }
};

template <typename real>
EngineBase *make_engine(int device, int *err) {
  Engine<real> *e = new Engine<real>(device);
  if (e->init(err) != FELBM_OK) { delete e; return nullptr; }
  return e;
}

}  // namespace fm
