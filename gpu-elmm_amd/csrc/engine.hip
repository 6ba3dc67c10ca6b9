// engine.hip - device-resident ELMM LES step: buffers, set-up of ordered
// correction lists, stage sequencing and the C ABI.
//
// One elmm_handle = one case with all fields resident in HBM.  A step is the
// reference's TMarchRK3 (src/tsteps.f90:24-243) issued as a fixed sequence
// of kernels on the handle's stream; nothing returns to the host between
// stages.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "case_setup.hpp"
#include "device_common.hpp"
#include "elmm_gpu.h"
#include "errors.hpp"
#include "kernels.hpp"

namespace elmm {

// ------------------------------------------- host operations vs. capture
namespace {
std::shared_mutex g_capture_mu;
thread_local bool t_capturing = false;
thread_local int t_host_ops = 0;

// held by a thread around one graph capture
struct CaptureLock {
  CaptureLock() {
    g_capture_mu.lock();
    t_capturing = true;
  }
  ~CaptureLock() {
    t_capturing = false;
    g_capture_mu.unlock();
  }
  CaptureLock(const CaptureLock &) = delete;
  CaptureLock &operator=(const CaptureLock &) = delete;
};
}  // namespace

HostOpLock::HostOpLock() : own_(!t_capturing && t_host_ops == 0) {
  if (own_) g_capture_mu.lock_shared();
  t_host_ops++;
}
HostOpLock::~HostOpLock() {
  t_host_ops--;
  if (own_) g_capture_mu.unlock_shared();
}



#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)


struct HipError {
  std::string msg;
  explicit HipError(std::string m) : msg(std::move(m)) {}
};
struct ArgError {
  std::string msg;
  explicit ArgError(std::string m) : msg(std::move(m)) {}
};

namespace {
// A thread started by a thread that holds a HostOpLock, and joined before
// that lock is released, works under it.
struct HostOpInherit {
  HostOpInherit() { t_host_ops++; }
  ~HostOpInherit() { t_host_ops--; }
};

// Runs the tasks on threads of their own, on device dev and under the
// calling thread's HostOpLock, and rethrows the first task's exception.
void run_parallel(std::vector<std::function<void()>> &tasks, int dev) {
  std::vector<std::exception_ptr> err(tasks.size());
  std::vector<std::thread> th;
  th.reserve(tasks.size());
  for (size_t q = 0; q < tasks.size(); q++)
    th.emplace_back([&, q] {
      HostOpInherit inherit;
      try {
        HIPCHK(hipSetDevice(dev));
        tasks[q]();
      } catch (...) {
        err[q] = std::current_exception();
      }
    });
  for (auto &t : th) t.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}
}  // namespace

int current_device() {
  int d = 0;
  return hipGetDevice(&d) == hipSuccess ? d : 0;
}
void use_device(int d) { (void)hipSetDevice(d); }

static const double RK_alpha[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};
static const double RK_beta[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double RK_rho[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

// Device allocation owned by the engine.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t b) {
    bytes = b ? b : 16;
    HIPCHK(hipMalloc(&p, bytes));
    HIPCHK(hipMemset(p, 0, bytes));
  }
  template <typename T> T *as() const { return (T *)p; }
};

template <typename T>
static void upload_vec(DevBuf &b, const std::vector<T> &v) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

struct DevTerms {
  DevBuf cell, ptr, term;
  int n = 0;
  TermList view() const { return TermList{n, cell.as<int32_t>(), ptr.as<int32_t>(), term.as<int32_t>()}; }
};

// Set-up phase times on stderr when ELMM_SETUP_TRACE is set (each phase ends
// with a device synchronisation, so its time includes its device work).
struct SetupTrace {
  bool on = std::getenv("ELMM_SETUP_TRACE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char *what) {
    if (!on) return;
    (void)hipDeviceSynchronize();
    const auto n = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[elmm setup] %-14s %8.2f ms\n", what, std::chrono::duration<double, std::milli>(n - t).count());
    t = n;
  }
};

// Host-side accumulation of ordered per-cell terms, keyed by target cell
// (the element index of the cell in the field layout, 0 <= key < nkeys):
// cells in order of first appearance, each cell's terms in insertion order.
// The key -> cell map is a dense array from calloc, whose untouched pages
// cost nothing; terms are kept flat and grouped per cell by a stable
// counting sort.
struct TermBuilder {
  using Term = std::array<int32_t, 4>;
  std::vector<std::array<int, 3>> cells;
  std::vector<long> keys;                    // per cell
  std::vector<std::pair<int, Term>> items;   // (cell, term), insertion order
  long nkeys;
  std::unique_ptr<int, decltype(&std::free)> slot{nullptr, &std::free};  // key -> cell + 1
  explicit TermBuilder(long n) : nkeys(n) {}
  int find(long key) const { return slot ? slot.get()[key] - 1 : -1; }
  int cell_of(long key, int i, int j, int k) {
    if (key < 0 || key >= nkeys) throw ArgError("correction term outside the field allocation");
    if (!slot) {
      slot.reset(static_cast<int *>(std::calloc((size_t)nkeys, sizeof(int))));
      if (!slot) throw std::bad_alloc();
    }
    int &s = slot.get()[key];
    if (s == 0) {
      cells.push_back({i, j, k});
      keys.push_back(key);
      s = (int)cells.size();
    }
    return s - 1;
  }
  void add(long key, int i, int j, int k, const Term &t) { items.emplace_back(cell_of(key, i, j, k), t); }
  // CSR form: ptr (cells + 1) and the terms grouped per cell
  void csr(std::vector<int32_t> &ptr, std::vector<int32_t> &flat) const {
    ptr.assign(cells.size() + 1, 0);
    for (auto &it : items) ptr[it.first + 1]++;
    for (size_t q = 0; q < cells.size(); q++) ptr[q + 1] += ptr[q];
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    flat.assign(4 * items.size(), 0);
    for (auto &it : items) std::memcpy(&flat[4 * (size_t)pos[it.first]++], it.second.data(), sizeof(Term));
  }
  // Uploads the cells in increasing field offset (consecutive threads of the
  // kernels that apply the terms then touch neighbouring cells; every cell
  // is applied independently of the others), each cell's terms in order.
  void upload(DevTerms &d) const {
    std::vector<int32_t> p, t;
    csr(p, t);
    std::vector<int> ord(cells.size());
    for (size_t q = 0; q < ord.size(); q++) ord[q] = (int)q;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    std::vector<int32_t> c(3 * cells.size()), p2{0}, t2;
    t2.reserve(t.size());
    for (size_t r = 0; r < ord.size(); r++) {
      const int q = ord[r];
      for (int e = 0; e < 3; e++) c[3 * r + e] = cells[q][e];
      t2.insert(t2.end(), t.begin() + 4 * (size_t)p[q], t.begin() + 4 * (size_t)p[q + 1]);
      p2.push_back((int32_t)(t2.size() / 4));
    }
    d.n = (int)cells.size();
    upload_vec(d.cell, c);
    upload_vec(d.ptr, p2);
    upload_vec(d.term, t2);
  }
};

// Type-independent part of an engine: configuration, stream, state flags
// and the per-group kernel timers.  The C ABI talks to this interface.
struct EngineIface {
  elmm_config cfg{};
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at creation (owns the stream and buffers)
  bool initialized = false, first_step_done = false;
  double last_div = 0, last_cfl = 0, last_res = 0, last_stored_div = 0, last_delta = 0;
  int last_refinements = 0;
  bool want_diag = false;
  // time stepping (src/main.f90:81-128): the current dt, the simulated time,
  // steps since creation, steps enqueued since the last synchronisation and
  // the first step diagnosed as diverging (-1 none, 0 none yet)
  double dt_cur = 0, time = 0;
  long long steps_total = 0;
  int steps_pending = 0;
  long long diverged_step = -1;
  std::string diverged_msg;
  // device record of the CFL monitor and the delta (k_monitor.hip)
  StepMon *mon = nullptr;
  // kernel timers (HIP events on the stream, read after each sync)
  bool timers_on = false;
  std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> event_pool;
  double tms[ELMM_T_NGROUPS] = {};
  int64_t tlaunch[ELMM_T_NGROUPS] = {};
  // two-step graphs per field-pointer parity (captured on first use)
  bool use_graphs = true;
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};

  virtual ~EngineIface() = default;
  virtual void create(const elmm_config &c, const elmm_lists &L) = 0;
  virtual void xfer(int field, int lb, const int32_t ext[3], double *host, bool up) = 0;
  virtual void init() = 0;
  virtual void step(int nsteps) = 0;
  virtual void enqueue_steps(int nsteps) = 0;
  virtual double divergence() = 0;
  // divergence() with caller-provided device memory for the block maxima
  // (at least partial_slots doubles) and the final maximum (result)
  virtual double divergence_into(double *partials, long partial_slots, double *result) = 0;
  virtual long divergence_blocks() const = 0;
  // WMPoints%ustar (count; copied to out when out is not null and cap
  // suffices; synchronises the stream)
  virtual int point_ustar(double *out, long cap) = 0;
  // asynchronous field frame (elmm_frame_async / elmm_frame_wait)
  virtual void frame_async(const std::string &dir) = 0;
  virtual int frame_wait() = 0;
  virtual std::string frame_error_text() const = 0;
  // tile shapes and z chunks in use (elmm_get_tiles)
  virtual void tiles(elmm_tiles *out) const = 0;
  // Reads the device step record (synchronises the stream): CFL and delta
  // of the last step, divergence by CFL_max (src/main.f90:161-168).
  void read_monitor() {
    if (!mon) return;
    StepMon m{};
    HIPCHK(hipMemcpyAsync(&m, mon, sizeof m, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (cfg.cfl_check && !cfg.variable_dt) last_cfl = m.cfl_last;
    if (cfg.delta_diag) last_delta = m.delta_last;
    if (m.first_bad >= 0 && diverged_step < 0) {
      diverged_step = m.first_bad;
      diverged_msg = "Solution diverged. 2 (CFL " + std::to_string(m.cfl_last) + " above CFL_max at step " +
                     std::to_string(m.first_bad) + ")";
    }
  }
  // main.f90's time update after a step of length dt (src/main.f90:116-127)
  void advance_time(double dt) {
    steps_total++;
    if (cfg.variable_dt) {
      time = time + dt;
      if (cfg.end_time > 0 && std::fabs(time - cfg.end_time) < 5 * spacing(cfg.end_time)) time = cfg.end_time;
    } else {
      time = cfg.start_time + (double)steps_total * dt;
      if (cfg.end_time > 0 && std::fabs(time - cfg.end_time) < dt / 10) time = cfg.end_time;
    }
  }
  static double spacing(double x) { return std::nextafter(std::fabs(x), INFINITY) - std::fabs(x); }
  void release_graphs() {
    for (auto &e : graph_exec)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  }

  hipEvent_t get_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    HostOpLock host;
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  // Runs f (which enqueues exactly one kernel) between two events when the
  // timers are on.
  template <typename F> void timed(int grp, F &&f) {
    if (!timers_on) {
      f();
      return;
    }
    hipEvent_t a = get_event(), b = get_event();
    HIPCHK(hipEventRecord(a, stream));
    f();
    HIPCHK(hipEventRecord(b, stream));
    pending.emplace_back(grp, a, b);
  }
  void collect_timers() {
    for (auto &t : pending) {
      hipEvent_t a = std::get<1>(t), b = std::get<2>(t);
      HIPCHK(hipEventSynchronize(b));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, a, b));
      tms[std::get<0>(t)] += ms;
      tlaunch[std::get<0>(t)] += 1;
      event_pool.push_back(a);
      event_pool.push_back(b);
    }
    pending.clear();
  }
};

// T: field storage and stencil arithmetic type.  double: the FP64 form (the
// reference's arithmetic); float: the precision form (FP32 fields and
// stencils, FP64 divergence/RHS, Phi, residual, reductions and corrections,
// FP32 spectral solves inside FP64 iterative refinement).
template <typename T>
struct Engine : EngineIface {
  std::vector<double> zw_copy;
  Geo g{};
  Grid1D grid;
  int nsc = 0;  // scalar slots: the passive scalars, the temperature (buoyancy), the moisture
  int tslot = -1;  // slot of the temperature, -1 without buoyancy
  int mslot = -1;  // slot of the moisture, -1 without it
  Geo gtemp{};     // g with the temperature's boundary conditions
  Geo gmoist{};    // g with the moisture's boundary conditions
  // g with the boundary conditions of scalar slot s
  const Geo &gscal(int s) const { return s == tslot ? gtemp : s == mslot ? gmoist : g; }
  DevBuf tfix_ijk;  // solid cells whose temperature is the mean of fluid neighbours
  int tfix_n = 0;
  // fields
  DevBuf U, V, W, U2, V2, W2, Us, Vs, Ws, Pr, Visc, TDiff, Q, Phi, RHS;
  std::vector<std::unique_ptr<DevBuf>> S, Snew, Sadv;
  T *u, *v, *w, *u2, *v2, *w2;  // current rotation of U/U2
  std::vector<T *> s_cur, s_nxt;
  // masks and codes
  DevBuf code[3], code_s, mask[3], prtype;
  // coordinates
  DevBuf coords;
  // wall model
  DevBuf wm_ijk, wm_list, wm_dist, wm_z0, wm_ustar, wm_fluxp, wm_fluxm, wm_logdz0;
  DevBuf wm_tclass, wm_tval, wm_tflux, wmp_tclass, wmp_tval, wmp_tflux, wmp_af;
  DevBuf wm_qval, wm_qflux, wmp_qval, wmp_qflux;
  bool mo_any = false;  // some wall-model point has a thermal class (buoyancy)
  DevBuf width_tab, width_idx, sgs_rdz, sgs_rz2, sgs_rdzf, sgs_rz2f;
  SgsWidth sgs_width{nullptr, nullptr, 0};
  int wm_n = 0;
  DevBuf wmp_ijk, wmp_div, wmp_dist, wmp_z0, wmp_logdz0, wmp_ustar;
  int wmp_n = 0;
  MOParams mo_params(const DevBuf &tc, const DevBuf &tv, const DevBuf &tf, const DevBuf &qv,
                     const DevBuf &qf) const {
    return MOParams{cfg.enable_buoyancy ? 1 : 0, tc.as<int8_t>(), tv.as<double>(), tf.as<double>(), cfg.grav_acc,
                    cfg.temperature_ref, cfg.nu_scalar, mslot >= 0 ? 1 : 0, qv.as<double>(), qf.as<double>()};
  }
  // the temperature's wall source tflux * af goes to src_flux[src_n + p], the
  // moisture's qflux * af to src_flux[src_n + wmp_n + p]
  WMUstarPoints wmpdev() const {
    return WMUstarPoints{wmp_n,
                         wmp_ijk.as<int32_t>(),
                         wmp_dist.as<double>(),
                         wmp_z0.as<double>(),
                         wmp_logdz0.as<double>(),
                         wmp_ustar.as<double>(),
                         mo_params(wmp_tclass, wmp_tval, wmp_tflux, wmp_qval, wmp_qflux),
                         wmp_af.as<double>(),
                         tslot >= 0 ? src_flux.as<double>() + src_n : nullptr,
                         mslot >= 0 ? src_flux.as<double>() + src_n + wmp_n : nullptr};
  }
  int src_n = 0;
  const T *temp_cur() const { return tslot >= 0 ? s_cur[tslot] : nullptr; }
  const T *moist_cur() const { return mslot >= 0 ? s_cur[mslot] : nullptr; }
  DevTerms mom_terms[3];
  std::vector<std::unique_ptr<DevTerms>> sc_terms;
  DevBuf src_flux;
  // immersed boundary
  DevBuf ib_tgt[4], ib_ptr[4], ib_src[4], ib_coef[4], ib_tmp;
  int ib_n[4] = {0, 0, 0, 0};
  std::vector<int32_t> visc_levels;
  DevTerms q_terms;
  DevBuf qmask;  // per (j, k) row: bit s set when cells 64s+1..64s+64 may hold a non-zero Q
  DevBuf null_ijk[3];
  int null_n[3] = {0, 0, 0};
  // pressure
  DevBuf X, twx, twy, lamx, lamy, az64, cz64;
  PoissonPlan<T> pp{};
  // precision form without refinement sweeps: the spectral solve runs in
  // FP64 (complex double array) on the FP32 fields' FP64 divergence
  bool solve64 = false;
  DevBuf X64, twx64, twy64;
  PoissonPlan<double> pp64{};
  // geometric multigrid (poisson_solver = ELMM_POISSON_MG, k_multigrid.hip):
  // dense level arrays, line-solve tables, the coarsest level's work arrays
  bool use_mg = false;
  std::vector<MgLevel> mg;
  std::vector<std::unique_ptr<DevBuf>> mg_bufs;
  DevBuf mg_tab, mg_dz, mg_cs, mg_lam, mg_Y, mg_X, mg_cp, mg_partial;
  MgCoarse mgc{};
  int mg_cycles = 0;
  double mg_res_last = 0;
  void setup_mg();
  double mg_resmax0();
  void mg_vcycle(int l);
  void mg_solve();
  DevBuf partial, cfl_partial, div_partial;
  int top_blocks = 0, red_blocks = 0, cfl_blocks = 0, cfl_cap = 0;
  int refine_iters = 0;
  // momentum kernel selection: the fused z-marching kernel on uniform grids
  // (ELMM_MOMENTUM=percomp selects the per-component kernels), its z chunk
  // length (ELMM_ZCHUNK, default 64 levels)
  bool fused_momentum = false;
  int zchunk = 64;
  // scalar kernel: z-marching on uniform grids (ELMM_SCALAR=cellwise selects
  // the one-thread-per-cell kernel), 16 levels per block (measured 7 % faster
  // than 64 in FP32: more blocks to hide the per-level load latency;
  // ELMM_ZCHUNK sets both chunk lengths)
  bool zmarch_scalar = false;
  int zchunk_scalar = 16;
  // fused momentum tile (ELMM_MOM_TILE=32x8 / 64x8 / 64x4 fixes it)
  int mom_tile = kMomTile64x4;
  // settings fixed by the environment are not tuned
  bool env_mom_tile = false, env_zchunk = false, env_scalar_zchunk = false, env_sgs_zchunk = false;
  bool tuned = false;
  void autotune();
  StageConst stage_const(int st) const;

  IBDev ibdev(int c) const {
    return IBDev{ib_n[c], ib_tgt[c].as<int32_t>(), ib_ptr[c].as<int32_t>(), ib_src[c].as<int32_t>(),
                 ib_coef[c].as<double>()};
  }
  WMPoints wmdev() {
    return WMPoints{wm_n,
                    wm_ijk.as<int32_t>(),
                    wm_list.as<int32_t>(),
                    wm_dist.as<double>(),
                    wm_z0.as<double>(),
                    wm_ustar.as<double>(),
                    wm_fluxp.as<double>(),
                    wm_fluxm.as<double>(),
                    wm_logdz0.as<double>(),
                    mo_params(wm_tclass, wm_tval, wm_tflux, wm_qval, wm_qflux)};
  }
  bool has_ib() const { return ib_n[0] + ib_n[1] + ib_n[2] > 0; }

  void create(const elmm_config &c, const elmm_lists &L) override;
  void setup_geo();
  void setup_lists(const elmm_lists &L);
  void setup_poisson();
  void xfer(int field, int lb, const int32_t ext[3], double *host, bool up) override;
  void bound_uvw(T *a, T *b, T *c) {
    timed(ELMM_T_GHOST, [&] { launch_bound_uvw<T>(g, a, b, c, stream); });
  }
  // BoundUVW after IBMomentum: only the ghost cells whose source is an IB
  // point change, so only those are rewritten (lists built at set-up;
  // ELMM_GHOST_DELTA=0 runs the full fill)
  DevBuf gdelta;
  int gdelta_n[3] = {0, 0, 0};
  bool use_gdelta = false;
  void bound_uvw_after_ib(T *a, T *b, T *c) {
    if (use_gdelta)
      timed(ELMM_T_GHOST, [&] { launch_ghost_delta<T>(a, b, c, gdelta.as<long long>(), gdelta_n, stream); });
    else
      bound_uvw(a, b, c);
  }
  void setup_ghost_delta();
  void ib_momentum(T *a, T *b, T *c);
  void pressure_correction(T *a, T *b, T *c, bool use_q, double coef, bool diag = false);
  void poisson_solve(int accumulate);
  void subgrid_stresses(bool cfl);
  // variable time steps: the first stage's SGS kernel (which also writes the
  // CFL block maxima) is launched before the step to find dt
  bool sgs_prelaunched = false;
  void choose_dt();
  DevBuf mon_buf, delta_partial;
  int dblocks = 0;
  void stage(int st, bool last);
  void init() override;
  void run_steps(int nsteps, bool sync_diag);
  void step(int nsteps) override;
  void enqueue_steps(int nsteps) override;
  void first_call();
  int replay_steps(int nsteps);
  double divergence() override;
  double divergence_into(double *partials, long partial_slots, double *result) override;
  long divergence_blocks() const override { return red_blocks; }
  int point_ustar(double *out, long cap) override {
    if (out && cap >= wmp_n && wmp_n > 0) {
      HIPCHK(hipMemcpyAsync(out, wmp_ustar.p, (size_t)wmp_n * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    return wmp_n;
  }
  // frames: pinned host staging of U, V, W, Pr and the scalars in the device
  // layout, filled by copies on the engine's stream; a writer thread waits
  // for them and writes the reference's VTK files
  std::vector<void *> frame_host;
  hipEvent_t frame_done = nullptr;
  std::thread frame_thread;
  int frame_status = 0;
  std::string frame_error;
  void frame_async(const std::string &dir) override;
  int frame_wait() override;
  std::string frame_error_text() const override { return frame_error; }
  ~Engine() override {
    if (frame_thread.joinable()) frame_thread.join();
    for (void *p : frame_host) (void)hipHostFree(p);
    if (frame_done) (void)hipEventDestroy(frame_done);
  }
  void tiles(elmm_tiles *out) const override {
    std::memset(out, 0, sizeof *out);
    out->mom_tx = !fused_momentum ? 0 : mom_tile == kMomTile32x8 ? 32 : 64;
    out->mom_ty = !fused_momentum ? 0 : mom_tile == kMomTile64x4 ? 4 : 8;
    if (sizeof(T) == 8 && mom_tile == kMomTile64x8) out->mom_tx = 32;
    out->mom_zchunk = fused_momentum ? zchunk : 0;
    out->sgs_zchunk = g.sgs_zchunk > 0 ? g.sgs_zchunk : 26;
    out->scalar_zchunk = zmarch_scalar ? zchunk_scalar : 0;
    out->tuned = tuned ? 1 : 0;
  }
  double max_div(T *a, T *b, T *c, bool use_q);
  double reduce_max(DevBuf &buf, int n);
};

template <typename T>
void Engine<T>::setup_geo() {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  component_extents(cfg, g.n);
  long row = kFrontPad + nx + 6;
  g.sy = (row + kRowAlign - 1) / kRowAlign * kRowAlign;
  g.sz = g.sy * (ny + 6);
  g.base = kFrontPad + 2 + 2 * g.sy + 2 * g.sz;
  g.ntot = g.sz * (nz + 6) + kRowAlign;
  grid = make_grid(cfg);
  g.dxmin = grid.dxmin;
  g.dymin = grid.dymin;
  g.dzmin = grid.dzmin;
  g.rdx = div_recip(g.dxmin);
  g.rdy = div_recip(g.dymin);
  g.rdz = div_recip(g.dzmin);
  g.rdxf = div_recipf((float)g.dxmin);
  g.rdyf = div_recipf((float)g.dymin);
  g.rdzf = div_recipf((float)g.dzmin);
  g.variable_z = cfg.stretched_z != 0;
  {
    const char *m = std::getenv("ELMM_MOMENTUM");
    fused_momentum = momentum_zm_supported(g) && !(m && std::string(m) == "percomp");
    const char *sg = std::getenv("ELMM_SGS");
    g.sgs_cellwise = sg && std::string(sg) == "cellwise";
    const char *sz = std::getenv("ELMM_SGS_ZCHUNK");
    g.sgs_zchunk = sz ? std::max(0, std::atoi(sz)) : 0;
    env_sgs_zchunk = g.sgs_zchunk > 0;
    if (const char *mt = std::getenv("ELMM_MOM_TILE")) {
      const std::string v = mt;
      mom_tile = v == "32x8" ? kMomTile32x8 : v == "64x8" ? kMomTile64x8 : kMomTile64x4;
      env_mom_tile = true;
    }
    const char *sk = std::getenv("ELMM_SCALAR");
    zmarch_scalar = !g.variable_z && !(sk && std::string(sk) == "cellwise");
    const char *gv = std::getenv("ELMM_GRAPH");
    use_graphs = !(gv && std::string(gv) == "0");
    const char *zc = std::getenv("ELMM_ZCHUNK");
    if (zc && std::atoi(zc) > 0) {
      zchunk = zchunk_scalar = std::atoi(zc);
      env_zchunk = env_scalar_zchunk = true;
    }
    const char *zs = std::getenv("ELMM_SCALAR_ZCHUNK");
    if (zs && std::atoi(zs) > 0) {
      zchunk_scalar = std::atoi(zs);
      env_scalar_zchunk = true;
    }
  }
  for (int s = 0; s < 6; s++) {
    g.bt[s] = cfg.btype[s];
    g.sbt[s] = cfg.scal_btype[s];
    g.side_scal[s] = cfg.side_scal[s];
  }
  // coordinate tables: dxPr dyPr dzPr dzW zPr zW, each n+8 with offset 3
  std::vector<double> tab;
  size_t off[6];
  auto push = [&](const std::vector<double> &v, int k) {
    off[k] = tab.size();
    tab.insert(tab.end(), v.begin(), v.end());
  };
  push(grid.dxPr, 0);
  push(grid.dyPr, 1);
  push(grid.dzPr, 2);
  push(grid.dzW, 3);
  push(grid.zPr, 4);
  push(grid.zW, 5);
  upload_vec(coords, tab);
  const double *base = coords.as<double>();
  g.dxPr = base + off[0] + 3;
  g.dyPr = base + off[1] + 3;
  g.dzPr = base + off[2] + 3;
  g.dzW = base + off[3] + 3;
  g.zPr = base + off[4] + 3;
  g.zW = base + off[5] + 3;
  // SGS filter width (src/subgrid.f90:62): filter_ratio *
  // (dxPr*dyPr*dzPr)**(1/3), from the host libm like the reference's.  The
  // products dxPr(i)*dyPr(j) take few distinct values (one on exactly
  // representable spacings), so the width is tabulated per distinct product
  // and level, with a column -> product index (uint8).
  {
    std::vector<double> prods;
    std::vector<uint8_t> idx((size_t)nx * ny, 0);
    bool fits = true;
    for (int j = 1; j <= ny && fits; j++)
      for (int i = 1; i <= nx && fits; i++) {
        const double d = Grid1D::at(grid.dxPr, i) * Grid1D::at(grid.dyPr, j);
        size_t q = 0;
        while (q < prods.size() && prods[q] != d) q++;
        if (q == prods.size()) {
          if (prods.size() == 255) {
            fits = false;
            break;
          }
          prods.push_back(d);
        }
        idx[(size_t)(i - 1) + (size_t)nx * (j - 1)] = (uint8_t)q;
      }
    if (fits) {
      const int ld = nz + 2;
      std::vector<double> wk(prods.size() * ld, 0.0);
      for (size_t q = 0; q < prods.size(); q++)
        for (int k = 1; k <= nz; k++)
          wk[q * ld + k] = cfg.filter_ratio * std::pow(prods[q] * Grid1D::at(grid.dzPr, k), 1.0 / 3.0);
      upload_vec(width_tab, wk);
      if (prods.size() > 1) upload_vec(width_idx, idx);
      sgs_width = SgsWidth{width_tab.as<double>(), prods.size() > 1 ? width_idx.as<uint8_t>() : nullptr, ld};
    }
  }
  // reciprocals of the strain divisors (FP64 form): used when every one
  // satisfies div_rn's condition, so that the quotients equal the divisions
  if (sgs_width.tab) {
    const int ld = nz + 2;
    std::vector<double> rdz(ld, 0.0), rz2(ld, 0.0);
    bool ok = g.rdx != 0 && g.rdy != 0;
    for (int k = 1; k <= nz && ok; k++) {
      rdz[k] = div_recip(Grid1D::at(grid.dzPr, k));
      rz2[k] = div_recip(2 * (Grid1D::at(grid.zPr, k + 1) - Grid1D::at(grid.zPr, k - 1)));
      ok = rdz[k] != 0 && rz2[k] != 0;
    }
    if (ok) {
      upload_vec(sgs_rdz, rdz);
      upload_vec(sgs_rz2, rz2);
      sgs_width.rdx = g.rdx;
      sgs_width.rdy = g.rdy;
      sgs_width.rdz = sgs_rdz.as<double>();
      sgs_width.rz2 = sgs_rz2.as<double>();
    }
    // the precision form's level reciprocals (single-precision divisions,
    // the values the per-cell kernel computes in place)
    std::vector<float> rdzf(ld, 0.f), rz2f(ld, 0.f);
    for (int k = 1; k <= nz; k++) {
      rdzf[k] = 1.f / (float)Grid1D::at(grid.dzPr, k);
      rz2f[k] = 1.f / (2 * ((float)Grid1D::at(grid.zPr, k + 1) - (float)Grid1D::at(grid.zPr, k - 1)));
    }
    upload_vec(sgs_rdzf, rdzf);
    upload_vec(sgs_rz2f, rz2f);
    sgs_width.rdzf = sgs_rdzf.as<float>();
    sgs_width.rz2f = sgs_rz2f.as<float>();
  }
}

template <typename T>
void Engine<T>::setup_poisson() {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  if (!use_mg) {
    if (!fft_len_ok(nx) || !fft_len_ok(ny)) throw ArgError("FFT pressure solver needs nx and ny of the form 2^a 3^b");
    if (nx > 1024 || ny > 1024) throw ArgError("FFT pressure solver supports nx, ny <= 1024");
    if (nz > 1024) throw ArgError("FFT pressure solver supports nz <= 1024");
  }
  const int M = nx / 2 + 1;
  // spectral rows padded to 128 bytes, so that the y pass's 8-mode segments
  // are whole cache lines
  const int Mp = solve64 || sizeof(T) == 8 ? (M + 7) / 8 * 8 : (M + 15) / 16 * 16;
  // exp(-2 pi i q/n), evaluated in long double and rounded once to R; q <
  // n/2 for the radix-2 passes, q < n for the mixed-radix ones (x passes:
  // when nx or ny is not a power of two; y passes: when ny is not)
  const bool mr_x = !fft_pow2(nx) || !fft_pow2(ny), mr_y = !fft_pow2(ny);
  auto tw = [](int n, bool full, auto one) {
    using R = decltype(one);
    const int len = full ? n : n / 2;
    std::vector<C2<R>> t(std::max(1, len));
    for (int q = 0; q < len; q++) {
      long double a = 2.0L * 3.14159265358979323846264338327950288L * q / n;
      t[q].x = (R)(double)cosl(a);
      t[q].y = (R)(double)-sinl(a);
    }
    return t;
  };
  using V2 = C2<T>;
  if (use_mg) {
    // no spectral plan: the FFT passes are not used
  } else if (!solve64) {
    X.alloc((size_t)Mp * ny * nz * sizeof(V2));
    upload_vec(twx, tw(nx, mr_x, (T)1));
    upload_vec(twy, tw(ny, mr_y, (T)1));
  } else {
    X64.alloc((size_t)Mp * ny * nz * sizeof(C2<double>));
    upload_vec(twx64, tw(nx, mr_x, 1.0));
    upload_vec(twy64, tw(ny, mr_y, 1.0));
  }
  std::vector<double> lx(M), ly(ny), a(nz), c(nz);
  for (int m = 0; m < M; m++) {
    double s = std::sin(M_PI * m / nx);
    lx[m] = -4 * s * s / (grid.dxmin * grid.dxmin);
  }
  for (int j = 0; j < ny; j++) {
    double s = std::sin(M_PI * j / ny);
    ly[j] = -4 * s * s / (grid.dymin * grid.dymin);
  }
  for (int k = 1; k <= nz; k++) {
    a[k - 1] = k > 1 ? 1 / (Grid1D::at(grid.dzW, k - 1) * Grid1D::at(grid.dzPr, k)) : 0;
    c[k - 1] = k < nz ? 1 / (Grid1D::at(grid.dzW, k) * Grid1D::at(grid.dzPr, k)) : 0;
  }
  upload_vec(lamx, lx);
  upload_vec(lamy, ly);
  upload_vec(az64, a);
  upload_vec(cz64, c);
  if (!solve64)
    pp = PoissonPlan<T>{nx, ny, nz, M, Mp, X.as<V2>(), twx.as<V2>(), twy.as<V2>(), lamx.as<double>(),
                        lamy.as<double>(), az64.as<double>(), cz64.as<double>(), g.dzPr, g.dzW};
  else
    pp64 = PoissonPlan<double>{nx, ny, nz, M, Mp, X64.as<C2<double>>(), twx64.as<C2<double>>(), twy64.as<C2<double>>(),
                               lamx.as<double>(), lamy.as<double>(), az64.as<double>(), cz64.as<double>(), g.dzPr,
                               g.dzW};
  // ELMM_ZPIPE=0/1 forces the z solve's load pipelining off/on (default: on
  // for grids with more than 1024 wavefronts of modes)
  if (const char *zp = std::getenv("ELMM_ZPIPE")) pp.zpipe = pp64.zpipe = std::atoi(zp) ? 1 : 0;
  top_blocks = (int)((nx * ny + 255) / 256);
  red_blocks = (int)(((long)nx * ny * nz + 255) / 256);
  {
    // room for the most partials any SGS z chunk the tuner tries gives
    const int saved = g.sgs_zchunk;
    if (cfg.autotune && !env_sgs_zchunk) g.sgs_zchunk = 13;
    cfl_cap = sgs_partials(g, sgs_width);
    g.sgs_zchunk = saved;
  }
  cfl_blocks = sgs_partials(g, sgs_width);
  cfl_cap = std::max(cfl_cap, cfl_blocks);
  partial.alloc((std::max(top_blocks, 1) + 1) * sizeof(double));
  div_partial.alloc(std::max(red_blocks, 1) * sizeof(double));
  cfl_partial.alloc(std::max(cfl_cap, 1) * sizeof(double));
  mon_buf.alloc(sizeof(StepMon));
  mon = mon_buf.as<StepMon>();
  StepMon m0{};
  m0.first_bad = -1;
  HIPCHK(hipMemcpy(mon, &m0, sizeof m0, hipMemcpyHostToDevice));
  if (cfg.delta_diag) {
    dblocks = delta_blocks(g);
    delta_partial.alloc((size_t)3 * dblocks * sizeof(double));
  }
  if (use_mg) setup_mg();
}

// Multigrid levels: x and y halved while both stay multiples of 2 and at
// least 4 cells, every z level kept; per level the line-solve tables (den_k
// = D_k - az_k e_{k-1}, e_k = cz_k / den_k with D_k = (((ax + ax) + (ay +
// ay)) + az_k) + cz_k) and, on the coarsest, the Fourier tables and the
// mode eigenvalues.  The tables are computed on the host exactly as the
// oracle computes them.
template <typename T>
void Engine<T>::setup_mg() {
  const int nz = cfg.nz;
  int nx = cfg.nx, ny = cfg.ny, L = 1;
  while (nx % 4 == 0 && ny % 4 == 0 && nx / 2 >= 4 && ny / 2 >= 4) {
    nx /= 2;
    ny /= 2;
    L++;
  }
  std::vector<double> az(nz), cz(nz);
  for (int k = 1; k <= nz; k++) {
    az[k - 1] = k > 1 ? 1 / (Grid1D::at(grid.dzW, k - 1) * Grid1D::at(grid.dzPr, k)) : 0;
    cz[k - 1] = k < nz ? 1 / (Grid1D::at(grid.dzW, k) * Grid1D::at(grid.dzPr, k)) : 0;
  }
  std::vector<double> tab(2 * (size_t)L * nz);
  mg.assign(L, MgLevel{});
  mg_bufs.clear();
  for (int l = 0; l < L; l++) {
    MgLevel &v = mg[l];
    v.nx = cfg.nx >> l;
    v.ny = cfg.ny >> l;
    v.nz = nz;
    const double dxl = grid.dxmin * (double)(1 << l), dyl = grid.dymin * (double)(1 << l);
    v.ax = 1.0 / (dxl * dxl);
    v.ay = 1.0 / (dyl * dyl);
    const size_t n = (size_t)v.nx * v.ny * nz;
    mg_bufs.emplace_back(new DevBuf);
    mg_bufs.back()->alloc(n * sizeof(double));
    v.phi = mg_bufs.back()->as<double>();
    mg_bufs.emplace_back(new DevBuf);
    mg_bufs.back()->alloc(n * sizeof(double));
    v.rhs = mg_bufs.back()->as<double>();
    double *den = &tab[(size_t)(2 * l) * nz], *e = &tab[(size_t)(2 * l + 1) * nz];
    double eprev = 0;
    for (int k = 1; k <= nz; k++) {
      const double D = (((v.ax + v.ax) + (v.ay + v.ay)) + az[k - 1]) + cz[k - 1];
      den[k - 1] = D - az[k - 1] * eprev;
      e[k - 1] = cz[k - 1] / den[k - 1];
      eprev = e[k - 1];
    }
  }
  upload_vec(mg_tab, tab);
  std::vector<double> dz(2 * (size_t)(nz + 1), 0.0);
  for (int k = 1; k <= nz; k++) {
    dz[k] = Grid1D::at(grid.dzPr, k);
    dz[nz + 1 + k] = Grid1D::at(grid.dzW, k);
  }
  upload_vec(mg_dz, dz);
  const MgLevel &c = mg[L - 1];
  const long plane = (long)c.nx * c.ny;
  std::vector<double> cs(2 * (size_t)(c.nx + c.ny));
  auto twid = [](int n, double *cv, double *sv) {
    for (int q = 0; q < n; q++) {
      long double t = 2.0L * 3.14159265358979323846264338327950288L * q / n;
      cv[q] = (double)cosl(t);
      sv[q] = -(double)sinl(t);
    }
  };
  twid(c.nx, &cs[0], &cs[c.nx]);
  twid(c.ny, &cs[2 * c.nx], &cs[2 * c.nx + c.ny]);
  upload_vec(mg_cs, cs);
  std::vector<double> lam(plane);
  for (int n = 0; n < c.ny; n++)
    for (int m = 0; m < c.nx; m++) {
      const double sxm = std::sin(M_PI * m / c.nx), syn = std::sin(M_PI * n / c.ny);
      lam[m + (long)n * c.nx] = -4 * sxm * sxm * c.ax - 4 * syn * syn * c.ay;
    }
  upload_vec(mg_lam, lam);
  mg_Y.alloc((size_t)plane * nz * 2 * sizeof(double));
  mg_X.alloc((size_t)plane * nz * 2 * sizeof(double));
  mg_cp.alloc((size_t)plane * nz * sizeof(double));
  mgc.Y = mg_Y.as<double2>();
  mgc.X = mg_X.as<double2>();
  mgc.cp = mg_cp.as<double>();
  mgc.cx = mg_cs.as<double>();
  mgc.sx = mg_cs.as<double>() + c.nx;
  mgc.cy = mg_cs.as<double>() + 2 * c.nx;
  mgc.sy = mg_cs.as<double>() + 2 * c.nx + c.ny;
  mgc.lam = mg_lam.as<double>();
  mg_partial.alloc(((size_t)mg_resmax_blocks(mg[0]) + 1) * sizeof(double));
}

// Max-norm residual of level 0 (synchronises the stream).
template <typename T>
double Engine<T>::mg_resmax0() {
  const int nb = mg_resmax_blocks(mg[0]);
  double *part = mg_partial.as<double>();
  launch_mg_resmax(mg[0], az64.as<double>(), cz64.as<double>(), part, stream);
  launch_max_final(part, nb, part + nb, stream);
  double r = 0;
  HIPCHK(hipMemcpyAsync(&r, part + nb, sizeof r, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return r;
}

// V-cycle from level l: 2 pre-smoothing sweeps (both colours), residual
// restricted to level l + 1, its cycle, prolongation, 2 post-smoothing
// sweeps; the coarsest level is solved exactly.
template <typename T>
void Engine<T>::mg_vcycle(int l) {
  const int L = (int)mg.size(), nz = cfg.nz;
  const double *az = az64.as<double>(), *cz = cz64.as<double>();
  if (l == L - 1) {
    const double *dz = mg_dz.as<double>();
    launch_mg_coarse(mg[l], mgc, az, cz, dz, dz + nz + 1, stream);
    return;
  }
  const double *den = mg_tab.as<double>() + (size_t)(2 * l) * nz, *e = den + nz;
  auto smooth = [&] {
    for (int s = 0; s < 2; s++) {
      launch_mg_sweep(mg[l], az, den, e, 0, stream);
      launch_mg_sweep(mg[l], az, den, e, 1, stream);
    }
  };
  smooth();
  launch_mg_restrict(mg[l], mg[l + 1], az, cz, stream);
  mg_vcycle(l + 1);
  launch_mg_prolong(mg[l + 1], mg[l], stream);
  smooth();
}

// PoissMG's loop (src/multigrid.f90:1161-1168): from P = 0, V-cycles until
// the max-norm residual is below poisson_tol or mg_cycles_max cycles ran.
template <typename T>
void Engine<T>::mg_solve() {
  launch_mg_zero(mg[0].phi, (long)cfg.nx * cfg.ny * cfg.nz, stream);
  const int maxc = cfg.mg_cycles_max > 0 ? cfg.mg_cycles_max : 50;
  const double tol = cfg.poisson_tol > 0 ? cfg.poisson_tol : 1e-10;
  int cyc = 0;
  double R = mg_resmax0();
  while (cyc < maxc && !(R < tol)) {
    mg_vcycle(0);
    cyc++;
    R = mg_resmax0();
  }
  mg_cycles = cyc;
  mg_res_last = R;
}

template <typename T>
void Engine<T>::setup_ghost_delta() {
  use_gdelta = false;
  const char *e = std::getenv("ELMM_GHOST_DELTA");
  if (!has_ib() || (e && std::string(e) == "0")) return;
  long cap = 1;
  for (int c = 0; c < 3; c++) cap = std::max(cap, ghost_shell_cells(g, c));
  DevBuf mark, tmp, cnt;
  mark.alloc((size_t)g.ntot);
  tmp.alloc((size_t)cap * 2 * sizeof(long long));
  cnt.alloc(sizeof(int));
  // the IB lists and these buffers were filled on the null stream; the
  // kernels below run on the handle's stream
  HIPCHK(hipStreamSynchronize(nullptr));
  std::vector<long long> all;
  for (int c = 0; c < 3; c++) {
    const int n = ghost_delta_build(g, c, ib_tgt[c].as<int32_t>(), ib_n[c], mark.as<uint8_t>(), tmp.as<long long>(),
                                    (int)cap, cnt.as<int>(), stream);
    HIPCHK(hipGetLastError());
    if (n < 0 || n > cap) throw HipError("ghost delta list construction failed");
    const size_t off = all.size();
    all.resize(off + 2 * (size_t)n);
    if (n) HIPCHK(hipMemcpy(all.data() + off, tmp.p, 2 * (size_t)n * sizeof(long long), hipMemcpyDeviceToHost));
    gdelta_n[c] = n;
  }
  upload_vec(gdelta, all);
  use_gdelta = true;
}

template <typename T>
void Engine<T>::setup_lists(const elmm_lists &L) {
  const long ntot = g.ntot;
  if (g.variable_z) {
    // the stretched-z path carries the z-wall lists only (periodic x and y,
    // no solid bodies)
    bool extra = L.ib_n[0] || L.ib_n[1] || L.ib_n[2];
    for (int l = 0; l < 18; l++) extra = extra || (l % 6 < 4 && L.wm_n[l]);
    if (extra) throw ArgError("solid bodies on stretched z grids are not built in this version");
  }
  SetupTrace tr;
  auto lin = [&](int i, int j, int k) { return gix(g, i, j, k); };
  // host copies of masks in the device layout: the input arrays span
  // -2..n+3 in each direction, x fastest, so every (j, k) row is one
  // contiguous run in both layouts
  auto remap = [&](auto *out, const auto *src, auto f) {
    const long hy = cfg.nx + 6, hz = (long)(cfg.nx + 6) * (cfg.ny + 6);
    for (int k = -2; k <= cfg.nz + 3; k++)
      for (int j = -2; j <= cfg.ny + 3; j++) {
        const auto *s = src + (j + 2) * hy + (k + 2) * hz;
        auto *o = out + lin(-2, j, k);
        for (int i = 0; i < cfg.nx + 6; i++) o[i] = f(s[i]);
      }
  };
  // wall-model points of the 18 momentum lists, concatenated in list order
  std::vector<int32_t> wijk, wlist;
  std::vector<double> wdist, wz0, wustar, wtval, wtflux, wqval, wqflux;
  std::vector<int8_t> wtclass;
  std::vector<int> list_off(19, 0);
  for (int l = 0; l < 18; l++) {
    list_off[l] = (int)wlist.size();
    for (int p = 0; p < L.wm_n[l]; p++) {
      wijk.insert(wijk.end(), {L.wm_ijk[l][3 * p], L.wm_ijk[l][3 * p + 1], L.wm_ijk[l][3 * p + 2]});
      wlist.push_back(l);
      wdist.insert(wdist.end(), {L.wm_dist[l][3 * p], L.wm_dist[l][3 * p + 1], L.wm_dist[l][3 * p + 2]});
      wz0.push_back(L.wm_z0[l][p]);
      wustar.push_back(1.0);  // WMpointUVW%ustar = 1 (src/wallmodels.f90:79)
      // thermal class; temperature_flux = sideTemp on a flux wall, else 0
      const int8_t tc = L.wm_tclass[l] ? L.wm_tclass[l][p] : 0;
      const double tv = L.wm_tclass[l] ? L.wm_tval[l][p] : 0.0;
      wtclass.push_back(tc);
      wtval.push_back(tv);
      wtflux.push_back(tc == ELMM_WM_SURFACE_FLUX ? tv : 0.0);
      wqval.push_back(L.wm_qval[l] ? L.wm_qval[l][p] : 0.0);
      wqflux.push_back(L.wm_qval[l] ? L.wm_qflux[l][p] : 0.0);
      if (tc != ELMM_WM_NEUTRAL && cfg.enable_buoyancy) mo_any = true;
    }
  }
  upload_vec(wm_tclass, wtclass);
  upload_vec(wm_tval, wtval);
  upload_vec(wm_tflux, wtflux);
  upload_vec(wm_qval, wqval);
  upload_vec(wm_qflux, wqflux);
  list_off[18] = (int)wlist.size();
  wm_n = (int)wlist.size();
  std::vector<double> wlog(wm_n, 0.0);
  for (int p = 0; p < wm_n; p++) {
    const double *d = &wdist[3 * p];
    double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (wz0[p] > 0) wlog[p] = std::log(dist / wz0[p]);
  }
  upload_vec(wm_logdz0, wlog);
  upload_vec(wm_ijk, wijk);
  upload_vec(wm_list, wlist);
  upload_vec(wm_dist, wdist);
  upload_vec(wm_z0, wz0);
  upload_vec(wm_ustar, wustar);
  wm_fluxp.alloc(std::max(wm_n, 1) * sizeof(double));
  wm_fluxm.alloc(std::max(wm_n, 1) * sizeof(double));

  tr.mark(" wall model points");

  // the rest on threads of their own: per velocity component its advection
  // codes and ordered wall-model terms; Prtype and the scalar flux masks; the
  // scalar code and ordered scalar terms; the IB lists; the mass-source terms
  // and nullified points (disjoint members, read-only inputs)
  std::vector<uint8_t> codes[3];
  const int32_t *types[3] = {L.utype, L.vtype, L.wtype};
  std::vector<std::function<void()>> tasks;
  for (int c = 0; c < 3; c++)
    tasks.emplace_back([&, c] {
      codes[c].assign(ntot, 0);
      remap(codes[c].data(), types[c], [](int32_t t) -> uint8_t { return t == 0 ? 0 : t < 0 ? 1 : 2; });
      // ordered momentum corrections (wmfluxes-nobranch-inc.f90:49-208)
      if (cfg.wall_model) {
        {  // (one component's builder)
          TermBuilder tb(ntot);
          const int *n = g.n[c];
          for (int dir = 0; dir < 6; dir++) {
            const int l = c * 6 + dir, ax = dir / 2, plus = dir % 2;
            for (int pass = 0; pass < 2; pass++)
              for (int p = 0; p < L.wm_n[l]; p++) {
                int q[3] = {L.wm_ijk[l][3 * p], L.wm_ijk[l][3 * p + 1], L.wm_ijk[l][3 * p + 2]};
                int f[3] = {q[0], q[1], q[2]}, t[3] = {q[0], q[1], q[2]};
                if (!plus) {
                  f[ax] -= 1;
                  if (pass == 1) t[ax] -= 1;
                } else if (pass == 0) {
                  t[ax] += 1;
                }
                const int neg = pass == 1;
                if (t[0] < 1 || t[0] > n[0] || t[1] < 1 || t[1] > n[1] || t[2] < 1 || t[2] > n[2]) continue;
                long key = lin(t[0], t[1], t[2]);
                // last slot: level of the face's lower cell (stretched-z terms)
                tb.add(key, t[0], t[1], t[2],
                       {ax | (neg << 2), list_off[l] + p, (int32_t)lin(f[0], f[1], f[2]), f[2]});
                codes[c][key] |= CODE_CORRECTED;
              }
          }
          tb.upload(mom_terms[c]);
        }
      }
      upload_vec(code[c], codes[c]);
    });
  tasks.emplace_back([&] {
    std::vector<int32_t> pt(ntot, 0);
    remap(pt.data(), L.prtype, [](int32_t t) { return t; });
    upload_vec(prtype, pt);
    for (int d = 0; d < 3; d++) {
      std::vector<uint8_t> m(ntot, 0);
      remap(m.data(), L.scfl_mask[d], [](uint8_t t) { return t; });
      upload_vec(mask[d], m);
    }
  });
  tasks.emplace_back([&] {
    std::vector<uint8_t> scode(ntot, 0);
    remap(scode.data(), L.prtype, [](int32_t t) -> uint8_t { return t > 0 ? CODE_SOLID : 0; });
    // scalar wall-model points and ordered scalar terms
    wmp_n = cfg.wall_model ? L.wmp_n : 0;
    std::vector<int32_t> pijk;
    for (int p = 0; p < wmp_n; p++) pijk.insert(pijk.end(), {L.wmp_ijk[3 * p], L.wmp_ijk[3 * p + 1], L.wmp_ijk[3 * p + 2]});
    upload_vec(wmp_ijk, pijk);
    wmp_div.alloc(std::max(wmp_n, 1) * sizeof(double));
    {
      // WMPoints' distance vectors, roughness, log(|dist|/z0) and u* (initially
      // 1, the WMpoint default, src/wallmodels.f90:40)
      std::vector<double> pd(L.wmp_dist, L.wmp_dist + 3 * (size_t)wmp_n), pz(L.wmp_z0, L.wmp_z0 + wmp_n),
          plog(wmp_n, 0.0), pu(wmp_n, 1.0);
      for (int p = 0; p < wmp_n; p++) {
        const double *d = &pd[3 * p];
        const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (pz[p] > 0) plog[p] = std::log(dist / pz[p]);
      }
      upload_vec(wmp_dist, pd);
      upload_vec(wmp_z0, pz);
      upload_vec(wmp_logdz0, plog);
      upload_vec(wmp_ustar, pu);
      // thermal class, value, persistent heat flux and ComputeAreaFactor
      // (src/wallmodels.f90:170-203; norm2 as the square root of the sum of
      // squares): the wall area over the cell volume from the dominant
      // component of the wall normal
      std::vector<int8_t> tc(wmp_n, 0);
      std::vector<double> tv(wmp_n, 0.0), tf(wmp_n, 0.0), af(wmp_n, 0.0), qv(wmp_n, 0.0), qf(wmp_n, 0.0);
      const double dxm = g.dxmin, dym = g.dymin, dzm = g.dzmin;
      for (int p = 0; p < wmp_n; p++) {
        if (L.wmp_tclass) {
          tc[p] = L.wmp_tclass[p];
          tv[p] = L.wmp_tval[p];
          tf[p] = tc[p] == ELMM_WM_SURFACE_FLUX ? tv[p] : 0.0;
          if (tc[p] != ELMM_WM_NEUTRAL && cfg.enable_buoyancy) mo_any = true;
        }
        if (L.wmp_qval) {
          qv[p] = L.wmp_qval[p];
          qf[p] = L.wmp_qflux[p];
        }
        const double *d = &pd[3 * p];
        if (-d[0] == 0 && -d[1] == 0 && -d[2] == 0) continue;
        const double V = dxm * dym * dzm;
        const double nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double n0 = std::fabs(-d[0] / nrm), n1 = std::fabs(-d[1] / nrm), n2 = std::fabs(-d[2] / nrm);
        double on, dd;
        if (n0 > n1 && n0 > n2) {
          on = n0;
          dd = std::sqrt(dym * dzm);
        } else if (n1 > n0 && n1 > n2) {
          on = n1;
          dd = std::sqrt(dxm * dzm);
        } else {
          on = n2;
          dd = std::sqrt(dxm * dym);
        }
        const double l = dd / on;
        const double S = dd * l;
        af[p] = S / V;
      }
      upload_vec(wmp_tclass, tc);
      upload_vec(wmp_tval, tv);
      upload_vec(wmp_tflux, tf);
      upload_vec(wmp_af, af);
      upload_vec(wmp_qval, qv);
      upload_vec(wmp_qflux, qf);
    }
    // point sources, then (with buoyancy) one slot per wall-model point for
    // the temperature's wall source written by k_wm_ustar every stage
    src_n = L.src_n;
    std::vector<double> sflux(L.src_flux, L.src_flux + L.src_n);
    if (tslot >= 0) sflux.resize(src_n + (mslot >= 0 ? 2 : 1) * (size_t)wmp_n, 0.0);
    upload_vec(src_flux, sflux);
    if (tslot >= 0) {
      // solid cells of -1..n+2 with a fluid neighbour there: their temperature
      // is the mean of those neighbours after each stage
      const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
      const long hy = nx + 6, hz = (long)(nx + 6) * (ny + 6);
      auto pt = [&](int i, int j, int k) { return L.prtype[(i + 2) + (j + 2) * hy + (k + 2) * hz]; };
      std::vector<int32_t> fix;
      for (int k = -1; k <= nz + 2; k++)
        for (int j = -1; j <= ny + 2; j++)
          for (int i = -1; i <= nx + 2; i++) {
            if (pt(i, j, k) <= 0) continue;
            const bool any = (i > -1 && pt(i - 1, j, k) <= 0) || (i < nx + 2 && pt(i + 1, j, k) <= 0) ||
                             (j > -1 && pt(i, j - 1, k) <= 0) || (j < ny + 2 && pt(i, j + 1, k) <= 0) ||
                             (k > -1 && pt(i, j, k - 1) <= 0) || (k < nz + 2 && pt(i, j, k + 1) <= 0);
            if (any) fix.insert(fix.end(), {i, j, k});
          }
      tfix_n = (int)(fix.size() / 3);
      upload_vec(tfix_ijk, fix);
    }
    for (int s = 0; s < nsc; s++) sc_terms.emplace_back(new DevTerms);
    if (nsc > 0) {
      std::vector<TermBuilder> tbs;
      for (int s = 0; s < nsc; s++) tbs.emplace_back(ntot);
      auto interior = [&](int i, int j, int k) {
        return i >= 1 && i <= cfg.nx && j >= 1 && j <= cfg.ny && k >= 1 && k <= cfg.nz;
      };
      for (int s = 0; s < nsc; s++) {
        TermBuilder &tb = tbs[s];
        for (int d = 0; d < 3; d++)
          for (int pass = 0; pass < 2; pass++)
            for (int p = 0; p < L.scfl_n[d]; p++) {
              int q[3] = {L.scfl_ijk[d][3 * p], L.scfl_ijk[d][3 * p + 1], L.scfl_ijk[d][3 * p + 2]};
              int t[3] = {q[0], q[1], q[2]};
              if (pass == 1) t[d] += 1;
              if (!interior(t[0], t[1], t[2])) continue;
              tb.add(lin(t[0], t[1], t[2]), t[0], t[1], t[2],
                     {(pass == 0 ? 0 : 1) | (d << 2), (int32_t)lin(q[0], q[1], q[2]), 0, 0});
            }
        for (int p = 0; p < L.src_n; p++) {
          if (L.src_scalar[p] != s) continue;
          int q[3] = {L.src_ijk[3 * p], L.src_ijk[3 * p + 1], L.src_ijk[3 * p + 2]};
          tb.add(lin(q[0], q[1], q[2]), q[0], q[1], q[2], {2, p, 0, 0});
        }
        for (int p = 0; p < wmp_n; p++) {
          int q[3] = {L.wmp_ijk[3 * p], L.wmp_ijk[3 * p + 1], L.wmp_ijk[3 * p + 2]};
          // TemperatureExtra: the surface heat flux times the area factor,
          // then the divergence term (src/scalars.f90:322-328)
          if (s == tslot) tb.add(lin(q[0], q[1], q[2]), q[0], q[1], q[2], {2, L.src_n + p, 0, 0});
          // MoistureExtra: the moisture flux likewise (src/scalars.f90:350-356)
          if (s == mslot) tb.add(lin(q[0], q[1], q[2]), q[0], q[1], q[2], {2, L.src_n + wmp_n + p, 0, 0});
          tb.add(lin(q[0], q[1], q[2]), q[0], q[1], q[2], {3, p, 0, 0});
        }
      }
      // every scalar lists the union of corrected cells (same cell order)
      TermBuilder uni(ntot);
      for (int s = 0; s < nsc; s++)
        for (size_t q = 0; q < tbs[s].cells.size(); q++) {
          auto &cc = tbs[s].cells[q];
          uni.cell_of(tbs[s].keys[q], cc[0], cc[1], cc[2]);
        }
      for (int s = 0; s < nsc; s++) {
        std::vector<int32_t> sp, st;
        tbs[s].csr(sp, st);
        TermBuilder out(ntot);
        for (size_t q = 0; q < uni.cells.size(); q++) {
          const auto &cc = uni.cells[q];
          const long key = uni.keys[q];
          const int oc = out.cell_of(key, cc[0], cc[1], cc[2]);
          const int ic = tbs[s].find(key);
          if (ic < 0) continue;
          for (int t = sp[ic]; t < sp[ic + 1]; t++)
            out.items.emplace_back(oc, TermBuilder::Term{st[4 * t], st[4 * t + 1], st[4 * t + 2], st[4 * t + 3]});
        }
        out.upload(*sc_terms[s]);
      }
      for (long key : uni.keys) scode[key] |= CODE_CORRECTED;
    }
    upload_vec(code_s, scode);
  });
  tasks.emplace_back([&] {
    // immersed-boundary lists
    for (int c = 0; c < 4; c++) {
      ib_n[c] = L.ib_n[c];
      std::vector<int32_t> tg(L.ib_ijk[c], L.ib_ijk[c] + 3 * (size_t)ib_n[c]);
      std::vector<int32_t> pt(L.ib_ptr[c], L.ib_ptr[c] + (ib_n[c] ? ib_n[c] + 1 : 0));
      if (pt.empty()) pt.push_back(0);
      int nt = pt.back();
      std::vector<int32_t> sr(L.ib_src_ijk[c], L.ib_src_ijk[c] + 3 * (size_t)nt);
      std::vector<double> cf(L.ib_coef[c], L.ib_coef[c] + nt);
      if (c == 3 && ib_n[3] > 0) {
        // dependency levels of the sequential ScalFlIBPoints loop
        std::unordered_map<long, int> wlev, rlev;
        std::vector<int> lev(ib_n[3]);
        for (int p = 0; p < ib_n[3]; p++) {
          long tk = lin(tg[3 * p], tg[3 * p + 1], tg[3 * p + 2]);
          int lv = 0;
          auto bump = [&](std::unordered_map<long, int> &m, long key) {
            auto it = m.find(key);
            if (it != m.end()) lv = std::max(lv, it->second + 1);
          };
          for (int t = pt[p]; t < pt[p + 1]; t++) bump(wlev, lin(sr[3 * t], sr[3 * t + 1], sr[3 * t + 2]));
          bump(rlev, tk);
          bump(wlev, tk);
          lev[p] = lv;
          wlev[tk] = std::max(wlev.count(tk) ? wlev[tk] : 0, lv);
          for (int t = pt[p]; t < pt[p + 1]; t++) {
            long sk = lin(sr[3 * t], sr[3 * t + 1], sr[3 * t + 2]);
            rlev[sk] = std::max(rlev.count(sk) ? rlev[sk] : 0, lv);
          }
        }
        int nl = *std::max_element(lev.begin(), lev.end()) + 1;
        std::vector<int> order(ib_n[3]);
        for (int p = 0; p < ib_n[3]; p++) order[p] = p;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lev[a] < lev[b]; });
        std::vector<int32_t> tg2, pt2{0}, sr2;
        std::vector<double> cf2;
        visc_levels.assign(nl + 1, 0);
        for (int q = 0; q < ib_n[3]; q++) {
          int p = order[q];
          tg2.insert(tg2.end(), {tg[3 * p], tg[3 * p + 1], tg[3 * p + 2]});
          for (int t = pt[p]; t < pt[p + 1]; t++) {
            sr2.insert(sr2.end(), {sr[3 * t], sr[3 * t + 1], sr[3 * t + 2]});
            cf2.push_back(cf[t]);
          }
          pt2.push_back((int32_t)cf2.size());
          visc_levels[lev[p] + 1] = q + 1;
        }
        for (int l = 1; l <= nl; l++) visc_levels[l] = std::max(visc_levels[l], visc_levels[l - 1]);
        tg = tg2;
        pt = pt2;
        sr = sr2;
        cf = cf2;
      }
      upload_vec(ib_tgt[c], tg);
      upload_vec(ib_ptr[c], pt);
      upload_vec(ib_src[c], sr);
      upload_vec(ib_coef[c], cf);
    }
    ib_tmp.alloc((size_t)std::max(ib_n[0] + ib_n[1] + ib_n[2], 1) * sizeof(double));
  });
  tasks.emplace_back([&] {
    // IBMassSources gather terms (src/tsteps.f90:490-531) + Bound_Q fold targets
    if (cfg.ib_mass_sources && L.ib_n[0] + L.ib_n[1] + L.ib_n[2] > 0) {
      TermBuilder tb(ntot);
      for (int c = 0; c < 3; c++)
        for (int p = 0; p < L.ib_n[c]; p++) {
          int q[3] = {L.ib_ijk[c][3 * p], L.ib_ijk[c][3 * p + 1], L.ib_ijk[c][3 * p + 2]};
          int32_t src = (int32_t)lin(q[0], q[1], q[2]);
          tb.add(lin(q[0], q[1], q[2]), q[0], q[1], q[2], {c, src, q[c], 0});
          int r[3] = {q[0], q[1], q[2]};
          r[c] += 1;
          tb.add(lin(r[0], r[1], r[2]), r[0], r[1], r[2], {c | (1 << 2), src, q[c] + 1, 0});
        }
      // fold destinations must be rewritten every stage (their pre-fold value
      // is their own direct sum, zero when they have none)
      const int N[3] = {cfg.nx, cfg.ny, cfg.nz};
      std::vector<std::array<int, 3>> targets(tb.cells.begin(), tb.cells.end());
      // Bound_Q reads a ghost cell only when exactly one of its coordinates is
      // outside the interior (src/boundaries.f90:2094-2192)
      for (auto &q : targets) {
        int nghost = 0, a = -1;
        for (int b = 0; b < 3; b++)
          if (q[b] < 1 || q[b] > N[b]) {
            nghost++;
            a = b;
          }
        if (nghost != 1) continue;
        bool per = cfg.btype[2 * a] == ELMM_BC_PERIODIC;
        std::array<int, 3> d = q;
        if (q[a] <= 0) d[a] = per ? N[a] : 1;
        else d[a] = per ? 1 : N[a];
        tb.cell_of(lin(d[0], d[1], d[2]), d[0], d[1], d[2]);
      }
      tb.upload(q_terms);
      // rows (j, k) and 64-cell segments of x where Q may be non-zero: the
      // listed interior cells (everything else stays zero)
      std::vector<uint32_t> qm((size_t)cfg.ny * cfg.nz, 0);
      for (auto &q : tb.cells)
        if (q[0] >= 1 && q[0] <= cfg.nx && q[1] >= 1 && q[1] <= cfg.ny && q[2] >= 1 && q[2] <= cfg.nz)
          qm[(size_t)(q[1] - 1) + (size_t)cfg.ny * (q[2] - 1)] |= 1u << ((q[0] - 1) >> 6);
      upload_vec(qmask, qm);
    }
    for (int c = 0; c < 3; c++) {
      null_n[c] = L.null_n[c];
      std::vector<int32_t> v(L.null_ijk[c], L.null_ijk[c] + 3 * (size_t)null_n[c]);
      upload_vec(null_ijk[c], v);
    }
  });
  run_parallel(tasks, device);
  tr.mark(" lists");
  setup_ghost_delta();
  tr.mark(" ghost delta");
}

template <typename T>
void Engine<T>::create(const elmm_config &c, const elmm_lists &L) {
  cfg = c;
  std::string err = validate_config(cfg);
  if (!err.empty()) throw ArgError(err);
  const int want = std::is_same<T, double>::value ? ELMM_FP64 : ELMM_FP32_REFINED;
  if (cfg.precision != want) throw ArgError("engine/precision mismatch");
  // FP32 form: refine_iters > 0 -> FP32 spectral solves + that many FP64
  // refinement sweeps; 0 -> one FP64 spectral solve
  refine_iters = want == ELMM_FP32_REFINED ? (cfg.refine_iters > 0 ? cfg.refine_iters : 0) : 0;
  solve64 = want == ELMM_FP32_REFINED && refine_iters == 0;
  if (cfg.poisson_solver != ELMM_POISSON_FFT && cfg.poisson_solver != ELMM_POISSON_MG)
    throw ArgError("poisson_solver must be ELMM_POISSON_FFT or ELMM_POISSON_MG");
  use_mg = cfg.poisson_solver == ELMM_POISSON_MG;
  if (use_mg) {
    if (cfg.nx % 2 || cfg.ny % 2) throw ArgError("the multigrid solver needs even nx and ny");
    if (refine_iters > 0) throw ArgError("the multigrid solver runs in FP64: refine_iters must be 0");
  }
  for (int s = 4; s < 6; s++)
    if (cfg.nscalars > 0 && cfg.scal_btype[s] == ELMM_BC_PERIODIC) throw ArgError("periodic scalar z boundaries");
  if (cfg.stretched_z) {
    zw_copy.assign(cfg.zw, cfg.zw + cfg.nz + 1);
    cfg.zw = zw_copy.data();
  }
  if (cfg.enable_moisture && !cfg.enable_buoyancy) throw ArgError("moisture needs enable_buoyancy");
  nsc = cfg.nscalars + (cfg.enable_buoyancy ? 1 : 0) + (cfg.enable_moisture ? 1 : 0);
  tslot = cfg.enable_buoyancy ? cfg.nscalars : -1;
  mslot = cfg.enable_moisture ? cfg.nscalars + 1 : -1;
  dt_cur = cfg.variable_dt ? cfg.dt_max : cfg.dt;
  time = cfg.start_time;
  SetupTrace tr;
  HIPCHK(hipGetDevice(&device));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  tr.mark("stream");
  setup_geo();
  if (cfg.enable_buoyancy) {
    // temperature: periodic x/y, Dirichlet or Neumann walls in z; the fused
    // momentum kernel carries the buoyancy force
    if (g.variable_z) throw ArgError("buoyancy needs a uniform grid (scalars on stretched z are not built)");
    if (!fused_momentum) throw ArgError("buoyancy needs the fused momentum kernel (not ELMM_MOMENTUM=percomp)");
    if (!(cfg.temperature_ref > 0)) throw ArgError("temperature_ref must be positive");
    if (!(cfg.nu > 0)) throw ArgError("buoyancy needs a positive molecular viscosity");
    for (int s = 0; s < 4; s++)
      if (cfg.temp_btype[s] != ELMM_BC_PERIODIC) throw ArgError("the temperature must be periodic in x and y");
    for (int s = 4; s < 6; s++)
      if (cfg.temp_btype[s] != ELMM_BC_DIRICHLET && cfg.temp_btype[s] != ELMM_BC_NEUMANN &&
          !(cfg.temp_btype[s] == ELMM_BC_CONSTFLUX && mo_wall(cfg, s)))
        throw ArgError("temperature z boundaries must be Dirichlet, Neumann, or a flux on a wall-model wall");
    gtemp = g;
    for (int s = 0; s < 6; s++) {
      // a wall that takes the temperature into its wall-model points has
      // TempBtype = 0: the mirror ghost rule (src/wallmodels.f90:618-620)
      gtemp.sbt[s] = mo_wall(cfg, s) ? 0 : cfg.temp_btype[s];
      gtemp.side_scal[s] = cfg.side_temp[s];
    }
  }
  if (cfg.enable_moisture) {
    for (int s = 0; s < 4; s++)
      if (cfg.moist_btype[s] != ELMM_BC_PERIODIC) throw ArgError("the moisture must be periodic in x and y");
    for (int s = 4; s < 6; s++)
      if (cfg.moist_btype[s] != ELMM_BC_DIRICHLET && cfg.moist_btype[s] != ELMM_BC_NEUMANN &&
          !mo_moist_flux_wall(cfg, s))
        throw ArgError("moisture z boundaries must be Dirichlet, Neumann, or a flux on a wall-model wall");
    gmoist = g;
    for (int s = 0; s < 6; s++) {
      // a flux moisture on a wall-model wall: MoistBtype = 0, the mirror
      // (src/wallmodels.f90:627-629)
      gmoist.sbt[s] = mo_moist_flux_wall(cfg, s) ? 0 : cfg.moist_btype[s];
      gmoist.side_scal[s] = cfg.side_moist[s];
    }
  }
  tr.mark("geometry");
  const size_t fb = g.ntot * sizeof(T), fb64 = g.ntot * sizeof(double);
  for (DevBuf *b : {&U, &V, &W, &U2, &V2, &W2, &Us, &Vs, &Ws, &Pr, &Visc, &TDiff}) b->alloc(fb);
  for (DevBuf *b : {&Q, &Phi, &RHS}) b->alloc(fb64);
  for (int s = 0; s < nsc; s++) {
    S.emplace_back(new DevBuf);
    Snew.emplace_back(new DevBuf);
    Sadv.emplace_back(new DevBuf);
    S.back()->alloc(fb);
    Snew.back()->alloc(fb);
    Sadv.back()->alloc(fb);
    s_cur.push_back(S.back()->as<T>());
    s_nxt.push_back(Snew.back()->as<T>());
  }
  u = U.as<T>();
  v = V.as<T>();
  w = W.as<T>();
  u2 = U2.as<T>();
  v2 = V2.as<T>();
  w2 = W2.as<T>();
  tr.mark("fields");
  setup_lists(L);
  tr.mark("lists");
  setup_poisson();
  // the set-up's fills and uploads run on the null stream: drain it (not the
  // device, whose other handles may be stepping)
  HIPCHK(hipStreamSynchronize(nullptr));
  tr.mark("poisson");
}

template <typename T>
void Engine<T>::xfer(int field, int lb, const int32_t ext[3], double *host, bool up) {
  void *d = nullptr;
  bool is64 = false;  // Phi and Q are FP64 in every form
  switch (field) {
    case ELMM_F_U: d = u; break;
    case ELMM_F_V: d = v; break;
    case ELMM_F_W: d = w; break;
    case ELMM_F_PR: d = Pr.p; break;
    case ELMM_F_VISC: d = Visc.p; break;
    case ELMM_F_TDIFF: d = TDiff.p; break;
    case ELMM_F_PHI: d = Phi.p; is64 = true; break;
    case ELMM_F_Q: d = Q.p; is64 = true; break;
    default:
      if (field >= ELMM_F_SCALAR && field < ELMM_F_SCALAR + cfg.nscalars) d = s_cur[field - ELMM_F_SCALAR];
      if (field == ELMM_F_TEMPERATURE && tslot >= 0) d = s_cur[tslot];
      if (field == ELMM_F_MOISTURE && mslot >= 0) d = s_cur[mslot];
  }
  if (!d) throw ArgError("unknown field id " + std::to_string(field));
  if (ext[0] <= 0 || ext[1] <= 0 || ext[2] <= 0) throw ArgError("extents must be positive");
  HIPCHK(hipStreamSynchronize(stream));
  const size_t es = is64 ? sizeof(double) : sizeof(T);
  const int n[3] = {cfg.nx, cfg.ny, cfg.nz};
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = std::max(lb, -2);
    hi[a] = std::min(lb + ext[a] - 1, n[a] + 3);
  }
  // An upload of the whole box -2..n+3 replaces every element any kernel
  // reads or writes; the rest of the allocation (row padding) is never
  // written after the allocation's zero fill.  Such an upload starts from
  // zeros instead of the field's current contents.
  const bool whole = up && lo[0] == -2 && lo[1] == -2 && lo[2] == -2 && hi[0] == n[0] + 3 && hi[1] == n[1] + 3 &&
                     hi[2] == n[2] + 3;
  // host image of the allocation: FP64 for Phi and Q, the field type else
  std::vector<double> stage(is64 ? g.ntot : 0);
  std::vector<T> staget(is64 ? 0 : g.ntot);
  double *st64 = stage.data();
  void *himg = is64 ? (void *)stage.data() : (void *)staget.data();
  if (!whole) HIPCHK(hipMemcpy(himg, d, g.ntot * es, hipMemcpyDeviceToHost));
  // rows (j, k) of the box, split over threads
  const long hy = ext[0], hz = (long)ext[0] * ext[1];
  const int nj = hi[1] - lo[1] + 1, nk = hi[2] - lo[2] + 1;
  const int nth = (int)std::min<long>(8, std::max<long>(1, (long)nj * nk * (hi[0] - lo[0] + 1) / 262144));
  if (nj > 0 && nk > 0 && hi[0] >= lo[0]) run_tasks(nth, [&](int t) {
    for (int r = t; r < nj * nk; r += nth) {
      const int j = lo[1] + r % nj, k = lo[2] + r / nj;
      const long h0 = (lo[0] - lb) + (j - lb) * hy + (k - lb) * hz;
      const long p0 = gix(g, lo[0], j, k);
      const int w = hi[0] - lo[0] + 1;
      if (is64) {
        if (up)
          std::memcpy(st64 + p0, host + h0, w * sizeof(double));
        else
          std::memcpy(host + h0, st64 + p0, w * sizeof(double));
      } else {
        T *f = staget.data() + p0;
        if (up)
          for (int i = 0; i < w; i++) f[i] = (T)host[h0 + i];
        else
          for (int i = 0; i < w; i++) host[h0 + i] = (double)f[i];
      }
    }
  });
  if (up) {
    const void *src = himg;
    HIPCHK(hipMemcpy(d, src, g.ntot * es, hipMemcpyHostToDevice));
    // scalar buffers rotate: both copies start from the uploaded state so
    // ghost cells that no fill touches keep the uploaded value
    if (field >= ELMM_F_SCALAR && field < ELMM_F_SCALAR + cfg.nscalars)
      HIPCHK(hipMemcpy(s_nxt[field - ELMM_F_SCALAR], src, g.ntot * es, hipMemcpyHostToDevice));
    if (field == ELMM_F_TEMPERATURE && tslot >= 0)
      HIPCHK(hipMemcpy(s_nxt[tslot], src, g.ntot * es, hipMemcpyHostToDevice));
    if (field == ELMM_F_MOISTURE && mslot >= 0)
      HIPCHK(hipMemcpy(s_nxt[mslot], src, g.ntot * es, hipMemcpyHostToDevice));
  }
}

template <typename T>
void Engine<T>::ib_momentum(T *a, T *b, T *c) {
  const IBDev ib[3] = {ibdev(0), ibdev(1), ibdev(2)};
  timed(ELMM_T_IB, [&] { launch_ib_momentum(g, ib, a, b, c, ib_tmp.as<double>(), stream); });
}

template <typename T>
double Engine<T>::reduce_max(DevBuf &buf, int n) {
  std::vector<double> h(n);
  HIPCHK(hipMemcpyAsync(h.data(), buf.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  // a NaN partial makes the maximum +inf (std::max would drop it)
  double m = 0;
  for (double x : h) m = x != x ? INFINITY : std::max(m, x);
  return m;
}

template <typename T>
double Engine<T>::max_div(T *a, T *b, T *c, bool use_q) {
  launch_max_div(g, a, b, c, use_q ? Q.as<double>() : nullptr, div_partial.as<double>(), red_blocks, stream);
  return reduce_max(div_partial, red_blocks);
}

// PressureCorrection, src/pressure.f90:169-266.  PrePoisson's BoundUVW
// (pressure.f90:290) repeats the fill the caller has just done and is
// omitted; the zeroing of open-boundary faces inside solids
// (pressure.f90:301-342) only acts on non-periodic, non-noslip sides, where
// the normal ghost is already 0 for every supported boundary type.
// Spectral passes 2-5 of a solve whose x-transformed right-hand side is in
// X: y FFT, z tridiagonal, inverse y, inverse x into Phi (or added to Phi).
template <typename T>
void Engine<T>::poisson_solve(int accumulate) {
  auto passes = [&](const auto &plan) {
    timed(ELMM_T_POISSON_Y, [&] { launch_fft_y(plan, 0, stream); });
    timed(ELMM_T_POISSON_Z, [&] { launch_poisson_solve_z(g, plan, stream); });
    timed(ELMM_T_POISSON_Y, [&] { launch_fft_y(plan, 1, stream); });
    timed(ELMM_T_POISSON_X_INV, [&] { launch_fft_x_inv(g, plan, Phi.as<double>(), accumulate, stream); });
  };
  if (solve64)
    passes(pp64);
  else
    passes(pp);
}

template <typename T>
void Engine<T>::pressure_correction(T *a, T *b, T *c, bool use_q, double coef, bool diag) {
  const double dt2 = coef, dt2_rec = 1.0 / dt2;
  double *q = use_q ? Q.as<double>() : nullptr;
  const uint32_t *qm = use_q ? qmask.as<uint32_t>() : nullptr;
  // first solve; the precision form keeps the FP64 right-hand side for the
  // refinement sweeps or for the residual diagnostic
  double *rhs_keep = refine_iters > 0 || (solve64 && diag) ? RHS.as<double>() : nullptr;
  if (use_mg) {
    timed(ELMM_T_POISSON_X_FWD, [&] {
      launch_mg_rhs<T>(g, a, b, c, q, dt2_rec, mg[0].rhs, RHS.as<double>(), stream);
    });
    timed(ELMM_T_POISSON_Z, [&] { mg_solve(); });
    timed(ELMM_T_POISSON_X_INV, [&] { launch_mg_to_phi(g, mg[0].phi, Phi.as<double>(), stream); });
    if (diag) {
      launch_refine_residual(g, Phi.as<double>(), RHS.as<double>(), az64.as<double>(), cz64.as<double>(), nullptr,
                             div_partial.as<double>(), red_blocks, stream);
      last_res = reduce_max(div_partial, red_blocks);
      last_div = dt2 * last_res;
      last_refinements = mg_cycles;
    }
    timed(ELMM_T_REDUCE, [&] {
      launch_top_sum(g, Pr.as<T>(), Phi.as<double>(), partial.as<double>(), top_blocks, stream);
    });
    timed(ELMM_T_PROJECTION, [&] {
      launch_projection(g, a, b, c, Pr.as<T>(), Phi.as<double>(), partial.as<double>(), top_blocks, dt2, stream);
    });
    timed(ELMM_T_GHOST, [&] { launch_bound_pr<T>(g, Pr.as<T>(), stream); });
    return;
  }
  timed(ELMM_T_POISSON_X_FWD, [&] {
    if (solve64)
      launch_fft_x_fwd(g, pp64, a, b, c, q, qm, nullptr, nullptr, dt2_rec, rhs_keep, stream);
    else
      launch_fft_x_fwd(g, pp, a, b, c, q, qm, nullptr, nullptr, dt2_rec, rhs_keep, stream);
  });
  poisson_solve(0);
  // iterative refinement: the x pass evaluates r = RHS - A Phi in FP64 and
  // transforms it; FP32 solve of A e = r; Phi += e (FP64)
  for (int it = 0; it < refine_iters; it++) {
    timed(ELMM_T_REFINE, [&] {
      launch_fft_x_fwd(g, pp, a, b, c, nullptr, nullptr, RHS.as<double>(), Phi.as<double>(), 1.0, nullptr, stream);
    });
    poisson_solve(1);
  }
  if (diag && (refine_iters > 0 || solve64)) {
    // FP64 residual of the final Phi; the divergence of the FP64-corrected
    // velocity equals coef * r before the correction is rounded to T
    launch_refine_residual(g, Phi.as<double>(), RHS.as<double>(), az64.as<double>(), cz64.as<double>(), nullptr,
                           div_partial.as<double>(), red_blocks, stream);
    last_res = reduce_max(div_partial, red_blocks);
    last_div = dt2 * last_res;
    last_refinements = refine_iters;
  }
  timed(ELMM_T_REDUCE, [&] {
    launch_top_sum(g, Pr.as<T>(), Phi.as<double>(), partial.as<double>(), top_blocks, stream);
  });
  timed(ELMM_T_PROJECTION, [&] {
    launch_projection(g, a, b, c, Pr.as<T>(), Phi.as<double>(), partial.as<double>(), top_blocks, dt2, stream);
  });
  timed(ELMM_T_GHOST, [&] { launch_bound_pr<T>(g, Pr.as<T>(), stream); });
}

// SubgridStresses, src/dynamics.f90:869-908 (ComputeViscsWM only writes
// values that SubgridModel overwrites plus the diagnostic WMPoints%ustar;
// BoundViscosity(TDiff) is idempotent on the frozen TDiff and done at init).
template <typename T>
void Engine<T>::subgrid_stresses(bool cfl) {
  if (!sgs_prelaunched)
    timed(ELMM_T_SGS, [&] {
      launch_sgs(g, u, v, w, Visc.as<T>(), cfg.nu, cfg.c_smag, cfg.filter_ratio, sgs_width, cfg.sgs_model,
                 cfl ? cfl_partial.as<double>() : nullptr, cfl_blocks, stream);
    });
  sgs_prelaunched = false;
  // ComputeViscsWM: WMPoints%ustar (the viscosity it writes is overwritten
  // by SubgridModel)
  if (cfg.wall_model && wmp_n)
    timed(ELMM_T_WALL, [&] { launch_wm_ustar<T>(g, u, v, w, temp_cur(), moist_cur(), wmpdev(), cfg.nu, (T *)nullptr, stream); });
  // constant dt with the CFL check: the step's CFL number and the divergence
  // test, on the device (graph-safe)
  if (cfl && !cfg.variable_dt)
    timed(ELMM_T_REDUCE, [&] {
      launch_cfl_monitor(cfl_partial.as<double>(), cfl_blocks, dt_cur, cfg.cfl_max, mon, stream);
    });
  if (cfg.wall_model && wm_n)
    timed(ELMM_T_WALL, [&] { launch_wm_fluxes(g, u, v, w, temp_cur(), moist_cur(), wmdev(), cfg.nu, g.variable_z, stream); });
  timed(ELMM_T_GHOST, [&] { launch_bound_visc<T>(g, Visc.as<T>(), stream); });
  if (ib_n[3])
    timed(ELMM_T_IB, [&] {
      launch_ib_visc_serial(g, ibdev(3), Visc.as<T>(), visc_levels.data(), (int)visc_levels.size() - 1, stream);
    });
  // ComputeTDiff (src/scalar_diffusion.f90:983-1009) with buoyancy, then
  // BoundViscosity(TDiff); without it TDiff keeps its initial values
  if (cfg.enable_buoyancy)
    timed(ELMM_T_OTHER, [&] {
      launch_tdiff_init(g, Visc.as<T>(), TDiff.as<T>(), cfg.nu, cfg.nu_scalar, stream);
      launch_bound_visc<T>(g, TDiff.as<T>(), stream);
    });
}

template <typename T>
StageConst Engine<T>::stage_const(int st) const {
  const double dt = dt_cur;
  StageConst sc{};
  sc.stage = st;
  sc.dt = dt;
  sc.rho = RK_rho[st];
  sc.beta = RK_beta[st];
  sc.rho_dt = RK_rho[st] * dt;
  sc.beta_dt = RK_beta[st] * dt;
  sc.coef = 2 * RK_alpha[st] * dt;
  sc.A = -sc.coef;
  sc.Ax = -sc.coef / g.dxmin;
  sc.Ay = -sc.coef / g.dymin;
  sc.Az = -sc.coef / g.dzmin;
  sc.prgx = cfg.pr_gradient_x;
  sc.prgy = cfg.pr_gradient_y;
  if (cfg.enable_buoyancy) {
    // A = grav_acc / temperature_ref, A2 = A / 2, and A temperature_ref
    const double A = cfg.grav_acc / cfg.temperature_ref;
    sc.buoy = 1;
    sc.moist = cfg.enable_moisture ? 1 : 0;
    sc.bA2 = A / 2.0;
    sc.bATref = A * cfg.temperature_ref;
  }
  return sc;
}

template <typename T>
void Engine<T>::stage(int st, bool last) {
  const double dt = dt_cur;
  const StageConst sc = stage_const(st);

  subgrid_stresses(st == 0 && (cfg.cfl_check || cfg.variable_dt));
  T *ustar[3] = {Us.as<T>(), Vs.as<T>(), Ws.as<T>()};
  T *U2p[3] = {u2, v2, w2};
  if (fused_momentum) {
    const uint8_t *codes[3] = {code[0].as<uint8_t>(), code[1].as<uint8_t>(), code[2].as<uint8_t>()};
    timed(ELMM_T_MOMENTUM, [&] {
      launch_momentum_zm<T>(g, u, v, w, Visc.as<T>(), Pr.as<T>(), tslot >= 0 ? s_cur[tslot] : nullptr, mslot >= 0 ? s_cur[mslot] : nullptr, ustar, U2p,
                            codes, sc, zchunk, mom_tile,
                                 stream);
    });
  }
  if (!fused_momentum)
    for (int c = 0; c < 3; c++)
      timed(ELMM_T_MOMENTUM, [&] {
        launch_momentum(g, c, u, v, w, Visc.as<T>(), Pr.as<T>(), ustar[c], U2p[c], code[c].as<uint8_t>(),
                        sc, stream);
      });
  if (mom_terms[0].n + mom_terms[1].n + mom_terms[2].n) {
    // the three components' ordered corrections read only the stage-start
    // fields and write disjoint arrays: one launch
    const TermList tl3[3] = {mom_terms[0].view(), mom_terms[1].view(), mom_terms[2].view()};
    timed(ELMM_T_MOMENTUM_CORR, [&] {
      launch_momentum_corr3(g, u, v, w, Visc.as<T>(), Pr.as<T>(), ustar, U2p, sc, tl3, wm_fluxp.as<double>(),
                            wm_fluxm.as<double>(), stream);
    });
  }
  if (nsc > 0) {
    // ScalarRK3 (src/scalars.f90:88-299) with the stage-start velocity;
    // the stage-start BoundScalar repeats the previous stage's final fill.
    if (wmp_n)
      timed(ELMM_T_WALL, [&] {
        launch_wm_div(g, WMScalarPoints{wmp_n, wmp_ijk.as<int32_t>(), wmp_div.as<double>()}, prtype.as<int32_t>(), u,
                      v, w, stream);
      });
    for (int s = 0; s < nsc; s++) {
      timed(ELMM_T_SCALAR, [&] {
        if (zmarch_scalar)
          launch_scalar_zm(g, u, v, w, TDiff.as<T>(), s_cur[s], s_nxt[s], Sadv[s]->as<T>(), code_s.as<uint8_t>(),
                           mask[0].as<uint8_t>(), mask[1].as<uint8_t>(), mask[2].as<uint8_t>(), sc, zchunk_scalar,
                           stream);
        else
          launch_scalar(g, u, v, w, TDiff.as<T>(), s_cur[s], s_nxt[s], Sadv[s]->as<T>(), code_s.as<uint8_t>(),
                        mask[0].as<uint8_t>(), mask[1].as<uint8_t>(), mask[2].as<uint8_t>(), sc, stream);
      });
      if (sc_terms[s]->n)
        timed(ELMM_T_SCALAR_CORR, [&] {
          launch_scalar_corr(g, u, v, w, TDiff.as<T>(), s_cur[s], s_nxt[s], Sadv[s]->as<T>(),
                             code_s.as<uint8_t>(), mask[0].as<uint8_t>(), mask[1].as<uint8_t>(),
                             mask[2].as<uint8_t>(), sc, sc_terms[s]->view(), src_flux.as<double>(),
                             wmp_div.as<double>(), stream);
        });
      timed(ELMM_T_GHOST, [&] { launch_bound_scalar<T>(gscal(s), s_nxt[s], stream); });
      if (s == tslot && tfix_n)
        timed(ELMM_T_SCALAR_CORR, [&] {
          launch_temp_solid_fix<T>(g, tfix_ijk.as<int32_t>(), tfix_n, prtype.as<int32_t>(), s_nxt[s], stream);
        });
      if (s == mslot)
        timed(ELMM_T_SCALAR_CORR,
              [&] { launch_moist_fill<T>(g, code_s.as<uint8_t>(), s_nxt[s], cfg.moisture_ref, stream); });
      std::swap(s_cur[s], s_nxt[s]);
    }
  }
  bound_uvw(u2, v2, w2);
  const bool use_q = cfg.ib_mass_sources && has_ib();
  if (has_ib()) {
    ib_momentum(u2, v2, w2);
    bound_uvw_after_ib(u2, v2, w2);
  }
  if (use_q) timed(ELMM_T_IB, [&] { launch_mass_sources(g, q_terms.view(), u2, v2, w2, Q.as<double>(), stream); });
  pressure_correction(u2, v2, w2, use_q, sc.coef, last && want_diag);
  if (cfg.delta_diag)
    timed(ELMM_T_REDUCE, [&] {
      launch_delta<T>(g, u, u2, v, v2, w, w2, delta_partial.as<double>(), dblocks, st, dt, mon, stream);
    });
  if (last && want_diag) {
    bound_uvw(u2, v2, w2);
    last_stored_div = max_div(u2, v2, w2, use_q);
    if (refine_iters == 0 && !solve64) last_div = last_stored_div;
  }
  std::swap(u, u2);
  std::swap(v, v2);
  std::swap(w, w2);
  if (null_n[0] + null_n[1] + null_n[2]) {
    const int32_t *const nl[3] = {null_ijk[0].as<int32_t>(), null_ijk[1].as<int32_t>(), null_ijk[2].as<int32_t>()};
    timed(ELMM_T_IB, [&] { launch_null(g, null_n, nl, u, v, w, stream); });
  }
  bound_uvw(u, v, w);
  if (has_ib()) {
    ib_momentum(u, v, w);
    bound_uvw_after_ib(u, v, w);
  }
}

template <typename T>
void Engine<T>::init() {
  // src/initial.f90:4284-4342
  SetupTrace tr;
  bound_uvw(u, v, w);
  launch_bound_pr<T>(g, Pr.as<T>(), stream);
  tr.mark("init bounds");
  pressure_correction(u, v, w, false, 1.0);
  tr.mark("init projection");
  launch_sgs(g, u, v, w, Visc.as<T>(), cfg.nu, cfg.c_smag, cfg.filter_ratio,
             sgs_width, cfg.sgs_model, nullptr, 0, stream);
  launch_bound_visc<T>(g, Visc.as<T>(), stream);
  if (nsc > 0) {
    launch_tdiff_init(g, Visc.as<T>(), TDiff.as<T>(), cfg.nu, cfg.nu_scalar, stream);
    launch_bound_visc<T>(g, TDiff.as<T>(), stream);
  }
  for (int s = 0; s < nsc; s++) launch_bound_scalar<T>(gscal(s), s_cur[s], stream);
  if (cfg.wall_model && wmp_n) launch_wm_ustar<T>(g, u, v, w, temp_cur(), moist_cur(), wmpdev(), cfg.nu, Visc.as<T>(), stream);
  if (cfg.wall_model) launch_wm_fluxes(g, u, v, w, temp_cur(), moist_cur(), wmdev(), cfg.nu, g.variable_z, stream);
  launch_bound_visc<T>(g, Visc.as<T>(), stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  tr.mark("init rest");
  initialized = true;
}

// first call of TMarchRK3 (src/tsteps.f90:54-77)
template <typename T>
void Engine<T>::first_call() {
  if (first_step_done) return;
  bound_uvw(u, v, w);
  if (has_ib()) ib_momentum(u, v, w);
  first_step_done = true;
}

// Enqueues the largest even number of plain steps (<= nsteps) as replays of a
// captured graph of two steps and returns that number.  Every stage swaps
// the U/U2 (and scalar) pointers, so two steps restore them: one graph per
// pointer parity covers every call.  The captured work is kernels only (no
// memset or copy nodes); timed runs and ELMM_GRAPH=0 enqueue directly.
template <typename T>
int Engine<T>::replay_steps(int nsteps) {
  // a variable dt changes the stage constants every step: no replay
  // nor for the multigrid solve, whose cycle count depends on the data
  if (!use_graphs || timers_on || nsteps < 2 || !first_step_done || cfg.variable_dt || use_mg) return 0;
  const int par = u == U.as<T>() ? 0 : 1;
  if (!graph_exec[par]) {
    // (a host operation of this thread still open would block the capture
    // lock: enqueue directly then)
    if (t_host_ops > 0) return 0;
    CaptureLock capture;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < 2; r++)
      for (int st = 0; st < 3; st++) stage(st, false);
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamEndCapture(stream, &gr));
    const hipError_t e = hipGraphInstantiate(&graph_exec[par], gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIPCHK(e);
  }
  const int n2 = nsteps / 2;
  for (int r = 0; r < n2; r++) HIPCHK(hipGraphLaunch(graph_exec[par], stream));
  return 2 * n2;
}

// TimeStepLength with variable time steps (src/dynamics.f90:517-549): the
// first stage's SGS kernel writes the CFL block maxima of the step-start
// velocity; their maximum gives dt = min(CFL / max, dt_max) (dt_max for a
// field at rest, the smallest normal number for a NaN velocity), cut to reach
// end_time.  Synchronises the stream.
template <typename T>
void Engine<T>::choose_dt() {
  timed(ELMM_T_SGS, [&] {
    launch_sgs(g, u, v, w, Visc.as<T>(), cfg.nu, cfg.c_smag, cfg.filter_ratio, sgs_width, cfg.sgs_model,
               cfl_partial.as<double>(), cfl_blocks, stream);
  });
  sgs_prelaunched = true;
  const double m = reduce_max(cfl_partial, cfl_blocks);  // +inf for a NaN velocity
  double dt;
  if (!(m < INFINITY))
    dt = std::numeric_limits<double>::min();
  else if (m > 0)
    dt = std::min(cfg.cfl_target / m, cfg.dt_max);
  else
    dt = cfg.dt_max;
  if (cfg.end_time > 0 && dt + time > cfg.end_time) dt = cfg.end_time - time;
  dt_cur = dt;
  last_cfl = m * dt;
}

// ---------------------------------------------------------------- autotuner
// The reference sizes its CPU loop tiles from the L1 size and the number of
// arrays a loop touches (InitTiles, src/tiling.f90:16-111).  Here the tile
// shapes and z-chunk lengths of the z-marching kernels are measured: on the
// first step of a handle (before any step graph is captured) every candidate
// runs on the case's own fields with the first stage's constants, timed with
// HIP events (one warm-up launch, then three timed), and the fastest is
// kept.  Every candidate computes the same bits (only the work split
// differs).  The launches write only what the first stage then rewrites
// (Ustar and U2 interiors, the viscosity, the new scalar and its tendency),
// so tuning leaves the step unchanged.  Choices are cached per process by
// field type and grid shape; settings fixed by environment variables are
// kept.  ELMM_TUNE_LOG=1 prints the choices.
namespace {
struct TileChoice {
  int mom_tile, zchunk, sgs_zchunk, scalar_zchunk;
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, TileChoice> g_tune_cache;
}  // namespace

template <typename T>
void Engine<T>::autotune() {
  tuned = true;
  const auto key = std::make_tuple((int)sizeof(T), cfg.nx, cfg.ny, cfg.nz, nsc);
  TileChoice c{mom_tile, zchunk, g.sgs_zchunk, zchunk_scalar};
  bool hit = false;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      c = it->second;
      hit = true;
    }
  }
  if (!hit) {
    hipEvent_t e0, e1;
    {
      HostOpLock host;
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
    }
    auto time_of = [&](auto launch) {
      launch();
      HIPCHK(hipEventRecord(e0, stream));
      for (int r = 0; r < 3; r++) launch();
      HIPCHK(hipEventRecord(e1, stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    const StageConst sc0 = stage_const(0);
    if (fused_momentum) {
      T *ustar[3] = {Us.as<T>(), Vs.as<T>(), Ws.as<T>()};
      T *U2p[3] = {u2, v2, w2};
      const uint8_t *codes[3] = {code[0].as<uint8_t>(), code[1].as<uint8_t>(), code[2].as<uint8_t>()};
      std::vector<int> tiles = env_mom_tile ? std::vector<int>{mom_tile}
                                            : std::vector<int>{kMomTile64x4, kMomTile32x8};
      if (!env_mom_tile && sizeof(T) == 4) tiles.push_back(kMomTile64x8);
      const std::vector<int> chunks = env_zchunk ? std::vector<int>{zchunk} : std::vector<int>{64, 32};
      // candidates timed in three interleaved rounds, each keeping its best
      // time (a clock or cache drift between candidates timed one after the
      // other flipped close choices from run to run)
      std::vector<std::pair<int, int>> cand;
      for (int t : tiles)
        for (int ch : chunks) cand.emplace_back(t, ch);
      // Each timed launch follows the kernel that precedes it in the step
      // (the SGS z-march, which reads the same velocity fields), so the
      // caches hold what they hold in the step: timed back to back instead,
      // the tile that is slower in the step measured faster.
      std::vector<float> best_of(cand.size(), 1e30f);
      for (int round = 0; round < 3; round++)
        for (size_t k = 0; k < cand.size(); k++) {
          float ms = 0;
          for (int r = 0; r < 3; r++) {
            launch_sgs(g, u, v, w, Visc.as<T>(), cfg.nu, cfg.c_smag, cfg.filter_ratio, sgs_width, cfg.sgs_model,
                       (double *)nullptr, 0, stream);
            HIPCHK(hipEventRecord(e0, stream));
            launch_momentum_zm<T>(g, u, v, w, Visc.as<T>(), Pr.as<T>(), tslot >= 0 ? s_cur[tslot] : nullptr,
                                  mslot >= 0 ? s_cur[mslot] : nullptr, ustar, U2p, codes, sc0, cand[k].second,
                                  cand[k].first, stream);
            HIPCHK(hipEventRecord(e1, stream));
            HIPCHK(hipEventSynchronize(e1));
            float one = 0;
            HIPCHK(hipEventElapsedTime(&one, e0, e1));
            if (round > 0 || r > 0) ms += one;  // the very first launch of a candidate warms up
          }
          if (round > 0) best_of[k] = std::min(best_of[k], ms);
        }
      float best = 1e30f;
      for (size_t k = 0; k < cand.size(); k++)
        if (best_of[k] < best) {
          best = best_of[k];
          c.mom_tile = cand[k].first;
          c.zchunk = cand[k].second;
        }
    }
    if (!env_sgs_zchunk && !g.sgs_cellwise && !g.variable_z && sgs_width.tab) {
      const int saved = g.sgs_zchunk;
      float best = 1e30f;
      for (int ch : {26, 32, 16, 13}) {
        g.sgs_zchunk = ch;
        const float ms = time_of([&] {
          launch_sgs(g, u, v, w, Visc.as<T>(), cfg.nu, cfg.c_smag, cfg.filter_ratio, sgs_width, cfg.sgs_model,
                     (double *)nullptr, 0, stream);
        });
        if (ms < best) {
          best = ms;
          c.sgs_zchunk = ch;
        }
      }
      g.sgs_zchunk = saved;
    }
    if (zmarch_scalar && nsc > 0 && !env_scalar_zchunk) {
      float best = 1e30f;
      for (int ch : {16, 13, 26, 32}) {
        const float ms = time_of([&] {
          launch_scalar_zm(g, u, v, w, TDiff.as<T>(), s_cur[0], s_nxt[0], Sadv[0]->as<T>(), code_s.as<uint8_t>(),
                           mask[0].as<uint8_t>(), mask[1].as<uint8_t>(), mask[2].as<uint8_t>(), sc0, ch, stream);
        });
        if (ms < best) {
          best = ms;
          c.scalar_zchunk = ch;
        }
      }
    }
    {
      HostOpLock host;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[key] = c;
  }
  mom_tile = c.mom_tile;
  zchunk = c.zchunk;
  g.sgs_zchunk = c.sgs_zchunk;
  zchunk_scalar = c.scalar_zchunk;
  cfl_blocks = sgs_partials(g, sgs_width);
  if (std::getenv("ELMM_TUNE_LOG"))
    std::fprintf(stderr, "[elmm tune] %dx%dx%d %s%s: momentum tile %s, z chunk %d; sgs z chunk %d; scalar z chunk %d\n",
                 cfg.nx, cfg.ny, cfg.nz, sizeof(T) == 4 ? "fp32r" : "fp64", hit ? " (cached)" : "",
                 c.mom_tile == kMomTile64x4 ? "64x4" : c.mom_tile == kMomTile32x8 ? "32x8" : "64x8", c.zchunk,
                 c.sgs_zchunk > 0 ? c.sgs_zchunk : 26, c.scalar_zchunk);
}

// Steps of TMarchRK3 and the main loop's time update (src/main.f90:112-169).
// Constant dt: every step but the last (which carries the diagnostics) may
// replay a step graph; the CFL monitor runs on the device and is read once
// at the end.  Variable dt: one synchronisation per step to find dt; a dt
// below dt_min from step 3 on ends the call as divergence (main.f90:153-159).
template <typename T>
void Engine<T>::run_steps(int nsteps, bool sync_diag) {
  if (cfg.autotune && !tuned && nsteps > 0) autotune();
  int s = 0;
  auto one = [&](bool last) {
    if (cfg.variable_dt) choose_dt();
    for (int st = 0; st < 3; st++) stage(st, last && st == 2);
    advance_time(dt_cur);
    if (cfg.variable_dt && steps_total >= 3 && dt_cur < cfg.dt_min && diverged_step < 0) {
      diverged_step = steps_total;
      diverged_msg = "Solution diverged. 1 (dt " + std::to_string(dt_cur) + " below dt_min at step " +
                     std::to_string(steps_total) + ")";
    }
  };
  if (nsteps > 0 && !first_step_done) {
    first_call();
    one(sync_diag && nsteps == 1);
    s = 1;
  }
  if (nsteps - s > 1 && diverged_step < 0) {
    const int r = replay_steps(sync_diag ? nsteps - s - 1 : nsteps - s);
    for (int q = 0; q < r; q++) advance_time(dt_cur);
    s += r;
  }
  for (; s < nsteps && !(cfg.variable_dt && diverged_step >= 0); s++) one(sync_diag && s == nsteps - 1);
  steps_pending += s;
  HIPCHK(hipGetLastError());
}

template <typename T>
void Engine<T>::step(int nsteps) {
  run_steps(nsteps, true);
  HIPCHK(hipStreamSynchronize(stream));
  read_monitor();
  collect_timers();
}

// Steps enqueued on the engine's stream without a host synchronisation
// (a variable dt synchronises once per step).
template <typename T>
void Engine<T>::enqueue_steps(int nsteps) {
  want_diag = false;
  run_steps(nsteps, false);
}

template <typename T>
double Engine<T>::divergence() {
  bound_uvw(u, v, w);
  const bool use_q = cfg.ib_mass_sources && has_ib() && first_step_done;
  return max_div(u, v, w, use_q);
}

// divergence() over caller-provided device memory: the block maxima into
// `partials`, their maximum (a NaN counts as +inf) into `result` on the
// device, then 8 bytes copied to the host.  Launches kernels only.
template <typename T>
double Engine<T>::divergence_into(double *partials, long partial_slots, double *result) {
  if (red_blocks > partial_slots) throw ArgError("divergence scratch too small for the grid");
  bound_uvw(u, v, w);
  const bool use_q = cfg.ib_mass_sources && has_ib() && first_step_done;
  launch_max_div(g, u, v, w, use_q ? Q.as<double>() : nullptr, partials, red_blocks, stream);
  launch_max_final(partials, red_blocks, result, stream);
  double m = 0;
  HIPCHK(hipMemcpyAsync(&m, result, sizeof m, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return m;
}

// Field frame without a host synchronisation (the reference writes frames
// from pthreads, src/frames.f90:369-420): device-to-host copies of the
// current U, V, W, Pr and scalars into pinned staging on the engine's stream
// (so the frame is the state after every step enqueued before it), then a
// writer thread waits for the copies, converts them to ghosted FP64 arrays
// and writes out.vtk, scalars.vtk and U/V/W.vtk (elmm_vtk_write).  Steps
// enqueued afterwards run while the files are written.  A frame started
// while another is being written first joins it.
template <typename T>
void Engine<T>::frame_async(const std::string &dir) {
  if (frame_thread.joinable()) frame_wait();
  const int nf = 4 + cfg.nscalars;  // U, V, W, Pr and the passive scalars
  if (frame_host.empty()) {
    for (int f = 0; f < nf; f++) {
      void *p = nullptr;
      HIPCHK(hipHostMalloc(&p, (size_t)g.ntot * sizeof(T), hipHostMallocDefault));
      frame_host.push_back(p);
    }
    HIPCHK(hipEventCreateWithFlags(&frame_done, hipEventDisableTiming));
  }
  const void *src[12] = {u, v, w, Pr.p};
  for (int s = 0; s < cfg.nscalars; s++) src[4 + s] = s_cur[s];
  for (int f = 0; f < nf; f++)
    HIPCHK(hipMemcpyAsync(frame_host[f], src[f], (size_t)g.ntot * sizeof(T), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipEventRecord(frame_done, stream));
  frame_status = 0;
  frame_error.clear();
  frame_thread = std::thread([this, dir, nf] {
    if (hipEventSynchronize(frame_done) != hipSuccess) {
      frame_status = ELMM_E_HIP;
      frame_error = "frame copies failed";
      return;
    }
    const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
    const long hy = nx + 6, hz = (long)(nx + 6) * (ny + 6), hn = hz * (nz + 6);
    std::vector<std::vector<double>> host(nf, std::vector<double>(hn, 0.0));
    for (int f = 0; f < nf; f++) {
      const T *d = static_cast<const T *>(frame_host[f]);
      for (int k = -2; k <= nz + 3; k++)
        for (int j = -2; j <= ny + 3; j++)
          for (int i = -2; i <= nx + 3; i++)
            host[f][(i + 2) + (j + 2) * hy + (k + 2) * hz] = (double)d[gix(g, i, j, k)];
    }
    std::vector<const double *> sp;
    for (int s = 0; s < cfg.nscalars; s++) sp.push_back(host[4 + s].data());
    const int rc = elmm_vtk_write(dir.c_str(), &cfg, host[0].data(), host[1].data(), host[2].data(), host[3].data(),
                                  nullptr, cfg.nscalars, sp.empty() ? nullptr : sp.data());
    if (rc) {
      frame_status = rc;
      frame_error = elmm_last_error();
    }
  });
}

template <typename T>
int Engine<T>::frame_wait() {
  if (frame_thread.joinable()) frame_thread.join();
  return frame_status;
}

template struct Engine<double>;
template struct Engine<float>;

}  // namespace elmm

using elmm::ArgError;
using elmm::HipError;

struct elmm_handle {
  std::unique_ptr<elmm::EngineIface> e;
};

template <typename F>
static int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const ArgError &a) {
    return elmm::fail(ELMM_E_ARG, a.msg);
  } catch (const HipError &h) {
    return elmm::fail(ELMM_E_HIP, h.msg);
  } catch (const std::bad_alloc &) {
    return elmm::fail(ELMM_E_NOMEM, "out of host memory");
  }
}

static void release_engine(elmm::EngineIface *e) {
  if (!e || !e->stream) return;
  (void)hipStreamSynchronize(e->stream);
  e->release_graphs();
  for (auto &t : e->pending) {
    (void)hipEventDestroy(std::get<1>(t));
    (void)hipEventDestroy(std::get<2>(t));
  }
  e->pending.clear();
  for (auto ev : e->event_pool) (void)hipEventDestroy(ev);
  e->event_pool.clear();
  (void)hipStreamDestroy(e->stream);
  e->stream = nullptr;
}

// ------------------------------------------- allocation-free divergence
// elmm_divergence_noalloc keeps its block maxima and final maximum in this
// file-scope device array instead of handle memory: one block maximum per
// 256 cells, so grids of at most kNoallocMaxCells cells.  Every device the
// library runs on holds its own copy until the process ends; the host uses
// the per-device address from hipGetSymbolAddress (the array's name is a
// host-side stand-in, not a device address).  Every block maximum and the
// result are rewritten by each call, so nothing needs clearing; calls that
// overlap on one device race on the array.
constexpr long kNoallocSlots = 1L << 18;
constexpr long kNoallocMaxCells = kNoallocSlots * 256;
static __device__ __align__(256) double scratch[kNoallocSlots + 32];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (called from elmm_create).
int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return elmm::fail(ELMM_E_ARG, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return elmm::fail(ELMM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = static_cast<double *>(ptr);
  return 0;
}

double *scratch_of(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" int elmm_create(const elmm_config *cfg, const elmm_lists *lists, elmm_handle **out) {
  if (!cfg || !lists || !out) return elmm::fail(ELMM_E_ARG, "null argument to elmm_create");
  *out = nullptr;
  if (cfg->precision != ELMM_FP64 && cfg->precision != ELMM_FP32_REFINED)
    return elmm::fail(ELMM_E_ARG, "precision must be ELMM_FP64 or ELMM_FP32_REFINED");
  elmm::HostOpLock host;
  auto h = std::make_unique<elmm_handle>();
  if (cfg->precision == ELMM_FP64)
    h->e.reset(new elmm::Engine<double>());
  else
    h->e.reset(new elmm::Engine<float>());
  int rc = guarded([&] { h->e->create(*cfg, *lists); });
  if (rc == 0) rc = scratch_setup(h->e->device);
  if (rc == 0)
    *out = h.release();
  else
    release_engine(h->e.get());
  return rc;
}

extern "C" int elmm_create_case(const elmm_config *cfg, const elmm_case *c, elmm_handle **out) {
  if (!cfg || !c || !out) return elmm::fail(ELMM_E_ARG, "null argument to elmm_create_case");
  elmm_lists v;
  int rc = elmm_case_view(c, &v);
  if (rc) return rc;
  return elmm_create(cfg, &v, out);
}

extern "C" int elmm_upload(elmm_handle *h, int32_t field, int32_t lb, const int32_t ext[3], const double *src) {
  if (!h || !ext || !src) return elmm::fail(ELMM_E_ARG, "null argument to elmm_upload");
  elmm::HostOpLock host;
  return guarded([&] { h->e->xfer(field, lb, ext, const_cast<double *>(src), true); });
}

extern "C" int elmm_download(elmm_handle *h, int32_t field, int32_t lb, const int32_t ext[3], double *dst) {
  if (!h || !ext || !dst) return elmm::fail(ELMM_E_ARG, "null argument to elmm_download");
  elmm::HostOpLock host;
  return guarded([&] { h->e->xfer(field, lb, ext, dst, false); });
}

extern "C" int elmm_init(elmm_handle *h) {
  if (!h) return elmm::fail(ELMM_E_ARG, "null handle");
  elmm::HostOpLock host;
  return guarded([&] { h->e->init(); });
}

// The main loop's divergence tests (src/main.f90:153-168): a variable dt
// below dt_min, or a CFL number above CFL_max (a NaN velocity counts as
// above), from step 3 on.  The call completes its steps and reports the
// first step that failed.
static int check_diverged(elmm_handle *h) {
  if (h->e->diverged_step >= 0) return elmm::fail(ELMM_E_DIVERGED, h->e->diverged_msg);
  return 0;
}

static void fill_diag(elmm_handle *h, elmm_diag *diag) {
  std::memset(diag, 0, sizeof *diag);
  diag->cfl = h->e->last_cfl;
  diag->delta = h->e->last_delta;
  diag->max_div = h->e->last_div;
  diag->stored_div = h->e->last_stored_div;
  diag->poisson_res = h->e->last_res;
  diag->refinements = h->e->last_refinements;
  diag->steps_done = h->e->steps_pending;
  diag->dt = h->e->dt_cur;
  diag->time = h->e->time;
}

extern "C" int elmm_step(elmm_handle *h, int32_t nsteps, elmm_diag *diag) {
  if (!h) return elmm::fail(ELMM_E_ARG, "null handle");
  if (nsteps < 0) return elmm::fail(ELMM_E_ARG, "nsteps must be >= 0");
  if (!h->e->initialized) return elmm::fail(ELMM_E_STATE, "elmm_step before elmm_init");
  h->e->want_diag = diag != nullptr;
  h->e->steps_pending = 0;
  int rc = guarded([&] { h->e->step(nsteps); });
  if (rc) return rc;
  if (diag) fill_diag(h, diag);
  h->e->steps_pending = 0;
  return check_diverged(h);
}

extern "C" int elmm_step_async(elmm_handle *h, int32_t nsteps) {
  if (!h) return elmm::fail(ELMM_E_ARG, "null handle");
  if (!h->e->initialized) return elmm::fail(ELMM_E_STATE, "elmm_step_async before elmm_init");
  h->e->want_diag = false;
  if (nsteps < 0) return elmm::fail(ELMM_E_ARG, "nsteps must be >= 0");
  return guarded([&] { h->e->enqueue_steps(nsteps); });
}

// Joins the enqueued steps; the diagnostics are those of a call without
// diag (CFL, delta, dt, time, steps since the last join) and the divergence
// tests cover every joined step.
extern "C" int elmm_wait(elmm_handle *h, elmm_diag *diag) {
  if (!h) return elmm::fail(ELMM_E_ARG, "null handle");
  int rc = guarded([&] {
    HIPCHK(hipStreamSynchronize(h->e->stream));
    h->e->read_monitor();
  });
  if (rc) return rc;
  if (diag) fill_diag(h, diag);
  h->e->steps_pending = 0;
  return check_diverged(h);
}

extern "C" int elmm_query(elmm_handle *h, int32_t *idle) {
  if (!h || !idle) return elmm::fail(ELMM_E_ARG, "null argument to elmm_query");
  const hipError_t e = hipStreamQuery(h->e->stream);
  if (e == hipSuccess) {
    *idle = 1;
    return 0;
  }
  if (e == hipErrorNotReady) {
    *idle = 0;
    return 0;
  }
  return elmm::fail(ELMM_E_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

extern "C" int elmm_divergence(elmm_handle *h, double *max_div) {
  if (!h || !max_div) return elmm::fail(ELMM_E_ARG, "null argument to elmm_divergence");
  elmm::HostOpLock host;
  return guarded([&] { *max_div = h->e->divergence(); });
}

extern "C" int elmm_divergence_noalloc(elmm_handle *h, double *max_div) {
  if (!h || !max_div) return elmm::fail(ELMM_E_ARG, "null argument to elmm_divergence_noalloc");
  const elmm_config &c = h->e->cfg;
  const long cells = (long)c.nx * c.ny * c.nz;
  if (cells > kNoallocMaxCells)
    return elmm::fail(ELMM_E_ARG, "elmm_divergence_noalloc: grid of " + std::to_string(cells) +
                                      " cells exceeds its scratch limit of " + std::to_string(kNoallocMaxCells) +
                                      " cells (use elmm_divergence)");
  double *ptr = scratch_of(h->e->device);
  if (!ptr) return elmm::fail(ELMM_E_STATE, "elmm_divergence_noalloc: no scratch address for the handle's device");
  // slots [0, kNoallocSlots): block maxima; slot kNoallocSlots: the result
  elmm::HostOpLock host;
  return guarded([&] { *max_div = h->e->divergence_into(ptr, kNoallocSlots, ptr + kNoallocSlots); });
}

extern "C" int elmm_wm_ustar(elmm_handle *h, int32_t cap, double *ustar, int32_t *n) {
  if (!h || !n) return elmm::fail(ELMM_E_ARG, "null argument to elmm_wm_ustar");
  elmm::HostOpLock host;
  return guarded([&] {
    const int m = h->e->point_ustar(nullptr, 0);
    *n = m;
    if (ustar) {
      if (cap < m) throw elmm::ArgError("elmm_wm_ustar: cap " + std::to_string(cap) + " below the " + std::to_string(m) +
                                        " wall-model points");
      h->e->point_ustar(ustar, cap);
    }
  });
}

extern "C" int elmm_timers_enable(elmm_handle *h, int32_t enable) {
  if (!h) return elmm::fail(ELMM_E_ARG, "null handle");
  elmm::HostOpLock host;
  return guarded([&] {
    HIPCHK(hipStreamSynchronize(h->e->stream));
    h->e->collect_timers();
    h->e->timers_on = enable != 0;
  });
}

extern "C" int elmm_timers_read(elmm_handle *h, elmm_timers *out, int32_t reset) {
  if (!h || !out) return elmm::fail(ELMM_E_ARG, "null argument to elmm_timers_read");
  elmm::HostOpLock host;
  return guarded([&] {
    HIPCHK(hipStreamSynchronize(h->e->stream));
    h->e->collect_timers();
    for (int g = 0; g < ELMM_T_NGROUPS; g++) {
      out->ms[g] = h->e->tms[g];
      out->launches[g] = h->e->tlaunch[g];
      if (reset) {
        h->e->tms[g] = 0;
        h->e->tlaunch[g] = 0;
      }
    }
  });
}

extern "C" const char *elmm_timer_name(int32_t g) {
  static const char *names[ELMM_T_NGROUPS] = {
      "sgs", "wall_model", "momentum", "momentum_corr", "scalar", "scalar_corr", "ghost", "immersed_boundary",
      "poisson_x_fwd", "poisson_y", "poisson_z", "poisson_x_inv", "projection", "reduce", "other", "refine"};
  return (g >= 0 && g < ELMM_T_NGROUPS) ? names[g] : "invalid";
}

extern "C" int elmm_frame_async(elmm_handle *h, const char *dir) {
  if (!h || !dir) return elmm::fail(ELMM_E_ARG, "null argument to elmm_frame_async");
  if (h->e->cfg.nscalars > 8) return elmm::fail(ELMM_E_ARG, "frames hold at most 8 scalars");
  elmm::HostOpLock host;
  return guarded([&] { h->e->frame_async(dir); });
}

extern "C" int elmm_frame_wait(elmm_handle *h) {
  if (!h) return elmm::fail(ELMM_E_ARG, "null handle");
  elmm::HostOpLock host;
  const int rc = h->e->frame_wait();
  if (rc) return elmm::fail(rc, static_cast<elmm::EngineIface *>(h->e.get())->frame_error_text());
  return 0;
}

extern "C" int elmm_get_tiles(elmm_handle *h, elmm_tiles *out) {
  if (!h || !out) return elmm::fail(ELMM_E_ARG, "null argument to elmm_get_tiles");
  h->e->tiles(out);
  return 0;
}

extern "C" int elmm_destroy(elmm_handle *h) {
  if (!h) return 0;
  elmm::HostOpLock host;  // frees
  release_engine(h->e.get());
  delete h;
  return 0;
}
