// server.cpp - continuous-batching request loop over device-resident cases.
//
// Requests (a case, its initial fields and a run length) are queued and
// admitted in arrival order while their HBM footprint fits the budget and a
// slot is free.  Every admitted case keeps its fields resident on its own
// stream; progress() enqueues the next chunk of steps of every idle case, so
// the work of all active cases overlaps on the device, retires finished cases
// (results copied to the caller's buffers, memory returned to the budget) and
// admits queued requests as memory frees up.  Latency = completion - submit.
//
// Admission (handle creation with its list set-up, upload, initial
// projection, first chunk) and retirement (downloads, divergence, release)
// run on worker threads, at most kWorkers of each at a time, so that the
// host work of one request overlaps the steps of the others instead of
// holding up their next chunks.  A handle is used by one thread at a time:
// its worker until the task ends, then the thread that calls progress().
//
// No counterpart in the reference, which runs one case per process
// (src/main.f90); this is the serving form of the step.
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "case_setup.hpp"
#include "elmm_gpu.h"
#include "errors.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  elmm_config cfg{};
  std::vector<double> zw;
  const elmm_case *c = nullptr;
  int32_t nsteps = 0;
  // initial state (copied at submit), ghosted with lower bound -2
  std::vector<double> U, V, W, Pr;
  std::vector<std::vector<double>> S;
  // output buffers owned by the caller
  double *oU = nullptr, *oV = nullptr, *oW = nullptr, *oPr = nullptr;
  std::vector<double *> oS;
  size_t footprint = 0;
  Clock::time_point submitted;
  // run state
  elmm_handle *h = nullptr;
  int32_t done = 0;
  bool finished = false;
  int status = 0;
  double latency = 0;
  double max_div = 0;
  std::string error;      // message of a failure (taken on the failing thread)
  std::future<int> task;  // admission or retirement in flight on a worker
};

constexpr int kWorkers = 4;

bool task_ready(Request &r) { return r.task.wait_for(std::chrono::seconds(0)) == std::future_status::ready; }

}  // namespace

struct elmm_server {
  size_t budget = 0, used = 0;
  int device = 0;  // current at elmm_server_create; the workers run on it
  int max_active = 0, chunk = 0;
  int64_t next_id = 1;
  std::deque<std::unique_ptr<Request>> queue;
  std::vector<std::unique_ptr<Request>> preparing, active, retiring;
  std::unordered_map<int64_t, std::unique_ptr<Request>> done;
  int64_t cell_steps = 0;
};

namespace {

int fail_req(Request &r, int rc) {
  r.finished = true;
  r.status = rc;
  r.error = elmm::last_error();
  if (r.h) {
    elmm_destroy(r.h);
    r.h = nullptr;
  }
  return rc;
}

// create + upload + init, then the first chunk (a worker thread)
int admit(int chunk, Request &r) {
  int rc = elmm_create_case(&r.cfg, r.c, &r.h);
  if (rc) return fail_req(r, rc);
  const int32_t e[3] = {r.cfg.nx + 6, r.cfg.ny + 6, r.cfg.nz + 6};
  if ((rc = elmm_upload(r.h, ELMM_F_U, -2, e, r.U.data())) || (rc = elmm_upload(r.h, ELMM_F_V, -2, e, r.V.data())) ||
      (rc = elmm_upload(r.h, ELMM_F_W, -2, e, r.W.data())) || (rc = elmm_upload(r.h, ELMM_F_PR, -2, e, r.Pr.data())))
    return fail_req(r, rc);
  for (size_t q = 0; q < r.S.size(); q++)
    if ((rc = elmm_upload(r.h, ELMM_F_SCALAR + (int)q, -2, e, r.S[q].data()))) return fail_req(r, rc);
  // the host copies are no longer needed
  r.U = r.V = r.W = r.Pr = std::vector<double>();
  r.S.clear();
  if ((rc = elmm_init(r.h))) return fail_req(r, rc);
  const int32_t n = std::min(chunk, r.nsteps);
  if (n > 0 && (rc = elmm_step_async(r.h, n))) return fail_req(r, rc);
  r.done = n;
  return 0;
}

// results to the caller's buffers, divergence, release (a worker thread)
int retire(Request &r) {
  const int32_t e[3] = {r.cfg.nx + 6, r.cfg.ny + 6, r.cfg.nz + 6};
  int rc = elmm_wait(r.h, nullptr);
  if (!rc && r.oU) rc = elmm_download(r.h, ELMM_F_U, -2, e, r.oU);
  if (!rc && r.oV) rc = elmm_download(r.h, ELMM_F_V, -2, e, r.oV);
  if (!rc && r.oW) rc = elmm_download(r.h, ELMM_F_W, -2, e, r.oW);
  if (!rc && r.oPr) rc = elmm_download(r.h, ELMM_F_PR, -2, e, r.oPr);
  for (size_t q = 0; !rc && q < r.oS.size(); q++)
    if (r.oS[q]) rc = elmm_download(r.h, ELMM_F_SCALAR + (int)q, -2, e, r.oS[q]);
  if (!rc) rc = elmm_divergence(r.h, &r.max_div);
  // a NaN or infinite divergence is a diverged run even without the CFL test
  if (!rc && !std::isfinite(r.max_div)) rc = elmm::fail(ELMM_E_DIVERGED, "Solution diverged (non-finite divergence)");
  r.latency = std::chrono::duration<double>(Clock::now() - r.submitted).count();
  if (rc) r.error = elmm::last_error();
  elmm_destroy(r.h);
  r.h = nullptr;
  r.finished = true;
  r.status = rc;
  return rc;
}

}  // namespace

extern "C" {

int elmm_footprint(const elmm_config *cfg, int64_t *bytes) {
  if (!cfg || !bytes) return elmm::fail(ELMM_E_ARG, "null argument to elmm_footprint");
  // device layout: rows padded to a multiple of 16 after 13 front elements
  const long row = 13 + cfg->nx + 6;
  const long sy = (row + 15) / 16 * 16;
  const long ntot = sy * (cfg->ny + 6) * (cfg->nz + 6) + 16;
  const int e = cfg->precision == ELMM_FP32_REFINED ? 4 : 8;
  // 12 velocity/pressure/viscosity fields of the field type, 3 per scalar,
  // Q, Phi and RHS in FP64; per cell the advection codes of U, V, W and the
  // scalar code, the three scalar flux masks (7 bytes) and Prtype (int32);
  // the one-byte mark array of the set-up (freed after elmm_create, counted
  // because it is live at the peak)
  const long M = cfg->nx / 2 + 1;
  const int es = cfg->precision == ELMM_FP32_REFINED && cfg->refine_iters > 0 ? 4 : 8;
  const long Mp = es == 8 ? (M + 7) / 8 * 8 : (M + 15) / 16 * 16;  // 128-byte rows
  // the temperature and the moisture are one more slot each
  const int nslots = cfg->nscalars + (cfg->enable_buoyancy ? 1 : 0) + (cfg->enable_moisture ? 1 : 0);
  int64_t b = (int64_t)ntot * (12 * e + 3 * e * nslots + 3 * 8 + 7 + 4 + 1);
  if (cfg->poisson_solver == ELMM_POISSON_MG) {
    // multigrid: phi and rhs per level (x and y halved: at most 4/3 of the
    // cells) and the coarsest level's work arrays (below one level 0 array)
    b += (int64_t)cfg->nx * cfg->ny * cfg->nz * 2 * 8 * 4 / 3 + (int64_t)cfg->nx * cfg->ny * cfg->nz * 8;
  } else {
    // spectral array: complex FP32 with refinement sweeps, complex FP64 otherwise
    b += (int64_t)Mp * cfg->ny * cfg->nz * 2 * es;
  }
  // block partials of the reductions (divergence, CFL, delta: at most four
  // doubles per 256 cells) and the allocation granularity of the ~60
  // buffers of a handle (2 MiB each at most)
  b += (int64_t)ntot / 256 * 4 * 8 + (int64_t)60 * (2 << 20);
  *bytes = b;
  return 0;
}



int elmm_case_footprint(const elmm_config *cfg, const elmm_case *c, int64_t *bytes) {
  if (!cfg || !c || !bytes) return elmm::fail(ELMM_E_ARG, "null argument to elmm_case_footprint");
  int rc = elmm_footprint(cfg, bytes);
  if (rc) return rc;
  *bytes += elmm::case_list_bytes(c);
  return 0;
}

int elmm_server_create(int64_t hbm_budget, int32_t max_active, int32_t chunk_steps, elmm_server **out) {
  if (!out || hbm_budget <= 0 || max_active <= 0 || chunk_steps <= 0)
    return elmm::fail(ELMM_E_ARG, "elmm_server_create: budget, slots and chunk must be positive");
  auto s = std::make_unique<elmm_server>();
  s->budget = (size_t)hbm_budget;
  s->max_active = max_active;
  s->chunk = chunk_steps;
  s->device = elmm::current_device();
  *out = s.release();
  return 0;
}

int elmm_server_submit(elmm_server *s, const elmm_request *q, int64_t *id) {
  if (!s || !q || !q->cfg || !q->c || !q->U || !q->V || !q->W || !q->Pr || q->nsteps < 0)
    return elmm::fail(ELMM_E_ARG, "elmm_server_submit: incomplete request");
  auto r = std::make_unique<Request>();
  r->cfg = *q->cfg;
  if (r->cfg.stretched_z && r->cfg.zw) {
    r->zw.assign(r->cfg.zw, r->cfg.zw + r->cfg.nz + 1);
    r->cfg.zw = r->zw.data();
  }
  std::string err = elmm::validate_config(r->cfg);
  if (!err.empty()) return elmm::fail(ELMM_E_ARG, err);
  r->c = q->c;
  r->nsteps = q->nsteps;
  const size_t n = (size_t)(r->cfg.nx + 6) * (r->cfg.ny + 6) * (r->cfg.nz + 6);
  r->U.assign(q->U, q->U + n);
  r->V.assign(q->V, q->V + n);
  r->W.assign(q->W, q->W + n);
  r->Pr.assign(q->Pr, q->Pr + n);
  for (int k = 0; k < r->cfg.nscalars; k++) {
    if (q->scalars && q->scalars[k])
      r->S.emplace_back(q->scalars[k], q->scalars[k] + n);
    else
      r->S.emplace_back(n, 0.0);
  }
  r->oU = q->out_U;
  r->oV = q->out_V;
  r->oW = q->out_W;
  r->oPr = q->out_Pr;
  for (int k = 0; k < r->cfg.nscalars; k++) r->oS.push_back(q->out_scalars ? q->out_scalars[k] : nullptr);
  int64_t fb = 0;
  elmm_case_footprint(&r->cfg, r->c, &fb);
  r->footprint = (size_t)fb;
  if (r->footprint > s->budget) return elmm::fail(ELMM_E_NOMEM, "request larger than the server's HBM budget");
  r->submitted = Clock::now();
  r->id = s->next_id++;
  if (id) *id = r->id;
  s->queue.push_back(std::move(r));
  return 0;
}

int elmm_server_progress(elmm_server *s, int32_t *n_active, int32_t *n_queued) {
  if (!s) return elmm::fail(ELMM_E_ARG, "null server");
  // a failed request's message becomes this thread's last error
  auto finish = [&](std::unique_ptr<Request> r) {
    if (r->status) (void)elmm::fail(r->status, "request " + std::to_string(r->id) + ": " + r->error);
    s->used -= r->footprint;
    s->done.emplace(r->id, std::move(r));
  };
  // admissions that have ended
  for (size_t a = 0; a < s->preparing.size();) {
    if (!task_ready(*s->preparing[a])) {
      a++;
      continue;
    }
    std::unique_ptr<Request> r = std::move(s->preparing[a]);
    s->preparing.erase(s->preparing.begin() + a);
    r->task.get();
    if (r->finished)
      finish(std::move(r));
    else
      s->active.push_back(std::move(r));
  }
  // advance the active cases or hand them to a retirement worker
  for (size_t a = 0; a < s->active.size();) {
    Request &r = *s->active[a];
    int32_t idle = 0;
    int rc = elmm_query(r.h, &idle);
    if (rc) {
      fail_req(r, rc);
    } else if (idle) {
      if (r.done >= r.nsteps) {
        if ((int)s->retiring.size() >= kWorkers) {
          a++;
          continue;
        }
        Request *rp = &r;
        const int dev = s->device;
        r.task = std::async(std::launch::async, [rp, dev] {
          elmm::use_device(dev);
          return retire(*rp);
        });
        s->retiring.push_back(std::move(s->active[a]));
        s->active.erase(s->active.begin() + a);
        continue;
      }
      const int32_t k = std::min(s->chunk, r.nsteps - r.done);
      rc = elmm_step_async(r.h, k);
      if (rc)
        fail_req(r, rc);
      else
        r.done += k;
    }
    if (r.finished) {
      finish(std::move(s->active[a]));
      s->active.erase(s->active.begin() + a);
    } else {
      a++;
    }
  }
  // retirements that have ended
  for (size_t a = 0; a < s->retiring.size();) {
    if (!task_ready(*s->retiring[a])) {
      a++;
      continue;
    }
    std::unique_ptr<Request> r = std::move(s->retiring[a]);
    s->retiring.erase(s->retiring.begin() + a);
    r->task.get();
    s->cell_steps += (int64_t)r->cfg.nx * r->cfg.ny * r->cfg.nz * r->nsteps;
    finish(std::move(r));
  }
  // admit in arrival order while memory, slots and workers allow
  while (!s->queue.empty() && (int)(s->preparing.size() + s->active.size() + s->retiring.size()) < s->max_active &&
         (int)s->preparing.size() < kWorkers && s->used + s->queue.front()->footprint <= s->budget) {
    std::unique_ptr<Request> r = std::move(s->queue.front());
    s->queue.pop_front();
    s->used += r->footprint;
    Request *rp = r.get();
    const int chunk = s->chunk, dev = s->device;
    r->task = std::async(std::launch::async, [rp, chunk, dev] {
      elmm::use_device(dev);
      return admit(chunk, *rp);
    });
    s->preparing.push_back(std::move(r));
  }
  if (n_active) *n_active = (int32_t)(s->preparing.size() + s->active.size() + s->retiring.size());
  if (n_queued) *n_queued = (int32_t)s->queue.size();
  return 0;
}

int elmm_server_status(elmm_server *s, int64_t id, elmm_request_status *st) {
  if (!s || !st) return elmm::fail(ELMM_E_ARG, "null argument to elmm_server_status");
  std::memset(st, 0, sizeof *st);
  auto it = s->done.find(id);
  if (it == s->done.end()) {
    st->state = 0;
    for (auto *list : {&s->preparing, &s->active, &s->retiring})
      for (auto &r : *list)
        if (r->id == id) {
          st->state = 1;
          st->steps_done = r->done;
        }
    return 0;
  }
  const Request &r = *it->second;
  st->state = r.status ? -1 : 2;
  st->status = r.status;
  st->steps_done = r.nsteps;
  st->latency_s = r.latency;
  st->max_div = r.max_div;
  return 0;
}

int elmm_server_stats(elmm_server *s, int64_t *cell_steps, int64_t *hbm_in_use) {
  if (!s) return elmm::fail(ELMM_E_ARG, "null server");
  if (cell_steps) *cell_steps = s->cell_steps;
  if (hbm_in_use) *hbm_in_use = (int64_t)s->used;
  return 0;
}

int elmm_server_destroy(elmm_server *s) {
  if (!s) return 0;
  // workers first (a retirement releases its own handle)
  for (auto &r : s->preparing) r->task.wait();
  for (auto &r : s->retiring) r->task.wait();
  for (auto *list : {&s->preparing, &s->active})
    for (auto &r : *list)
      if (r->h) {
        elmm_wait(r->h, nullptr);
        elmm_destroy(r->h);
      }
  delete s;
  return 0;
}

}  // extern "C"
