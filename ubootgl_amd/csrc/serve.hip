// serve.hip -- served refinement solves (ubg_server_*): a request loop with
// continuous batching on top of the batch engine (batch.hip).
//
// One worker thread owns the engine.  Each turn it waits for the refinement
// step in flight, enqueues the next one (the members still iterating, plus
// the requests made ready during the last turn, which take their first
// update), and while that step runs on the device it retires the members
// that finished (the next timestep of a request starts warm from its
// solution, a request whose last timestep finished copies its solution out
// and releases its memory) and admits queued requests in arrival order while
// the device arena and the active limit allow.  Those transfers use a second
// stream and touch only domains outside the step in flight, so the host work
// overlaps the device work; a started timestep joins the step after next.
// No counterpart in the reference.
#include "batch_engine.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

namespace {

struct SrvGeom {
  ubg::BGeom *g = nullptr;
  int W = 0, H = 0;
  float h = 0.0f;
  std::vector<uint8_t> f8;
  size_t dom_bytes = 0;
};

struct SrvReq {
  uint64_t id = 0;
  int geom = 0, T = 1;
  std::vector<const float *> rhs;
  std::vector<double> fnorm;
  double *x_out = nullptr;
  int state = 0; // 0 queued, 1 running, 2 done, 3 failed
  int t = 0, iters = 0;
  bool conv = true;
  double rel = 0.0;
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
  ubg::BDomain *dom = nullptr;
};

} // namespace

struct ubg_server {
  ubg_server_params prm{};
  ubg::BatchEngine eng;
  ubg::Arena arena;
  hipStream_t stream = nullptr;
  std::vector<SrvGeom> geoms;
  std::mutex eng_mu; // engine + arena + geometries (taken before mu)
  std::mutex mu;     // queue, request table, stats
  std::condition_variable cv;
  std::deque<SrvReq *> queue;
  std::unordered_map<uint64_t, std::unique_ptr<SrvReq>> reqs;
  hipStream_t cstream = nullptr; // transfers of domains outside the step in flight
  // worker only: members of the step in flight (engine order: running, then
  // fresh), requests ready to join the next step, still-iterating requests
  std::vector<SrvReq *> members, ready, running;
  bool inflight = false;
  std::thread worker;
  bool stop = false;
  uint64_t next_id = 1;
  ubg_server_stats st{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void finish(ubg_server *S, SrvReq *r, int state) {
  if (r->dom) {
    S->eng.domain_release(S->arena, r->dom);
    r->dom = nullptr;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  r->state = state;
  r->t_done = S->now();
  if (state == 2) S->st.completed++;
  else S->st.failed++;
  S->st.bytes_in_use = S->arena.used();
  S->cv.notify_all();
}

// start timestep r->t of a request in no step in flight: rhs upload and a
// warm start from its solution, on the copy stream (domain_start syncs it)
int start_timestep(ubg_server *S, SrvReq *r) {
  const SrvGeom &G = S->geoms[r->geom];
  UBG_HIP(hipMemcpyAsync(r->dom->rhs, r->rhs[r->t], (size_t)G.W * G.H * 4, hipMemcpyHostToDevice,
                         S->cstream));
  return S->eng.domain_start(r->dom, S->prm.tol, S->prm.max_iter, r->fnorm[r->t], S->cstream);
}

// one worker turn; returns false when there is nothing left to do and stop is set
bool worker_turn(ubg_server *S) {
  if (!S->inflight && S->running.empty() && S->ready.empty()) { // idle: wait for a request
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->queue.empty() && S->stop) return false;
    if (S->queue.empty()) {
      S->cv.wait_for(lk, std::chrono::milliseconds(20));
      return true;
    }
  }
  std::lock_guard<std::mutex> elk(S->eng_mu);
  const double t_turn = S->now();
  // 1. the step in flight: wait for it, then split its members into those
  // still iterating and those whose timestep ended
  std::vector<SrvReq *> finished;
  if (S->inflight) {
    S->inflight = false;
    const int rc = S->eng.sync();
    std::vector<SrvReq *> members;
    members.swap(S->members);
    S->running.clear();
    if (rc != UBG_OK) {
      for (SrvReq *r : members) finish(S, r, 3);
      return true;
    }
    std::vector<ubg::BDomain *> act = S->eng.active_members();
    size_t ai = 0;
    for (SrvReq *r : members) {
      if (ai < act.size() && act[ai] == r->dom) { // still iterating
        ai++;
        S->running.push_back(r);
      } else {
        finished.push_back(r);
      }
    }
  }
  const double t_synced = S->now();
  // 2. the next step: the running members, and the ready ones with their
  // first update (one step with them as fresh, the rest of the turn with
  // everyone running; a domain whose first update already met the tolerance
  // is inactive and skipped on the device)
  std::vector<SrvReq *> fresh;
  fresh.swap(S->ready);
  if (!S->running.empty() || !fresh.empty()) {
    std::vector<ubg::BDomain *> rd, fd;
    for (SrvReq *r : S->running) rd.push_back(r->dom);
    for (SrvReq *r : fresh) fd.push_back(r->dom);
    int rc = S->eng.set_members(rd, fd);
    if (rc == UBG_OK) rc = S->eng.enqueue_steps(1);
    if (rc == UBG_OK && S->prm.steps_per_turn > 1) {
      std::vector<ubg::BDomain *> all = rd;
      all.insert(all.end(), fd.begin(), fd.end());
      rc = S->eng.set_members(all, {});
      if (rc == UBG_OK) rc = S->eng.enqueue_steps(S->prm.steps_per_turn - 1);
    }
    std::vector<SrvReq *> members = S->running;
    members.insert(members.end(), fresh.begin(), fresh.end());
    S->running.clear();
    if (rc != UBG_OK) {
      (void)S->eng.sync();
      for (SrvReq *r : members) finish(S, r, 3);
    } else {
      S->members = members;
      S->inflight = true;
      std::lock_guard<std::mutex> lk(S->mu);
      S->st.steps++;
    }
  }
  const double t_launched = S->now();
  // 3. while the step runs: retire the finished members of the last one ...
  for (SrvReq *r : finished) {
    ubg::DState ds;
    if (S->eng.domain_read_state(r->dom, &ds, S->cstream) != UBG_OK || ds.nonfinite) {
      finish(S, r, 3);
      continue;
    }
    r->dom->xcur = ds.xcur;
    r->iters += ds.iters;
    r->rel = ds.rel;
    r->conv = r->conv && ds.converged;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->st.solves++;
    }
    if (r->t + 1 < r->T) {
      r->t++;
      if (start_timestep(S, r) != UBG_OK) {
        finish(S, r, 3);
        continue;
      }
      S->ready.push_back(r);
      continue;
    }
    if (r->x_out && S->eng.domain_get_solution(r->dom, r->x_out, S->cstream) != UBG_OK) {
      finish(S, r, 3);
      continue;
    }
    finish(S, r, 2);
  }
  const double t_retired = S->now();
  // ... and admit queued requests, in arrival order
  std::vector<SrvReq *> admitted;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    while (!S->queue.empty() &&
           (int)(S->members.size() + S->ready.size() + admitted.size()) < S->prm.max_active) {
      SrvReq *r = S->queue.front();
      const SrvGeom &G = S->geoms[r->geom];
      ubg::BDomain *d = nullptr;
      int rc = S->eng.domain_create(S->arena, G.g, G.h, &d, S->cstream);
      if (rc != UBG_OK) {
        if (rc == UBG_E_BUDGET && !(S->members.empty() && S->ready.empty() && admitted.empty()))
          break; // wait for memory to free up
        S->queue.pop_front();
        r->state = 3;
        r->t_done = S->now();
        S->st.failed++;
        S->cv.notify_all();
        continue;
      }
      S->queue.pop_front();
      r->dom = d;
      r->state = 1;
      r->t_admit = S->now();
      admitted.push_back(r);
    }
    S->st.queued = (int)S->queue.size();
    S->st.bytes_in_use = S->arena.used();
    S->st.peak_bytes = std::max(S->st.peak_bytes, S->arena.used());
  }
  for (SrvReq *r : admitted) {
    if (start_timestep(S, r) != UBG_OK) {
      finish(S, r, 3);
      continue;
    }
    S->ready.push_back(r);
  }
  std::lock_guard<std::mutex> lk(S->mu);
  const double t_end = S->now();
  S->st.busy_s += t_end - t_turn;
  S->st.wait_s += t_synced - t_turn;
  S->st.setup_s += t_launched - t_synced;
  S->st.retire_s += t_retired - t_launched;
  S->st.admit_s += t_end - t_retired;
  S->st.active = (int)(S->members.size() + S->ready.size());
  S->st.peak_active = std::max(S->st.peak_active, S->st.active);
  return true;
}

void worker_main(ubg_server *S) {
  while (worker_turn(S)) {
  }
}

} // namespace

extern "C" {

UBG_API void ubg_server_params_default(ubg_server_params *out) {
  if (!out) return;
  out->hbm_budget = (size_t)200 * 1000 * 1000 * 1000;
  out->max_active = 1024;
  out->tol = 1e-6;
  out->vcycles_per_iter = 4;
  out->max_iter = 200;
  out->steps_per_turn = 1;
}

UBG_API int ubg_server_create(const ubg_server_params *prm, const ubg_params *mg,
                              ubg_server **out) {
  if (!out) return fail(UBG_E_INVALID, "null output pointer");
  int nd = 0;
  ubg_device_count(&nd);
  if (nd <= 0) return fail(UBG_E_NO_DEVICE, "no HIP device visible");
  ubg_server_params P;
  ubg_server_params_default(&P);
  if (prm) P = *prm;
  if (!(P.tol > 0.0 && P.tol < 1.0) || P.vcycles_per_iter < 1 || P.vcycles_per_iter > 256 ||
      P.max_iter < 1 || P.max_active < 1 || P.hbm_budget < ((size_t)1 << 20) ||
      P.steps_per_turn < 1 || P.steps_per_turn > 64)
    return fail(UBG_E_INVALID, "server parameters out of range");
  ubg_params M;
  default_params(&M);
  if (mg) {
    UBG_TRY(check_params(mg));
    M = *mg;
  }
  auto *S = new ubg_server();
  S->prm = P;
  hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    e = hipStreamCreateWithFlags(&S->cstream, hipStreamNonBlocking);
    if (e != hipSuccess) (void)hipStreamDestroy(S->stream);
  }
  if (e != hipSuccess) {
    delete S;
    return fail(UBG_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  int rc = S->arena.init(P.hbm_budget);
  if (rc == UBG_OK) rc = S->eng.init(S->stream, M, P.vcycles_per_iter);
  if (rc != UBG_OK) {
    S->arena.release();
    (void)hipStreamDestroy(S->stream);
    (void)hipStreamDestroy(S->cstream);
    delete S;
    return rc;
  }
  S->worker = std::thread(worker_main, S);
  *out = S;
  return UBG_OK;
}

UBG_API int ubg_server_destroy(ubg_server *S) {
  if (!S) return UBG_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    for (SrvReq *r : S->queue) {
      r->state = 3;
      r->t_done = S->now();
      S->st.failed++;
    }
    S->queue.clear();
    S->cv.notify_all();
  }
  if (S->worker.joinable()) S->worker.join();
  (void)hipStreamSynchronize(S->stream);
  (void)hipStreamSynchronize(S->cstream);
  for (SrvGeom &g : S->geoms) S->eng.geom_release(S->arena, g.g);
  S->geoms.clear();
  S->arena.release();
  hipStream_t s = S->stream, cs = S->cstream;
  delete S;
  (void)hipStreamDestroy(s);
  (void)hipStreamDestroy(cs);
  return UBG_OK;
}

UBG_API int ubg_server_add_geometry(ubg_server *S, int W, int H, const float *flag, float h,
                                    int *geom_id) {
  if (!S || !flag || !geom_id) return fail(UBG_E_INVALID, "null argument");
  if (!(h > 0.0f) || !std::isfinite(h)) return fail(UBG_E_INVALID, "h must be positive");
  ubg::BShape sh;
  UBG_TRY(ubg::batch_shape(W, H, &sh));
  std::lock_guard<std::mutex> elk(S->eng_mu);
  SrvGeom G;
  G.W = W;
  G.H = H;
  G.h = h;
  G.dom_bytes = sh.dom_bytes;
  UBG_TRY(validate_flag(flag, (size_t)W * H, G.f8));
  UBG_TRY(S->eng.geom_create(S->arena, W, H, flag, &G.g));
  std::lock_guard<std::mutex> lk(S->mu);
  S->geoms.push_back(std::move(G));
  *geom_id = (int)S->geoms.size() - 1;
  S->st.bytes_in_use = S->arena.used();
  S->st.peak_bytes = std::max(S->st.peak_bytes, S->arena.used());
  return UBG_OK;
}

UBG_API int ubg_server_submit(ubg_server *S, int geom_id, int timesteps, const float *const *rhs,
                              const double *fnorm, double *x_out, uint64_t *id) {
  if (!S || !rhs || !id) return fail(UBG_E_INVALID, "null argument");
  if (timesteps < 1 || timesteps > 1024) return fail(UBG_E_INVALID, "timesteps out of range");
  int W, H;
  size_t need;
  const uint8_t *f8;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(UBG_E_STATE, "server is stopping");
    if (geom_id < 0 || geom_id >= (int)S->geoms.size()) return fail(UBG_E_INVALID, "bad geometry id");
    const SrvGeom &G = S->geoms[geom_id];
    W = G.W;
    H = G.H;
    need = G.dom_bytes;
    f8 = G.f8.data();
  }
  if (need > S->arena.capacity()) return fail(UBG_E_BUDGET, "request exceeds the HBM budget");
  auto r = std::make_unique<SrvReq>();
  r->geom = geom_id;
  r->T = timesteps;
  r->x_out = x_out;
  for (int t = 0; t < timesteps; t++) {
    if (!rhs[t]) return fail(UBG_E_INVALID, "null rhs");
    r->rhs.push_back(rhs[t]);
    if (fnorm) {
      if (!(fnorm[t] >= 0.0) || !std::isfinite(fnorm[t])) return fail(UBG_E_INVALID, "bad fnorm");
      r->fnorm.push_back(fnorm[t]);
      continue;
    }
    // ||F f|| over interior cells, the residual of x = 0 with a zero ring
    double s = 0.0;
    for (int y = 1; y < H - 1; y++)
      for (int x = 1; x < W - 1; x++) {
        const size_t i = (size_t)y * W + x;
        if (f8[i]) s += (double)rhs[t][i] * (double)rhs[t][i];
      }
    if (!std::isfinite(s)) return fail(UBG_E_NONFINITE, "rhs has a non-finite value");
    r->fnorm.push_back(std::sqrt(s));
  }
  std::lock_guard<std::mutex> lk(S->mu);
  r->id = S->next_id++;
  r->t_submit = S->now();
  *id = r->id;
  S->queue.push_back(r.get());
  S->reqs.emplace(r->id, std::move(r));
  S->st.submitted++;
  S->st.queued = (int)S->queue.size();
  S->cv.notify_all();
  return UBG_OK;
}

UBG_API int ubg_server_wait(ubg_server *S, uint64_t id, double timeout_s, ubg_request_info *info) {
  if (!S) return fail(UBG_E_INVALID, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(UBG_E_INVALID, "unknown request id");
  SrvReq *r = it->second.get();
  auto done = [&] { return r->state >= 2; };
  if (timeout_s < 0) S->cv.wait(lk, done);
  else S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  if (info) {
    info->id = r->id;
    info->state = r->state;
    info->timesteps_done = r->state == 2 ? r->T : r->t;
    info->iterations = r->iters;
    info->converged = r->conv ? 1 : 0;
    info->rel_residual = r->rel;
    info->t_submit = r->t_submit;
    info->t_admit = r->t_admit;
    info->t_done = r->t_done;
  }
  if (!done()) return fail(UBG_E_STATE, "request still in flight");
  if (r->state == 3) return fail(UBG_E_HIP, "request failed");
  return UBG_OK;
}

UBG_API int ubg_server_drain(ubg_server *S) {
  if (!S) return fail(UBG_E_INVALID, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv.wait(lk, [&] { return S->st.completed + S->st.failed >= S->st.submitted; });
  return UBG_OK;
}

UBG_API int ubg_server_stats_get(ubg_server *S, ubg_server_stats *st) {
  if (!S || !st) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->st;
  return UBG_OK;
}

UBG_API double ubg_server_clock(ubg_server *S) { return S ? S->now() : 0.0; }

} // extern "C"
