// capi_common.hpp -- internal to libubg.so: error plumbing, the device
// memory pool, parameter checks and the MG handle (V-cycle driver) shared by
// the C ABI translation units.
#pragma once

#include "ubg_internal.hpp"
#include "../../include/ubg_c.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace ubg {


inline thread_local std::string g_err; // ubg_last_error()

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define UBG_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? UBG_E_OOM : UBG_E_HIP,                   \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                  \
  } while (0)

#define UBG_TRY(expr)                                                                  \
  do {                                                                                 \
    int rc_ = (expr);                                                                  \
    if (rc_ != UBG_OK) return rc_;                                                     \
  } while (0)

inline int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(UBG_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return UBG_OK;
}

inline size_t align_up(size_t n, size_t a = 256) { return (n + a - 1) / a * a; }

// levels with at least this many cells run the row-streamed level kernels
// (below, a level has too few strip x row-block units to fill the chip and
// the LDS tiles are faster)
constexpr size_t kFuseMinCells = (size_t)1 << 22;
// the single-workgroup LDS V-cycle starts at the first level with at most
// this many cells (one CU's lanes do all of its work, so larger levels run
// faster as tiles spread over the chip; measured: starting at 8192 cells
// slows the 4096x2048 step by ~2%)
constexpr size_t kCoarseMaxCells = 4096;

// A simple bump allocator over one hipMalloc.
struct Pool {
  char *base = nullptr;
  size_t size = 0, off = 0;
  int init(size_t bytes) {
    size = bytes;
    UBG_HIP(hipMalloc((void **)&base, bytes));
    UBG_HIP(hipMemset(base, 0, bytes));
    // hipMemset runs on the null stream, which the handles' non-blocking
    // streams do not wait for: finish it before any of them writes the pool
    UBG_HIP(hipStreamSynchronize(nullptr));
    return UBG_OK;
  }
  template <typename T> T *take(size_t n) {
    T *p = (T *)(base + off);
    off += align_up(n * sizeof(T));
    return p;
  }
  void release() {
    if (base) (void)hipFree(base);
    base = nullptr;
  }
};

inline void level_sizes(int W, int H, std::vector<int> &ws, std::vector<int> &hs) {
  ws.clear();
  hs.clear();
  int cw = W, ch = H;
  while (cw > 3 && ch > 3) { // pressure_solver.hpp:20-30
    ws.push_back(cw);
    hs.push_back(ch);
    cw /= 2;
    ch /= 2;
  }
}

inline int validate_flag(const float *flag, size_t n, std::vector<uint8_t> &out) {
  out.resize(n);
  for (size_t i = 0; i < n; i++) {
    float v = flag[i];
    if (v == 0.0f) out[i] = 0;
    else if (v == 1.0f) out[i] = 1;
    else
      return fail(UBG_E_FLAG_VALUE, "flag value at index " + std::to_string(i) + " is " +
                                        std::to_string(v) + " (flags must be 0 or 1)");
  }
  return UBG_OK;
}

struct Timer {
  bool on = false;
  std::vector<hipEvent_t> ev; // pairs
  size_t used = 0;
  double total_ms = 0.0;
  long launches = 0;
};

// run `fn` (one kernel launch through launch_k) with the timer's next event
// pair attached to its dispatch, when the timer is given and enabled
template <typename Fn> void with_timer(Timer *t, Fn fn) {
  const bool on = t && t->on;
  if (on) {
    if (t->used + 2 > t->ev.size()) {
      size_t n = t->ev.size();
      t->ev.resize(n + 512);
      for (size_t i = n; i < t->ev.size(); i++) (void)hipEventCreate(&t->ev[i]);
    }
    g_launch_ev = LaunchEvents{t->ev[t->used], t->ev[t->used + 1]};
  }
  fn();
  if (on) {
    if (!g_launch_ev.start) t->used += 2; // the launch took the pair
    g_launch_ev = LaunchEvents{};
  }
}

inline void default_params(ubg_params *p) {
  p->nu_pre = 3;
  p->nu_post = 3;
  p->nu_coarse = 5;
  p->vcycles_per_step = 2;
  p->alpha = 1.0f;
  p->use_graph = 1;
  p->save_current = 1;
}

inline int check_params(const ubg_params *p) {
  if (!p) return fail(UBG_E_INVALID, "null params");
  if (p->nu_pre < 0 || p->nu_post < 0 || p->nu_coarse < 0 || p->vcycles_per_step < 0 ||
      p->nu_pre > 64 || p->nu_post > 64 || p->nu_coarse > 256 || p->vcycles_per_step > 1024)
    return fail(UBG_E_INVALID, "smoothing counts out of range");
  if (!(p->alpha > 0.0f && p->alpha < 2.0f)) return fail(UBG_E_INVALID, "alpha must be in (0,2)");
  return UBG_OK;
}


} // namespace ubg

using namespace ubg;

struct IRState; // FP64 iterative refinement buffers (refine.hip)

// ===========================================================================
// MG
// ===========================================================================
namespace ubg {
// An event another stream recorded after work on a handle's buffers (a
// batched step of several simulations).  Shared by the batch that records
// it and the handles that must wait for it; destroyed with its last holder.
struct ExtEvent {
  hipEvent_t ev = nullptr;
  ~ExtEvent() {
    if (ev) (void)hipEventDestroy(ev);
  }
};
// Before a handle enqueues work on `stream`: wait for the pending external
// event, if any, and note that the handle's stream has work of its own.
inline void enter_stream(hipStream_t stream, std::shared_ptr<ExtEvent> &ext, bool &own_work) {
  if (ext) {
    (void)hipStreamWaitEvent(stream, ext->ev, 0);
    ext.reset();
  }
  own_work = true;
}
} // namespace ubg

struct ubg_mg {
  int W = 0, H = 0, levels = 0;
  // a simulation's MG: its owner's external event and work flag (null for a
  // stand-alone MG)
  std::shared_ptr<ubg::ExtEvent> *owner_ext = nullptr;
  bool *owner_work = nullptr;
  void enter() {
    if (owner_ext) ubg::enter_stream(stream, *owner_ext, *owner_work);
  }
  std::vector<Level> lv;
  Pool pool;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  ubg_params prm{};
  double *d_acc = nullptr;
  IRState *ir = nullptr;   // allocated by the first refinement solve
  unsigned flag_gen = 0;   // bumped whenever the level-0 flag may have changed
  Timer *timers = nullptr; // borrowed from the owning simulation (UBG_TIMER_COUNT)
  std::recursive_mutex mu;

  // run `fn` (one kernel launch) with timer `id`'s next event pair attached
  // to its dispatch when the timer is enabled and the work is on level 0
  template <typename Fn> void timed(int id, int l, Fn fn) {
    Timer *t = timers ? &timers[id] : nullptr;
    with_timer(t && l == 0 ? t : nullptr, fn);
  }

  // n sweeps at level l, one launch each
  void sweeps(int l, int n, float h, bool zg) {
    Level &L = lv[l];
    for (int i = 0; i < n; i++)
      timed(UBG_TIMER_SMOOTH_L0, l, [&] { launch_rbgs_fused(L, h, prm.alpha, zg, stream); });
  }

  // Whether solve_level(l) (l > 0: zero initial guess, no ZG) reads its p
  // before writing it.  The LDS coarse V-cycle and the tile pre pass do not,
  // so the restriction above may leave that p as it is instead of zeroing it
  // (the branches of solve_level decide alike).
  bool reads_zero_p(int l) {
    const size_t n = (size_t)lv[l].w * lv[l].h;
    if (l == levels - 2) return true; // sweeps on the coarsest solved level
    if (n <= kCoarseMaxCells && coarse_fits_lds(lv.data(), l, levels)) return false;
    if (n < kFuseMinCells && tile_pass_ok(prm.nu_pre) && tile_pass_ok(prm.nu_post)) return false;
    return true;
  }

  // MG::solveLevel (pressure_solver.cpp:201-248).  A level costs two fused
  // passes: (pre-smoothing, residual, restriction) and (prolongation,
  // correction, post-smoothing); setZeroGradientBC happens inside them.
  void solve_level(int l, float h, bool zg) {
    const bool z = zg && l == 0;
    if (l == levels - 2) {
      sweeps(l, prm.nu_coarse, h, false); // no ZG on the coarsest solved level
      return;
    }
    if (l > 0 && (size_t)lv[l].w * lv[l].h <= kCoarseMaxCells &&
        coarse_fits_lds(lv.data(), l, levels)) {
      launch_coarse_vcycle(lv.data(), l, levels, h, prm.nu_pre, prm.nu_post, prm.nu_coarse,
                           prm.alpha, stream);
      return;
    }
    Level &L = lv[l], &C = lv[l + 1];
    if ((size_t)L.w * L.h < kFuseMinCells) {
      float hc = h * (L.w - 1.0f) / (C.w - 1.0f); // pressure_solver.cpp:229
      if (!z && tile_pass_ok(prm.nu_pre) && tile_pass_ok(prm.nu_post)) {
        // small level: two LDS-tiled launches instead of 2 + nu_pre + nu_post
        launch_tile_pre(L, C, prm.nu_pre, h, prm.alpha, l > 0, reads_zero_p(l + 1), stream);
        solve_level(l + 1, hc, false);
        launch_tile_post(L, C, prm.nu_post, h, prm.alpha, stream);
        return;
      }
      sweeps(l, prm.nu_pre, h, z);
      launch_residual_restrict(L, C, h, stream);
      solve_level(l + 1, hc, false);
      launch_prolong_correct(L, C, stream);
      if (z) launch_zero_gradient(L.p, L.w, L.h, stream);
      sweeps(l, prm.nu_post, h, z);
      return;
    }
    // the last pre-smoothing sweep and the first post-smoothing sweep ride in
    // the grid-transfer passes
    const int pre_fused = std::min(prm.nu_pre, 1);
    sweeps(l, prm.nu_pre - pre_fused, h, z);
    timed(UBG_TIMER_L0_PRE, l, [&] {
      launch_presmooth_restrict(L, C, pre_fused, h, prm.alpha, z, reads_zero_p(l + 1), stream);
    });
    float hc = h * (L.w - 1.0f) / (C.w - 1.0f); // pressure_solver.cpp:229
    solve_level(l + 1, hc, false);
    const int post_fused = std::min(prm.nu_post, 1);
    timed(UBG_TIMER_L0_POST, l, [&] { launch_prolong_postsmooth(L, C, post_fused, h, prm.alpha, z, stream); });
    sweeps(l, prm.nu_post - post_fused, h, z);
  }

  // coarse flags of levels 1..L-1 from the level-0 flag (updateFields)
  void classify() {
    flag_gen++;
    launch_build_mask(lv[0].flag, lv[0].mask, lv[0].w, lv[0].h, stream);
    for (int l = 1; l < levels; l++) {
      launch_coarsen(lv[l - 1].flag, lv[l - 1].w, lv[l].flag, lv[l].w, lv[l].h, stream);
      launch_build_mask(lv[l].flag, lv[l].mask, lv[l].w, lv[l].h, stream);
    }
  }
};

namespace ubg {

// ---- a simulation's step around its pressure solve (batched steps) ------
struct SimLevel0 {
  int W = 0, H = 0;
  float h = 0.0f;
  float *p = nullptr, *f = nullptr; // current pressure, divergence rhs (device)
  const uint8_t *flag = nullptr;    // 0/1 level-0 flag (device)
  hipStream_t stream = nullptr;
  ubg_params prm{};
  unsigned flag_gen = 0;            // moves whenever the flag may have changed
};
SimLevel0 sim_level0(ubg_sim *s);
// step() up to the V-cycles (eager, on the simulation's stream) and the
// stages after them (setPBC + correction + setVBCs + saveCurrent)
int sim_step_before_solve(ubg_sim *s, float dt);
int sim_step_after_solve(ubg_sim *s);
// the whole-batch form: the applyAccumulatedVelocity / vy-boundary /
// face-factor prelude on the simulation's stream, then the records of the
// four per-cell stages of its step (two diffuse iterations, advect, setVBCs
// + divergence and correction tail); the simulation's buffer roles advance
// as its step would leave them
int sim_batched_records(ubg_sim *s, float dt, SimRec rec[4], bool *has_sinks);
int sim_apply_sinks_on(ubg_sim *s, hipStream_t st);
// whether the simulation's stream has had work since the last batched step
// (clears the mark)
bool sim_take_own_work(ubg_sim *s);
// the event the simulation's next own work waits for
void sim_set_ext_event(ubg_sim *s, const std::shared_ptr<ExtEvent> &ev);

// Build an MG hierarchy.  If level-0 buffers are given they are borrowed
// (the simulation's p, f, flag, mask); otherwise the MG owns them.
int mg_create(int W, int H, hipStream_t stream, float *p0, float *p0_alt, float *f0,
              uint8_t *flag0, uint8_t *mask0, ubg_mg **out);
int mg_upload_flag(ubg_mg *m, const float *flag);
// releases the refinement buffers of an MG handle (refine.hip)
void ir_release(ubg_mg *m);

} // namespace ubg
