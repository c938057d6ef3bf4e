// capi.hip -- the C ABI of libubg.so (include/ubg_c.h): handle lifetime,
// device memory, host copies, the MG V-cycle driver and Simulation::step.
//
// Host-side control flow follows the reference:
//   MG::MG / updateFields / solve / solveLevel   pressure_solver.hpp:16-62,
//                                                pressure_solver.cpp:201-248
//   Simulation ctor / step / diffuse / advect /  simulation.hpp:32-67,
//   project / setVBCs / setPBC / accumulate      simulation.cpp:16-208, 241-396
// Every kernel is issued on the handle's stream; a whole step can be captured
// once per (front-buffer parity, dt) and replayed as a hipGraph.
#include "capi_common.hpp"

#include <unordered_map>

namespace ubg {

int mg_create(int W, int H, hipStream_t stream, float *p0, float *p0_alt, float *f0,
              uint8_t *flag0, uint8_t *mask0, ubg_mg **out) {
  if (!out) return fail(UBG_E_INVALID, "null output pointer");
  if (W < 8 || H < 8)
    return fail(UBG_E_DOMAIN_TOO_SMALL, "MG needs at least 8 cells in each dimension");
  auto *m = new ubg_mg();
  m->W = W;
  m->H = H;
  default_params(&m->prm);
  std::vector<int> ws, hs;
  level_sizes(W, H, ws, hs);
  m->levels = (int)ws.size();
  size_t bytes = 256;
  for (int l = 0; l < m->levels; l++) {
    size_t n = (size_t)ws[l] * hs[l];
    bool own = (l > 0) || !p0;
    bytes += align_up(n * 4);              // r
    if (own) bytes += 3 * align_up(n * 4) + 2 * align_up(n); // p, p_alt, f, flag, mask
  }
  int rc = m->pool.init(bytes);
  if (rc) {
    delete m;
    return rc;
  }
  m->d_acc = m->pool.take<double>(4);
  m->lv.resize(m->levels);
  for (int l = 0; l < m->levels; l++) {
    Level &L = m->lv[l];
    L.w = ws[l];
    L.h = hs[l];
    size_t n = (size_t)L.w * L.h;
    L.r = m->pool.take<float>(n);
    if (l == 0 && p0) {
      L.p = p0;
      L.p_alt = p0_alt;
      L.f = f0;
      L.flag = flag0;
      L.mask = mask0;
    } else {
      L.p = m->pool.take<float>(n);
      L.p_alt = m->pool.take<float>(n);
      L.f = m->pool.take<float>(n);
      L.flag = m->pool.take<uint8_t>(n);
      L.mask = m->pool.take<uint8_t>(n);
    }
  }
  if (stream) {
    m->stream = stream;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      m->pool.release();
      delete m;
      return fail(UBG_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    m->own_stream = true;
  }
  // MG(w,h): every coarse flag starts as 1 (flagcs fill(1.0), hpp:26);
  // level 0 too until updateFields.
  for (int l = 0; l < m->levels; l++) {
    if (l == 0 && p0) continue;
    UBG_HIP(hipMemsetAsync(m->lv[l].flag, 1, (size_t)m->lv[l].w * m->lv[l].h, m->stream));
    launch_build_mask(m->lv[l].flag, m->lv[l].mask, m->lv[l].w, m->lv[l].h, m->stream);
  }
  rc = check_launch();
  if (rc) return rc;
  UBG_HIP(hipStreamSynchronize(m->stream));
  *out = m;
  return UBG_OK;
}

int mg_upload_flag(ubg_mg *m, const float *flag) {
  std::vector<uint8_t> f8;
  UBG_TRY(validate_flag(flag, (size_t)m->W * m->H, f8));
  UBG_HIP(hipMemcpyAsync(m->lv[0].flag, f8.data(), f8.size(), hipMemcpyHostToDevice, m->stream));
  m->classify();
  UBG_TRY(check_launch());
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

} // namespace ubg

extern "C" {

UBG_API int ubg_version(void) { return 100; }
UBG_API const char *ubg_last_error(void) { return g_err.c_str(); }
UBG_API void ubg_params_default(ubg_params *out) {
  if (out) default_params(out);
}
UBG_API int ubg_device_count(int *count) {
  if (!count) return fail(UBG_E_INVALID, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return UBG_OK;
}
UBG_API int ubg_mg_level_sizes(int W, int H, int *widths, int *heights, int cap) {
  std::vector<int> ws, hs;
  level_sizes(W, H, ws, hs);
  for (int i = 0; i < (int)ws.size() && i < cap; i++) {
    if (widths) widths[i] = ws[i];
    if (heights) heights[i] = hs[i];
  }
  return (int)ws.size();
}

UBG_API int ubg_mg_create(int width, int height, ubg_mg **out) {
  int n = 0;
  ubg_device_count(&n);
  if (n <= 0) return fail(UBG_E_NO_DEVICE, "no HIP device visible");
  return mg_create(width, height, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, out);
}

UBG_API int ubg_mg_destroy(ubg_mg *m) {
  if (!m) return UBG_OK;
  (void)hipStreamSynchronize(m->stream);
  ir_release(m);
  if (m->own_stream) (void)hipStreamDestroy(m->stream);
  m->pool.release();
  delete m;
  return UBG_OK;
}

UBG_API int ubg_mg_set_params(ubg_mg *m, const ubg_params *prm) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  UBG_TRY(check_params(prm));
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  m->prm = *prm;
  return UBG_OK;
}

UBG_API int ubg_mg_num_levels(const ubg_mg *m, int *levels) {
  if (!m || !levels) return fail(UBG_E_INVALID, "null argument");
  *levels = m->levels;
  return UBG_OK;
}

UBG_API int ubg_mg_update_fields(ubg_mg *m, const float *flag) {
  if (!m || !flag) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  return mg_upload_flag(m, flag);
}

static int level_ptr(ubg_mg *m, int level, int which, float **ptr) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  if (level < 0 || level >= m->levels) return fail(UBG_E_INVALID, "level out of range");
  Level &L = m->lv[level];
  if (which == UBG_LEVEL_P) *ptr = L.p;
  else if (which == UBG_LEVEL_F) *ptr = L.f;
  else if (which == UBG_LEVEL_R) *ptr = L.r;
  else return fail(UBG_E_INVALID, "unknown level field");
  return UBG_OK;
}

UBG_API int ubg_mg_set_level_field(ubg_mg *m, int level, int which, const float *src) {
  float *d = nullptr;
  UBG_TRY(level_ptr(m, level, which, &d));
  if (!src) return fail(UBG_E_INVALID, "null source");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  size_t n = (size_t)m->lv[level].w * m->lv[level].h;
  UBG_HIP(hipMemcpyAsync(d, src, n * 4, hipMemcpyHostToDevice, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

UBG_API int ubg_mg_get_level_field(ubg_mg *m, int level, int which, float *dst) {
  float *d = nullptr;
  UBG_TRY(level_ptr(m, level, which, &d));
  if (!dst) return fail(UBG_E_INVALID, "null destination");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  size_t n = (size_t)m->lv[level].w * m->lv[level].h;
  UBG_HIP(hipMemcpyAsync(dst, d, n * 4, hipMemcpyDeviceToHost, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

UBG_API int ubg_mg_get_level_flag(ubg_mg *m, int level, float *dst) {
  if (!m || !dst) return fail(UBG_E_INVALID, "null argument");
  if (level < 0 || level >= m->levels) return fail(UBG_E_INVALID, "level out of range");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  const Level &L = m->lv[level];
  size_t n = (size_t)L.w * L.h;
  std::vector<uint8_t> tmp(n);
  UBG_HIP(hipMemcpyAsync(tmp.data(), L.flag, n, hipMemcpyDeviceToHost, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  for (size_t i = 0; i < n; i++) dst[i] = (float)tmp[i];
  return UBG_OK;
}

UBG_API int ubg_mg_level_op(ubg_mg *m, int op, int level, float h, float alpha) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  if (level < 0 || level >= m->levels) return fail(UBG_E_INVALID, "level out of range");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  const Level &L = m->lv[level];
  switch (op) {
  case UBG_MGOP_RBGS:
    launch_rbgs_fused(m->lv[level], h, alpha, false, m->stream);
    break;
  case UBG_MGOP_ZERO_GRADIENT:
    launch_zero_gradient(L.p, L.w, L.h, m->stream);
    break;
  case UBG_MGOP_RESIDUAL:
    launch_fill_f32(L.r, (size_t)L.w * L.h, 0.0f, m->stream);
    launch_residual(L, h, m->stream);
    break;
  case UBG_MGOP_RESTRICT_RESIDUAL:
    if (level + 1 >= m->levels) return fail(UBG_E_INVALID, "no coarser level");
    launch_residual_restrict(L, m->lv[level + 1], h, m->stream);
    break;
  case UBG_MGOP_PROLONG_CORRECT:
    if (level + 1 >= m->levels) return fail(UBG_E_INVALID, "no coarser level");
    launch_prolong_correct(L, m->lv[level + 1], m->stream);
    break;
  default:
    return fail(UBG_E_INVALID, "unknown MG operation");
  }
  UBG_TRY(check_launch());
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

UBG_API int ubg_mg_solve(ubg_mg *m, float h, int zg, int ncycles) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  if (ncycles < 0) return fail(UBG_E_INVALID, "negative cycle count");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  for (int c = 0; c < ncycles; c++) m->solve_level(0, h, zg != 0);
  UBG_TRY(check_launch());
  return UBG_OK;
}

UBG_API int ubg_mg_solve_host(ubg_mg *m, float *p, const float *f, const float *flag, float h,
                              int zg) {
  if (!m || !p || !f) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  const Level &L = m->lv[0];
  size_t n = (size_t)L.w * L.h;
  if (flag) {
    std::vector<uint8_t> f8;
    UBG_TRY(validate_flag(flag, n, f8));
    UBG_HIP(hipMemcpyAsync(L.flag, f8.data(), n, hipMemcpyHostToDevice, m->stream));
    launch_build_mask(L.flag, L.mask, L.w, L.h, m->stream);
    UBG_HIP(hipStreamSynchronize(m->stream));
  }
  UBG_HIP(hipMemcpyAsync(L.p, p, n * 4, hipMemcpyHostToDevice, m->stream));
  UBG_HIP(hipMemcpyAsync(L.f, f, n * 4, hipMemcpyHostToDevice, m->stream));
  m->solve_level(0, h, zg != 0);
  UBG_TRY(check_launch());
  UBG_HIP(hipMemcpyAsync(p, L.p, n * 4, hipMemcpyDeviceToHost, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

UBG_API int ubg_mg_rbgs(ubg_mg *m, float h, float alpha, int nsweeps) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  float saved = m->prm.alpha;
  m->prm.alpha = alpha;
  m->sweeps(0, nsweeps, h, false);
  m->prm.alpha = saved;
  UBG_TRY(check_launch());
  return UBG_OK;
}

UBG_API int ubg_mg_gs(ubg_mg *m, float h, float alpha, int nsweeps) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  if (nsweeps < 0) return fail(UBG_E_INVALID, "negative sweep count");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  const Level &L = m->lv[0];
  for (int i = 0; i < nsweeps; i++) launch_gs(L.p, L.f, L.mask, L.w, L.h, h, alpha, m->stream);
  UBG_TRY(check_launch());
  return UBG_OK;
}

UBG_API int ubg_mg_residual(ubg_mg *m, float h, float *r_dst, double *norm) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  const Level &L = m->lv[0];
  size_t n = (size_t)L.w * L.h;
  launch_fill_f32(L.r, n, 0.0f, m->stream);
  launch_residual(L, h, m->stream);
  launch_fill_f64(m->d_acc, 1, 0.0, m->stream);
  launch_sumsq(L.r, n, m->d_acc, m->stream);
  UBG_TRY(check_launch());
  double acc = 0.0;
  UBG_HIP(hipMemcpyAsync(&acc, m->d_acc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (r_dst) UBG_HIP(hipMemcpyAsync(r_dst, L.r, n * 4, hipMemcpyDeviceToHost, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  if (norm) *norm = std::sqrt(acc);
  if (!std::isfinite(acc)) return fail(UBG_E_NONFINITE, "residual norm is not finite");
  return UBG_OK;
}

UBG_API int ubg_mg_sync(ubg_mg *m) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  {
    std::lock_guard<std::recursive_mutex> g(m->mu);
    m->enter();
  }
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

} // extern "C"

// ===========================================================================
// Simulation
// ===========================================================================
struct SinkH {
  float x, y, z;
};

struct ubg_sim {
  int W = 0, H = 0;
  float pwidth = 0, mu = 0, h = 0, dt = 0;
  BCs bc;
  float *vx[2] = {nullptr, nullptr}, *vy[2] = {nullptr, nullptr};
  int fx = 0, fy = 0;
  float *vx_acc = nullptr, *vy_acc = nullptr, *vx_cur = nullptr, *vy_cur = nullptr;
  float *p_buf[2] = {nullptr, nullptr}; // pressure ping-pong pair, owned; the
                                        // current one is mg->lv[0].p
  float *f = nullptr, *r = nullptr;
  uint8_t *flag = nullptr, *mask = nullptr;
  // per-cell face factors of the vx / vy stencils (flag products, see
  // launch_face_masks), rebuilt when the MG's flag generation moves
  uint16_t *fmask = nullptr;
  unsigned fmask_gen = ~0u;
  bool accum_dirty = false;
  // vy's boundary cells (both buffers) hold what setVBCs computes from vy's
  // front interior; anything else that writes velocities clears it
  bool vy_bc_fresh = false;
  ubg_mg *mg = nullptr;
  Pool pool;
  hipStream_t stream = nullptr;
  ubg_params prm{};
  std::vector<SinkH> sinks;
  int *d_sink_xy = nullptr;
  float *d_sink_val = nullptr;
  int sink_cap = 0;
  // device staging for the entry points that take host arrays; each of them
  // ends with a stream synchronize, so one buffer serves all (grows only)
  char *stage = nullptr;
  size_t stage_cap = 0;
  int staging(size_t bytes, void **out) {
    if (bytes > stage_cap) {
      UBG_HIP(hipStreamSynchronize(stream));
      if (stage) (void)hipFree(stage);
      stage = nullptr;
      stage_cap = 0;
      UBG_HIP(hipMalloc((void **)&stage, bytes));
      stage_cap = bytes;
    }
    *out = stage;
    return UBG_OK;
  }
  // tracers are stored sorted by 16x16-cell tile of their position (gathers
  // of one wave then touch few cache lines); tr_perm[slot] = tracer id.
  float *tr_xy = nullptr, *tr_age = nullptr;
  int *tr_perm = nullptr;
  std::vector<int> tr_perm_h;
  int ntr = 0, tr_cap = 0;
  uint32_t tr_seed = 1;
  Timer timers[UBG_TIMER_COUNT];
  // captured step graphs keyed by the buffer-role state they start from
  struct State {
    int fx = 0, fy = 0;
    std::vector<float *> lp; // p / p_alt of every MG level
    bool operator==(const State &o) const { return fx == o.fx && fy == o.fy && lp == o.lp; }
  };
  struct G {
    hipGraphExec_t exec = nullptr;
    float dt = 0;
    State before, after;
  };
  std::vector<G> graphs;
  std::recursive_mutex mu_;
  // batched steps (ubg_sims_step) run this simulation's passes on a batch's
  // stream: the event the batch recorded after them (waited for before this
  // handle's next own work), and whether this handle's stream has had work
  // of its own since the last batched step
  std::shared_ptr<ExtEvent> ext_ev;
  bool own_work = true;
  void enter() { enter_stream(stream, ext_ev, own_work); }

  float *P() const { return mg->lv[0].p; }

  State state() const {
    State st;
    st.fx = fx;
    st.fy = fy;
    for (auto &L : mg->lv) {
      st.lp.push_back(L.p);
      st.lp.push_back(L.p_alt);
    }
    return st;
  }
  void set_state(const State &st) {
    fx = st.fx;
    fy = st.fy;
    for (size_t l = 0; l < mg->lv.size(); l++) {
      mg->lv[l].p = st.lp[2 * l];
      mg->lv[l].p_alt = st.lp[2 * l + 1];
    }
  }

  VelView vel() const { return VelView{W, H, vx[fx], vx[1 - fx], vy[fy], vy[1 - fy]}; }

  void set_vbcs() {
    launch_set_vbcs(vel(), bc, stream);
    vy_bc_fresh = true;
  }

  // face factors current for the flag field (outside any captured graph)
  void ensure_fmask() {
    if (fmask_gen == mg->flag_gen) return;
    launch_face_masks(flag, fmask, W, H, stream);
    fmask_gen = mg->flag_gen;
  }

  // The reference's setVBCs after the first vx iteration of diffuse leaves
  // vy's boundary current for vy's interior before vy is diffused; here the
  // vx and vy iterations are paired, so that refresh (vy only: vx must keep
  // its old boundary for its first iteration) happens before, when needed.
  void refresh_vy_bcs() {
    if (!vy_bc_fresh) launch_set_vbcs_vy(vel(), bc, stream);
    vy_bc_fresh = true;
  }

  // simulation.cpp:104-162: iteration k of the vx loop and of the vy loop
  // in one pass, each with the setVBCs that follows it.  Needs vy_bc_fresh.
  void diffuse_fused() {
    const float a = dt * mu * (W - 1.0f) / pwidth;
    for (int it = 1; it < 3; it++) {
      launch_diffuse_both(vel(), fmask, a, bc, stream);
      fx = 1 - fx;
      fy = 1 - fy;
    }
    vy_bc_fresh = true;
  }
  void diffuse() {
    ensure_fmask();
    refresh_vy_bcs();
    diffuse_fused();
  }

  // simulation.cpp:241-354
  void advect() {
    const float ih = 1.0f / h;
    with_timer(&timers[UBG_TIMER_ADVECT],
               [&] { launch_advect(vel(), flag, 0.5f * dt * ih, dt * ih, stream); });
    fx = 1 - fx;
    fy = 1 - fy;
    vy_bc_fresh = false;
  }

  // sinks (simulation.cpp:173-187), stamped in list order: where the 3x3
  // patches of several sinks overlap the last one wins, so each cell's final
  // value is settled here and the device writes every cell once.  A patch
  // cell is (int)(g + d) of the float centre g, as the reference indexes it.
  int apply_sinks() { return apply_sinks_on(stream); }
  int apply_sinks_on(hipStream_t stream) {
    if (sinks.empty()) return UBG_OK;
    std::vector<int> cells;
    std::vector<float> val;
    std::unordered_map<long long, size_t> slot;
    for (auto &s : sinks) {
      float gx = s.x / h + 0.5f, gy = s.y / h + 0.5f;
      if (gx <= 3 || gx > W - 3 || gy <= 3 || gy > H - 3) continue;
      for (int yy = -1; yy <= 1; yy++)
        for (int xx = -1; xx <= 1; xx++) {
          const long long c = (long long)(int)(gy + yy) * W + (int)(gx + xx);
          auto it = slot.find(c);
          if (it == slot.end()) {
            slot.emplace(c, val.size());
            cells.push_back((int)c);
            val.push_back(s.z);
          } else {
            val[it->second] = s.z;
          }
        }
      s.z = (float)(s.z * std::pow(0.000001, (double)(dt * 50)));
    }
    sinks.erase(std::remove_if(sinks.begin(), sinks.end(),
                               [](const SinkH &t) { return t.z < 0.05f; }),
                sinks.end());
    int n = (int)val.size();
    if (n == 0) return UBG_OK;
    if (n > sink_cap) {
      if (d_sink_xy) (void)hipFree(d_sink_xy);
      if (d_sink_val) (void)hipFree(d_sink_val);
      sink_cap = std::max(n, 576);
      UBG_HIP(hipMalloc((void **)&d_sink_xy, sizeof(int) * sink_cap));
      UBG_HIP(hipMalloc((void **)&d_sink_val, sizeof(float) * sink_cap));
    }
    UBG_HIP(hipMemcpyAsync(d_sink_xy, cells.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
    UBG_HIP(hipMemcpyAsync(d_sink_val, val.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream));
    launch_stamp_sinks(f, d_sink_xy, d_sink_val, n, stream);
    // the host arrays must outlive the async copies
    UBG_HIP(hipStreamSynchronize(stream));
    return UBG_OK;
  }

  void divergence() { launch_divergence(vel(), f, 1.0f / h, stream); }

  void correct_velocity() {
    launch_correct_velocity(vel(), P(), flag, 1.0f / h, stream);
    vy_bc_fresh = false;
  }

  void mg_solve() { mg->solve_level(0, h, true); }

  // simulation.cpp:164-208
  int project() {
    divergence();
    UBG_TRY(apply_sinks());
    for (int c = 0; c < prm.vcycles_per_step; c++) mg_solve();
    launch_set_pbc(P(), W, H, bc, stream);
    correct_velocity();
    return UBG_OK;
  }

  void save_current() {
    (void)hipMemcpyAsync(vx_cur, vx[fx], sizeof(float) * (W - 1) * H, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(vy_cur, vy[fy], sizeof(float) * W * (H - 1), hipMemcpyDeviceToDevice, stream);
  }

  // The part of step() after applyAccumulatedVelocity (simulation.cpp:
  // 356-374) in fused passes: diffuse (2 passes), advect, setVBCs +
  // divergence, [sinks,] the V-cycles, setPBC + correction + setVBCs +
  // saveCurrentVelocityFields.  Needs vy_bc_fresh.
  int step_body(bool with_sinks) {
    UBG_TRY(body_before_solve(with_sinks));
    for (int c = 0; c < prm.vcycles_per_step; c++) mg_solve();
    body_after_solve();
    return UBG_OK;
  }
  // the step's stages before and after the pressure solve's V-cycles
  int body_before_solve(bool with_sinks) {
    diffuse_fused();
    advect();
    launch_vbcs_divergence(vel(), bc, f, 1.0f / h, stream);
    if (with_sinks) UBG_TRY(apply_sinks());
    return UBG_OK;
  }
  void body_after_solve() {
    launch_correct_tail(vel(), P(), fmask, 1.0f / h, bc, prm.save_current ? vx_cur : nullptr,
                        prm.save_current ? vy_cur : nullptr, stream);
    vy_bc_fresh = true;
  }
  // step() up to the V-cycles, eager (no graph): applyAccumulatedVelocity,
  // diffuse, advect, setVBCs + divergence, sinks (simulation.cpp:356-369,
  // 164-187); the pressure solve is then the caller's
  int step_before_solve(float timestep) {
    dt = timestep;
    if (accum_dirty) {
      launch_accumulate(vel(), vx_acc, vy_acc, stream);
      accum_dirty = false;
      vy_bc_fresh = false;
    }
    refresh_vy_bcs();
    ensure_fmask();
    return body_before_solve(true);
  }

  // simulation.cpp:356-374
  int step(float timestep) {
    dt = timestep;
    if (accum_dirty) {
      launch_accumulate(vel(), vx_acc, vy_acc, stream);
      accum_dirty = false;
      vy_bc_fresh = false;
    }
    refresh_vy_bcs(); // outside any captured graph: it depends on the history
    ensure_fmask();
    bool any_timer = false;
    for (auto &t : timers) any_timer = any_timer || t.on;
    bool graphable = prm.use_graph && sinks.empty() && !any_timer;
    if (graphable) {
      State st = state();
      G *g = nullptr;
      for (auto &cand : graphs)
        if (cand.dt == dt && cand.before == st) g = &cand;
      if (!g) {
        if (graphs.size() >= 8) invalidate_graphs();
        G ng;
        ng.dt = dt;
        ng.before = st;
        hipGraph_t graph;
        UBG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        (void)step_body(false);
        hipError_t e = hipStreamEndCapture(stream, &graph);
        ng.after = state();
        set_state(st); // capture executed nothing: restore the buffer roles
        if (e != hipSuccess) return fail(UBG_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(UBG_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        graphs.push_back(ng);
        g = &graphs.back();
      }
      UBG_HIP(hipGraphLaunch(g->exec, stream));
      set_state(g->after);
      vy_bc_fresh = true;
    } else {
      UBG_TRY(step_body(true));
    }
    return check_launch();
  }

  void invalidate_graphs() {
    for (auto &g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  float *field(int id) {
    switch (id) {
    case UBG_FIELD_VX: return vx[fx];
    case UBG_FIELD_VX_BACK: return vx[1 - fx];
    case UBG_FIELD_VY: return vy[fy];
    case UBG_FIELD_VY_BACK: return vy[1 - fy];
    case UBG_FIELD_VX_ACCUM: return vx_acc;
    case UBG_FIELD_VY_ACCUM: return vy_acc;
    case UBG_FIELD_VX_CURRENT: return vx_cur;
    case UBG_FIELD_VY_CURRENT: return vy_cur;
    case UBG_FIELD_P: return P();
    case UBG_FIELD_F: return f;
    case UBG_FIELD_R: return r;
    }
    return nullptr;
  }
  void shape(int id, int *w, int *hh) const {
    switch (id) {
    case UBG_FIELD_VX: case UBG_FIELD_VX_BACK: case UBG_FIELD_VX_ACCUM: case UBG_FIELD_VX_CURRENT:
      *w = W - 1; *hh = H; return;
    case UBG_FIELD_VY: case UBG_FIELD_VY_BACK: case UBG_FIELD_VY_ACCUM: case UBG_FIELD_VY_CURRENT:
      *w = W; *hh = H - 1; return;
    default:
      *w = W; *hh = H;
    }
  }
};

// Simulation(flag, pwidth, mu) (simulation.hpp:32-67), or with blank = true
// Simulation(pwidth, mu, width, height) (simulation.hpp:20-30): every field
// zero (the flag too), no inflow column, and MG(width, height), whose coarse
// flags stay 1 until updateFields.
static int sim_create(const float *flag, int width, int height, float pwidth, float mu,
                      bool blank, ubg_sim **out) {
  if ((!flag && !blank) || !out) return fail(UBG_E_INVALID, "null argument");
  if (width < 8 || height < 8)
    return fail(UBG_E_DOMAIN_TOO_SMALL, "Simulation needs at least 8 cells in each dimension");
  int ndev = 0;
  ubg_device_count(&ndev);
  if (ndev <= 0) return fail(UBG_E_NO_DEVICE, "no HIP device visible");
  std::vector<uint8_t> f8;
  if (blank) f8.assign((size_t)width * height, 0);
  else UBG_TRY(validate_flag(flag, (size_t)width * height, f8));
  auto *s = new ubg_sim();
  s->W = width;
  s->H = height;
  s->pwidth = pwidth;
  s->mu = mu;
  default_params(&s->prm);
  const size_t nc = (size_t)width * height, nx = (size_t)(width - 1) * height,
               ny = (size_t)width * (height - 1);
  size_t bytes = 256 + 4 * align_up(nx * 4) + 4 * align_up(ny * 4) + 4 * align_up(nc * 4) +
                 2 * align_up(nc) + align_up(nc * 2);
  int rc = s->pool.init(bytes);
  if (rc) {
    delete s;
    return rc;
  }
  for (int b = 0; b < 2; b++) {
    s->vx[b] = s->pool.take<float>(nx);
    s->vy[b] = s->pool.take<float>(ny);
  }
  s->vx_acc = s->pool.take<float>(nx);
  s->vy_acc = s->pool.take<float>(ny);
  s->vx_cur = s->pool.take<float>(nx);
  s->vy_cur = s->pool.take<float>(ny);
  s->p_buf[0] = s->pool.take<float>(nc);
  s->p_buf[1] = s->pool.take<float>(nc);
  s->f = s->pool.take<float>(nc);
  s->r = s->pool.take<float>(nc);
  s->flag = s->pool.take<uint8_t>(nc);
  s->mask = s->pool.take<uint8_t>(nc);
  s->fmask = s->pool.take<uint16_t>(nc);
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    s->pool.release();
    delete s;
    return fail(UBG_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  // simulation.hpp:55-60: vy(x,0) = 0 and vx(0,y) = 1 in both buffers (the
  // pool starts zeroed)
  if (!blank) {
    std::vector<float> col((size_t)(width - 1) * height, 0.0f);
    for (int y = 0; y < height; y++) col[(size_t)y * (width - 1)] = 1.0f;
    for (int b = 0; b < 2; b++)
      (void)hipMemcpyAsync(s->vx[b], col.data(), col.size() * 4, hipMemcpyHostToDevice, s->stream);
  }
  (void)hipMemcpyAsync(s->flag, f8.data(), nc, hipMemcpyHostToDevice, s->stream);
  (void)hipStreamSynchronize(s->stream);
  rc = mg_create(width, height, s->stream, s->p_buf[0], s->p_buf[1], s->f, s->flag, s->mask,
                 &s->mg);
  if (rc == UBG_OK) {
    s->mg->owner_ext = &s->ext_ev;
    s->mg->owner_work = &s->own_work;
  }
  if (rc) {
    (void)hipStreamDestroy(s->stream);
    s->pool.release();
    delete s;
    return rc;
  }
  s->mg->timers = s->timers;
  if (blank) // MG(w, h): the level-0 mask of the zero flag, coarse flags 1
    launch_build_mask(s->flag, s->mask, width, height, s->stream);
  else
    s->mg->classify(); // MG(flag) -> updateFields (simulation.hpp:62)
  rc = check_launch();
  if (rc) return rc;
  s->h = pwidth / (width - 1.0f); // simulation.hpp:29, :63
  UBG_HIP(hipStreamSynchronize(s->stream));
  *out = s;
  return UBG_OK;
}

extern "C" {

UBG_API int ubg_sim_create(const float *flag, int width, int height, float pwidth, float mu,
                           ubg_sim **out) {
  return sim_create(flag, width, height, pwidth, mu, false, out);
}

UBG_API int ubg_sim_create_blank(int width, int height, float pwidth, float mu, ubg_sim **out) {
  return sim_create(nullptr, width, height, pwidth, mu, true, out);
}

UBG_API int ubg_sim_destroy(ubg_sim *s) {
  if (!s) return UBG_OK;
  s->enter(); // batched passes on this simulation's buffers end first
  (void)hipStreamSynchronize(s->stream);
  s->invalidate_graphs();
  for (auto &t : s->timers)
    for (auto ev : t.ev) (void)hipEventDestroy(ev);
  ubg_mg_destroy(s->mg);
  if (s->d_sink_xy) (void)hipFree(s->d_sink_xy);
  if (s->d_sink_val) (void)hipFree(s->d_sink_val);
  if (s->tr_xy) (void)hipFree(s->tr_xy);
  if (s->tr_age) (void)hipFree(s->tr_age);
  if (s->tr_perm) (void)hipFree(s->tr_perm);
  if (s->stage) (void)hipFree(s->stage);
  (void)hipStreamDestroy(s->stream);
  s->pool.release();
  delete s;
  return UBG_OK;
}

UBG_API int ubg_sim_set_params(ubg_sim *s, const ubg_params *prm) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  UBG_TRY(check_params(prm));
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  s->prm = *prm;
  s->mg->prm = *prm;
  s->invalidate_graphs();
  return UBG_OK;
}

UBG_API int ubg_sim_get_params(const ubg_sim *s, ubg_params *out) {
  if (!s || !out) return fail(UBG_E_INVALID, "null argument");
  *out = s->prm;
  return UBG_OK;
}

UBG_API int ubg_sim_dims(const ubg_sim *s, int *w, int *h, float *hh, float *pw, float *mu) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  if (w) *w = s->W;
  if (h) *h = s->H;
  if (hh) *hh = s->h;
  if (pw) *pw = s->pwidth;
  if (mu) *mu = s->mu;
  return UBG_OK;
}

UBG_API int ubg_sim_field_shape(const ubg_sim *s, int field, int *w, int *h) {
  if (!s || !w || !h) return fail(UBG_E_INVALID, "null argument");
  if (field < 0 || field >= UBG_FIELD_COUNT) return fail(UBG_E_INVALID, "unknown field");
  s->shape(field, w, h);
  return UBG_OK;
}

UBG_API int ubg_sim_set_field(ubg_sim *s, int field, const float *src) {
  if (!s || !src) return fail(UBG_E_INVALID, "null argument");
  if (field < 0 || field >= UBG_FIELD_COUNT) return fail(UBG_E_INVALID, "unknown field");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  s->vy_bc_fresh = false;
  int w, h;
  s->shape(field, &w, &h);
  size_t n = (size_t)w * h;
  if (field == UBG_FIELD_FLAG) {
    UBG_TRY(mg_upload_flag(s->mg, src)); // memcpy + mg.updateFields
    return UBG_OK;
  }
  UBG_HIP(hipMemcpyAsync(s->field(field), src, n * 4, hipMemcpyHostToDevice, s->stream));
  UBG_HIP(hipStreamSynchronize(s->stream));
  if (field == UBG_FIELD_VX_ACCUM || field == UBG_FIELD_VY_ACCUM) s->accum_dirty = true;
  return UBG_OK;
}

UBG_API int ubg_sim_get_field(ubg_sim *s, int field, float *dst) {
  if (!s || !dst) return fail(UBG_E_INVALID, "null argument");
  if (field < 0 || field >= UBG_FIELD_COUNT) return fail(UBG_E_INVALID, "unknown field");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  int w, h;
  s->shape(field, &w, &h);
  size_t n = (size_t)w * h;
  if (field == UBG_FIELD_FLAG) {
    std::vector<uint8_t> tmp(n);
    UBG_HIP(hipMemcpyAsync(tmp.data(), s->flag, n, hipMemcpyDeviceToHost, s->stream));
    UBG_HIP(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < n; i++) dst[i] = (float)tmp[i];
    return UBG_OK;
  }
  UBG_HIP(hipMemcpyAsync(dst, s->field(field), n * 4, hipMemcpyDeviceToHost, s->stream));
  UBG_HIP(hipStreamSynchronize(s->stream));
  return UBG_OK;
}

UBG_API int ubg_sim_field_device_ptr(ubg_sim *s, int field, void **dptr) {
  if (!s || !dptr) return fail(UBG_E_INVALID, "null argument");
  if (field < 0 || field >= UBG_FIELD_COUNT) return fail(UBG_E_INVALID, "unknown field");
  *dptr = field == UBG_FIELD_FLAG ? (void *)s->flag : (void *)s->field(field);
  return UBG_OK;
}

UBG_API int ubg_sim_step(ubg_sim *s, float dt) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  return s->step(dt);
}

UBG_API int ubg_sim_advect_tracers(ubg_sim *s, float dt, uint32_t seed) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  launch_tracers(s->vel(), s->flag, s->h, s->pwidth, dt, seed, s->tr_xy, s->tr_age, s->tr_perm,
                 s->ntr, s->stream);
  return check_launch();
}

UBG_API int ubg_sim_run(ubg_sim *s, float dt, int nsteps) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  for (int i = 0; i < nsteps; i++) {
    UBG_TRY(s->step(dt));
    if (s->ntr > 0) {
      launch_tracers(s->vel(), s->flag, s->h, s->pwidth, dt, s->tr_seed++, s->tr_xy, s->tr_age,
                     s->tr_perm, s->ntr, s->stream);
    }
  }
  return check_launch();
}

UBG_API int ubg_sim_stage(ubg_sim *s, int stage, float dt) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  s->dt = dt;
  switch (stage) {
  case UBG_STAGE_ACCUM:
    launch_accumulate(s->vel(), s->vx_acc, s->vy_acc, s->stream);
    s->accum_dirty = false;
    s->vy_bc_fresh = false;
    break;
  case UBG_STAGE_DIFFUSE: s->diffuse(); break;
  case UBG_STAGE_ADVECT: s->advect(); break;
  case UBG_STAGE_SET_VBCS: s->set_vbcs(); break;
  case UBG_STAGE_PROJECT: UBG_TRY(s->project()); break;
  case UBG_STAGE_SET_PBC: launch_set_pbc(s->P(), s->W, s->H, s->bc, s->stream); break;
  case UBG_STAGE_SAVE_CURRENT: s->save_current(); break;
  case UBG_STAGE_DIVERGENCE:
    s->divergence();
    UBG_TRY(s->apply_sinks());
    break;
  case UBG_STAGE_CORRECT_VELOCITY: s->correct_velocity(); break;
  case UBG_STAGE_MG_SOLVE: s->mg_solve(); break;
  default: return fail(UBG_E_INVALID, "unknown stage");
  }
  UBG_TRY(check_launch());
  UBG_HIP(hipStreamSynchronize(s->stream));
  return UBG_OK;
}

UBG_API int ubg_sim_set_grids(ubg_sim *s, int x, int y, float val) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  if (val != 0.0f && val != 1.0f) return fail(UBG_E_FLAG_VALUE, "flag value must be 0 or 1");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  // simulation.hpp:82-98: out-of-range cells are ignored
  if (!(x >= 0 && y >= 0 && x < s->W && y < s->H)) return UBG_OK;
  s->vy_bc_fresh = false;
  const int W = s->W;
  uint8_t v8 = val != 0.0f ? 1 : 0;
  UBG_HIP(hipMemcpyAsync(s->flag + (size_t)y * W + x, &v8, 1, hipMemcpyHostToDevice, s->stream));
  if (val == 0.0f) {
    const float z = 0.0f;
    if (x < W - 1) UBG_HIP(hipMemcpyAsync(s->vx[s->fx] + (size_t)y * (W - 1) + x, &z, 4, hipMemcpyHostToDevice, s->stream));
    if (x > 0) UBG_HIP(hipMemcpyAsync(s->vx[s->fx] + (size_t)y * (W - 1) + x - 1, &z, 4, hipMemcpyHostToDevice, s->stream));
    if (y < s->H - 1) UBG_HIP(hipMemcpyAsync(s->vy[s->fy] + (size_t)y * W + x, &z, 4, hipMemcpyHostToDevice, s->stream));
    if (y > 0) UBG_HIP(hipMemcpyAsync(s->vy[s->fy] + (size_t)(y - 1) * W + x, &z, 4, hipMemcpyHostToDevice, s->stream));
    UBG_HIP(hipMemcpyAsync(s->P() + (size_t)y * W + x, &z, 4, hipMemcpyHostToDevice, s->stream));
  }
  // the level-0 neighbour mask around the cell
  launch_build_mask(s->flag, s->mask, s->W, s->H, s->stream);
  s->mg->flag_gen++;
  UBG_TRY(check_launch());
  UBG_HIP(hipStreamSynchronize(s->stream));
  return UBG_OK;
}

UBG_API int ubg_sim_set_bcs(ubg_sim *s, int west, int east, int north, int south) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  for (int b : {west, east, north, south})
    if (b < BC_INFLOW || b > BC_NOSLIP) return fail(UBG_E_INVALID, "unknown boundary condition");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  if (s->bc.west == west && s->bc.east == east && s->bc.north == north && s->bc.south == south)
    return UBG_OK;
  s->bc.west = west;
  s->bc.east = east;
  s->bc.north = north;
  s->bc.south = south;
  s->vy_bc_fresh = false;
  s->invalidate_graphs();
  return UBG_OK;
}

UBG_API int ubg_sim_get_bcs(const ubg_sim *s, int *west, int *east, int *north, int *south) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  if (west) *west = s->bc.west;
  if (east) *east = s->bc.east;
  if (north) *north = s->bc.north;
  if (south) *south = s->bc.south;
  return UBG_OK;
}

UBG_API int ubg_sim_set_sinks(ubg_sim *s, const float *xyz, int n) {
  if (!s || n < 0 || (n > 0 && !xyz)) return fail(UBG_E_INVALID, "bad argument");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  s->sinks.clear();
  for (int i = 0; i < n; i++) s->sinks.push_back({xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
  return UBG_OK;
}

UBG_API int ubg_sim_add_sink(ubg_sim *s, float x, float y, float z) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  s->sinks.push_back({x, y, z});
  return UBG_OK;
}

UBG_API int ubg_sim_get_sinks(const ubg_sim *s, float *xyz, int cap, int *count) {
  if (!s || !count) return fail(UBG_E_INVALID, "null argument");
  *count = (int)s->sinks.size();
  for (int i = 0; i < (int)s->sinks.size() && i < cap && xyz; i++) {
    xyz[3 * i] = s->sinks[i].x;
    xyz[3 * i + 1] = s->sinks[i].y;
    xyz[3 * i + 2] = s->sinks[i].z;
  }
  return UBG_OK;
}

UBG_API int ubg_sim_get_mg(ubg_sim *s, ubg_mg **mg) {
  if (!s || !mg) return fail(UBG_E_INVALID, "null argument");
  *mg = s->mg;
  return UBG_OK;
}

UBG_API int ubg_sim_bilinear_vel(ubg_sim *s, const float *c, float *out, int n) {
  if (!s || !c || !out || n < 0) return fail(UBG_E_INVALID, "bad argument");
  if (n == 0) return UBG_OK;
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  float *d = nullptr;
  UBG_TRY(s->staging(sizeof(float) * 4 * n, (void **)&d));
  UBG_HIP(hipMemcpyAsync(d, c, sizeof(float) * 2 * n, hipMemcpyHostToDevice, s->stream));
  launch_bilinear_vel(s->vel(), d, d + 2 * n, n, s->stream);
  int rc = check_launch();
  hipError_t e = hipMemcpyAsync(out, d + 2 * n, sizeof(float) * 2 * n, hipMemcpyDeviceToHost,
                                s->stream);
  const hipError_t es = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = es;
  if (rc) return rc;
  if (e != hipSuccess) return fail(UBG_E_HIP, hipGetErrorString(e));
  return UBG_OK;
}

UBG_API int ubg_sim_sample_flag(ubg_sim *s, int mode, const float *pc, float *out, int n) {
  if (!s || !pc || !out || n < 0) return fail(UBG_E_INVALID, "bad argument");
  if (mode < UBG_FLAG_NEAREST || mode > UBG_FLAG_NORMAL) return fail(UBG_E_INVALID, "unknown mode");
  if (n == 0) return UBG_OK;
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  const int nout = mode == UBG_FLAG_NORMAL ? 2 * n : n;
  float *d = nullptr;
  UBG_TRY(s->staging(sizeof(float) * (2 * (size_t)n + nout), (void **)&d));
  UBG_HIP(hipMemcpyAsync(d, pc, sizeof(float) * 2 * n, hipMemcpyHostToDevice, s->stream));
  launch_sample_flag(s->flag, s->W, s->H, s->pwidth, s->h, mode, d, d + 2 * n, n, s->stream);
  int rc = check_launch();
  hipError_t e = hipMemcpyAsync(out, d + 2 * n, sizeof(float) * nout, hipMemcpyDeviceToHost,
                                s->stream);
  const hipError_t es = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = es;
  if (rc) return rc;
  if (e != hipSuccess) return fail(UBG_E_HIP, hipGetErrorString(e));
  return UBG_OK;
}

UBG_API int ubg_sim_get_dt(ubg_sim *s, float *dt) {
  if (!s || !dt) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  const size_t n = (size_t)(s->W - 2) * (s->H - 2);
  const int nb = (int)std::min<size_t>(1024, (n + 255) / 256);
  float *d = nullptr;
  UBG_TRY(s->staging(sizeof(float) * nb, (void **)&d));
  launch_max_vel(s->vel(), d, nb, s->stream);
  UBG_TRY(check_launch());
  std::vector<float> part(nb);
  UBG_HIP(hipMemcpyAsync(part.data(), d, sizeof(float) * nb, hipMemcpyDeviceToHost, s->stream));
  UBG_HIP(hipStreamSynchronize(s->stream));
  // simulation.cpp:227-234: the scan starts at 1e-7; glm::max keeps the
  // larger value, so the order of the partial maxima does not matter
  float max_vel = 1.0e-7f;
  for (float v : part)
    if (max_vel < v) max_vel = v;
  *dt = s->pwidth / (s->W - 1.0f) / max_vel * 2.5f; // simulation.cpp:236
  return UBG_OK;
}

UBG_API int ubg_sim_center_p(ubg_sim *s) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  const size_t n = (size_t)(s->W - 2) * (s->H - 2);
  const int nb = (int)std::min<size_t>(1024, (n + 255) / 256);
  double *d = nullptr;
  UBG_TRY(s->staging(sizeof(double) * nb, (void **)&d));
  launch_sum_interior(s->P(), s->W, s->H, d, nb, s->stream);
  UBG_TRY(check_launch());
  std::vector<double> part(nb);
  UBG_HIP(hipMemcpyAsync(part.data(), d, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
  UBG_HIP(hipStreamSynchronize(s->stream));
  double sum = 0.0;
  for (double v : part) sum += v;
  // simulation.cpp:218-223: the float sum divided by width, then height
  const float avgP = (float)sum;
  launch_sub_interior(s->P(), s->W, s->H, avgP / s->W / s->H, s->stream);
  UBG_TRY(check_launch());
  UBG_HIP(hipStreamSynchronize(s->stream));
  return UBG_OK;
}

UBG_API int ubg_sim_advect_items(ubg_sim *s, ubg_item *items, int n, float game_dt, int simple) {
  if (!s || n < 0 || (n > 0 && !items)) return fail(UBG_E_INVALID, "bad argument");
  if (!std::isfinite(game_dt)) return fail(UBG_E_INVALID, "game_dt must be finite");
  if (n == 0) return UBG_OK;
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  const size_t bytes = sizeof(ubg_item) * n + sizeof(float) * 2 * n;
  char *d = nullptr;
  UBG_TRY(s->staging(bytes, (void **)&d));
  ubg_item *di = (ubg_item *)d;
  float *snap = (float *)(d + sizeof(ubg_item) * n);
  std::vector<float> pos(2 * (size_t)n);
  for (int i = 0; i < n; i++) {
    pos[2 * i] = items[i].pos[0];
    pos[2 * i + 1] = items[i].pos[1];
  }
  UBG_HIP(hipMemcpyAsync(di, items, sizeof(ubg_item) * n, hipMemcpyHostToDevice, s->stream));
  UBG_HIP(hipMemcpyAsync(snap, pos.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice,
                         s->stream));
  launch_items(di, snap, n, s->vx[s->fx], s->vy[s->fy], s->P(), s->flag, s->vx_acc, s->vy_acc, s->W,
               s->H, s->h, s->pwidth, game_dt, simple != 0, s->stream);
  s->accum_dirty = true; // the next step adds the items' reaction to the velocity
  int rc = check_launch();
  hipError_t e = hipMemcpyAsync(items, di, sizeof(ubg_item) * n, hipMemcpyDeviceToHost, s->stream);
  const hipError_t es = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = es;
  if (rc) return rc;
  if (e != hipSuccess) return fail(UBG_E_HIP, hipGetErrorString(e));
  return UBG_OK;
}

UBG_API int ubg_sim_shift_map(ubg_sim *s, const float *new_flag) {
  if (!s || !new_flag) return fail(UBG_E_INVALID, "null argument");
  if (s->W > 16384) return fail(UBG_E_INVALID, "map shift needs W <= 16384");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  std::vector<uint8_t> f8;
  UBG_TRY(validate_flag(new_flag, (size_t)s->W * s->H, f8));
  uint8_t *d = nullptr;
  UBG_TRY(s->staging(f8.size(), (void **)&d));
  UBG_HIP(hipMemcpyAsync(d, f8.data(), f8.size(), hipMemcpyHostToDevice, s->stream));
  launch_shift_map(s->vel(), s->P(), d, s->flag, s->stream);
  s->vy_bc_fresh = false;
  s->save_current();              // ubootgl_app.cpp:292
  s->mg->classify();              // ubootgl_app.cpp:295, mg.updateFields
  int rc = check_launch();
  hipError_t e = hipStreamSynchronize(s->stream);
  if (rc) return rc;
  if (e != hipSuccess) return fail(UBG_E_HIP, hipGetErrorString(e));
  return UBG_OK;
}

UBG_API int ubg_terrain_flags(const uint8_t *rgba, int width, int height, int scale,
                              float *flag_full, float *flag_sim) {
  if (!rgba || !flag_sim || width <= 0 || height <= 0 || scale <= 0)
    return fail(UBG_E_INVALID, "bad argument");
  int nd = 0;
  ubg_device_count(&nd);
  if (nd <= 0) return fail(UBG_E_NO_DEVICE, "no HIP device visible");
  const size_t npx = (size_t)width * height;
  const int sw = width / scale, sh = height / scale;
  const size_t nsim = (size_t)sw * sh;
  char *d = nullptr;
  UBG_HIP(hipMalloc((void **)&d, 4 * npx + 4 * npx + 4 * nsim + 256));
  uint8_t *drgba = (uint8_t *)d;
  float *dfull = (float *)(d + 4 * npx);
  float *dsim = (float *)(d + 8 * npx);
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) {
    e = hipMemcpyAsync(drgba, rgba, 4 * npx, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      launch_terrain_flags(drgba, width, height, scale, dfull, dsim, st);
      e = hipGetLastError();
    }
    if (e == hipSuccess && flag_full)
      e = hipMemcpyAsync(flag_full, dfull, 4 * npx, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && nsim) e = hipMemcpyAsync(flag_sim, dsim, 4 * nsim, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  (void)hipFree(d);
  if (e != hipSuccess) return fail(UBG_E_HIP, hipGetErrorString(e));
  return UBG_OK;
}

UBG_API int ubg_sim_collocate(ubg_sim *s, float *vxy, float *mag, int on_device) {
  if (!s || !vxy || !mag) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  const size_t n = (size_t)(2 * s->W - 1) * (2 * s->H - 1);
  if (on_device) {
    launch_collocate(s->vx_cur, s->vy_cur, vxy, mag, s->W, s->H, s->stream);
    UBG_TRY(check_launch());
    return UBG_OK;
  }
  float *d = nullptr;
  UBG_TRY(s->staging(sizeof(float) * 3 * n, (void **)&d));
  launch_collocate(s->vx_cur, s->vy_cur, d, d + 2 * n, s->W, s->H, s->stream);
  int rc = check_launch();
  hipError_t e = hipMemcpyAsync(vxy, d, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(mag, d + 2 * n, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream);
  const hipError_t es = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = es;
  if (rc) return rc;
  if (e != hipSuccess) return fail(UBG_E_HIP, hipGetErrorString(e));
  return UBG_OK;
}

UBG_API int ubg_sim_set_tracers(ubg_sim *s, const float *xy, const float *age, int n) {
  if (!s || n < 0 || (n > 0 && !xy)) return fail(UBG_E_INVALID, "bad argument");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  if (n > s->tr_cap) {
    if (s->tr_xy) (void)hipFree(s->tr_xy);
    if (s->tr_age) (void)hipFree(s->tr_age);
    if (s->tr_perm) (void)hipFree(s->tr_perm);
    s->tr_xy = s->tr_age = nullptr;
    s->tr_perm = nullptr;
    s->tr_cap = 0;
    UBG_HIP(hipMalloc((void **)&s->tr_xy, sizeof(float) * 2 * (size_t)n));
    UBG_HIP(hipMalloc((void **)&s->tr_age, sizeof(float) * (size_t)n));
    UBG_HIP(hipMalloc((void **)&s->tr_perm, sizeof(int) * (size_t)n));
    s->tr_cap = n;
  }
  s->ntr = n;
  if (n == 0) return UBG_OK;
  // sort by 16x16 tile of the grid cell holding the point
  const int tw = (s->W + 15) / 16;
  std::vector<int64_t> key(n);
  for (int i = 0; i < n; i++) {
    float cx = xy[2 * i] / s->h, cy = xy[2 * i + 1] / s->h;
    int ix = std::isfinite(cx) ? (int)std::min(std::max(cx, 0.0f), (float)(s->W - 1)) : 0;
    int iy = std::isfinite(cy) ? (int)std::min(std::max(cy, 0.0f), (float)(s->H - 1)) : 0;
    key[i] = ((int64_t)(iy >> 4) * tw + (ix >> 4)) * 256 + (iy & 15) * 16 + (ix & 15);
  }
  std::vector<int> &perm = s->tr_perm_h;
  perm.resize(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<float> sxy(2 * (size_t)n), sage((size_t)n, 0.0f);
  for (int j = 0; j < n; j++) {
    sxy[2 * (size_t)j] = xy[2 * (size_t)perm[j]];
    sxy[2 * (size_t)j + 1] = xy[2 * (size_t)perm[j] + 1];
    if (age) sage[j] = age[perm[j]];
  }
  UBG_HIP(hipMemcpyAsync(s->tr_xy, sxy.data(), sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice, s->stream));
  UBG_HIP(hipMemcpyAsync(s->tr_age, sage.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice, s->stream));
  UBG_HIP(hipMemcpyAsync(s->tr_perm, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s->stream));
  UBG_HIP(hipStreamSynchronize(s->stream));
  return UBG_OK;
}

UBG_API int ubg_sim_get_tracers(ubg_sim *s, float *xy, float *age, int n) {
  if (!s || n < 0 || n > s->ntr) return fail(UBG_E_INVALID, "bad argument");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  if (n == 0) return UBG_OK;
  const int N = s->ntr;
  std::vector<float> sxy(2 * (size_t)N), sage((size_t)N);
  UBG_HIP(hipMemcpyAsync(sxy.data(), s->tr_xy, sizeof(float) * 2 * (size_t)N, hipMemcpyDeviceToHost, s->stream));
  UBG_HIP(hipMemcpyAsync(sage.data(), s->tr_age, sizeof(float) * (size_t)N, hipMemcpyDeviceToHost, s->stream));
  UBG_HIP(hipStreamSynchronize(s->stream));
  const std::vector<int> &perm = s->tr_perm_h;
  for (int j = 0; j < N; j++) {
    int i = perm[j];
    if (i >= n) continue;
    if (xy) {
      xy[2 * (size_t)i] = sxy[2 * (size_t)j];
      xy[2 * (size_t)i + 1] = sxy[2 * (size_t)j + 1];
    }
    if (age) age[i] = sage[j];
  }
  return UBG_OK;
}

UBG_API int ubg_sim_sync(ubg_sim *s) {
  if (!s) return fail(UBG_E_INVALID, "null sim");
  {
    std::lock_guard<std::recursive_mutex> g(s->mu_);
    s->enter();
  }
  UBG_HIP(hipStreamSynchronize(s->stream));
  return UBG_OK;
}

UBG_API int ubg_sim_timer_enable(ubg_sim *s, int timer, int on) {
  if (!s || timer < 0 || timer >= UBG_TIMER_COUNT) return fail(UBG_E_INVALID, "bad argument");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  UBG_HIP(hipStreamSynchronize(s->stream));
  Timer &t = s->timers[timer];
  t.on = on != 0;
  t.used = 0;
  t.total_ms = 0.0;
  t.launches = 0;
  return UBG_OK;
}

UBG_API int ubg_sim_timer_read(ubg_sim *s, int timer, double *total_ms, long *launches) {
  if (!s || timer < 0 || timer >= UBG_TIMER_COUNT) return fail(UBG_E_INVALID, "bad argument");
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  UBG_HIP(hipStreamSynchronize(s->stream));
  Timer &t = s->timers[timer];
  for (size_t i = 0; i + 1 < t.used; i += 2) {
    float ms = 0.0f;
    UBG_HIP(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
    t.total_ms += ms;
    t.launches++;
  }
  t.used = 0;
  if (total_ms) *total_ms = t.total_ms;
  if (launches) *launches = t.launches;
  return UBG_OK;
}

} // extern "C"

// the step around its pressure solve, for batched steps (batch.hip)
namespace ubg {

int sim_step_before_solve(ubg_sim *s, float dt) {
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  UBG_TRY(s->step_before_solve(dt));
  return check_launch();
}

int sim_step_after_solve(ubg_sim *s) {
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->enter();
  s->body_after_solve();
  return check_launch();
}

int sim_batched_records(ubg_sim *s, float dt, SimRec rec[4], bool *has_sinks) {
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  if (s->accum_dirty || !s->vy_bc_fresh || s->fmask_gen != s->mg->flag_gen)
    s->enter(); // the prelude below enqueues on the simulation's stream
  s->dt = dt;
  if (s->accum_dirty) {
    launch_accumulate(s->vel(), s->vx_acc, s->vy_acc, s->stream);
    s->accum_dirty = false;
    s->vy_bc_fresh = false;
  }
  s->refresh_vy_bcs();
  s->ensure_fmask();
  UBG_TRY(check_launch());
  const float ih = 1.0f / s->h;
  auto rec_of = [&](SimRec &r) {
    const VelView v = s->vel();
    r = SimRec{};
    r.vxf = v.vxf;
    r.vxb = v.vxb;
    r.vyf = v.vyf;
    r.vyb = v.vyb;
    r.fmask = s->fmask;
    r.flag = s->flag;
    r.f = s->f;
    r.p = s->P();
    r.xcur = s->prm.save_current ? s->vx_cur : nullptr;
    r.ycur = s->prm.save_current ? s->vy_cur : nullptr;
    r.W = s->W;
    r.H = s->H;
    r.west = s->bc.west;
    r.east = s->bc.east;
    r.north = s->bc.north;
    r.south = s->bc.south;
    r.a = dt * s->mu * (s->W - 1.0f) / s->pwidth; // as diffuse_fused
    r.hdt = 0.5f * dt * ih;                       // as advect
    r.fdt = dt * ih;
    r.ih = ih;
  };
  // the roles of step_body's passes: two diffuse iterations and advect each
  // flip both velocity buffers; setVBCs + divergence and the correction
  // tail see the roles after the third flip
  for (int k = 0; k < 4; k++) {
    rec_of(rec[k]);
    if (k < 3) {
      s->fx = 1 - s->fx;
      s->fy = 1 - s->fy;
    }
  }
  s->vy_bc_fresh = true; // after the correction tail
  *has_sinks = !s->sinks.empty();
  return UBG_OK;
}

int sim_apply_sinks_on(ubg_sim *s, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  return s->apply_sinks_on(st);
}

bool sim_take_own_work(ubg_sim *s) {
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  const bool w = s->own_work;
  s->own_work = false;
  return w;
}

void sim_set_ext_event(ubg_sim *s, const std::shared_ptr<ExtEvent> &ev) {
  std::lock_guard<std::recursive_mutex> g(s->mu_);
  s->ext_ev = ev;
  s->own_work = false;
}

SimLevel0 sim_level0(ubg_sim *s) {
  SimLevel0 o;
  o.W = s->W;
  o.H = s->H;
  o.h = s->h;
  o.p = s->P();
  o.f = s->f;
  o.flag = s->flag;
  o.stream = s->stream;
  o.prm = s->prm;
  o.flag_gen = s->mg->flag_gen;
  return o;
}

} // namespace ubg
