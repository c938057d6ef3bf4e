// simulation.hpp -- drop-in Simulation (simulation.hpp:18-141 of the
// reference) backed by libubg.so.
//
// The fields live on the GPU.  The public data members of the reference
// (vx, vy, p, f, flag, r, *_accum, *_current, sinks, bc*) are kept as host
// mirrors with the same names and types; the one behavioural difference is
// that the mirrors are exchanged explicitly:
//   syncToHost()    device -> mirrors (all fields, both vx/vy buffers)
//   syncToDevice()  mirrors -> device (all fields; flag re-runs updateFields)
// step()/diffuse()/advect()/project()/... run on the device; sinks and BCs are
// pushed before and pulled after each call, as the reference keeps them as
// plain members.  Writing the flag mirror and calling mg.updateFields(flag)
// (the game's per-frame pattern, ubootgl_app.cpp:111-112) uploads the flag
// and reclassifies the coarse levels.
#pragma once

#include "../ubg_c.h"
#include "db2dgrid.hpp"
#include "pressure_solver.hpp"

#include <array>
#include <mutex>
#include <random>
#include <sstream>
#include <type_traits>
#include <vector>

// The game's component types (components.hpp:6-42), viewed by
// advectFloatingItems / advectFloatingItemsSimple; complete wherever those
// are called.
struct CoItem;
struct CoKinematics;
struct CoKinematicsSimple;

class Simulation {
public:
  enum class BC { INFLOW, OUTFLOW, OUTFLOW_ZERO_PRESSURE, NOSLIP };

  struct Vec2 {
    float x, y;
  };
  // a sink (x, y, strength), as the reference's glm::vec3; converts from any
  // type with float members x, y, z, so sinks.push_back(glm::vec3(...))
  // (explosion.cpp:33) compiles unchanged
  struct Vec3 {
    float x, y, z;
    Vec3() = default;
    Vec3(float x, float y, float z) : x(x), y(y), z(z) {}
    template <class V, class = decltype(std::declval<const V &>().z)>
    Vec3(const V &v) : x((float)v.x), y((float)v.y), z((float)v.z) {}
  };

  // simulation.hpp:20-30: every field zero (the flag too), MG(width, height)
  // with all coarse flags 1 until mg.updateFields(flag).  The reference leaves
  // the boundary conditions uninitialised; here they start as the other
  // constructor sets them.
  Simulation(float pwidth, float mu, int width, int height)
      : pwidth(pwidth), mu(mu), width(width), height(height), vx(width - 1, height),
        vy(width, height - 1), vx_accum(width - 1, height), vy_accum(width, height - 1),
        vx_current(width - 1, height), vy_current(width, height - 1), p(width, height),
        f(width, height), flag(width, height), r(width, height), disx(0.0f, 1.0f),
        disy(0.0f, (float)height / width) {
    ubg_check(ubg_sim_create_blank(width, height, pwidth, mu, &h_));
    attach();
  }

  // simulation.hpp:32-67
  Simulation(const Single2DGrid &flagInput, float pwidth, float mu)
      : pwidth(pwidth), mu(mu), width(flagInput.width), height(flagInput.height),
        vx(width - 1, height), vy(width, height - 1), vx_accum(width - 1, height),
        vy_accum(width, height - 1), vx_current(width - 1, height),
        vy_current(width, height - 1), p(width, height), f(width, height), flag(flagInput),
        r(width, height), disx(0.0f, 1.0f), disy(0.0f, (float)height / width) {
    ubg_check(ubg_sim_create(flagInput.data(), width, height, pwidth, mu, &h_));
    attach();
  }
  Simulation(const Simulation &) = delete;
  Simulation &operator=(const Simulation &) = delete;
  ~Simulation() {
    if (h_) ubg_sim_destroy(h_);
  }

  void setParams(const ubg_params &prm) { ubg_check(ubg_sim_set_params(h_, &prm)); }

  // simulation.cpp:356-374
  void step(float timestep) {
    dt = timestep;
    pushBCs();
    pushSinks();
    ubg_check(ubg_sim_step(h_, timestep));
    pullSinks();
  }
  void diffuse() { stage(UBG_STAGE_DIFFUSE); }
  void advect() { stage(UBG_STAGE_ADVECT); }
  void project() { stage(UBG_STAGE_PROJECT); }
  void setVBCs() { stage(UBG_STAGE_SET_VBCS); }
  void setPBC() { stage(UBG_STAGE_SET_PBC); }
  void applyAccumulatedVelocity() {
    std::scoped_lock lock(accum_mutex);
    ubg_check(ubg_sim_set_field(h_, UBG_FIELD_VX_ACCUM, vx_accum.data()));
    ubg_check(ubg_sim_set_field(h_, UBG_FIELD_VY_ACCUM, vy_accum.data()));
    stage(UBG_STAGE_ACCUM);
    vx_accum.fill(0.0f);
    vy_accum.fill(0.0f);
  }
  void saveCurrentVelocityFields() { stage(UBG_STAGE_SAVE_CURRENT); }

  // simulation.cpp:23-34, 50-78: the boundary rules as value functions
  float singlePBC(BC bc, float a) { return bc == BC::OUTFLOW_ZERO_PRESSURE ? -a : a; }
  float VBCPar(BC bc, float a, float b) {
    if (bc == BC::INFLOW) return b;
    if (bc == BC::OUTFLOW || bc == BC::OUTFLOW_ZERO_PRESSURE) return a > 0.0f ? a : 0.0f;
    return 0.0f;
  }
  float VBCPer(BC bc, float a, float b) {
    if (bc == BC::INFLOW) return b;
    if (bc == BC::OUTFLOW || bc == BC::OUTFLOW_ZERO_PRESSURE) return a > 0.0f ? a : 0.0f;
    return -a;
  }

  // simulation.cpp:226-238 (on the device fields; the diagnostic line goes
  // to diag as in the reference)
  float getDT() {
    float rdt = 0.0f;
    ubg_check(ubg_sim_get_dt(h_, &rdt));
    diag << "SET_DT: dt=" << rdt << "\n";
    return rdt;
  }
  // simulation.cpp:210-224 (on the device pressure)
  void centerP() { ubg_check(ubg_sim_center_p(h_)); }

  // simulation.cpp:398-425: flag samplers at a physical position (any type
  // with float members x and y, e.g. glm::vec2)
  template <typename V> float psampleFlagNearest(const V &pc) {
    return sampleFlag1(UBG_FLAG_NEAREST, pc);
  }
  template <typename V> float psampleFlagLinear(const V &pc) {
    return sampleFlag1(UBG_FLAG_LINEAR, pc);
  }
  template <typename V> V psampleFlagNormal(const V &pc) {
    float in[2] = {(float)pc.x, (float)pc.y}, out[2];
    ubg_check(ubg_sim_sample_flag(h_, UBG_FLAG_NORMAL, in, out, 1));
    V r = pc;
    r.x = out[0];
    r.y = out[1];
    return r;
  }
  // n positions at once (one device call): pc = (x0, y0, x1, y1, ...)
  void psampleFlags(int mode, const float *pc, float *out, int n) {
    ubg_check(ubg_sim_sample_flag(h_, mode, pc, out, n));
  }

  // advect_floating_items.cpp:16-138 and :140-274 with the reference's
  // signatures (simulation.hpp:116-117): the registry (entt::registry in the
  // game) views the game's CoItem + CoKinematics / CoKinematicsSimple.  This
  // header does not include entt; the registry type is a template parameter,
  // so the call in ubootgl_app.cpp:129-130 compiles unchanged.
  template <class Registry> void advectFloatingItems(Registry &registry, float gameDT) {
    advectFloatingItemsView<CoItem, CoKinematics>(registry, gameDT, false);
  }
  template <class Registry> void advectFloatingItemsSimple(Registry &registry, float gameDT) {
    advectFloatingItemsView<CoItem, CoKinematicsSimple>(registry, gameDT, true);
  }
  // The view's items are packed in visiting order, integrated on the device
  // and written back; their reaction lands in the device-side accumulation
  // fields, which the next step applies (the host mirrors vx_accum / vy_accum
  // are not involved).
  template <class ItemC, class KinC, class Registry>
  void advectFloatingItemsView(Registry &registry, float gameDT, bool simple) {
    auto view = registry.template view<ItemC, KinC>();
    using Entity = std::decay_t<decltype(*view.begin())>;
    std::vector<Entity> ents;
    std::vector<ubg_item> items;
    for (auto e : view) {
      const ItemC &it = view.template get<ItemC>(e);
      const KinC &k = view.template get<KinC>(e);
      ubg_item u{};
      u.size[0] = it.size.x;
      u.size[1] = it.size.y;
      u.pos[0] = it.pos.x;
      u.pos[1] = it.pos.y;
      u.rotation = it.rotation;
      u.mass = k.mass;
      u.vel[0] = k.vel.x;
      u.vel[1] = k.vel.y;
      u.force[0] = k.force.x;
      u.force[1] = k.force.y;
      u.ang_vel = k.angVel;
      u.ang_force = k.angForce;
      u.bump_count = k.bumpCount;
      items.push_back(u);
      ents.push_back(e);
    }
    {
      std::scoped_lock lock(accum_mutex);
      ubg_check(ubg_sim_advect_items(h_, items.data(), (int)items.size(), gameDT, simple ? 1 : 0));
    }
    for (size_t i = 0; i < ents.size(); i++) {
      ItemC &it = view.template get<ItemC>(ents[i]);
      KinC &k = view.template get<KinC>(ents[i]);
      const ubg_item &u = items[i];
      it.pos.x = u.pos[0];
      it.pos.y = u.pos[1];
      it.rotation = u.rotation;
      k.vel.x = u.vel[0];
      k.vel.y = u.vel[1];
      k.force.x = u.force[0];
      k.force.y = u.force[1];
      k.angVel = u.ang_vel;
      k.angForce = u.ang_force;
      k.bumpCount = u.bump_count;
    }
  }

  // simulation.hpp:82-98
  template <typename V> void setGrids(const V &c, float val) {
    ubg_check(ubg_sim_set_grids(h_, (int)c.x, (int)c.y, val));
  }
  // advect_floating_items.cpp:11-14 (grid-space position)
  template <typename V> V bilinearVel(const V &c) {
    float in[2] = {(float)c.x, (float)c.y}, out[2];
    ubg_check(ubg_sim_bilinear_vel(h_, in, out, 1));
    V r = c;
    r.x = out[0];
    r.y = out[1];
    return r;
  }

  float *getFlag() { return flag.data(); }
  float *getP() { return p.data(); }
  float *getR() { return r.data(); }

  void syncToHost() {
    get(UBG_FIELD_VX, vx.data());
    get(UBG_FIELD_VX_BACK, vx.back_data());
    get(UBG_FIELD_VY, vy.data());
    get(UBG_FIELD_VY_BACK, vy.back_data());
    get(UBG_FIELD_VX_ACCUM, vx_accum.data());
    get(UBG_FIELD_VY_ACCUM, vy_accum.data());
    get(UBG_FIELD_VX_CURRENT, vx_current.data());
    get(UBG_FIELD_VY_CURRENT, vy_current.data());
    get(UBG_FIELD_P, p.data());
    get(UBG_FIELD_F, f.data());
    get(UBG_FIELD_FLAG, flag.data());
    get(UBG_FIELD_R, r.data());
  }
  void syncToDevice() {
    set(UBG_FIELD_VX, vx.data());
    set(UBG_FIELD_VX_BACK, vx.back_data());
    set(UBG_FIELD_VY, vy.data());
    set(UBG_FIELD_VY_BACK, vy.back_data());
    set(UBG_FIELD_VX_ACCUM, vx_accum.data());
    set(UBG_FIELD_VY_ACCUM, vy_accum.data());
    set(UBG_FIELD_VX_CURRENT, vx_current.data());
    set(UBG_FIELD_VY_CURRENT, vy_current.data());
    set(UBG_FIELD_P, p.data());
    set(UBG_FIELD_F, f.data());
    set(UBG_FIELD_FLAG, flag.data());
    set(UBG_FIELD_R, r.data());
    pushBCs();
    pushSinks();
  }
  ubg_sim *handle() { return h_; }
  friend class SimulationBatch;

  // public data members of the reference (simulation.hpp:119-140)
  float pwidth;
  float mu;
  int width, height;
  float dt = 0.0f;
  BC bcWest = BC::INFLOW, bcEast = BC::OUTFLOW_ZERO_PRESSURE, bcNorth = BC::NOSLIP,
     bcSouth = BC::NOSLIP;
  std::stringstream diag;
  DoubleBuffered2DGrid vx, vy;
  Single2DGrid vx_accum, vy_accum;
  Single2DGrid vx_current, vy_current;
  std::mutex accum_mutex;
  Single2DGrid p, f, flag, r;
  MG mg;
  float h = 0.0f;
  // simulation.hpp:136-138 (seeded and drawn by the game, not by the step)
  std::default_random_engine gen;
  std::uniform_real_distribution<float> disx;
  std::uniform_real_distribution<float> disy;
  std::vector<Vec3> sinks;

private:
  void attach() {
    float hh = 0;
    ubg_check(ubg_sim_dims(h_, nullptr, nullptr, &hh, nullptr, nullptr));
    h = hh;
    ubg_mg *m = nullptr;
    ubg_check(ubg_sim_get_mg(h_, &m));
    mg = MG::borrow(m, width, height);
    pullBCs();
    syncToHost();
  }
  template <typename V> float sampleFlag1(int mode, const V &pc) {
    float in[2] = {(float)pc.x, (float)pc.y}, out = 0.0f;
    ubg_check(ubg_sim_sample_flag(h_, mode, in, &out, 1));
    return out;
  }
  void stage(int st) {
    pushBCs();
    pushSinks();
    ubg_check(ubg_sim_stage(h_, st, dt));
    pullSinks();
  }
  void get(int id, float *dst) { ubg_check(ubg_sim_get_field(h_, id, dst)); }
  void set(int id, const float *src) { ubg_check(ubg_sim_set_field(h_, id, src)); }
  void pushBCs() {
    ubg_check(ubg_sim_set_bcs(h_, (int)bcWest, (int)bcEast, (int)bcNorth, (int)bcSouth));
  }
  void pullBCs() {
    int w, e, n, s;
    ubg_check(ubg_sim_get_bcs(h_, &w, &e, &n, &s));
    bcWest = (BC)w;
    bcEast = (BC)e;
    bcNorth = (BC)n;
    bcSouth = (BC)s;
  }
  void pushSinks() {
    ubg_check(ubg_sim_set_sinks(h_, sinks.empty() ? nullptr : &sinks[0].x, (int)sinks.size()));
  }
  void pullSinks() {
    int n = 0;
    ubg_check(ubg_sim_get_sinks(h_, nullptr, 0, &n));
    sinks.resize((size_t)n);
    if (n) ubg_check(ubg_sim_get_sinks(h_, &sinks[0].x, n, &n));
  }
  ubg_sim *h_ = nullptr;
};

// Several simulations stepped together, their pressure solves batched into
// one launch per multigrid stage (ubg_sims_step).  No counterpart in the
// reference, which steps one Simulation at a time; step(dts) leaves every
// simulation as sims[i]->step(dts[i]) would, bit for bit.  The simulations
// must share their multigrid parameters and be at least 16 x 16.
class SimulationBatch {
public:
  explicit SimulationBatch(std::vector<Simulation *> sims) : sims_(std::move(sims)) {
    for (Simulation *s : sims_) hs_.push_back(s->h_);
    ubg_check(ubg_batch_create_for_sims(hs_.data(), (int)hs_.size(), &b_));
  }
  ~SimulationBatch() {
    if (b_) ubg_batch_destroy(b_);
  }
  SimulationBatch(const SimulationBatch &) = delete;
  SimulationBatch &operator=(const SimulationBatch &) = delete;

  void step(const std::vector<float> &dts) {
    if (dts.size() != sims_.size()) throw ubg_error(UBG_E_INVALID, "one dt per simulation");
    for (size_t i = 0; i < sims_.size(); i++) {
      sims_[i]->dt = dts[i];
      sims_[i]->pushBCs();
      sims_[i]->pushSinks();
    }
    ubg_check(ubg_sims_step(b_, hs_.data(), (int)hs_.size(), dts.data()));
    for (Simulation *s : sims_) s->pullSinks();
  }

private:
  std::vector<Simulation *> sims_;
  std::vector<ubg_sim *> hs_;
  ubg_batch *b_ = nullptr;
};
