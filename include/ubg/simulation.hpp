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
#include <sstream>
#include <type_traits>
#include <vector>

class Simulation {
public:
  enum class BC { INFLOW, OUTFLOW, OUTFLOW_ZERO_PRESSURE, NOSLIP };

  struct Vec2 {
    float x, y;
  };
  struct Vec3 {
    float x, y, z;
  };

  // simulation.hpp:32-67
  Simulation(const Single2DGrid &flagInput, float pwidth, float mu)
      : pwidth(pwidth), mu(mu), width(flagInput.width), height(flagInput.height),
        vx(width - 1, height), vy(width, height - 1), vx_accum(width - 1, height),
        vy_accum(width, height - 1), vx_current(width - 1, height),
        vy_current(width, height - 1), p(width, height), f(width, height), flag(flagInput),
        r(width, height) {
    ubg_check(ubg_sim_create(flagInput.data(), width, height, pwidth, mu, &h_));
    float hh = 0;
    ubg_check(ubg_sim_dims(h_, nullptr, nullptr, &hh, nullptr, nullptr));
    h = hh;
    ubg_mg *m = nullptr;
    ubg_check(ubg_sim_get_mg(h_, &m));
    mg = MG::borrow(m, width, height);
    pullBCs();
    syncToHost();
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

  // advect_floating_items.cpp:16-274.  The reference takes an entt::registry
  // and views CoItem + CoKinematics (or CoKinematicsSimple); this header does
  // not depend on entt or the game's components, so the game keeps two
  // forwarders:
  //   void advectFloatingItems(entt::registry &r, float dt)
  //     { advectFloatingItemsView<CoItem, CoKinematics>(r, dt, false); }
  //   void advectFloatingItemsSimple(entt::registry &r, float dt)
  //     { advectFloatingItemsView<CoItem, CoKinematicsSimple>(r, dt, true); }
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
  std::vector<Vec3> sinks;

private:
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
