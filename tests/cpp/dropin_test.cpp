// Drop-in test: code written against the reference's Simulation / MG API
// (simulation.hpp, pressure_solver.hpp, db2dgrid.hpp) compiled unchanged
// against include/ubg/*.hpp and run on the GPU through libubg.so, checked
// against the CPU oracle (test infrastructure, linked only here).
#include "ubg/simulation.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <memory>
#include <vector>

extern "C" {
void *oracle_sim_create(const float *flag, int W, int H, float pwidth, float mu);
void oracle_sim_destroy(void *s);
int oracle_sim_set_field(void *s, int id, const float *src);
int oracle_sim_get_field(void *s, int id, float *dst);
void oracle_sim_step(void *s, float dt);
void oracle_sim_advect_items(void *s, void *items, int n, float gameDT, int simple);
void *oracle_mg_create(int W, int H, int fp64);
float oracle_flag_linear(const float *F, int W, int H, float pwidth, float px, float py);
float oracle_flag_nearest(const float *F, int W, int H, float pwidth, float px, float py);
void oracle_flag_normal(const float *F, int W, int H, float pwidth, float h, float px, float py,
                        float *out);
void oracle_gs_f32(float *p, const float *f, const float *fl, int W, int H, float h, float alpha);
void oracle_sim_center_p(void *s);
float oracle_sim_get_dt(void *s);
void oracle_mg_update_fields(void *m, const float *flag);
void oracle_mg_solve(void *m, float *p, const float *f, const float *flag, float h, int zg);
}

static double rel(const float *a, const float *b, size_t n) {
  double num = 0, den = 0;
  for (size_t i = 0; i < n; i++) {
    num += ((double)a[i] - b[i]) * ((double)a[i] - b[i]);
    den += (double)b[i] * b[i];
  }
  return std::sqrt(num / (den > 0 ? den : 1));
}

// A registry with the interface the drop-in's item forwarder uses
// (view<A, B>(), iteration over entities, view.get<T>(entity)); component
// layouts as components.hpp:6-27.
struct V2 {
  float x, y;
};
struct CoItem {
  V2 size, pos;
  float rotation;
};
struct CoKinematics {
  float mass;
  V2 vel, force;
  float angVel, angForce;
  int bumpCount;
};
struct CoKinematicsSimple {
  float mass;
  V2 vel, force;
  float angVel, angForce;
  int bumpCount;
};
// glm::vec3 stand-in for the sinks list (explosion.cpp:33 pushes glm::vec3)
struct V3 {
  float x, y, z;
};
template <class Kin> struct MiniRegistryT {
  std::vector<CoItem> items;
  std::vector<Kin> kins;
  struct View {
    MiniRegistryT *r;
    struct It {
      int i;
      int operator*() const { return i; }
      It &operator++() {
        ++i;
        return *this;
      }
      bool operator!=(const It &o) const { return i != o.i; }
    };
    It begin() const { return {0}; }
    It end() const { return {(int)r->items.size()}; }
    template <class T> T &get(int e) const {
      if constexpr (std::is_same_v<T, CoItem>) return r->items[e];
      else return r->kins[e];
    }
  };
  template <class A, class B> View view() { return View{this}; }
};
using MiniRegistry = MiniRegistryT<CoKinematics>;

static int fails = 0;
static void expect(bool ok, const char *what, double v) {
  std::printf("%-48s %s (%.3e)\n", what, ok ? "ok" : "FAIL", v);
  if (!ok) fails++;
}

int main() {
  // 1) MG on the harmonic Dirichlet problem of the reference's known-answer
  //    test (mgtest/mgtest.cpp:10-42), smaller grid.
  {
    const int N = 129;
    float h = 1.0 / (N - 1);
    Single2DGrid u(N, N), rhs(N, N), flag(N, N), r(N, N);
    u = 0.0;
    rhs = 0.0;
    flag = 1.0;
    for (int x = 1; x < N - 1; x++) u(x, N - 1) = std::sinh(M_PI) * std::sin(x / (N - 1.0) * M_PI);
    std::vector<float> uo(u.data(), u.data() + N * N);
    MG mg(N, N);
    void *om = oracle_mg_create(N, N, 0);
    oracle_mg_update_fields(om, flag.data());
    float prev = calculateResidualField(u, rhs, flag, r, h);
    bool mono = true;
    for (int i = 0; i < 5; i++) {
      mg.solve(u, rhs, flag, h);
      oracle_mg_solve(om, uo.data(), rhs.data(), flag.data(), h, 0);
      float res = calculateResidualField(u, rhs, flag, r, h);
      mono = mono && res < prev;
      prev = res;
    }
    expect(mono, "MG::solve residual decreases every V-cycle", prev);
    expect(rel(u.data(), uo.data(), (size_t)N * N) < 1e-5, "MG::solve matches the oracle",
           rel(u.data(), uo.data(), (size_t)N * N));
  }
  // 2) Simulation: construct from a flag, set a velocity, step twice.
  {
    const int W = 96, H = 48;
    Single2DGrid flag(W, H);
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        bool wall = x == 0 || y == 0 || x == W - 1 || y == H - 1;
        bool disk = (x - 40) * (x - 40) + (y - 20) * (y - 20) <= 36;
        flag(x, y) = (wall || disk) ? 0.0f : 1.0f;
      }
    Simulation sim(flag, (W - 1.0f) / W, 0.001f);
    for (int y = 0; y < H; y++)
      for (int x = 1; x < W - 2; x++)
        sim.vx.f(x, y) = sim.vx.b(x, y) =
            std::sin(0.2f * x) * std::cos(0.3f * y) * flag(x, y) * flag(x + 1, y);
    for (int y = 1; y < H - 2; y++)
      for (int x = 0; x < W; x++)
        sim.vy.f(x, y) = sim.vy.b(x, y) = 0.5f * std::cos(0.1f * x + 0.2f * y) * flag(x, y) * flag(x, y + 1);
    sim.syncToDevice();
    void *os = oracle_sim_create(flag.data(), W, H, (W - 1.0f) / W, 0.001f);
    oracle_sim_set_field(os, 0, sim.vx.data());
    oracle_sim_set_field(os, 1, sim.vx.back_data());
    oracle_sim_set_field(os, 2, sim.vy.data());
    oracle_sim_set_field(os, 3, sim.vy.back_data());
    const float dt = 0.002f;
    for (int i = 0; i < 2; i++) {
      sim.step(dt);
      oracle_sim_step(os, dt);
    }
    sim.syncToHost();
    std::vector<float> ovx((size_t)(W - 1) * H), op((size_t)W * H);
    oracle_sim_get_field(os, 0, ovx.data());
    oracle_sim_get_field(os, 8, op.data());
    double rv = rel(sim.vx.data(), ovx.data(), ovx.size());
    double rp = rel(sim.p.data(), op.data(), op.size());
    expect(rv < 1e-5, "Simulation::step vx matches the oracle", rv);
    expect(rp < 1e-5, "Simulation::step p matches the oracle", rp);
    // floating items through the registry forwarder, against the oracle on
    // the same fields (advect_floating_items.cpp:16-138)
    {
      MiniRegistry reg;
      std::vector<ubg_item> ref;
      for (int i = 0; i < 24; i++) {
        CoItem it{{0.009f, 0.0022f}, {(10.0f + 3.1f * i) / W, (8.0f + 1.3f * i) / W}, 0.3f * i};
        CoKinematics k{1.0f, {0.01f * (i % 5) - 0.02f, 0.005f * (i % 3)}, {0.0f, 0.0f}, 0.1f * i, 0.0f, 0};
        reg.items.push_back(it);
        reg.kins.push_back(k);
        ubg_item u{};
        u.size[0] = it.size.x;
        u.size[1] = it.size.y;
        u.pos[0] = it.pos.x;
        u.pos[1] = it.pos.y;
        u.rotation = it.rotation;
        u.mass = k.mass;
        u.vel[0] = k.vel.x;
        u.vel[1] = k.vel.y;
        u.ang_vel = k.angVel;
        ref.push_back(u);
      }
      for (int frame = 0; frame < 2; frame++) {
        sim.advectFloatingItemsView<CoItem, CoKinematics>(reg, 1.0f / 60, false);
        oracle_sim_advect_items(os, ref.data(), (int)ref.size(), 1.0f / 60, 0);
      }
      double worst = 0;
      int bumps = 0;
      for (size_t i = 0; i < ref.size(); i++) {
        worst = std::fmax(worst, std::fabs(reg.items[i].pos.x - ref[i].pos[0]));
        worst = std::fmax(worst, std::fabs(reg.items[i].pos.y - ref[i].pos[1]));
        worst = std::fmax(worst, std::fabs(reg.kins[i].vel.x - ref[i].vel[0]));
        worst = std::fmax(worst, std::fabs(reg.kins[i].vel.y - ref[i].vel[1]));
        bumps += reg.kins[i].bumpCount;
      }
      expect(worst < 1e-6, "advectFloatingItems (registry view) matches the oracle", worst);
      expect(bumps > 0, "items hit the disk or the walls", bumps);
    }
    // the game's per-frame flag update (ubootgl_app.cpp:111-112)
    sim.flag(10, 10) = 0.0f;
    sim.mg.updateFields(sim.flag);
    sim.setGrids(Simulation::Vec2{11.0f, 10.0f}, 0.0f);
    sim.syncToHost();
    expect(sim.flag(11, 10) == 0.0f && sim.vx.f(10, 10) == 0.0f, "setGrids / updateFields", 0.0);
    // the rest of the reference's public Simulation interface
    // (simulation.hpp:20-140), called as the game calls it
    {
      sim.syncToHost();
      oracle_sim_set_field(os, 0, sim.vx.data());
      oracle_sim_set_field(os, 2, sim.vy.data());
      oracle_sim_set_field(os, 8, sim.p.data());
      float dtg = sim.getDT(), dto = oracle_sim_get_dt(os);
      expect(dtg == dto, "getDT matches the oracle", std::fabs(dtg - dto));
      int bad = 0;
      for (int i = 0; i < 200; i++) {
        V2 pc{(i * 0.37f - 3.0f) / W, (i * 0.21f - 2.0f) / W};
        float fl = sim.psampleFlagLinear(pc), fn = sim.psampleFlagNearest(pc);
        V2 nrm = sim.psampleFlagNormal(pc);
        float on[2];
        oracle_flag_normal(sim.flag.data(), W, H, sim.pwidth, sim.h, pc.x, pc.y, on);
        bad += fl != oracle_flag_linear(sim.flag.data(), W, H, sim.pwidth, pc.x, pc.y);
        bad += fn != oracle_flag_nearest(sim.flag.data(), W, H, sim.pwidth, pc.x, pc.y);
        bad += nrm.x != on[0] || nrm.y != on[1];
        // ubootgl_app.cpp:144 / :166
        if (sim.psampleFlagLinear(pc) < 0.001) bad += fl >= 0.001f;
      }
      expect(bad == 0, "psampleFlagLinear / Nearest / Normal match the oracle", bad);
      sim.centerP();
      oracle_sim_center_p(os);
      sim.syncToHost();
      std::vector<float> op2((size_t)W * H);
      oracle_sim_get_field(os, 8, op2.data());
      double rc = rel(sim.p.data(), op2.data(), op2.size());
      expect(rc < 1e-6, "centerP matches the oracle", rc);
      expect(sim.singlePBC(Simulation::BC::OUTFLOW_ZERO_PRESSURE, 2.0f) == -2.0f &&
                 sim.singlePBC(Simulation::BC::NOSLIP, 2.0f) == 2.0f &&
                 sim.VBCPar(Simulation::BC::OUTFLOW, -1.0f, 5.0f) == 0.0f &&
                 sim.VBCPar(Simulation::BC::INFLOW, -1.0f, 5.0f) == 5.0f &&
                 sim.VBCPer(Simulation::BC::NOSLIP, 3.0f, 5.0f) == -3.0f,
             "singlePBC / VBCPar / VBCPer", 0.0);
      sim.sinks.push_back(V3{0.4f, 0.25f, 120.0f}); // explosion.cpp:33
      sim.project();
      expect(sim.sinks.size() == 1 && sim.sinks[0].z < 120.0f, "sinks list (glm::vec3 push)",
             sim.sinks.empty() ? 0.0 : sim.sinks[0].z);
      sim.sinks.clear();
      // advect_floating_items.cpp:16, :148 with the reference's signatures:
      // the same items as the explicit view calls (the item physics reads
      // only the fluid fields, which the calls do not change)
      MiniRegistry reg, reg2;
      MiniRegistryT<CoKinematicsSimple> regs, regs2;
      for (int i = 0; i < 8; i++) {
        reg.items.push_back(CoItem{{0.01f, 0.003f}, {(20.0f + 5 * i) / W, 20.0f / W}, 0.0f});
        reg.kins.push_back(CoKinematics{1.0f, {0.01f, -0.004f * i}, {0.0f, 0.0f}, 0.1f, 0.0f, 0});
        regs.items.push_back(reg.items.back());
        regs.kins.push_back(
            CoKinematicsSimple{1.0f, {0.01f, -0.004f * i}, {0.0f, 0.0f}, 0.1f, 0.0f, 0});
      }
      reg2 = reg;
      regs2 = regs;
      sim.advectFloatingItems(reg, 1.0f / 60);
      sim.advectFloatingItemsView<CoItem, CoKinematics>(reg2, 1.0f / 60, false);
      sim.advectFloatingItemsSimple(regs, 1.0f / 60);
      sim.advectFloatingItemsView<CoItem, CoKinematicsSimple>(regs2, 1.0f / 60, true);
      int same = 0;
      for (int i = 0; i < 8; i++) {
        same += reg.items[i].pos.x == reg2.items[i].pos.x && reg.items[i].pos.y == reg2.items[i].pos.y;
        same += regs.items[i].pos.x == regs2.items[i].pos.x && regs.items[i].pos.y == regs2.items[i].pos.y;
      }
      expect(same == 16 && reg.items[0].pos.x != 20.0f / W,
             "advectFloatingItems / Simple (registry)", same);
      float u = sim.disx(sim.gen), v = sim.disy(sim.gen);
      expect(u >= 0.0f && u <= 1.0f && v >= 0.0f && v <= (float)H / W, "gen / disx / disy", u);
      float *fp = sim.getFlag(), *pp = sim.getP(), *rp = sim.getR();
      expect(fp == sim.flag.data() && pp == sim.p.data() && rp == sim.r.data(),
             "getFlag / getP / getR", 0.0);
      sim.applyAccumulatedVelocity();
      sim.diffuse();
      sim.advect();
      sim.setVBCs();
      sim.setPBC();
      sim.saveCurrentVelocityFields();
    }
    oracle_sim_destroy(os);
  }
  // 3) Simulation(pwidth, mu, width, height) (simulation.hpp:20-30) and the
  //    free gs (pressure_solver.hpp:3-4)
  {
    const int W = 40, H = 24;
    Simulation blank((W - 1.0f) / W, 0.001f, W, H);
    blank.syncToHost();
    double s = 0;
    for (int i = 0; i < (W - 1) * H; i++) s += std::fabs(blank.vx.data()[i]);
    expect(s == 0.0 && blank.flag(3, 3) == 0.0f, "Simulation(pwidth, mu, W, H) starts zero", s);
    Single2DGrid p(W, H), f(W, H), flag(W, H);
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        flag(x, y) = (x == 0 || y == 0 || x == W - 1 || y == H - 1 || (x * 7 + y * 3) % 11 == 0)
                         ? 0.0f
                         : 1.0f;
        p(x, y) = std::sin(0.3f * x) + std::cos(0.2f * y);
        f(x, y) = 0.01f * ((x + 2 * y) % 5);
      }
    std::vector<float> po(p.data(), p.data() + W * H);
    gs(p, f, flag, 1.0f / W, 1.0f);
    oracle_gs_f32(po.data(), f.data(), flag.data(), W, H, 1.0f / W, 1.0f);
    expect(rel(p.data(), po.data(), po.size()) == 0.0, "gs matches the oracle bit for bit",
           rel(p.data(), po.data(), po.size()));
  }
  // 4) SimulationBatch: three simulations of different sizes stepped with
  //    their pressure solves batched, against the same three stepped alone
  {
    const int Ws[3] = {64, 96, 130}, Hs[3] = {40, 48, 70};
    std::vector<std::unique_ptr<Simulation>> alone, batched;
    for (int k = 0; k < 3; k++) {
      const int W = Ws[k], H = Hs[k];
      Single2DGrid flag(W, H);
      for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
          bool wall = x == 0 || y == 0 || x == W - 1 || y == H - 1;
          bool disk = (x - W / 3) * (x - W / 3) + (y - H / 2) * (y - H / 2) <= 25;
          flag(x, y) = (wall || disk) ? 0.0f : 1.0f;
        }
      for (auto *set : {&alone, &batched}) {
        set->emplace_back(new Simulation(flag, (W - 1.0f) / W, 0.001f));
        Simulation &s = *set->back();
        for (int y = 0; y < H; y++)
          for (int x = 1; x < W - 2; x++)
            s.vx.f(x, y) = s.vx.b(x, y) = std::sin(0.2f * x + k) * flag(x, y) * flag(x + 1, y);
        s.syncToDevice();
      }
    }
    alone[1]->sinks.push_back(Simulation::Vec3{0.3f, 0.2f, 5.0f});
    batched[1]->sinks.push_back(Simulation::Vec3{0.3f, 0.2f, 5.0f});
    SimulationBatch sb({batched[0].get(), batched[1].get(), batched[2].get()});
    const std::vector<float> dts = {0.002f, 0.0015f, 0.001f};
    for (int step = 0; step < 3; step++) {
      for (int k = 0; k < 3; k++) alone[k]->step(dts[k]);
      sb.step(dts);
    }
    double worst = 0;
    for (int k = 0; k < 3; k++) {
      alone[k]->syncToHost();
      batched[k]->syncToHost();
      const size_t nvx = (size_t)(Ws[k] - 1) * Hs[k], np = (size_t)Ws[k] * Hs[k];
      worst = std::fmax(worst, rel(batched[k]->vx.data(), alone[k]->vx.data(), nvx));
      worst = std::fmax(worst, rel(batched[k]->p.data(), alone[k]->p.data(), np));
    }
    expect(worst == 0.0, "SimulationBatch steps equal single steps bit for bit", worst);
    expect(batched[1]->sinks.size() == alone[1]->sinks.size(), "batched steps keep the sinks",
           (double)batched[1]->sinks.size());
  }
  std::printf("%s\n", fails ? "FAILED" : "ALL OK");
  return fails ? 1 : 0;
}
