// Drop-in test: code written against the reference's Simulation / MG API
// (simulation.hpp, pressure_solver.hpp, db2dgrid.hpp) compiled unchanged
// against include/ubg/*.hpp and run on the GPU through libubg.so, checked
// against the CPU oracle (test infrastructure, linked only here).
#include "ubg/simulation.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

extern "C" {
void *oracle_sim_create(const float *flag, int W, int H, float pwidth, float mu);
void oracle_sim_destroy(void *s);
int oracle_sim_set_field(void *s, int id, const float *src);
int oracle_sim_get_field(void *s, int id, float *dst);
void oracle_sim_step(void *s, float dt);
void oracle_sim_advect_items(void *s, void *items, int n, float gameDT, int simple);
void *oracle_mg_create(int W, int H, int fp64);
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
struct MiniRegistry {
  std::vector<CoItem> items;
  std::vector<CoKinematics> kins;
  struct View {
    MiniRegistry *r;
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
    oracle_sim_destroy(os);
  }
  std::printf("%s\n", fails ? "FAILED" : "ALL OK");
  return fails ? 1 : 0;
}
