// Drop-in test: code written against the reference's Simulation / MG API
// (simulation.hpp, pressure_solver.hpp, db2dgrid.hpp) compiled unchanged
// against include/ubg/*.hpp and run on the GPU through libubg.so, checked
// against the CPU oracle (test infrastructure, linked only here).
#include "ubg/simulation.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
void *oracle_sim_create(const float *flag, int W, int H, float pwidth, float mu);
void oracle_sim_destroy(void *s);
int oracle_sim_set_field(void *s, int id, const float *src);
int oracle_sim_get_field(void *s, int id, float *dst);
void oracle_sim_step(void *s, float dt);
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
