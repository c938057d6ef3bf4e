// sanitize_main.cpp -- TEST INFRASTRUCTURE: a host program that drives every
// entry point of the CPU oracle (ubg_oracle.cpp, compiled into it) on small
// seeded grids, built with AddressSanitizer and UndefinedBehaviorSanitizer
// (oracle/Makefile, target `sanitize`).  A clean run means no out-of-bounds
// access, use after free, overflow or other undefined behaviour on these
// paths; tests/test_oracle.py runs it.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
int oracle_set_threads(int n);
int oracle_mg_level_sizes(int W, int H, int *ws, int *hs, int cap);
void oracle_coarsen_flag(const float *f, int W, float *fc, int Wc, int Hc);
void oracle_rbgs_f32(float *p, const float *f, const float *fl, int W, int H, float h, float a);
double oracle_residual_f32(const float *p, const float *f, const float *fl, float *r, int W,
                           int H, float h);
double oracle_residual_f64(const double *p, const double *f, const double *fl, double *r, int W,
                           int H, double h);
void oracle_restrict_f32(const float *r, int W, float *rc, int Wc, int Hc);
void oracle_prolongate_f32(float *e, int W, int H, const float *ec, const float *flc, int Wc,
                           const float *fl);
void oracle_correct_f32(float *p, const float *e, int W, int H);
void oracle_zero_gradient_f32(float *p, int W, int H);
float oracle_bicubic(const float *g, int gw, int gh, float cx, float cy);
float oracle_bilinear(const float *g, int gw, int gh, float cx, float cy);
void *oracle_mg_create(int W, int H, int fp64);
void oracle_mg_destroy(void *m, int fp64);
void oracle_mg_update_fields(void *m, const float *flag);
void oracle_mg64_update_fields(void *m, const double *flag);
void oracle_mg_get_flag(void *m, int level, float *out);
void oracle_mg_solve(void *m, float *p, const float *f, const float *flag, float h, int zg);
void oracle_mg64_solve(void *m, double *p, const double *f, const double *flag, double h, int zg);
void *oracle_sim_create(const float *flag, int W, int H, float pwidth, float mu);
void oracle_sim_destroy(void *s);
int oracle_sim_set_field(void *s, int id, const float *src);
int oracle_sim_get_field(void *s, int id, float *dst);
float oracle_sim_h(void *s);
void oracle_sim_add_sink(void *s, float x, float y, float z);
void oracle_sim_step(void *s, float dt);
void oracle_sim_set_bcs(void *s, int west, int east, int north, int south);
int oracle_sim_stage(void *s, int stage, float dt);
void oracle_sim_bilinear_vel(void *s, const float *c, float *out, int n);
float oracle_flag_linear(const float *F, int W, int H, float pwidth, float px, float py);
float oracle_flag_nearest(const float *F, int W, int H, float pwidth, float px, float py);
void oracle_flag_normal(const float *F, int W, int H, float pwidth, float h, float px, float py,
                        float *out);
void oracle_gs_f32(float *p, const float *f, const float *fl, int W, int H, float h, float a);
void oracle_sim_center_p(void *s);
float oracle_sim_get_dt(void *s);
void oracle_sim_advect_tracers(void *s, float *xy, float *age, int n, float dt, uint32_t seed);
void oracle_sim_advect_items(void *s, void *items, int n, float gameDT, int simple);
}

struct Item { // the ubg_item layout
  float size[2], pos[2], rotation, mass, vel[2], force[2], ang_vel, ang_force;
  int bump_count;
};

static uint32_t st = 12345u;
static float rnd() { // a small LCG in [0, 1)
  st = 1664525u * st + 1013904223u;
  return (st >> 8) * (1.0f / 16777216.0f);
}

int main() {
  oracle_set_threads(2);
  for (int shape = 0; shape < 3; shape++) {
    const int W = shape == 0 ? 64 : (shape == 1 ? 37 : 8), H = shape == 0 ? 40 : (shape == 1 ? 23 : 8);
    std::vector<float> flag((size_t)W * H), vx((size_t)(W - 1) * H), vy((size_t)W * (H - 1));
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++)
        flag[(size_t)y * W + x] =
            (x == 0 || y == 0 || x == W - 1 || y == H - 1 || rnd() < 0.1f) ? 0.0f : 1.0f;
    for (auto &v : vx) v = rnd() - 0.5f;
    for (auto &v : vy) v = rnd() - 0.5f;
    // MG primitives and both solvers
    int ws[32], hs[32];
    const int nl = oracle_mg_level_sizes(W, H, ws, hs, 32);
    std::vector<float> p((size_t)W * H), f((size_t)W * H), r((size_t)W * H), e((size_t)W * H);
    for (auto &v : f) v = rnd() - 0.5f;
    const float h = 1.0f / W;
    oracle_rbgs_f32(p.data(), f.data(), flag.data(), W, H, h, 1.0f);
    oracle_gs_f32(p.data(), f.data(), flag.data(), W, H, h, 0.9f);
    oracle_residual_f32(p.data(), f.data(), flag.data(), r.data(), W, H, h);
    if (nl > 1) {
      std::vector<float> rc((size_t)ws[1] * hs[1]), fc((size_t)ws[1] * hs[1]);
      oracle_restrict_f32(r.data(), W, rc.data(), ws[1], hs[1]);
      oracle_coarsen_flag(flag.data(), W, fc.data(), ws[1], hs[1]);
      oracle_prolongate_f32(e.data(), W, H, rc.data(), fc.data(), ws[1], flag.data());
      oracle_correct_f32(p.data(), e.data(), W, H);
    }
    oracle_zero_gradient_f32(p.data(), W, H);
    void *m = oracle_mg_create(W, H, 0);
    oracle_mg_update_fields(m, flag.data());
    std::vector<float> lf((size_t)W * H);
    for (int l = 0; l < nl; l++) oracle_mg_get_flag(m, l, lf.data());
    oracle_mg_solve(m, p.data(), f.data(), flag.data(), h, 1);
    oracle_mg_destroy(m, 0);
    std::vector<double> p64((size_t)W * H), f64(f.begin(), f.end()), fl64(flag.begin(), flag.end()),
        r64((size_t)W * H);
    void *m64 = oracle_mg_create(W, H, 1);
    oracle_mg64_update_fields(m64, fl64.data());
    oracle_mg64_solve(m64, p64.data(), f64.data(), fl64.data(), h, 0);
    oracle_residual_f64(p64.data(), f64.data(), fl64.data(), r64.data(), W, H, h);
    oracle_mg_destroy(m64, 1);
    // samplers, including clamped and far-outside positions
    for (int i = 0; i < 200; i++) {
      const float cx = (rnd() * 1.4f - 0.2f) * W, cy = (rnd() * 1.4f - 0.2f) * H;
      float nrm[2];
      oracle_bicubic(vx.data(), W - 1, H, cx, cy);
      oracle_bilinear(vy.data(), W, H - 1, cx, cy);
      oracle_flag_linear(flag.data(), W, H, 1.0f, cx / W, cy / W);
      oracle_flag_nearest(flag.data(), W, H, 1.0f, cx / W, cy / W);
      oracle_flag_normal(flag.data(), W, H, 1.0f, h, cx / W, cy / W, nrm);
    }
    // the simulation: steps with sinks and every stage, tracers, items
    void *s = oracle_sim_create(flag.data(), W, H, (W - 1.0f) / W, 0.001f);
    oracle_sim_set_field(s, 0, vx.data());
    oracle_sim_set_field(s, 1, vx.data());
    oracle_sim_set_field(s, 2, vy.data());
    oracle_sim_set_field(s, 3, vy.data());
    const float hs_ = oracle_sim_h(s);
    oracle_sim_add_sink(s, W / 2 * hs_, H / 2 * hs_, 5.0f);
    oracle_sim_add_sink(s, (W / 2 + 1) * hs_, H / 2 * hs_, 7.0f);
    for (int k = 0; k < 3; k++) oracle_sim_step(s, 0.2f * hs_);
    oracle_sim_set_bcs(s, 1, 2, 3, 0);
    for (int stg = 0; stg < 10; stg++) oracle_sim_stage(s, stg, 0.2f * hs_);
    oracle_sim_center_p(s);
    oracle_sim_get_dt(s);
    std::vector<float> c(200), out(200), buf((size_t)W * H);
    for (auto &v : c) v = rnd() * W;
    oracle_sim_bilinear_vel(s, c.data(), out.data(), 100);
    for (int id = 0; id < 12; id++) oracle_sim_get_field(s, id, buf.data());
    std::vector<float> xy(2000), age(1000);
    for (int i = 0; i < 1000; i++) {
      xy[2 * i] = rnd() * 1.2f - 0.1f;
      xy[2 * i + 1] = (rnd() * 1.2f - 0.1f) * H / W;
      age[i] = rnd() * 7.0f;
    }
    oracle_sim_advect_tracers(s, xy.data(), age.data(), 1000, 0.01f, 3u);
    std::vector<Item> items(20);
    for (auto &it : items) {
      it = Item{{0.02f, 0.005f}, {rnd(), rnd() * H / W}, rnd() * 6.0f, 1.0f,
                {rnd() - 0.5f, rnd() - 0.5f}, {0.0f, 0.0f}, rnd(), 0.0f, 0};
    }
    oracle_sim_advect_items(s, items.data(), (int)items.size(), 1.0f / 60, 0);
    oracle_sim_advect_items(s, items.data(), (int)items.size(), 1.0f / 60, 1);
    oracle_sim_destroy(s);
  }
  std::printf("sanitize ok\n");
  return 0;
}
