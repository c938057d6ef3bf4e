// ubg_oracle.cpp -- CPU restatement of ubootgl's fluid step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: a clean-room, plain
// C++17/OpenMP restatement of the reference algorithm (te42kyfo/ubootgl), used
// to check the HIP implementation and, in bench.py, timed as the host CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  The product library (ubootgl_amd/libubg.so) never links or
// calls it.
//
// Every function cites the reference file:line it follows.  Loop bounds,
// evaluation order and the boundary-condition order are kept; the rbgs
// smoother always uses the reference's two-pass red/black semantics
// (pressure_solver.cpp:65-71), never its thread-count dependent wavefront
// branch (pressure_solver.cpp:72-87).
//
// Pinning (tests/test_oracle.py):
//   * MG::updateFields            closed-form integer rule, brute force
//   * rbgs / smoothingKernel      closed form per point, fixed point of an exact
//                                 discrete solution, colour-order independence
//   * residual/restrict/prolong   closed forms, linearity, constant fields
//   * MG V-cycle                  mgtest harmonic KAT (mgtest/mgtest.cpp:10-53),
//                                 fp64 direct solve on tiny grids
//   * bicubicSample               Catmull-Rom reproduces quadratics, node
//                                 interpolation, clamp rule
//   * advect                      zero velocity = identity, uniform shift
//   * diffuse                     mu=0 => masked identity, closed form
//   * project                     discrete identity div(v_new) = div(v) - L p
//   * tracers                     bilinear exact on linear fields; Wang hash and
//                                 the Numerical Recipes LCG as integer closed forms
//   * gs, centerP, getDT, flag    closed forms (lexicographic order on a point
//     samplers                    source, the W*H divisor, signed maximum,
//                                 linear ramps and the clamp / fract rules)
// Parity unpinned: the composition of advect's quirks under the reference's
// own -Ofast build (the reference cannot be compiled here: it needs glm and
// GLEW headers the image lacks), and glm::mix's exact rounding (x*(1-a)+y*a is
// assumed) inside psampleFlagLinear.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <vector>
#include <omp.h>
#if defined(UBG_ORACLE_TIMING) && defined(__AVX2__)
#include <immintrin.h>
#endif

#define ORACLE_API extern "C" __attribute__((visibility("default")))

namespace {

// ---------------------------------------------------------------------------
// Multigrid primitives, templated over the storage type (float = reference,
// double = fp64 oracle for the mixed-precision extension).
// ---------------------------------------------------------------------------

// pressure_solver.cpp:10-24 (smoothingKernel)
template <typename T>
inline T smooth_point(const T *p, const T *f, const T *fl, int W, int x, int y,
                      T h, T alpha) {
  const size_t i = (size_t)y * W + x;
  T sum = T(0);
  sum += fl[i - 1] + fl[i + 1] + fl[i + W] + fl[i - W];
  T val = T(0);
  val += p[i - 1] * fl[i - 1];
  val += p[i + 1] * fl[i + 1];
  val += p[i - W] * fl[i - W];
  val += p[i + W] * fl[i + W];
  val += f[i] * h * h;
  val = (sum == T(0)) ? T(0) : val / sum;
  return fl[i] * (alpha * val + (T(1) - alpha) * p[i]);
}

// pressure_solver.cpp:35-47 (red = (x+y) odd) and :65-71 (two passes)
template <typename T>
void rbgs(T *p, const T *f, const T *fl, int W, int H, T h, T alpha) {
  const bool par = H > 64;
#pragma omp parallel for schedule(static) if (par)
  for (int y = 1; y < H - 1; y++)
    for (int x = 1 + (y % 2); x < W - 1; x += 2)
      p[(size_t)y * W + x] = smooth_point(p, f, fl, W, x, y, h, alpha);
#pragma omp parallel for schedule(static) if (par)
  for (int y = 1; y < H - 1; y++)
    for (int x = 1 + ((y + 1) % 2); x < W - 1; x += 2)
      p[(size_t)y * W + x] = smooth_point(p, f, fl, W, x, y, h, alpha);
}

// pressure_solver.cpp:26-33 (gs): one lexicographic Gauss-Seidel sweep
template <typename T>
void gs(T *p, const T *f, const T *fl, int W, int H, T h, T alpha) {
  for (int y = 1; y < H - 1; y++)
    for (int x = 1; x < W - 1; x++) p[(size_t)y * W + x] = smooth_point(p, f, fl, W, x, y, h, alpha);
}

// pressure_solver.cpp:91-116 (calculateResidualField); the norm is summed in
// double (the reference's float OpenMP reduction has no fixed order).
template <typename T>
double residual(const T *p, const T *f, const T *fl, T *r, int W, int H, T h) {
  const T ihsq = T(1) / h / h;
  double l2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : l2) if (H > 64)
  for (int y = 1; y < H - 1; y++) {
    for (int x = 1; x < W - 1; x++) {
      const size_t i = (size_t)y * W + x;
      const T pc = p[i];
      T val = T(0);
      val += p[i - 1] * fl[i - 1] + pc * (T(1) - fl[i - 1]);
      val += p[i + 1] * fl[i + 1] + pc * (T(1) - fl[i + 1]);
      val += p[i - W] * fl[i - W] + pc * (T(1) - fl[i - W]);
      val += p[i + W] * fl[i + W] + pc * (T(1) - fl[i + W]);
      val += T(-4) * pc;
      val *= ihsq;
      r[i] = (f[i] + val) * fl[i];
      l2 += (double)r[i] * (double)r[i];
    }
  }
  return std::sqrt(l2);
}

// pressure_solver.cpp:118-132 (restrict: full weighting at (2x,2y))
template <typename T>
void restrict_fw(const T *r, int W, T *rc, int Wc, int Hc) {
#pragma omp parallel for schedule(static) if (Hc > 64)
  for (int y = 1; y < Hc - 1; y++) {
    for (int x = 1; x < Wc - 1; x++) {
      auto R = [&](int xx, int yy) { return r[(size_t)yy * W + xx]; };
      T v = T(0);
      v = R(2 * x - 1, 2 * y - 1) * T(1) + R(2 * x, 2 * y - 1) * T(2) +
          R(2 * x + 1, 2 * y - 1) * T(1);
      v += R(2 * x - 1, 2 * y) * T(2) + R(2 * x, 2 * y) * T(4) +
           R(2 * x + 1, 2 * y) * T(2);
      v += R(2 * x - 1, 2 * y + 1) * T(1) + R(2 * x, 2 * y + 1) * T(2) +
           R(2 * x + 1, 2 * y + 1) * T(1);
      rc[(size_t)y * Wc + x] = v * (T(1) / T(16));
    }
  }
}

// The prolongation denominator adds the double literal 0.0001 to a sum of
// flags and stores the result in the grid's type (pressure_solver.cpp:149).
template <typename T> inline T eps_sum(T s) { return (T)((double)s + 0.0001); }

// pressure_solver.cpp:134-172 (prolongate: e = 0, then four parity classes)
template <typename T>
void prolongate(T *e, int W, int H, const T *ec, const T *flc, int Wc,
                const T *fl) {
  std::fill(e, e + (size_t)W * H, T(0));
  auto EC = [&](int x, int y) { return ec[(size_t)y * Wc + x]; };
  auto FC = [&](int x, int y) { return flc[(size_t)y * Wc + x]; };
  const bool par = H > 64;
#pragma omp parallel for schedule(static) if (par)
  for (int y = 2; y < H - 1; y += 2)
    for (int x = 2; x < W - 1; x += 2)
      e[(size_t)y * W + x] = EC(x / 2, y / 2) * fl[(size_t)y * W + x];
#pragma omp parallel for schedule(static) if (par)
  for (int y = 2; y < H - 1; y += 2)
    for (int x = 1; x < W - 2; x += 2) {
      T sum = eps_sum<T>(FC(x / 2, y / 2) + FC(x / 2 + 1, y / 2));
      e[(size_t)y * W + x] =
          fl[(size_t)y * W + x] * (EC(x / 2, y / 2) + EC(x / 2 + 1, y / 2)) /
          sum;
    }
#pragma omp parallel for schedule(static) if (par)
  for (int y = 1; y < H - 2; y += 2)
    for (int x = 2; x < W - 1; x += 2) {
      T sum = eps_sum<T>(FC(x / 2, y / 2) + FC(x / 2, y / 2 + 1));
      e[(size_t)y * W + x] =
          fl[(size_t)y * W + x] * (EC(x / 2, y / 2) + EC(x / 2, y / 2 + 1)) /
          sum;
    }
#pragma omp parallel for schedule(static) if (par)
  for (int y = 1; y < H - 2; y += 2)
    for (int x = 1; x < W - 2; x += 2) {
      T sum = eps_sum<T>(FC(x / 2, y / 2) + FC(x / 2 + 1, y / 2 + 1) +
                         FC(x / 2 + 1, y / 2) + FC(x / 2, y / 2 + 1));
      e[(size_t)y * W + x] =
          fl[(size_t)y * W + x] *
          (EC(x / 2, y / 2) + EC(x / 2 + 1, y / 2 + 1) + EC(x / 2 + 1, y / 2) +
           EC(x / 2, y / 2 + 1)) /
          sum;
    }
}

// pressure_solver.cpp:174-181 (correct)
template <typename T> void correct(T *p, const T *e, int W, int H) {
#pragma omp parallel for schedule(static) if (H > 64)
  for (int y = 1; y < H - 1; y++)
    for (int x = 1; x < W - 1; x++) p[(size_t)y * W + x] += e[(size_t)y * W + x];
}

// pressure_solver.cpp:183-192 (setZeroGradientBC; corners untouched)
template <typename T> void zero_gradient(T *p, int W, int H) {
  for (int y = 1; y < H - 1; y++) {
    p[(size_t)y * W] = p[(size_t)y * W + 1];
    p[(size_t)y * W + W - 1] = p[(size_t)y * W + W - 2];
  }
  for (int x = 1; x < W - 1; x++) {
    p[x] = p[(size_t)W + x];
    p[(size_t)(H - 1) * W + x] = p[(size_t)(H - 2) * W + x];
  }
}

// pressure_solver.hpp:16-31: level sizes halve (integer) while both > 3.
void level_sizes(int W, int H, std::vector<int> &ws, std::vector<int> &hs) {
  ws.clear();
  hs.clear();
  int cw = W, ch = H;
  while (cw > 3 && ch > 3) {
    ws.push_back(cw);
    hs.push_back(ch);
    cw /= 2;
    ch /= 2;
  }
}

// pressure_solver.hpp:34-57 (updateFields): coarse flag is 1 iff the
// full-weighted fine-flag sum / 16 exceeds 0.2; the coarse boundary is 1.
template <typename T>
void coarsen_flag(const T *f, int W, T *fc, int Wc, int Hc) {
  std::fill(fc, fc + (size_t)Wc * Hc, T(1));
  for (int y = 1; y < Hc - 1; y++) {
    for (int x = 1; x < Wc - 1; x++) {
      auto F = [&](int xx, int yy) { return (float)f[(size_t)yy * W + xx]; };
      float v = 0.0f;
      v = F(2 * x - 1, 2 * y - 1) * 1 + F(2 * x, 2 * y - 1) * 2 +
          F(2 * x + 1, 2 * y - 1) * 1;
      v += F(2 * x - 1, 2 * y) * 2 + F(2 * x, 2 * y) * 4 +
           F(2 * x + 1, 2 * y) * 2;
      v += F(2 * x - 1, 2 * y + 1) * 1 + F(2 * x, 2 * y + 1) * 2 +
           F(2 * x + 1, 2 * y + 1) * 1;
      v *= 1.0f / 16.0f;
      fc[(size_t)y * Wc + x] = (v > 0.2) ? T(1) : T(0);
    }
  }
}

struct MGParams {
  int nu_pre = 3;    // pressure_solver.cpp:211
  int nu_post = 3;   // pressure_solver.cpp:243
  int nu_coarse = 5; // pressure_solver.cpp:204
  float alpha = 1.0f;
};

// pressure_solver.hpp:13-76 + pressure_solver.cpp:201-248
template <typename T> struct MG {
  int W, H, levels;
  std::vector<int> ws, hs;
  std::vector<std::vector<T>> rs, rcs, ecs, es, flagcs;
  MGParams prm;

  MG(int W_, int H_) : W(W_), H(H_) {
    level_sizes(W, H, ws, hs);
    levels = (int)ws.size();
    for (int l = 0; l < levels; l++) {
      size_t n = (size_t)ws[l] * hs[l];
      rs.emplace_back(n, T(0));
      rcs.emplace_back(n, T(0));
      ecs.emplace_back(n, T(0));
      es.emplace_back(n, T(0));
      flagcs.emplace_back(n, T(1));
    }
  }

  void update_fields(const T *flag) {
    std::copy(flag, flag + (size_t)W * H, flagcs[0].begin());
    for (int l = 1; l < levels; l++)
      coarsen_flag(flagcs[l - 1].data(), ws[l - 1], flagcs[l].data(), ws[l],
                   hs[l]);
  }

  void solve_level(T *p, const T *f, const T *flag, T h, int level, bool zg) {
    const int Wl = ws[level], Hl = hs[level];
    const T alpha = (T)prm.alpha;
    if (level == levels - 2) {
      for (int i = 0; i < prm.nu_coarse; i++) rbgs(p, f, flag, Wl, Hl, h, alpha);
      return;
    }
    for (int i = 0; i < prm.nu_pre; i++) {
      rbgs(p, f, flag, Wl, Hl, h, alpha);
      if (level == 0 && zg) zero_gradient(p, Wl, Hl);
    }
    T *r = rs[level].data();
    std::fill(rs[level].begin(), rs[level].end(), T(0));
    residual(p, f, flag, r, Wl, Hl, h);
    const int Wc = ws[level + 1], Hc = hs[level + 1];
    T *rc = rcs[level + 1].data();
    std::fill(rcs[level + 1].begin(), rcs[level + 1].end(), T(0));
    restrict_fw(r, Wl, rc, Wc, Hc);
    T *ec = ecs[level + 1].data();
    std::fill(ecs[level + 1].begin(), ecs[level + 1].end(), T(0));
    const T *flagc = flagcs[level + 1].data();
    // pressure_solver.cpp:229: h_c = h * (W_l - 1.0f) / (W_{l+1} - 1.0f)
    T hc = h * (T)(Wl - 1.0f) / (T)(Wc - 1.0f);
    solve_level(ec, rc, flagc, hc, level + 1, false);
    T *e = es[level].data();
    prolongate(e, Wl, Hl, ec, flagc, Wc, flag);
    correct(p, e, Wl, Hl);
    if (level == 0 && zg) zero_gradient(p, Wl, Hl);
    for (int i = 0; i < prm.nu_post; i++) {
      rbgs(p, f, flag, Wl, Hl, h, alpha);
      if (level == 0 && zg) zero_gradient(p, Wl, Hl);
    }
  }

  void solve(T *p, const T *f, const T *flag, T h, bool zg) {
    solve_level(p, f, flag, h, 0, zg);
  }
};

// ---------------------------------------------------------------------------
// Interpolators (interpolators.hpp)
// ---------------------------------------------------------------------------

// interpolators.hpp:78-85 (CubicHermite, Catmull-Rom)
inline float cubic_hermite(float t, float A, float B, float C, float D) {
  float a = -A / 2.0f + (3.0f * B) / 2.0f - (3.0f * C) / 2.0f + D / 2.0f;
  float b = A - (5.0f * B) / 2.0f + 2.0f * C - D / 2.0f;
  float c = -A / 2.0f + C / 2.0f;
  float d = B;
  return a * t * t * t + b * t * t + c * t + d;
}

// interpolators.hpp:92-206 (bicubicSample, one lane): clamp to [3, n-3],
// truncate, four vertical Catmull-Rom passes, then one horizontal pass.
inline float bicubic(const float *g, int gw, int gh, float cx, float cy) {
  cx = std::max(std::min(cx, gw - 3.0f), 3.0f);
  cy = std::max(std::min(cy, gh - 3.0f), 3.0f);
  int icx = (int)cx, icy = (int)cy;
  float stx = cx - std::trunc(cx);
  float sty = cy - std::trunc(cy);
  const float *base = g + (std::ptrdiff_t)(icy - 1) * gw + (icx - 1);
  float col[4];
  for (int m = 0; m < 4; m++)
    col[m] = cubic_hermite(sty, base[m], base[gw + m], base[2 * gw + m],
                           base[3 * gw + m]);
  return cubic_hermite(stx, col[0], col[1], col[2], col[3]);
}

#if defined(UBG_ORACLE_TIMING) && defined(__AVX2__)
// TIMING BUILD ONLY (bench.py's CPU baseline, compiled with the reference's
// -Ofast -march=native, never used for parity): eight lanes of an advect
// chunk at once, as the reference's AVX2 advect processes them
// (simulation.cpp:241-354, interpolators.hpp:78-206).  Each bicubic sample
// gathers the four rows of one stencil column for all lanes, runs the
// vertical Catmull-Rom passes eight wide, then the horizontal one.
inline __m256 cubic_hermite8(__m256 t, __m256 A, __m256 B, __m256 C, __m256 D) {
  __m256 a = -A / 2.0f + (3.0f * B) / 2.0f - (3.0f * C) / 2.0f + D / 2.0f;
  __m256 b = A - (5.0f * B) / 2.0f + 2.0f * C - D / 2.0f;
  __m256 c = -A / 2.0f + C / 2.0f;
  return a * t * t * t + b * t * t + c * t + B;
}
inline __m256 bicubic8(const float *g, int gw, int gh, __m256 cx, __m256 cy) {
  cx = _mm256_max_ps(_mm256_min_ps(cx, _mm256_set1_ps(gw - 3.0f)), _mm256_set1_ps(3.0f));
  cy = _mm256_max_ps(_mm256_min_ps(cy, _mm256_set1_ps(gh - 3.0f)), _mm256_set1_ps(3.0f));
  const __m256i icx = _mm256_cvttps_epi32(cx), icy = _mm256_cvttps_epi32(cy);
  const __m256 stx = cx - _mm256_round_ps(cx, _MM_FROUND_TO_ZERO | _MM_FROUND_NO_EXC);
  const __m256 sty = cy - _mm256_round_ps(cy, _MM_FROUND_TO_ZERO | _MM_FROUND_NO_EXC);
  const __m256i one = _mm256_set1_epi32(1), vw = _mm256_set1_epi32(gw);
  const __m256i i0 = _mm256_add_epi32(_mm256_sub_epi32(icx, one),
                                      _mm256_mullo_epi32(_mm256_sub_epi32(icy, one), vw));
  __m256 col[4];
  for (int m = 0; m < 4; m++) {
    const __m256i r0 = _mm256_add_epi32(i0, _mm256_set1_epi32(m));
    const __m256i r1 = _mm256_add_epi32(r0, vw), r2 = _mm256_add_epi32(r1, vw),
                  r3 = _mm256_add_epi32(r2, vw);
    col[m] = cubic_hermite8(sty, _mm256_i32gather_ps(g, r0, 4), _mm256_i32gather_ps(g, r1, 4),
                            _mm256_i32gather_ps(g, r2, 4), _mm256_i32gather_ps(g, r3, 4));
  }
  return cubic_hermite8(stx, col[0], col[1], col[2], col[3]);
}
inline __m256 lanes_x(int x0) {
  return _mm256_add_ps(_mm256_set1_ps((float)x0), _mm256_setr_ps(0, 1, 2, 3, 4, 5, 6, 7));
}
// the vx chunk (IS_VY false) or vy chunk starting at x0 in row y
template <bool IS_VY>
inline void advect_chunk8(const float *xf, const float *yf, float *xb, float *yb, const float *F,
                          int W, int H, int x0, int y, float hdt, float fdt) {
  const int Wx = W - 1, Hy = H - 1;
  const __m256 two = _mm256_set1_ps(2.0f), half = _mm256_set1_ps(0.5f);
  const size_t fr = (size_t)y * W + x0;
  const __m256 fa = _mm256_loadu_ps(F + fr - (IS_VY ? (size_t)W : 1)), fb = _mm256_loadu_ps(F + fr);
  if (_mm256_movemask_ps(_mm256_cmp_ps(_mm256_add_ps(fa, fb), two, _CMP_EQ_OQ)) == 0) return;
  const __m256 vh = _mm256_set1_ps(hdt), vf = _mm256_set1_ps(fdt);
  if (!IS_VY) {
    const __m256 posx = lanes_x(x0) + half, posy = _mm256_set1_ps((float)y);
    const __m256 u = _mm256_loadu_ps(xf + (size_t)y * Wx + x0);
    const __m256 w = ((_mm256_loadu_ps(yf + (size_t)y * W + x0) +
                       _mm256_loadu_ps(yf + (size_t)(y - 1) * W + x0)) +
                      (_mm256_loadu_ps(yf + (size_t)y * W + x0 + 1) +
                       _mm256_loadu_ps(yf + (size_t)(y - 1) * W + x0 + 1))) * 0.25f;
    const __m256 mx = posx - u * vh, my = posy - w * vh;
    const __m256 u2 = bicubic8(xf, Wx, H, mx - half, my), w2 = bicubic8(yf, W, Hy, mx, my - half);
    const __m256 ex = posx - u2 * vf, ey = posy - w2 * vf;
    __m256 v = bicubic8(xf, Wx, H, ex - half, ey);
    v = v * _mm256_loadu_ps(F + fr) * _mm256_loadu_ps(F + fr + 1);
    _mm256_storeu_ps(xb + (size_t)Wx * y + x0, v);
  } else {
    const __m256 posx = lanes_x(x0), posy = _mm256_set1_ps((float)y + 0.5f);
    const __m256 w = _mm256_loadu_ps(yf + (size_t)y * W + x0);
    const __m256 u = ((_mm256_loadu_ps(xf + (size_t)y * Wx + x0) +
                       _mm256_loadu_ps(xf + (size_t)(y - 1) * Wx + x0)) +
                      (_mm256_loadu_ps(xf + (size_t)y * Wx + x0 + 1) +
                       _mm256_loadu_ps(xf + (size_t)(y - 1) * Wx + x0 + 1))) * 0.25f;
    const __m256 mx = posx - u * vh, my = posy - w * vh;
    const __m256 u2 = bicubic8(xf, Wx, H, mx - half, my), w2 = bicubic8(yf, W, Hy, mx, my - half);
    const __m256 ex = posx - u2 * vf, ey = posy - w2 * vf;
    __m256 v = bicubic8(yf, W, Hy, ex, ey - half);
    v = v * _mm256_loadu_ps(F + fr) * _mm256_loadu_ps(F + fr + W);
    _mm256_storeu_ps(yb + (size_t)W * y + x0, v);
  }
}
#define UBG_ORACLE_AVX2_ADVECT 1
#endif

// interpolators.hpp:11-27 (scalar bilinearSample with clamp to [0, n-1.1])
inline float bilinear(const float *g, int gw, int gh, float cx, float cy) {
  cx = std::min(std::max(cx, 0.0f), gw - 1.1f);
  cy = std::min(std::max(cy, 0.0f), gh - 1.1f);
  int ix = (int)cx, iy = (int)cy;
  float sx = cx - std::floor(cx), sy = cy - std::floor(cy);
  float v1 = g[(size_t)iy * gw + ix];
  float v2 = g[(size_t)iy * gw + ix + 1];
  float v3 = g[(size_t)(iy + 1) * gw + ix];
  float v4 = g[(size_t)(iy + 1) * gw + ix + 1];
  float vm1 = v1 + (v2 - v1) * sx;
  float vm2 = v3 + (v4 - v3) * sx;
  return vm1 + (vm2 - vm1) * sy;
}

// ---------------------------------------------------------------------------
// Simulation (simulation.hpp / simulation.cpp)
// ---------------------------------------------------------------------------

enum BC { INFLOW = 0, OUTFLOW = 1, OUTFLOW_ZERO_PRESSURE = 2, NOSLIP = 3 };

struct Sink {
  float x, y, z;
};

struct Sim {
  int W, H, Wx, Hy; // cells; vx is Wx x H, vy is W x Hy
  float pwidth, mu, h, dt = 0.0f;
  int bcW = INFLOW, bcE = OUTFLOW_ZERO_PRESSURE, bcN = NOSLIP, bcS = NOSLIP;
  std::vector<float> vx[2], vy[2];
  int fx = 0, fy = 0; // front index of vx and of vy (db2dgrid.hpp:52-64)
  std::vector<float> vx_accum, vy_accum, vx_cur, vy_cur, p, f, flag, r;
  MG<float> mg;
  std::vector<Sink> sinks;
  int vcycles = 2; // simulation.cpp:189-190

  float *vxf() { return vx[fx].data(); }
  float *vxb() { return vx[1 - fx].data(); }
  float *vyf() { return vy[fy].data(); }
  float *vyb() { return vy[1 - fy].data(); }

  // simulation.hpp:32-67
  Sim(const float *flagIn, int W_, int H_, float pw, float mu_)
      : W(W_), H(H_), Wx(W_ - 1), Hy(H_ - 1), pwidth(pw), mu(mu_),
        mg(W_, H_) {
    for (int b = 0; b < 2; b++) {
      vx[b].assign((size_t)Wx * H, 0.0f);
      vy[b].assign((size_t)W * Hy, 0.0f);
    }
    vx_accum.assign((size_t)Wx * H, 0.0f);
    vy_accum.assign((size_t)W * Hy, 0.0f);
    vx_cur.assign((size_t)Wx * H, 0.0f);
    vy_cur.assign((size_t)W * Hy, 0.0f);
    p.assign((size_t)W * H, 0.0f);
    f.assign((size_t)W * H, 0.0f);
    r.assign((size_t)W * H, 0.0f);
    flag.assign(flagIn, flagIn + (size_t)W * H);
    for (int x = 0; x < W; x++) vy[0][x] = vy[1][x] = 0.0f;
    for (int y = 0; y < H; y++) vx[0][(size_t)y * Wx] = vx[1][(size_t)y * Wx] = 1.0f;
    mg.update_fields(flag.data());
    h = pwidth / (W - 1.0f);
  }

  // simulation.cpp:23-34
  float singlePBC(int bc, float a) { return bc == OUTFLOW_ZERO_PRESSURE ? -a : a; }
  // simulation.cpp:36-45
  void setPBC() {
    float *P = p.data();
    for (int y = 0; y < H; y++) {
      P[(size_t)y * W] = singlePBC(bcW, P[(size_t)y * W + 1]);
      P[(size_t)y * W + W - 1] = singlePBC(bcE, P[(size_t)y * W + W - 2]);
    }
    for (int x = 0; x < W; x++) {
      P[x] = singlePBC(bcS, P[(size_t)W + x]);
      P[(size_t)(H - 1) * W + x] = singlePBC(bcN, P[(size_t)(H - 2) * W + x]);
    }
  }
  // simulation.cpp:50-78
  static float VBCPar(int bc, float a, float b) {
    if (bc == INFLOW) return b;
    if (bc == OUTFLOW || bc == OUTFLOW_ZERO_PRESSURE) return std::fmax(a, 0.0f);
    return 0.0f;
  }
  static float VBCPer(int bc, float a, float b) {
    if (bc == INFLOW) return b;
    if (bc == OUTFLOW || bc == OUTFLOW_ZERO_PRESSURE) return std::fmax(a, 0.0f);
    return -a;
  }
  // simulation.cpp:80-102: sequential order defines the corners.
  void setVBCs() {
    float *xf = vxf(), *xb = vxb(), *yf = vyf(), *yb = vyb();
    for (int y = 0; y < H; y++) {
      size_t r0 = (size_t)y * Wx;
      xf[r0] = xb[r0] = VBCPar(bcW, xf[r0 + 1], xf[r0]);
      xf[r0 + Wx - 1] = xb[r0 + Wx - 1] = VBCPar(bcE, xf[r0 + Wx - 2], xf[r0 + Wx - 1]);
    }
    for (int x = 0; x < Wx; x++) {
      size_t top = (size_t)(H - 1) * Wx + x, top1 = (size_t)(H - 2) * Wx + x;
      xf[x] = xb[x] = VBCPer(bcS, xf[(size_t)Wx + x], xf[x]);
      xf[top] = xb[top] = VBCPer(bcN, xf[top1], xf[top]);
    }
    for (int y = 0; y < Hy; y++) {
      size_t r0 = (size_t)y * W;
      yf[r0] = yb[r0] = VBCPer(bcW, yf[r0 + 1], yf[r0]);
      yf[r0 + W - 1] = yb[r0 + W - 1] = VBCPer(bcE, yf[r0 + W - 2], yf[r0 + W - 1]);
    }
    for (int x = 0; x < W; x++) {
      size_t top = (size_t)(Hy - 1) * W + x, top1 = (size_t)(Hy - 2) * W + x;
      yf[x] = yb[x] = VBCPar(bcS, yf[(size_t)W + x], yf[x]);
      yf[top] = yb[top] = VBCPar(bcN, yf[top1], yf[top]);
    }
  }

  // simulation.cpp:376-396
  void applyAccumulatedVelocity() {
    float *xf = vxf(), *yf = vyf();
#pragma omp parallel for schedule(static)
    for (int y = 1; y < H - 1; y++)
      for (int x = 1; x < Wx - 1; x++) {
        size_t i = (size_t)y * Wx + x;
        xf[i] += vx_accum[i];
        vx_accum[i] = 0.0f;
      }
#pragma omp parallel for schedule(static)
    for (int y = 1; y < Hy - 1; y++)
      for (int x = 1; x < W - 1; x++) {
        size_t i = (size_t)y * W + x;
        yf[i] += vy_accum[i];
        vy_accum[i] = 0.0f;
      }
  }

  // simulation.cpp:104-162 (two Jacobi iterations for vx, then two for vy;
  // each followed by swap and setVBCs)
  void diffuse() {
    const float a = dt * mu * (W - 1.0f) / pwidth;
    const float *F = flag.data();
    auto FL = [&](int x, int y) { return F[(size_t)y * W + x]; };
    for (int it = 1; it < 3; it++) {
      const float *s = vxf();
      float *d = vxb();
#pragma omp parallel for schedule(static)
      for (int y = 1; y < H - 1; y++) {
        for (int x = 1; x < Wx - 1; x++) {
          auto V = [&](int xx, int yy) { return s[(size_t)yy * Wx + xx]; };
          float val = 0;
          val += V(x + 1, y) * FL(x + 1, y) * FL(x + 2, y);
          val += V(x - 1, y) * FL(x, y) * FL(x - 1, y);
          float fvn = FL(x, y + 1) * FL(x - 1, y + 1);
          val += V(x, y + 1) * fvn + (1.0f - fvn) * -V(x, y);
          float fvs = FL(x, y - 1) * FL(x - 1, y - 1);
          val += V(x, y - 1) * fvs + (1.0f - fvs) * -V(x, y);
          d[(size_t)y * Wx + x] =
              FL(x, y) * FL(x + 1, y) * (V(x, y) + a * val) / (1.0f + 4.0f * a);
        }
      }
      fx = 1 - fx; // vx.swap()
      setVBCs();
    }
    for (int it = 1; it < 3; it++) {
      const float *s = vyf();
      float *d = vyb();
#pragma omp parallel for schedule(static)
      for (int y = 1; y < Hy - 1; y++) {
        for (int x = 1; x < W - 1; x++) {
          auto V = [&](int xx, int yy) { return s[(size_t)yy * W + xx]; };
          float val = 0;
          val += V(x, y - 1) * FL(x, y) * FL(x, y - 1);
          val += V(x, y + 1) * FL(x, y + 1) * FL(x, y + 2);
          float fve = FL(x + 1, y) * FL(x + 1, y + 1);
          val += V(x + 1, y) * fve + (1.0f - fve) * -V(x, y);
          float fvw = FL(x - 1, y) * FL(x - 1, y + 1);
          val += V(x - 1, y) * fvw + (1.0f - fvw) * -V(x, y);
          d[(size_t)y * W + x] =
              FL(x, y) * FL(x, y + 1) * (V(x, y) + a * val) / (1.0f + 4.0f * a);
        }
      }
      fy = 1 - fy; // vy.swap()
      setVBCs();
    }
  }

  // simulation.cpp:241-354: RK2 backtrace over 8-lane chunks.  A chunk is
  // skipped unless some lane has flag(x-1)+flag(x)==2 (vx) or
  // flag(x,y)+flag(x,y-1)==2 (vy); chunks start at x=1 and stop 8 lanes short
  // of the grid width, so the tail columns keep the back buffer.  All loads use
  // flat row-major offsets exactly as the 8-wide vector loads do.
  void advect() {
    const float ih = 1.0f / h;
    const float hdt = 0.5f * dt * ih, fdt = dt * ih;
    const float *F = flag.data();
    const float *xf = vxf(), *yf = vyf();
    float *xb = vxb(), *yb = vyb();
#ifdef UBG_ORACLE_AVX2_ADVECT
#pragma omp parallel for schedule(dynamic, 16)
    for (int y = 1; y < H - 1; y++) {
      for (int x0 = 1; x0 < Wx - 8; x0 += 8)
        advect_chunk8<false>(xf, yf, xb, yb, F, W, H, x0, y, hdt, fdt);
      for (int x0 = 1; x0 < W - 8; x0 += 8)
        advect_chunk8<true>(xf, yf, xb, yb, F, W, H, x0, y, hdt, fdt);
    }
    fx = 1 - fx;
    fy = 1 - fy;
    return;
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int y = 1; y < H - 1; y++) {
      for (int x0 = 1; x0 < Wx - 8; x0 += 8) {
        bool any = false;
        for (int k = 0; k < 8; k++) {
          size_t fi = (size_t)y * W + x0 + k;
          if (F[fi - 1] + F[fi] == 2.0f) any = true;
        }
        if (!any) continue;
        for (int k = 0; k < 8; k++) {
          const int x = x0 + k;
          float posx = (float)x + 0.5f, posy = (float)y;
          float vx1 = xf[(size_t)x + (size_t)y * Wx];
          float vy1 = ((yf[x + (size_t)y * W] + yf[x + (size_t)(y - 1) * W]) +
                       (yf[x + 1 + (size_t)y * W] + yf[x + 1 + (size_t)(y - 1) * W])) *
                      0.25f;
          float midx = posx - vx1 * hdt, midy = posy - vy1 * hdt;
          float vx2 = bicubic(xf, Wx, H, midx - 0.5f, midy);
          float vy2 = bicubic(yf, W, Hy, midx, midy - 0.5f);
          float endx = posx - vx2 * fdt, endy = posy - vy2 * fdt;
          float v = bicubic(xf, Wx, H, endx - 0.5f, endy);
          v = v * F[(size_t)y * W + x] * F[(size_t)y * W + x + 1];
          xb[(size_t)Wx * y + x] = v;
        }
      }
      for (int x0 = 1; x0 < W - 8; x0 += 8) {
        bool any = false;
        for (int k = 0; k < 8; k++) {
          size_t fi = (size_t)y * W + x0 + k;
          if (F[fi] + F[fi - W] == 2.0f) any = true;
        }
        if (!any) continue;
        for (int k = 0; k < 8; k++) {
          const int x = x0 + k;
          float posy = (float)y + 0.5f, posx = (float)x;
          float vy1 = yf[(size_t)x + (size_t)y * W];
          float vx1 = ((xf[x + (size_t)y * Wx] + xf[x + (size_t)(y - 1) * Wx]) +
                       (xf[x + 1 + (size_t)y * Wx] + xf[x + 1 + (size_t)(y - 1) * Wx])) *
                      0.25f;
          float midx = posx - vx1 * hdt, midy = posy - vy1 * hdt;
          float vx2 = bicubic(xf, Wx, H, midx - 0.5f, midy);
          float vy2 = bicubic(yf, W, Hy, midx, midy - 0.5f);
          float endx = posx - vx2 * fdt, endy = posy - vy2 * fdt;
          float v = bicubic(yf, W, Hy, endx, endy - 0.5f);
          v = v * F[(size_t)y * W + x] * F[(size_t)(y + 1) * W + x];
          yb[(size_t)W * y + x] = v;
        }
      }
    }
    fx = 1 - fx; // vx.swap()
    fy = 1 - fy; // vy.swap()
  }

  // simulation.cpp:166-171
  void divergence() {
    const float ih = 1.0f / h;
    const float *xf = vxf(), *yf = vyf();
#pragma omp parallel for schedule(static)
    for (int y = 1; y < H - 1; y++)
      for (int x = 1; x < W - 1; x++)
        f[(size_t)y * W + x] =
            -ih * (xf[(size_t)y * Wx + x] - xf[(size_t)y * Wx + x - 1] +
                   yf[(size_t)y * W + x] - yf[(size_t)(y - 1) * W + x]);
  }

  // simulation.cpp:173-187
  void apply_sinks() {
    for (auto &s : sinks) {
      float gx = s.x / h + 0.5f, gy = s.y / h + 0.5f;
      if (gx <= 3 || gx > W - 3 || gy <= 3 || gy > H - 3) continue;
      for (int yy = -1; yy <= 1; yy++)
        for (int xx = -1; xx <= 1; xx++)
          f[(size_t)(int)(gy + yy) * W + (int)(gx + xx)] = s.z;
      s.z = (float)(s.z * std::pow(0.000001, (double)(dt * 50)));
    }
    sinks.erase(std::remove_if(sinks.begin(), sinks.end(),
                               [](const Sink &t) { return t.z < 0.05f; }),
                sinks.end());
  }

  // simulation.cpp:196-207
  void correct_velocity() {
    const float ih = 1.0f / h;
    const float *F = flag.data();
    float *xf = vxf(), *yf = vyf();
#pragma omp parallel for schedule(static)
    for (int y = 1; y < H - 1; y++)
      for (int x = 1; x < W - 2; x++)
        xf[(size_t)y * Wx + x] -= F[(size_t)y * W + x] * F[(size_t)y * W + x + 1] *
                                  ih * (p[(size_t)y * W + x + 1] - p[(size_t)y * W + x]);
#pragma omp parallel for schedule(static)
    for (int y = 1; y < H - 2; y++)
      for (int x = 1; x < W - 1; x++)
        yf[(size_t)y * W + x] -= F[(size_t)y * W + x] * F[(size_t)(y + 1) * W + x] *
                                 ih * (p[(size_t)(y + 1) * W + x] - p[(size_t)y * W + x]);
  }

  // simulation.cpp:164-208
  void project() {
    divergence();
    apply_sinks();
    for (int c = 0; c < vcycles; c++) mg.solve(p.data(), f.data(), flag.data(), h, true);
    setPBC();
    correct_velocity();
  }

  // simulation.cpp:16-19
  void saveCurrent() {
    std::copy(vx[fx].begin(), vx[fx].end(), vx_cur.begin());
    std::copy(vy[fy].begin(), vy[fy].end(), vy_cur.begin());
  }

  // simulation.cpp:210-224 (centerP).  The reference's float OpenMP
  // reduction has no fixed order; this sums row by row in float, the order
  // of a one-thread run.
  void centerP() {
    float avgP = 0;
    for (int y = 1; y < H - 1; y++)
      for (int x = 1; x < W - 1; x++) avgP += p[(size_t)y * W + x];
    for (int y = 1; y < H - 1; y++)
      for (int x = 1; x < W - 1; x++) p[(size_t)y * W + x] -= avgP / W / H;
  }

  // simulation.cpp:226-238 (getDT): signed maximum of vx(x, y) and vy(x, y)
  // over x in [1, W-2], y in [1, H-2], starting at 1e-7
  float getDT() {
    float max_vel = 1.0e-7f;
    const float *xf = vxf(), *yf = vyf();
    for (int y = 1; y < H - 1; y++)
      for (int x = 1; x < W - 1; x++) {
        const float a = xf[(size_t)y * Wx + x], b = yf[(size_t)y * W + x];
        const float m = (a < b) ? b : a;
        max_vel = (max_vel < m) ? m : max_vel;
      }
    return pwidth / (W - 1.0f) / max_vel * 2.5f;
  }

  // simulation.cpp:356-374
  void step(float timestep) {
    dt = timestep;
    applyAccumulatedVelocity();
    diffuse();
    advect();
    setVBCs();
    project();
    setVBCs();
    saveCurrent();
  }
};

// advect_tracer_points.cs:21-39: Thomas Wang's 32-bit hash and the Numerical
// Recipes LCG.
inline uint32_t wang_hash(uint32_t seed) {
  seed = (seed ^ 61u) ^ (seed >> 16);
  seed *= 9u;
  seed = seed ^ (seed >> 4);
  seed *= 0x27d4eb2du;
  seed = seed ^ (seed >> 15);
  return seed;
}
inline uint32_t lcg(uint32_t &s) {
  s = 1664525u * s + 1013904223u;
  return s;
}
inline float randf(uint32_t &s) { return (float)(lcg(s) % 100000u) / 100000.0f; }

// simulation.cpp:398-411 (psampleFlagLinear); glm::mix(x,y,a) = x*(1-a)+y*a
inline float flag_linear(const float *F, int W, int H, float pwidth, float px,
                         float py) {
  float cx = px / (pwidth / W) - 0.5f, cy = py / (pwidth / W) - 0.5f;
  int ix = (int)cx, iy = (int)cy;
  ix = std::max(0, std::min(W - 2, ix));
  iy = std::max(0, std::min(H - 2, iy));
  float sx = cx - std::floor(cx), sy = cy - std::floor(cy);
  float p01 = F[(size_t)(iy + 1) * W + ix], p11 = F[(size_t)(iy + 1) * W + ix + 1];
  float p00 = F[(size_t)iy * W + ix], p10 = F[(size_t)iy * W + ix + 1];
  auto mix = [](float a, float b, float t) { return a * (1.0f - t) + b * t; };
  return mix(mix(p00, p10, sx), mix(p01, p11, sx), sy);
}

// ---------------------------------------------------------------------------
// Floating items (advect_floating_items.cpp:16-274)
// ---------------------------------------------------------------------------

// glm's vec2 as used there: component-wise ops evaluated left to right;
// glm::min/max are comparison based, normalize = v * (1 / sqrt(dot)),
// reflect = I - N * dot(N, I) * 2, rotate = (x c - y s, x s + y c).
struct Vec2 {
  float x, y;
};
inline Vec2 operator+(Vec2 a, Vec2 b) { return {a.x + b.x, a.y + b.y}; }
inline Vec2 operator-(Vec2 a, Vec2 b) { return {a.x - b.x, a.y - b.y}; }
inline Vec2 operator*(Vec2 a, Vec2 b) { return {a.x * b.x, a.y * b.y}; }
inline Vec2 operator*(Vec2 a, float k) { return {a.x * k, a.y * k}; }
inline Vec2 operator*(float k, Vec2 a) { return {k * a.x, k * a.y}; }
inline Vec2 operator/(Vec2 a, float k) { return {a.x / k, a.y / k}; }
inline float vdot(Vec2 a, Vec2 b) { return a.x * b.x + a.y * b.y; }
inline float vlength(Vec2 a) { return std::sqrt(vdot(a, a)); }
inline Vec2 vnormalize(Vec2 a) { return a * (1.0f / std::sqrt(vdot(a, a))); }
inline Vec2 vreflect(Vec2 i, Vec2 n) { return i - n * vdot(n, i) * 2.0f; }
inline float gmin(float a, float b) { return (b < a) ? b : a; }
inline float gmax(float a, float b) { return (a < b) ? b : a; }

// sin / cos by the Cephes single-precision method (S. L. Moshier, Cephes
// Math Library sinf.c / cosf.c): reduction by pi/4 in three parts, then the
// degree-7 sine or degree-8 cosine polynomial of the reduced argument.
const float kFOPI = 1.27323954473516f;
const float kDP1 = 0.78515625f, kDP2 = 2.4187564849853515625e-4f, kDP3 = 3.77489497744594108e-8f;
inline float cephes_sin_poly(float x, float z) {
  return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
inline float cephes_cos_poly(float z) {
  return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z *
             z - 0.5f * z + 1.0f;
}
inline float cephes_reduce(float ax, int &j) {
  j = (int)(kFOPI * ax);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  return ((ax - y * kDP1) - y * kDP2) - y * kDP3;
}
inline float cephes_sin(float x) {
  int j;
  const float r = cephes_reduce(std::fabs(x), j);
  bool neg = x < 0.0f;
  if (j > 3) {
    neg = !neg;
    j -= 4;
  }
  const float z = r * r;
  const float v = (j == 1 || j == 2) ? cephes_cos_poly(z) : cephes_sin_poly(r, z);
  return neg ? -v : v;
}
inline float cephes_cos(float x) {
  int j;
  const float r = cephes_reduce(std::fabs(x), j);
  bool neg = false;
  if (j > 3) {
    j -= 4;
    neg = !neg;
  }
  if (j > 1) neg = !neg;
  const float z = r * r;
  const float v = (j == 1 || j == 2) ? cephes_sin_poly(r, z) : cephes_cos_poly(z);
  return neg ? -v : v;
}
inline Vec2 vrotate(Vec2 v, float a) {
  const float c = cephes_cos(a), sn = cephes_sin(a);
  return {v.x * c - v.y * sn, v.x * sn + v.y * c};
}

// components.hpp:6-42 (CoItem + CoKinematics / CoKinematicsSimple), the
// layout of ubg_item
struct OItem {
  float size[2], pos[2], rotation, mass, vel[2], force[2], ang_vel, ang_force;
  int bump_count;
};

struct ItemFluid {
  Sim *s;
  Vec2 vel(Vec2 c) const { // bilinearVel (advect_floating_items.cpp:11-14)
    return {bilinear(s->vxf(), s->Wx, s->H, c.x - 0.5f, c.y),
            bilinear(s->vyf(), s->W, s->Hy, c.x, c.y - 0.5f)};
  }
  float flag(Vec2 pc) const { return flag_linear(s->flag.data(), s->W, s->H, s->pwidth, pc.x, pc.y); }
  Vec2 normal(Vec2 pc) const { // psampleFlagNormal (simulation.cpp:413-420)
    const float h = s->h;
    const float p01 = flag({pc.x - h, pc.y + h}), p11 = flag({pc.x + h, pc.y + h});
    const float p00 = flag({pc.x - h, pc.y - h}), p10 = flag({pc.x + h, pc.y - h});
    return {p11 + p10 - p01 - p00, p01 + p11 - p00 - p10};
  }
  // bilinearScatter (interpolators.hpp:29-40)
  static void scatter(std::vector<float> &g, int gw, int gh, Vec2 c, float v) {
    c.x = gmin(gmax(c.x, 0.0f), gw - 1.1f);
    c.y = gmin(gmax(c.y, 0.0f), gh - 1.1f);
    const int ix = (int)c.x, iy = (int)c.y;
    const float sx = c.x - std::floor(c.x), sy = c.y - std::floor(c.y);
    const float ax = 1.0f - sx, ay = 1.0f - sy;
    g[(size_t)(iy + 1) * gw + ix + 1] += sx * sy * v;
    g[(size_t)(iy + 1) * gw + ix] += ax * sy * v;
    g[(size_t)iy * gw + ix + 1] += sx * ay * v;
    g[(size_t)iy * gw + ix] += ax * ay * v;
  }
  bool skip(Vec2 pos) const { // advect_floating_items.cpp:38-44
    const Vec2 gp = pos / (s->pwidth / (float)s->W);
    // a non-finite position counts as outside (the reference indexes the
    // grid at (int)NaN for such an item)
    return !(gp.x < s->W - 2 && gp.x > 1.0 && gp.y < s->H - 2 && gp.y > 1.0);
  }
};

inline int item_substeps(Vec2 vel, float dt, float h) { // :24-26 (2.5 is a double)
  return (int)std::fmin(15.0f, std::fmax(1.0f, gmax(std::fabs(vel.x), std::fabs(vel.y)) * dt / h * 2.5));
}

// Simulation::advectFloatingItems (advect_floating_items.cpp:16-138)
void advect_item(const ItemFluid &F, OItem &o, float gameDT) {
  Sim &S = *F.s;
  const float h = S.h;
  Vec2 size{o.size[0], o.size[1]}, pos{o.pos[0], o.pos[1]}, vel{o.vel[0], o.vel[1]},
      force{o.force[0], o.force[1]};
  const int steps = item_substeps(vel, gameDT, h);
  const float subDT = gameDT / steps;
  for (int step = 0; step < steps; step++) {
    const Vec2 before = pos;
    pos = pos + subDT * vel;
    o.rotation = (float)std::fmod(o.rotation + subDT * o.ang_vel + 2 * M_PI, 2 * M_PI);
    if (F.skip(pos)) continue;
    force = force + Vec2{0.0f, -0.5f} * o.mass;
    const Vec2 pts[5] = {{1, 1}, {-1, 1}, {1, -1}, {-1, -1}, {0, 0}};
    for (const Vec2 &q : pts) {
      const Vec2 sp = vrotate(q * 0.5f * size, o.rotation);
      if (F.flag(pos + sp) < 0.5f) {
        Vec2 n = vnormalize(F.normal(0.5f * (before + pos) + sp));
        if (F.flag(before + sp) > 0.5f) pos = before;
        if (vlength(n) > 0.0f) {
          n = n / vlength(n);
          if (vdot(vel, n) < 0.0) vel = vreflect(vel, n) * 0.7f;
          vel = vel + n * 0.07f * F.flag(pos + sp);
          if (vdot(force, n) < 0.0f) force = force + 1.1f * vdot(force, n) * n;
        }
        o.bump_count++;
      }
    }
    const Vec2 external = Vec2{0.0f, -0.5f} * o.mass + force;
    Vec2 central{0.0f, 0.0f};
    const float angForce = o.ang_force;
    const Vec2 surf[4] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    const float side[4] = {size.y, size.y, size.x, size.x};
    for (int i = 0; i < 4; i++) {
      const int nSP = (int)std::fmax(2.0f, side[i] / h);
      for (int k = 0; k < nSP; k++) {
        const Vec2 sp = surf[i] + Vec2{std::fabs(surf[i].y), std::fabs(surf[i].x)} *
                                      (1.0f - k * 2.0f / (nSP - 1));
        Vec2 t = sp * size * 0.5f;
        t = vrotate(t, o.rotation);
        t = t + pos;
        const Vec2 dv = F.vel(t / h) - (vel + 3.141f * vrotate(sp * size * 0.5f, o.rotation + 0.5f * 3.141f) * o.ang_vel);
        const Vec2 nrm = vrotate(vnormalize(sp), o.rotation);
        const Vec2 fc = nrm * gmin(0.0f, vdot(dv, vrotate(surf[i], o.rotation)));
        central = central + fc * 400000.0f * (0.003f + side[i]) * side[i] / (float)nSP;
        if (t.x / h < 1.0f || t.x / h > S.Wx - 2.0f || t.y / h < 1.0f || t.y / h > S.H - 2.0f)
          continue;
        const Vec2 dvec = fc * (0.003f + side[i]) * side[i] / (float)nSP * subDT * 18000000.0f;
        ItemFluid::scatter(S.vx_accum, S.Wx, S.H, t / h - Vec2{0.5f, 0.0f}, -dvec.x);
        ItemFluid::scatter(S.vy_accum, S.W, S.Hy, t / h - Vec2{0.0f, 0.5f}, -dvec.y);
      }
    }
    central = central + 1000.0f * (size.x + size.y) * (F.vel(pos / h) - vel);
    vel = vel + subDT * (external + central) / o.mass;
    const float angMass = size.x * size.y * o.mass * (1.0f / 12.0f);
    o.ang_vel += subDT * angForce / angMass;
    o.ang_vel *= 0.98;
  }
  o.pos[0] = pos.x;
  o.pos[1] = pos.y;
  o.vel[0] = vel.x;
  o.vel[1] = vel.y;
  o.ang_force = 0.0f;
  o.force[0] = o.force[1] = 0.0f;
}

// Simulation::advectFloatingItemsSimple (advect_floating_items.cpp:140-274);
// snap: the positions the 100 bins hold (taken before any item moves)
void advect_item_simple(const ItemFluid &F, OItem &o, const std::vector<Vec2> &snap, float gameDT) {
  Sim &S = *F.s;
  const float h = S.h;
  Vec2 size{o.size[0], o.size[1]}, pos{o.pos[0], o.pos[1]}, vel{o.vel[0], o.vel[1]},
      force{o.force[0], o.force[1]};
  Vec2 rep{0.0f, 0.0f};
  int contacts = 0;
  const int bin = ((int)(pos.x * 100)) % 100;
  for (const Vec2 &p2 : snap) {
    if (((int)(p2.x * 100)) % 100 != bin) continue;
    const Vec2 dis = pos - p2;
    if (dis.x * dis.x + dis.y * dis.y < (size.x * size.y * 0.4f)) {
      const float len = gmax(0.1f * size.x, vlength(dis));
      rep = rep + 0.0001f * (dis / len / len);
      contacts++;
    }
  }
  rep = rep / (float)std::max(1, contacts);
  force = force + rep * 10.0f;
  const int steps = item_substeps(vel, gameDT, h);
  const float subDT = gameDT / steps;
  for (int step = 0; step < steps; step++) {
    pos = pos + subDT * vel;
    o.rotation = (float)std::fmod(o.rotation + subDT * o.ang_vel + 2 * M_PI, 2 * M_PI);
    const Vec2 gridPos = pos / (S.pwidth / (float)S.W);
    if (F.skip(pos)) continue;
    if (F.flag(pos) < 1.0f) {
      Vec2 n = F.normal(pos);
      if (vlength(n) > 0.0f) {
        n = n / vlength(n);
        if (vdot(vel, n) < 0.0) vel = vreflect(vel, n) * 0.7f;
        if (vdot(force, n) < 0.0f) force = vreflect(force, n) * 0.7f;
        const Vec2 lat = vrotate(n, 0.5f * (float)M_PI);
        const float latVel = vdot(lat, vel);
        const float rotVel = o.ang_vel * (size.x + size.y) * 0.5f;
        const float latDiff = latVel - rotVel;
        force = force + latDiff * lat * 1.0f;
      } else {
        vel = Vec2{0.0f, 0.0f};
      }
      o.bump_count++;
    }
    Vec2 external = Vec2{0.0f, -0.5f} * o.mass + force;
    const Vec2 dvel = F.vel(pos / h) - vel;
    external = external + 2000.0f * (size.x + size.y) * dvel;
    if (pos.x / h > 1.0f && pos.x / h < S.Wx - 2.0f && pos.y / h < 1.0f && pos.y / h < S.H - 2.0f) {
      const Vec2 dvec = dvel * size.x * size.y;
      ItemFluid::scatter(S.vx_accum, S.Wx, S.H, pos / h - Vec2{0.5f, 0.0f}, -dvec.x);
      ItemFluid::scatter(S.vy_accum, S.W, S.Hy, pos / h - Vec2{0.0f, 0.5f}, -dvec.y);
    }
    vel = vel + subDT * external / o.mass;
    const int gx = (int)gridPos.x, gy = (int)gridPos.y;
    const float *vx = S.vxf(), *vy = S.vyf(), *p = S.p.data();
    const float fluidAng = -((vx[(size_t)gy * S.Wx + gx] - vx[(size_t)(gy - 1) * S.Wx + gx]) -
                             (vy[(size_t)gy * S.W + gx] - vy[(size_t)gy * S.W + gx - 1])) / h / 2;
    const float angMass = size.x * size.y * o.mass * (1.0f / 12.0f);
    const float k = gmin(1.0f, subDT / angMass * 0.0005f * (size.x + size.y) / 4.0f);
    o.ang_vel += k * (fluidAng - o.ang_vel);
    o.ang_vel += subDT / angMass * o.ang_force;
    S.vx_accum[(size_t)gy * S.Wx + gx] -= o.ang_vel * k * 0.01f * p[(size_t)gy * S.W + gx];
    S.vx_accum[(size_t)(gy - 1) * S.Wx + gx] += o.ang_vel * k * 0.01f * p[(size_t)(gy - 1) * S.W + gx];
    S.vy_accum[(size_t)gy * S.W + gx] += o.ang_vel * k * 0.01f * p[(size_t)gy * S.W + gx];
    S.vy_accum[(size_t)gy * S.W + gx - 1] -= o.ang_vel * k * 0.01f * p[(size_t)gy * S.W + gx - 1];
  }
  o.pos[0] = pos.x;
  o.pos[1] = pos.y;
  o.vel[0] = vel.x;
  o.vel[1] = vel.y;
  o.ang_force = 0.0f;
  o.force[0] = o.force[1] = 0.0f;
}

} // namespace

// ===========================================================================
// C ABI of the oracle (loaded by ctypes from tests/ and bench.py only)
// ===========================================================================

ORACLE_API int oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

ORACLE_API int oracle_mg_level_sizes(int W, int H, int *ws, int *hs, int cap) {
  std::vector<int> a, b;
  level_sizes(W, H, a, b);
  for (int i = 0; i < (int)a.size() && i < cap; i++) {
    ws[i] = a[i];
    hs[i] = b[i];
  }
  return (int)a.size();
}

ORACLE_API void oracle_coarsen_flag(const float *f, int W, float *fc, int Wc, int Hc) {
  coarsen_flag<float>(f, W, fc, Wc, Hc);
}
ORACLE_API void oracle_rbgs_f32(float *p, const float *f, const float *fl, int W, int H,
                                float h, float alpha) {
  rbgs<float>(p, f, fl, W, H, h, alpha);
}
ORACLE_API void oracle_rbgs_f64(double *p, const double *f, const double *fl, int W,
                                int H, double h, double alpha) {
  rbgs<double>(p, f, fl, W, H, h, alpha);
}
ORACLE_API double oracle_residual_f32(const float *p, const float *f, const float *fl,
                                      float *r, int W, int H, float h) {
  return residual<float>(p, f, fl, r, W, H, h);
}
ORACLE_API double oracle_residual_f64(const double *p, const double *f, const double *fl,
                                      double *r, int W, int H, double h) {
  return residual<double>(p, f, fl, r, W, H, h);
}
ORACLE_API void oracle_restrict_f32(const float *r, int W, float *rc, int Wc, int Hc) {
  restrict_fw<float>(r, W, rc, Wc, Hc);
}
ORACLE_API void oracle_prolongate_f32(float *e, int W, int H, const float *ec,
                                      const float *flc, int Wc, const float *fl) {
  prolongate<float>(e, W, H, ec, flc, Wc, fl);
}
ORACLE_API void oracle_correct_f32(float *p, const float *e, int W, int H) {
  correct<float>(p, e, W, H);
}
ORACLE_API void oracle_zero_gradient_f32(float *p, int W, int H) {
  zero_gradient<float>(p, W, H);
}
ORACLE_API float oracle_bicubic(const float *g, int gw, int gh, float cx, float cy) {
  return bicubic(g, gw, gh, cx, cy);
}
ORACLE_API float oracle_bilinear(const float *g, int gw, int gh, float cx, float cy) {
  return bilinear(g, gw, gh, cx, cy);
}
ORACLE_API float oracle_cubic_hermite(float t, float A, float B, float C, float D) {
  return cubic_hermite(t, A, B, C, D);
}
ORACLE_API uint32_t oracle_wang_hash(uint32_t s) { return wang_hash(s); }

// Checks the identity the GPU smoother uses to divide by the neighbour count
// s in {1..4} without a division: with inv = RN(1/s) and q = RN(v*inv),
// RN(q + RN?(v - s*q)*inv) (two fused multiply-adds) equals RN(v/s).  Scans
// float bit patterns 0, stride, 2*stride, ... (all finite v) and returns the
// number of mismatches other than v = -0.
ORACLE_API long oracle_check_count_division(uint32_t stride) {
  long bad = 0;
  const float invs[5] = {0.0f, 1.0f, 0.5f, 1.0f / 3.0f, 0.25f};
  for (int s = 1; s <= 4; s++) {
    const float inv = invs[s], sf = (float)s;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long i = 0; i < (1L << 32); i += stride) {
      uint32_t u = (uint32_t)i;
      float v;
      std::memcpy(&v, &u, 4);
      if (!std::isfinite(v) || u == 0x80000000u) continue;
      float ref = v / sf;
      float q = v * inv;
      float r = std::fma(-sf, q, v);
      float q2 = std::fma(r, inv, q);
      uint32_t a, b;
      std::memcpy(&a, &ref, 4);
      std::memcpy(&b, &q2, 4);
      if (a != b) bad++;
    }
  }
  return bad;
}
ORACLE_API uint32_t oracle_lcg(uint32_t s) { return lcg(s); }

// ---- MG handle (fp32 = reference, fp64 = the mixed-precision oracle) ----
ORACLE_API void *oracle_mg_create(int W, int H, int fp64) {
  if (fp64) return new MG<double>(W, H);
  return new MG<float>(W, H);
}
ORACLE_API void oracle_mg_destroy(void *m, int fp64) {
  if (fp64) delete (MG<double> *)m;
  else delete (MG<float> *)m;
}
ORACLE_API void oracle_mg_set_params(void *m, int fp64, int nu_pre, int nu_post,
                                     int nu_coarse, float alpha) {
  MGParams p;
  p.nu_pre = nu_pre;
  p.nu_post = nu_post;
  p.nu_coarse = nu_coarse;
  p.alpha = alpha;
  if (fp64) ((MG<double> *)m)->prm = p;
  else ((MG<float> *)m)->prm = p;
}
ORACLE_API void oracle_mg_update_fields(void *m, const float *flag) {
  ((MG<float> *)m)->update_fields(flag);
}
ORACLE_API void oracle_mg64_update_fields(void *m, const double *flag) {
  ((MG<double> *)m)->update_fields(flag);
}
ORACLE_API void oracle_mg_get_flag(void *m, int level, float *out) {
  auto *g = (MG<float> *)m;
  std::copy(g->flagcs[level].begin(), g->flagcs[level].end(), out);
}
ORACLE_API void oracle_mg_solve(void *m, float *p, const float *f, const float *flag,
                                float h, int zg) {
  ((MG<float> *)m)->solve(p, f, flag, h, zg != 0);
}
ORACLE_API void oracle_mg64_solve(void *m, double *p, const double *f,
                                  const double *flag, double h, int zg) {
  ((MG<double> *)m)->solve(p, f, flag, h, zg != 0);
}

// ---- Simulation handle ----
enum {
  OF_VX = 0, OF_VX_BACK, OF_VY, OF_VY_BACK, OF_VX_ACCUM, OF_VY_ACCUM,
  OF_VX_CURRENT, OF_VY_CURRENT, OF_P, OF_F, OF_FLAG, OF_R
};

static float *sim_field(Sim *s, int id) {
  switch (id) {
  case OF_VX: return s->vxf();
  case OF_VX_BACK: return s->vxb();
  case OF_VY: return s->vyf();
  case OF_VY_BACK: return s->vyb();
  case OF_VX_ACCUM: return s->vx_accum.data();
  case OF_VY_ACCUM: return s->vy_accum.data();
  case OF_VX_CURRENT: return s->vx_cur.data();
  case OF_VY_CURRENT: return s->vy_cur.data();
  case OF_P: return s->p.data();
  case OF_F: return s->f.data();
  case OF_FLAG: return s->flag.data();
  case OF_R: return s->r.data();
  }
  return nullptr;
}
static size_t sim_field_size(Sim *s, int id) {
  switch (id) {
  case OF_VX: case OF_VX_BACK: case OF_VX_ACCUM: case OF_VX_CURRENT:
    return (size_t)s->Wx * s->H;
  case OF_VY: case OF_VY_BACK: case OF_VY_ACCUM: case OF_VY_CURRENT:
    return (size_t)s->W * s->Hy;
  default: return (size_t)s->W * s->H;
  }
}

ORACLE_API void *oracle_sim_create(const float *flag, int W, int H, float pwidth, float mu) {
  return new Sim(flag, W, H, pwidth, mu);
}
ORACLE_API void oracle_sim_destroy(void *s) { delete (Sim *)s; }
ORACLE_API void oracle_sim_set_params(void *sp, int nu_pre, int nu_post, int nu_coarse,
                                      int vcycles, float alpha) {
  Sim *s = (Sim *)sp;
  s->mg.prm.nu_pre = nu_pre;
  s->mg.prm.nu_post = nu_post;
  s->mg.prm.nu_coarse = nu_coarse;
  s->mg.prm.alpha = alpha;
  s->vcycles = vcycles;
}
ORACLE_API int oracle_sim_set_field(void *sp, int id, const float *src) {
  Sim *s = (Sim *)sp;
  float *d = sim_field(s, id);
  if (!d) return -1;
  std::memcpy(d, src, sim_field_size(s, id) * sizeof(float));
  if (id == OF_FLAG) s->mg.update_fields(s->flag.data());
  return 0;
}
ORACLE_API int oracle_sim_get_field(void *sp, int id, float *dst) {
  Sim *s = (Sim *)sp;
  float *d = sim_field(s, id);
  if (!d) return -1;
  std::memcpy(dst, d, sim_field_size(s, id) * sizeof(float));
  return 0;
}
ORACLE_API float oracle_sim_h(void *sp) { return ((Sim *)sp)->h; }
ORACLE_API void oracle_sim_add_sink(void *sp, float x, float y, float z) {
  ((Sim *)sp)->sinks.push_back({x, y, z});
}
ORACLE_API int oracle_sim_num_sinks(void *sp) { return (int)((Sim *)sp)->sinks.size(); }
ORACLE_API void oracle_sim_step(void *sp, float dt) { ((Sim *)sp)->step(dt); }
// the public bcWest / bcEast / bcNorth / bcSouth members (simulation.hpp:127-128)
ORACLE_API void oracle_sim_set_bcs(void *sp, int west, int east, int north, int south) {
  Sim *s = (Sim *)sp;
  s->bcW = west;
  s->bcE = east;
  s->bcN = north;
  s->bcS = south;
}

enum { OS_ACCUM = 0, OS_DIFFUSE, OS_ADVECT, OS_SET_VBCS, OS_PROJECT, OS_SET_PBC,
       OS_SAVE_CURRENT, OS_DIVERGENCE, OS_CORRECT_VELOCITY, OS_MG_SOLVE };
ORACLE_API int oracle_sim_stage(void *sp, int stage, float dt) {
  Sim *s = (Sim *)sp;
  s->dt = dt;
  switch (stage) {
  case OS_ACCUM: s->applyAccumulatedVelocity(); break;
  case OS_DIFFUSE: s->diffuse(); break;
  case OS_ADVECT: s->advect(); break;
  case OS_SET_VBCS: s->setVBCs(); break;
  case OS_PROJECT: s->project(); break;
  case OS_SET_PBC: s->setPBC(); break;
  case OS_SAVE_CURRENT: s->saveCurrent(); break;
  case OS_DIVERGENCE: s->divergence(); s->apply_sinks(); break;
  case OS_CORRECT_VELOCITY: s->correct_velocity(); break;
  case OS_MG_SOLVE: s->mg.solve(s->p.data(), s->f.data(), s->flag.data(), s->h, true); break;
  default: return -1;
  }
  return 0;
}

// advect_floating_items.cpp:11-14 (bilinearVel), for n points in grid space.
ORACLE_API void oracle_sim_bilinear_vel(void *sp, const float *c, float *out, int n) {
  Sim *s = (Sim *)sp;
  for (int i = 0; i < n; i++) {
    out[2 * i] = bilinear(s->vxf(), s->Wx, s->H, c[2 * i] - 0.5f, c[2 * i + 1]);
    out[2 * i + 1] = bilinear(s->vyf(), s->W, s->Hy, c[2 * i], c[2 * i + 1] - 0.5f);
  }
}

// simulation.cpp:422-425 (psampleFlagNearest): the flag at the truncated
// grid position; a cell outside the grid gives 0 (the reference indexes out
// of bounds there)
ORACLE_API float oracle_flag_nearest(const float *F, int W, int H, float pwidth, float px,
                                     float py) {
  const float gx = px / (pwidth / W), gy = py / (pwidth / W);
  if (!(gx > -1.0f && gx < (float)W && gy > -1.0f && gy < (float)H)) return 0.0f;
  return F[(size_t)(int)gy * W + (int)gx];
}
// simulation.cpp:413-420 (psampleFlagNormal)
ORACLE_API void oracle_flag_normal(const float *F, int W, int H, float pwidth, float h, float px,
                                   float py, float *out) {
  const float p01 = flag_linear(F, W, H, pwidth, px - h, py + h);
  const float p11 = flag_linear(F, W, H, pwidth, px + h, py + h);
  const float p00 = flag_linear(F, W, H, pwidth, px - h, py - h);
  const float p10 = flag_linear(F, W, H, pwidth, px + h, py - h);
  out[0] = p11 + p10 - p01 - p00;
  out[1] = p01 + p11 - p00 - p10;
}
ORACLE_API void oracle_gs_f32(float *p, const float *f, const float *fl, int W, int H, float h,
                              float alpha) {
  gs(p, f, fl, W, H, h, alpha);
}
ORACLE_API void oracle_sim_center_p(void *sp) { ((Sim *)sp)->centerP(); }
ORACLE_API float oracle_sim_get_dt(void *sp) { return ((Sim *)sp)->getDT(); }

ORACLE_API float oracle_flag_linear(const float *F, int W, int H, float pwidth, float px,
                                    float py) {
  return flag_linear(F, W, H, pwidth, px, py);
}

// One tracer step: RK2 midpoint with bilinearVel (advect_floating_items.cpp:11-14),
// freeze/age rule and Wang/LCG respawn (advect_tracer_points.cs:42-82).  The
// velocity is the front buffer after the step, positions are in physical units.
ORACLE_API void oracle_sim_advect_tracers(void *sp, float *xy, float *age, int n, float dt,
                                          uint32_t seed) {
  Sim *s = (Sim *)sp;
  const float h = s->h;
  const float pdx = s->pwidth, pdy = s->pwidth * s->H / s->W;
  const float *xf = s->vxf(), *yf = s->vyf();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    float px = xy[2 * i], py = xy[2 * i + 1];
    float cx = px / h, cy = py / h;
    float v1x = bilinear(xf, s->Wx, s->H, cx - 0.5f, cy);
    float v1y = bilinear(yf, s->W, s->Hy, cx, cy - 0.5f);
    float mx = px + v1x * dt * 0.5f, my = py + v1y * dt * 0.5f;
    float mcx = mx / h, mcy = my / h;
    float v2x = bilinear(xf, s->Wx, s->H, mcx - 0.5f, mcy);
    float v2y = bilinear(yf, s->W, s->Hy, mcx, mcy - 0.5f);
    float nx = px + v2x * dt, ny = py + v2y * dt;
    float a = age[i];
    if (px < 0.0f || py < 0.0f || px > pdx || py > pdy ||
        flag_linear(s->flag.data(), s->W, s->H, s->pwidth, px, py) < 0.6f) {
      nx = px;
      ny = py;
      a += 0.1f;
    }
    if (a > 2.0f * 3.141f) {
      uint32_t st = wang_hash((uint32_t)i + seed);
      float rx = randf(st);
      float ry = randf(st);
      nx = rx * pdx;
      ny = ry * pdy;
      a = 0.0f;
    } else {
      a = a + 0.02f;
    }
    xy[2 * i] = nx;
    xy[2 * i + 1] = ny;
    age[i] = a;
  }
}

// advectFloatingItems / advectFloatingItemsSimple on the oracle simulation:
// items (ubg_item layout) in order, accumulation into vx_accum / vy_accum.
ORACLE_API void oracle_sim_advect_items(void *sp, void *items, int n, float gameDT, int simple) {
  Sim *s = (Sim *)sp;
  OItem *it = (OItem *)items;
  ItemFluid F{s};
  std::vector<Vec2> snap;
  for (int i = 0; i < n; i++) snap.push_back({it[i].pos[0], it[i].pos[1]});
  for (int i = 0; i < n; i++) {
    if (simple) advect_item_simple(F, it[i], snap, gameDT);
    else advect_item(F, it[i], gameDT);
  }
}

ORACLE_API float oracle_cephes_sin(float x) { return cephes_sin(x); }
ORACLE_API float oracle_cephes_cos(float x) { return cephes_cos(x); }

// Connected components of the fluid cells (flag 1) under 4-neighbour
// connectivity, the coupling of calculateResidualField
// (pressure_solver.cpp:100-108): labels 0..n-1 in raster order of each
// component's first cell, -1 on solid cells.  Breadth-first search from
// each unlabelled fluid cell in raster order.  Returns n.
ORACLE_API int oracle_components(const float *fl, int W, int H, int *lab) {
  const long n = (long)W * H;
  for (long i = 0; i < n; i++) lab[i] = -1;
  std::vector<long> queue;
  int nc = 0;
  for (long s = 0; s < n; s++) {
    if (fl[s] == 0.0f || lab[s] >= 0) continue;
    queue.clear();
    queue.push_back(s);
    lab[s] = nc;
    for (size_t k = 0; k < queue.size(); k++) {
      const long i = queue[k];
      const int x = (int)(i % W), y = (int)(i / W);
      const long nb[4] = {x > 0 ? i - 1 : -1, x + 1 < W ? i + 1 : -1, y > 0 ? i - W : -1,
                          y + 1 < H ? i + W : -1};
      for (long j : nb)
        if (j >= 0 && fl[j] != 0.0f && lab[j] < 0) {
          lab[j] = nc;
          queue.push_back(j);
        }
    }
    nc++;
  }
  return nc;
}

// Subtract from f, on every fluid cell, the mean of f over the cell's
// component.  The component sums are 64-bit fixed point: each f scaled by
// 2^s (s = 62 - e - en with max|f| < 2^e and n < 2^en, so no sum overflows)
// and rounded to the nearest integer, which makes them independent of the
// summation order; mean = sum 2^-s / count in float64; f = (float)(f - mean).
ORACLE_API int oracle_remove_component_means(const float *fl, int W, int H, float *f) {
  const long n = (long)W * H;
  std::vector<int> lab((size_t)n);
  const int nc = oracle_components(fl, W, H, lab.data());
  float amax = 0.0f;
  for (long i = 0; i < n; i++)
    if (lab[i] >= 0) amax = std::fmax(amax, std::fabs(f[i]));
  if (nc == 0 || amax == 0.0f) return nc;
  int e = 0, en = 0;
  (void)std::frexp((double)amax, &e);
  (void)std::frexp((double)n, &en);
  const int sh = 62 - e - en;
  const double scale = std::ldexp(1.0, sh), inv = std::ldexp(1.0, -sh);
  std::vector<long long> sum((size_t)nc, 0);
  std::vector<int> cnt((size_t)nc, 0);
  for (long i = 0; i < n; i++)
    if (lab[i] >= 0) {
      sum[lab[i]] += std::llrint((double)f[i] * scale);
      cnt[lab[i]]++;
    }
  for (long i = 0; i < n; i++)
    if (lab[i] >= 0) {
      const double mean = (double)sum[lab[i]] * inv / (double)cnt[lab[i]];
      f[i] = (float)((double)f[i] - mean);
    }
  return nc;
}
