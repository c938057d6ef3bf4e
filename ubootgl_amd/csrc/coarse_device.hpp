// coarse_device.hpp -- the single-workgroup LDS V-cycle of the coarse
// levels (MG::solveLevel, pressure_solver.cpp:201-248, from a level whose
// whole sub-hierarchy fits in LDS down to the coarsest solved level L-2),
// shared by the single-domain and the batched solvers.
#pragma once

#include "mg_device.hpp"

namespace ubg {

constexpr int MAXL = 16;
constexpr int TPB_C = 1024;
constexpr int WAVES_C = TPB_C / 64;
constexpr size_t LDS_BUDGET = 152 * 1024;

struct CoarseArgs {
  int n; // levels first .. L-2
  int w[MAXL], h[MAXL];
  float hh[MAXL];
  int op[MAXL], of[MAXL], om[MAXL]; // LDS byte offsets of p, f, mask
  int orr;                           // residual scratch (first level's size)
  const uint8_t *mask[MAXL];
  const float *f0;
  float *p0;
  int nu_pre, nu_post, nu_coarse;
  float alpha;
};

// 1024 threads as 16 waves x 64 lanes.  A level row with at most `count`
// items (cells of a colour, cells, blocks) takes L lanes, L the smallest
// power of two >= count (at most 64), so a wave holds G = 64 / L rows at
// once: lane group g of wave ty starts at row offset G ty + g and steps by
// 16 G.  On the narrow coarse levels this packs 2-8 consecutive rows into
// one wave, so a level of few rows keeps few waves busy (and few per SIMD)
// instead of one partly filled wave per row.
struct RowGroups {
  int lr;   // lane within the row group
  int r0;   // first row offset (G * wave + group)
  int step; // row step, 16 * G
  int L;    // lanes per row
};
__device__ __forceinline__ RowGroups row_groups(int count) {
  int lg = 6;
  while (lg > 0 && (1 << (lg - 1)) >= count) lg--;
  const int tx = threadIdx.x & 63;
  const int ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  RowGroups r;
  r.L = 1 << lg;
  r.lr = tx & (r.L - 1);
  r.r0 = (ty << (6 - lg)) + (tx >> lg);
  r.step = WAVES_C << (6 - lg);
  return r;
}

// A cell's smoothing constants, decoded once from its mask byte and f for
// sweeps that update the same cell again and again.  smooth_rule performs
// the operations of smooth() in the same order (the division by the count
// as div_count does it), so the results are identical.
struct CellRule {
  int sw, se, ss, sn; // 0 / -1: the neighbour's value is kept, or +0 added
  float inv, ns;      // RN(1 / s) and -s for the neighbour count s
  float F, fhh;       // own flag (0 / 1) and (f * h) * h
  bool empty;         // s == 0: the quotient is +0
};
__device__ __forceinline__ CellRule cell_rule(unsigned mk, float fv, float h) {
  CellRule c;
  c.sw = -(int)((mk >> 1) & 1u);
  c.se = -(int)((mk >> 2) & 1u);
  c.ss = -(int)((mk >> 3) & 1u);
  c.sn = -(int)((mk >> 4) & 1u);
  const unsigned s = mk >> 5;
  const float pow2 = __int_as_float((int)((127u - (s >> 1)) << 23));
  c.inv = (s == 3u) ? (1.0f / 3.0f) : pow2;
  c.ns = -(float)s;
  c.empty = s == 0u;
  c.F = (float)(mk & MK_SELF);
  c.fhh = fv * h * h;
  return c;
}
__device__ __forceinline__ float smooth_rule(float pc, float pw, float pe, float ps, float pn,
                                             const CellRule &c, float alpha) {
  float val = 0.0f;
  val += __int_as_float(__float_as_int(pw) & c.sw);
  val += __int_as_float(__float_as_int(pe) & c.se);
  val += __int_as_float(__float_as_int(ps) & c.ss);
  val += __int_as_float(__float_as_int(pn) & c.sn);
  val += c.fhh;
  const float q = val * c.inv;
  const float r = fmaf(c.ns, q, val);
  const float q2 = fmaf(r, c.inv, q);
  val = c.empty ? 0.0f : q2;
  return c.F * (alpha * val + (1.0f - alpha) * pc);
}

// nu red+black sweeps of a w x h level in LDS.  A row's cells of a colour
// are at most (w - 1) / 2; rows of one lane group share the colour parity
// (their offsets differ by multiples of 16).  When the lane groups cover
// every interior row and every cell of a colour at once (levels of at most
// 130 x 18, 66 x 34 or 34 x 66 cells), each thread owns at most one cell of
// each colour for the whole smoothing: its cells' rules
// are decoded once, and a colour phase is the five reads of p, the update
// and one write.
__device__ __forceinline__ void lds_sweeps(float *P, const float *Fv, const uint8_t *M, int w,
                                           int h, float hh, float alpha, int nu) {
  const RowGroups q = row_groups((w - 1) >> 1);
  if (q.step >= h - 2 && 2 * q.L >= w - 2) {
    const int y = 1 + q.r0;
    const int xa = 1 + ((1 + q.r0) & 1) + 2 * q.lr; // colour 0
    const int xb = 1 + ((2 + q.r0) & 1) + 2 * q.lr; // colour 1
    const bool oka = y < h - 1 && xa < w - 1, okb = y < h - 1 && xb < w - 1;
    const int ia = y * w + xa, ib = y * w + xb;
    CellRule ra{}, rb{};
    if (oka) ra = cell_rule(M[ia], Fv[ia], hh);
    if (okb) rb = cell_rule(M[ib], Fv[ib], hh);
    for (int s = 0; s < nu; s++) {
      if (oka) P[ia] = smooth_rule(P[ia], P[ia - 1], P[ia + 1], P[ia - w], P[ia + w], ra, alpha);
      __syncthreads();
      if (okb) P[ib] = smooth_rule(P[ib], P[ib - 1], P[ib + 1], P[ib - w], P[ib + w], rb, alpha);
      __syncthreads();
    }
    return;
  }
  for (int s = 0; s < nu; s++)
    for (int color = 0; color < 2; color++) {
      const int x0 = 1 + ((1 + q.r0 + color) & 1) + 2 * q.lr;
      for (int y = 1 + q.r0; y < h - 1; y += q.step) {
        for (int x = x0; x < w - 1; x += 2 * q.L) {
          int i = y * w + x;
          P[i] = smooth(P[i], P[i - 1], P[i + 1], P[i - w], P[i + w], Fv[i], M[i], hh, alpha);
        }
      }
      __syncthreads();
    }
}

__device__ __forceinline__ float lds_residual(const float *P, const float *Fv, const uint8_t *M,
                                              int i, int w, float ihsq) {
  return residual_of(P[i], P[i - 1], P[i + 1], P[i - w], P[i + w], Fv[i], M[i], ihsq);
}

// The whole V-cycle of levels first..L-2 of one domain in LDS, by one
// 1024-thread workgroup; f0 is the first level's rhs, the first level's
// solution is written to p0.
__device__ __forceinline__ void coarse_vcycle_body(const CoarseArgs &a, const float *f0,
                                                  float *p0) {
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x;
  // load: f of the first level, masks of all levels; p = 0 everywhere.
  // Each round issues KC loads per thread before its LDS stores, so a round
  // costs one memory latency instead of one per element; the first
  // TPB_C-element round of levels 1..KR rides in level 0's first round, so
  // the small levels' masks cost no latency of their own.
  constexpr int KC = 8, KR = 4;
  // elements [from, n) of level l, KC per thread per round
  auto stage = [&](int l, int from) {
    float *P = (float *)(smem + a.op[l]);
    float *Fv = (float *)(smem + a.of[l]);
    uint8_t *M = (uint8_t *)(smem + a.om[l]);
    const uint8_t *mg = a.mask[l];
    const int n = a.w[l] * a.h[l];
    for (int base = from; base < n; base += KC * TPB_C) {
      float vf[KC];
      uint8_t vm[KC];
#pragma unroll
      for (int k = 0; k < KC; k++) {
        const int i = base + tid + k * TPB_C;
        if (i < n) {
          vf[k] = (l == 0) ? f0[i] : 0.0f;
          vm[k] = mg[i];
        }
      }
#pragma unroll
      for (int k = 0; k < KC; k++) {
        const int i = base + tid + k * TPB_C;
        if (i < n) {
          P[i] = 0.0f;
          Fv[i] = vf[k];
          M[i] = vm[k];
        }
      }
    }
  };
  {
    uint8_t rm[KR];
#pragma unroll
    for (int r = 0; r < KR; r++)
      if (1 + r < a.n && tid < a.w[1 + r] * a.h[1 + r]) rm[r] = a.mask[1 + r][tid];
    stage(0, 0);
#pragma unroll
    for (int r = 0; r < KR; r++)
      if (1 + r < a.n && tid < a.w[1 + r] * a.h[1 + r]) {
        const int l = 1 + r;
        ((float *)(smem + a.op[l]))[tid] = 0.0f;
        ((float *)(smem + a.of[l]))[tid] = 0.0f;
        ((uint8_t *)(smem + a.om[l]))[tid] = rm[r];
      }
  }
  for (int l = 1; l < a.n; l++) stage(l, l <= KR ? TPB_C : 0);
  __syncthreads();

  // descend
  for (int l = 0; l < a.n - 1; l++) {
    float *P = (float *)(smem + a.op[l]);
    float *Fv = (float *)(smem + a.of[l]);
    const uint8_t *M = (const uint8_t *)(smem + a.om[l]);
    const int w = a.w[l], h = a.h[l];
    lds_sweeps(P, Fv, M, w, h, a.hh[l], a.alpha, a.nu_pre);
    // residual (pressure_solver.cpp:91-116) of every interior cell once,
    // then restrict (:118-132) into level l+1, whose p starts at 0
    float *Rr = (float *)(smem + a.orr);
    float *Fc = (float *)(smem + a.of[l + 1]);
    float *Pc = (float *)(smem + a.op[l + 1]);
    const int wc = a.w[l + 1], hc = a.h[l + 1];
    const float ihsq = 1.0f / a.hh[l] / a.hh[l];
    const RowGroups qr = row_groups(w - 2);
    for (int y = 1 + qr.r0; y < h - 1; y += qr.step)
      for (int x = 1 + qr.lr; x < w - 1; x += qr.L)
        Rr[y * w + x] = lds_residual(P, Fv, M, y * w + x, w, ihsq);
    __syncthreads();
    const RowGroups qc = row_groups(wc - 2);
    for (int y = 1 + qc.r0; y < hc - 1; y += qc.step)
    for (int x = 1 + qc.lr; x < wc - 1; x += qc.L) {
      int c = (2 * y) * w + 2 * x;
      auto R = [&](int dx, int dy) { return Rr[c + dy * w + dx]; };
      float v = 0.0f;
      v = R(-1, -1) * 1.0f + R(0, -1) * 2.0f + R(1, -1) * 1.0f;
      v += R(-1, 0) * 2.0f + R(0, 0) * 4.0f + R(1, 0) * 2.0f;
      v += R(-1, 1) * 1.0f + R(0, 1) * 2.0f + R(1, 1) * 1.0f;
      Fc[y * wc + x] = v * (1.0f / 16.0f);
      Pc[y * wc + x] = 0.0f;
    }
    __syncthreads();
  }
  // coarsest solved level
  {
    int l = a.n - 1;
    float *P = (float *)(smem + a.op[l]);
    float *Fv = (float *)(smem + a.of[l]);
    const uint8_t *M = (const uint8_t *)(smem + a.om[l]);
    lds_sweeps(P, Fv, M, a.w[l], a.h[l], a.hh[l], a.alpha, a.nu_coarse);
  }
  // ascend
  for (int l = a.n - 2; l >= 0; l--) {
    float *P = (float *)(smem + a.op[l]);
    float *Fv = (float *)(smem + a.of[l]);
    const uint8_t *M = (const uint8_t *)(smem + a.om[l]);
    const int w = a.w[l], h = a.h[l];
    const float *Pc = (const float *)(smem + a.op[l + 1]);
    const uint8_t *Mc = (const uint8_t *)(smem + a.om[l + 1]);
    const int wc = a.w[l + 1];
    // prolongate + correct (pressure_solver.cpp:134-181) by 2x2 blocks of
    // fine cells over coarse cell (ax, by), each by its parity rule
    auto E = [&](int xx, int yy) { return Pc[yy * wc + xx]; };
    auto FC = [&](int xx, int yy) { return (int)(Mc[yy * wc + xx] & MK_SELF); };
    const RowGroups qp = row_groups(((w - 2) >> 1) + 1);
    for (int by = qp.r0; by <= (h - 2) >> 1; by += qp.step)
    for (int ax = qp.lr; ax <= (w - 2) >> 1; ax += qp.L) {
      const float e00 = E(ax, by);
      const int f00 = FC(ax, by);
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
          const int x = 2 * ax + dx, y = 2 * by + dy;
          if (x < 1 || x > w - 2 || y < 1 || y > h - 2) continue;
          const int i = y * w + x;
          const float F = (M[i] & MK_SELF) ? 1.0f : 0.0f;
          float e = 0.0f;
          if (dx == 0 && dy == 0) {
            e = e00 * F;
          } else if (dy == 0) {
            if (x <= w - 3) e = F * (e00 + E(ax + 1, by)) / eps_count(f00 + FC(ax + 1, by));
          } else if (dx == 0) {
            if (y <= h - 3) e = F * (e00 + E(ax, by + 1)) / eps_count(f00 + FC(ax, by + 1));
          } else {
            if (x <= w - 3 && y <= h - 3)
              e = F * (e00 + E(ax + 1, by + 1) + E(ax + 1, by) + E(ax, by + 1)) /
                  eps_count(f00 + FC(ax + 1, by + 1) + FC(ax + 1, by) + FC(ax, by + 1));
          }
          P[i] += e;
        }
    }
    __syncthreads();
    lds_sweeps(P, Fv, M, w, h, a.hh[l], a.alpha, a.nu_post);
  }
  // write the first level's solution
  {
    const float *P = (const float *)(smem + a.op[0]);
    int n = a.w[0] * a.h[0];
    for (int i = tid; i < n; i += TPB_C) p0[i] = P[i];
  }
}

// LDS layout (p, f, mask per level, 16-byte aligned) of levels first..L-2
// for level sizes ws/hs; fills the offsets of `a` if given.  Returns the
// byte count, or SIZE_MAX if there are more than MAXL levels.
inline size_t coarse_lds_layout(const int *ws, const int *hs, int first, int levels,
                                CoarseArgs *a) {
  size_t off = 0;
  int k = 0;
  for (int l = first; l <= levels - 2; l++, k++) {
    if (k >= MAXL) return (size_t)-1;
    size_t n = (size_t)ws[l] * hs[l];
    if (a) {
      a->op[k] = (int)off;
      a->of[k] = (int)(off + ((n * 4 + 15) & ~size_t(15)));
      a->om[k] = (int)(off + 2 * ((n * 4 + 15) & ~size_t(15)));
    }
    off += 2 * ((n * 4 + 15) & ~size_t(15)) + ((n + 15) & ~size_t(15));
  }
  if (a) a->orr = (int)off;
  off += ((size_t)ws[first] * hs[first] * 4 + 15) & ~size_t(15);
  return off;
}

inline size_t coarse_lds_bytes(const Level *lv, int first, int levels, CoarseArgs *a) {
  int ws[64], hs[64];
  for (int l = 0; l < levels && l < 64; l++) {
    ws[l] = lv[l].w;
    hs[l] = lv[l].h;
  }
  return coarse_lds_layout(ws, hs, first, levels, a);
}

} // namespace ubg
