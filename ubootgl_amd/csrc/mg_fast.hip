// mg_fast.hip -- fused multigrid kernels for gfx950.
//
// 1. k_rbgs_tile: one full red+black sweep of rbgs (pressure_solver.cpp:49-89,
//    two-pass semantics) in one launch.  A workgroup loads a TX x TY tile of p
//    with a two-cell halo into LDS (f and the flag mask with a one-cell halo),
//    updates the red cells of the tile plus a one-cell ring (from the old black
//    values, exactly what the first pass computes there), then the black cells
//    of the tile from those red values, and writes the tile once.  Red cells
//    of the ring are recomputed by the neighbouring tile with identical
//    inputs, so the result equals the two-pass sweep.  With zero_gradient the
//    setZeroGradientBC copy (pressure_solver.cpp:183-192) is applied to the
//    tile's boundary cells from their updated interior neighbours; the tile
//    grid is shifted by one cell where needed so that every boundary cell
//    shares its tile with that neighbour.
//    HBM traffic per sweep: p read + write, f read, 1-byte mask read =
//    13 bytes/cell; halo re-reads are served by L2, and tiles are ordered so
//    that each XCD (own L2) works on a contiguous band of the grid.
//
// 2. k_coarse_vcycle: the V-cycle of every level from `first` down to the
//    coarsest solved level L-2 (pressure_solver.cpp:201-248), run by one
//    workgroup entirely in LDS (p, f and mask of all those levels fit in
//    160 KB once the level is <= ~128x64).  Replaces ~40 latency-bound
//    launches per V-cycle with one.
#include "ubg_internal.hpp"

namespace ubg {

namespace {

constexpr int TX = 64, TY = 32, TPB_T = 256;

__device__ __forceinline__ float inv_count(int s) {
  // 1/sum of neighbour flags; sum in {1..4}
  return s == 1 ? 1.0f : (s == 2 ? 0.5f : (s == 3 ? (1.0f / 3.0f) : 0.25f));
}

// smoothingKernel (pressure_solver.cpp:10-24) from neighbour values and mask
__device__ __forceinline__ float smooth(float pc, float pw, float pe, float ps, float pn,
                                        float fv, unsigned mk, float h, float alpha) {
  int s = ((mk >> 1) & 1) + ((mk >> 2) & 1) + ((mk >> 3) & 1) + ((mk >> 4) & 1);
  float val = 0.0f;
  val += (mk & MK_W) ? pw : 0.0f;
  val += (mk & MK_E) ? pe : 0.0f;
  val += (mk & MK_S) ? ps : 0.0f;
  val += (mk & MK_N) ? pn : 0.0f;
  val += fv * h * h;
  val = (s == 0) ? 0.0f : val * inv_count(s);
  float F = (mk & MK_SELF) ? 1.0f : 0.0f;
  return F * (alpha * val + (1.0f - alpha) * pc);
}

// Blocks are dispatched round-robin over the 8 XCDs; hand each XCD a
// contiguous range of tiles so halo re-reads hit its own L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int q = nb >> 3, rem = nb & 7, k = b & 7;
  return k * q + min(k, rem) + (b >> 3);
}

__global__ void __launch_bounds__(TPB_T)
    k_rbgs_tile(float *__restrict__ p, const float *__restrict__ f,
                const uint8_t *__restrict__ m, int w, int h, float hh, float alpha, int zg,
                int sx, int sy, int ntx) {
  __shared__ float sp[TY + 4][TX + 4]; // origin (x0-2, y0-2)
  __shared__ float sf[TY + 2][TX + 2]; // origin (x0-1, y0-1)
  __shared__ uint8_t sm[TY + 2][TX + 2];
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int x0 = (t % ntx) * TX - sx, y0 = (t / ntx) * TY - sy;
  const int tid = threadIdx.x;

  // Issue every global load of this thread before the first LDS store so the
  // loads overlap (compile-time trip counts, fully unrolled).
  constexpr int NP = (TY + 4) * (TX + 4), NPI = (NP + TPB_T - 1) / TPB_T;
  constexpr int NF = (TY + 2) * (TX + 2), NFI = (NF + TPB_T - 1) / TPB_T;
  float vp[NPI], vf[NFI];
  uint8_t vm[NFI];
#pragma unroll
  for (int k = 0; k < NPI; k++) {
    int li = tid + k * TPB_T;
    int r = li / (TX + 4), c = li - r * (TX + 4);
    int gx = x0 - 2 + c, gy = y0 - 2 + r;
    vp[k] = (li < NP && gx >= 0 && gx < w && gy >= 0 && gy < h) ? p[(size_t)gy * w + gx] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < NFI; k++) {
    int li = tid + k * TPB_T;
    int r = li / (TX + 2), c = li - r * (TX + 2);
    int gx = x0 - 1 + c, gy = y0 - 1 + r;
    bool in = li < NF && gx >= 1 && gx < w - 1 && gy >= 1 && gy < h - 1;
    size_t gi = (size_t)gy * w + gx;
    vf[k] = in ? f[gi] : 0.0f;
    vm[k] = in ? m[gi] : 0;
  }
#pragma unroll
  for (int k = 0; k < NPI; k++) {
    int li = tid + k * TPB_T;
    if (li < NP) (&sp[0][0])[li] = vp[k];
  }
#pragma unroll
  for (int k = 0; k < NFI; k++) {
    int li = tid + k * TPB_T;
    if (li < NF) {
      (&sf[0][0])[li] = vf[k];
      (&sm[0][0])[li] = vm[k];
    }
  }
  __syncthreads();

  // red: (x+y) odd over the tile plus a one-cell ring
  constexpr int RH = (TX + 2) / 2;
  for (int li = tid; li < (TY + 2) * RH; li += TPB_T) {
    int r = li / RH, k = li - r * RH;
    int c = 2 * k + ((x0 + y0 + r + 1) & 1);
    int gx = x0 - 1 + c, gy = y0 - 1 + r;
    if (gx >= 1 && gx < w - 1 && gy >= 1 && gy < h - 1) {
      int R = r + 1, Cc = c + 1;
      sp[R][Cc] = smooth(sp[R][Cc], sp[R][Cc - 1], sp[R][Cc + 1], sp[R - 1][Cc], sp[R + 1][Cc],
                         sf[r][c], sm[r][c], hh, alpha);
    }
  }
  __syncthreads();

  // black: (x+y) even over the tile
  constexpr int BH = TX / 2;
  for (int li = tid; li < TY * BH; li += TPB_T) {
    int r = li / BH, k = li - r * BH;
    int c = 2 * k + ((x0 + y0 + r) & 1);
    int gx = x0 + c, gy = y0 + r;
    if (gx >= 1 && gx < w - 1 && gy >= 1 && gy < h - 1) {
      int R = r + 2, Cc = c + 2;
      sp[R][Cc] = smooth(sp[R][Cc], sp[R][Cc - 1], sp[R][Cc + 1], sp[R - 1][Cc], sp[R + 1][Cc],
                         sf[r + 1][c + 1], sm[r + 1][c + 1], hh, alpha);
    }
  }
  __syncthreads();

  if (zg) {
    // setZeroGradientBC on the tile's boundary cells (corners untouched)
    for (int li = tid; li < TY * TX; li += TPB_T) {
      int r = li / TX, c = li - r * TX;
      int gx = x0 + c, gy = y0 + r;
      if (gx < 0 || gx >= w || gy < 0 || gy >= h) continue;
      int R = r + 2, Cc = c + 2;
      if (gy >= 1 && gy < h - 1) {
        if (gx == 0) sp[R][Cc] = sp[R][Cc + 1];
        else if (gx == w - 1) sp[R][Cc] = sp[R][Cc - 1];
      }
      if (gx >= 1 && gx < w - 1) {
        if (gy == 0) sp[R][Cc] = sp[R + 1][Cc];
        else if (gy == h - 1) sp[R][Cc] = sp[R - 1][Cc];
      }
    }
    __syncthreads();
  }

  for (int li = tid; li < TY * TX; li += TPB_T) {
    int r = li / TX, c = li - r * TX;
    int gx = x0 + c, gy = y0 + r;
    if (gx >= 0 && gx < w && gy >= 0 && gy < h) p[(size_t)gy * w + gx] = sp[r + 2][c + 2];
  }
}

// ---------------------------------------------------------------------------
// single-workgroup V-cycle in LDS
constexpr int MAXL = 16;
constexpr int TPB_C = 1024;
constexpr size_t LDS_BUDGET = 152 * 1024;

struct CoarseArgs {
  int n; // levels first .. L-2
  int w[MAXL], h[MAXL];
  float hh[MAXL];
  int op[MAXL], of[MAXL], om[MAXL]; // LDS byte offsets of p, f, mask
  const uint8_t *mask[MAXL];
  const float *f0;
  float *p0;
  int nu_pre, nu_post, nu_coarse;
  float alpha;
};

__device__ __forceinline__ void lds_sweep(float *P, const float *Fv, const uint8_t *M, int w,
                                          int h, float hh, float alpha) {
  const int half = (w - 1) / 2;
  const int n = (h - 2) * half;
  for (int color = 0; color < 2; color++) {
    for (int li = threadIdx.x; li < n; li += TPB_C) {
      int r = li / half, k = li - r * half;
      int y = r + 1;
      int x = 1 + ((y + color) & 1) + 2 * k;
      if (x < w - 1) {
        int i = y * w + x;
        P[i] = smooth(P[i], P[i - 1], P[i + 1], P[i - w], P[i + w], Fv[i], M[i], hh, alpha);
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ float lds_residual(const float *P, const float *Fv, const uint8_t *M,
                                              int i, int w, float ihsq) {
  unsigned mk = M[i];
  float pc = P[i];
  float val = 0.0f;
  val += (mk & MK_W) ? P[i - 1] : pc;
  val += (mk & MK_E) ? P[i + 1] : pc;
  val += (mk & MK_S) ? P[i - w] : pc;
  val += (mk & MK_N) ? P[i + w] : pc;
  val += -4.0f * pc;
  val *= ihsq;
  return (mk & MK_SELF) ? (Fv[i] + val) : 0.0f;
}

__device__ __forceinline__ float eps_sum(int s) { return (float)((double)s + 0.0001); }

__global__ void __launch_bounds__(TPB_C) k_coarse_vcycle(CoarseArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x;
  // load: f of the first level, masks of all levels; p = 0 everywhere
  for (int l = 0; l < a.n; l++) {
    float *P = (float *)(smem + a.op[l]);
    float *Fv = (float *)(smem + a.of[l]);
    uint8_t *M = (uint8_t *)(smem + a.om[l]);
    int n = a.w[l] * a.h[l];
    for (int i = tid; i < n; i += TPB_C) {
      P[i] = 0.0f;
      Fv[i] = (l == 0) ? a.f0[i] : 0.0f;
      M[i] = a.mask[l][i];
    }
  }
  __syncthreads();

  // descend
  for (int l = 0; l < a.n - 1; l++) {
    float *P = (float *)(smem + a.op[l]);
    float *Fv = (float *)(smem + a.of[l]);
    const uint8_t *M = (const uint8_t *)(smem + a.om[l]);
    const int w = a.w[l], h = a.h[l];
    for (int s = 0; s < a.nu_pre; s++) lds_sweep(P, Fv, M, w, h, a.hh[l], a.alpha);
    // residual + restrict into level l+1 (its p starts at 0)
    float *Fc = (float *)(smem + a.of[l + 1]);
    float *Pc = (float *)(smem + a.op[l + 1]);
    const int wc = a.w[l + 1], hc = a.h[l + 1];
    const float ihsq = 1.0f / a.hh[l] / a.hh[l];
    const int nc = (wc - 2) * (hc - 2);
    for (int li = tid; li < nc; li += TPB_C) {
      int y = li / (wc - 2) + 1, x = li % (wc - 2) + 1;
      int c = (2 * y) * w + 2 * x;
      auto R = [&](int dx, int dy) { return lds_residual(P, Fv, M, c + dy * w + dx, w, ihsq); };
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
    for (int s = 0; s < a.nu_coarse; s++) lds_sweep(P, Fv, M, a.w[l], a.h[l], a.hh[l], a.alpha);
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
    const int n = (w - 2) * (h - 2);
    for (int li = tid; li < n; li += TPB_C) {
      int y = li / (w - 2) + 1, x = li % (w - 2) + 1;
      int i = y * w + x;
      float F = (M[i] & MK_SELF) ? 1.0f : 0.0f;
      int ax = x >> 1, by = y >> 1;
      bool xe = !(x & 1), ye = !(y & 1);
      auto E = [&](int xx, int yy) { return Pc[yy * wc + xx]; };
      auto FC = [&](int xx, int yy) { return (int)(Mc[yy * wc + xx] & MK_SELF); };
      float e = 0.0f;
      if (xe && ye) {
        e = E(ax, by) * F;
      } else if (!xe && ye) {
        if (x <= w - 3) e = F * (E(ax, by) + E(ax + 1, by)) / eps_sum(FC(ax, by) + FC(ax + 1, by));
      } else if (xe && !ye) {
        if (y <= h - 3) e = F * (E(ax, by) + E(ax, by + 1)) / eps_sum(FC(ax, by) + FC(ax, by + 1));
      } else {
        if (x <= w - 3 && y <= h - 3)
          e = F * (E(ax, by) + E(ax + 1, by + 1) + E(ax + 1, by) + E(ax, by + 1)) /
              eps_sum(FC(ax, by) + FC(ax + 1, by + 1) + FC(ax + 1, by) + FC(ax, by + 1));
      }
      P[i] += e;
    }
    __syncthreads();
    for (int s = 0; s < a.nu_post; s++) lds_sweep(P, Fv, M, w, h, a.hh[l], a.alpha);
  }
  // write the first level's solution
  {
    const float *P = (const float *)(smem + a.op[0]);
    int n = a.w[0] * a.h[0];
    for (int i = tid; i < n; i += TPB_C) a.p0[i] = P[i];
  }
}

size_t coarse_lds_bytes(const Level *lv, int first, int levels, CoarseArgs *a) {
  size_t off = 0;
  int k = 0;
  for (int l = first; l <= levels - 2; l++, k++) {
    if (k >= MAXL) return (size_t)-1;
    size_t n = (size_t)lv[l].w * lv[l].h;
    if (a) {
      a->op[k] = (int)off;
      a->of[k] = (int)(off + ((n * 4 + 15) & ~size_t(15)));
      a->om[k] = (int)(off + 2 * ((n * 4 + 15) & ~size_t(15)));
    }
    off += 2 * ((n * 4 + 15) & ~size_t(15)) + ((n + 15) & ~size_t(15));
  }
  return off;
}

} // namespace

void launch_rbgs_fused(const Level &L, float h, float alpha, bool zero_gradient, hipStream_t s) {
  if (L.w < 3 || L.h < 3) return;
  int sx = ((L.w - 1) % TX == 0) ? 1 : 0;
  int sy = ((L.h - 1) % TY == 0) ? 1 : 0;
  int ntx = (L.w + sx + TX - 1) / TX, nty = (L.h + sy + TY - 1) / TY;
  hipLaunchKernelGGL(k_rbgs_tile, dim3(ntx * nty), dim3(TPB_T), 0, s, L.p, L.f, L.mask, L.w, L.h,
                     h, alpha, zero_gradient ? 1 : 0, sx, sy, ntx);
}

bool coarse_fits_lds(const Level *lv, int first, int levels) {
  if (first < 1 || first > levels - 2) return false;
  size_t b = coarse_lds_bytes(lv, first, levels, nullptr);
  return b <= LDS_BUDGET;
}

void launch_coarse_vcycle(const Level *lv, int first, int levels, float h, int nu_pre,
                          int nu_post, int nu_coarse, float alpha, hipStream_t s) {
  static bool attr_set = false;
  CoarseArgs a{};
  size_t bytes = coarse_lds_bytes(lv, first, levels, &a);
  a.n = levels - 1 - first;
  float hl = h;
  for (int k = 0; k < a.n; k++) {
    const Level &L = lv[first + k];
    a.w[k] = L.w;
    a.h[k] = L.h;
    a.mask[k] = L.mask;
    a.hh[k] = hl;
    if (k + 1 < a.n) hl = hl * (L.w - 1.0f) / (lv[first + k + 1].w - 1.0f); // pressure_solver.cpp:229
  }
  a.f0 = lv[first].f;
  a.p0 = lv[first].p;
  a.nu_pre = nu_pre;
  a.nu_post = nu_post;
  a.nu_coarse = nu_coarse;
  a.alpha = alpha;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void *)k_coarse_vcycle,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BUDGET);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_coarse_vcycle, dim3(1), dim3(TPB_C), bytes, s, a);
}

} // namespace ubg
