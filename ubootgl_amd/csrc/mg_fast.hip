// mg_fast.hip -- fused multigrid kernels for gfx950.
//
// 1. k_rbgs_stream: one full red+black sweep of rbgs (pressure_solver.cpp:49-89,
//    two-pass semantics) in one launch.  Each wave streams down a strip of
//    columns holding rows of p in registers; the red cells of row y+1 are
//    updated before the black cells of row y, which is exactly the order the
//    two-pass sweep implies (red reads only old black values, black reads only
//    new red values).  With zero_gradient the setZeroGradientBC copy
//    (pressure_solver.cpp:183-192) is applied as each row becomes final.  The
//    sweep reads p from one buffer and writes the other (the level swaps
//    them), so neighbouring waves never read each other's new values.
//    HBM traffic per sweep: p read + write, f read, 1-byte mask read =
//    13 bytes/cell; halo rows/columns are re-read from L2, and waves are ordered
//    so each XCD (own L2) works on a contiguous band of the grid.
//
// 2. k_coarse_vcycle: the V-cycle of every level from `first` down to the
//    coarsest solved level L-2 (pressure_solver.cpp:201-248), run by one
//    workgroup entirely in LDS (p, f and mask of all those levels fit in
//    160 KB once the level is <= ~128x64).  Replaces ~40 latency-bound
//    launches per V-cycle with one.
#include "coarse_device.hpp"
#include "stream_device.hpp"

namespace ubg {

namespace {

constexpr int TPB_T = 256;

// one wave per unit (strip x row block) of the level; body in stream_device.hpp
template <bool ZG, bool AL>
__global__ void __launch_bounds__(TPB_T)
    k_rbgs_pair(const float *__restrict__ src, float *__restrict__ dst,
                const float *__restrict__ f, const uint8_t *__restrict__ m, int w, int h,
                float hh, float alpha, int nstrips, int R, int nunits) {
  // wave-uniform values are made scalar so branches on them are uniform
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int unit = xcd_tile(blockIdx.x, gridDim.x) * (TPB_T / 64) + wid;
  if (unit >= nunits) return; // whole wave leaves together
  rbgs_pair_unit<ZG, AL>(src, dst, f, m, w, h, hh, alpha, nstrips, R, unit);
}

// ---------------------------------------------------------------------------
// single-workgroup V-cycle in LDS
__global__ void __launch_bounds__(TPB_C) k_coarse_vcycle(CoarseArgs a) {
  coarse_vcycle_body(a, a.f0, a.p0);
}

} // namespace

void launch_rbgs_fused(Level &L, float h, float alpha, bool zero_gradient, hipStream_t s) {
  if (L.w < 3 || L.h < 3) return;
  const int nstrips = (L.w + SW2 - 1) / SW2;
  // rows per wave: enough waves to fill the chip, at most 64 rows
  int R = (int)((long)nstrips * (L.h - 2) / 4096);
  R = R < 4 ? 4 : (R > 64 ? 64 : R);
  const int nrb = (L.h - 2 + R - 1) / R;
  const int nunits = nstrips * nrb;
  const int nblocks = (nunits + (TPB_T / 64) - 1) / (TPB_T / 64);
  const bool al = (L.w % 2) == 0;
  auto k = zero_gradient ? (al ? k_rbgs_pair<true, true> : k_rbgs_pair<true, false>)
                         : (al ? k_rbgs_pair<false, true> : k_rbgs_pair<false, false>);
  launch_k(k, dim3(nblocks), dim3(TPB_T), 0, s, L.p, L.p_alt, L.f, L.mask, L.w, L.h, h,
                     alpha, nstrips, R, nunits);
  float *t = L.p;
  L.p = L.p_alt;
  L.p_alt = t;
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
