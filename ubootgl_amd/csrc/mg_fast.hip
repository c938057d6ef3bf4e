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

namespace ubg {

namespace {

constexpr int TPB_T = 256;
constexpr int SW2 = 124;    // output columns per wave in the streaming smoother
constexpr int PREFETCH = 4; // rows loaded ahead by each streaming wave

// One wave = one strip of columns; lane l owns the horizontal cell pair
// (x0, x0+1), x0 = SW2*s - 2 + 2l, so each row of a pair holds exactly one
// red and one black cell and every smoother evaluation is useful.  Lanes
// 1..62 own their pair (SW2 = 124 columns per wave); lanes 0 and 63 are the
// halo the red update needs.  The wave streams down R rows keeping rows
// y-1..y+1 in registers (plus a D-row prefetch queue): at row y it updates
// the red cell of row y+1 (old black neighbours) and the black cell of row y
// (neighbours: red cells of rows y-1, y+1 and the row's own red cells, all
// already updated), which is exactly the two-pass order; row y is then final
// and stored.  The horizontal neighbour outside the pair comes from the
// adjacent lane (one __shfl per colour and row).  Reads come from `src`,
// writes go to `dst` (ping-pong), so no wave sees another wave's new values.
// AL: w is even, so every pair is 8-byte aligned and loads/stores are float2.
template <bool ZG, bool AL>
__global__ void __launch_bounds__(TPB_T)
    k_rbgs_pair(const float *__restrict__ src, float *__restrict__ dst,
                const float *__restrict__ f, const uint8_t *__restrict__ m, int w, int h,
                float hh, float alpha, int nstrips, int R, int nunits) {
  const int lane = threadIdx.x & 63;
  // wave-uniform values are made scalar so branches on them are uniform
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int unit = xcd_tile(blockIdx.x, gridDim.x) * (TPB_T / 64) + wid;
  if (unit >= nunits) return; // whole wave leaves together
  const int s = unit % nstrips, r = unit / nstrips;
  const int x0 = SW2 * s - 2 + 2 * lane, x1 = x0 + 1; // x0 even for every lane
  const bool in0 = x0 >= 0 && x0 < w, in1 = x1 >= 0 && x1 < w;
  const bool int0 = x0 >= 1 && x0 <= w - 2, int1 = x1 >= 1 && x1 <= w - 2;
  const bool own = lane >= 1 && lane <= 62;
  const int ya = 1 + R * r;
  const int yb = min(ya + R - 1, h - 2);
  // Loads use clamped (always valid) addresses; values at clamped positions
  // only ever feed cells that are not updated (halo lanes, boundary rows).
  const int c0 = min(max(x0, 0), w - 1), c1 = min(max(x1, 0), w - 1);
  const int cp = min(max(x0, 0), w - 2); // aligned pair start when AL
  auto rowoff = [&](int yy) { return (size_t)min(max(yy, 0), h - 1) * w; };
  auto ldp = [&](int yy, float &a, float &b) {
    const float *row = src + rowoff(yy);
    if (AL) {
      float2 v = *(const float2 *)(row + cp);
      a = v.x;
      b = v.y;
    } else {
      a = row[c0];
      b = row[c1];
    }
  };
  auto ldfm = [&](int yy, float &fa, float &fb, unsigned &ma, unsigned &mb) {
    const size_t o = rowoff(yy);
    if (AL) {
      float2 v = *(const float2 *)(f + o + cp);
      uchar2 q = *(const uchar2 *)(m + o + cp);
      fa = v.x;
      fb = v.y;
      ma = q.x;
      mb = q.y;
    } else {
      fa = f[o + c0];
      fb = f[o + c1];
      ma = m[o + c0];
      mb = m[o + c1];
    }
  };
  // New value of cell a (x0) or cell b (x1) of the pair in an interior row;
  // the horizontal neighbour outside the pair comes from the adjacent lane.
  auto relax_a = [&](float a, float b, float dn, float up, float fv, unsigned mk) {
    float west = from_left(b);
    float na = smooth(a, west, b, dn, up, fv, mk, hh, alpha);
    return int0 ? na : a;
  };
  auto relax_b = [&](float a, float b, float dn, float up, float fv, unsigned mk) {
    float east = from_right(a);
    float nb = smooth(b, a, east, dn, up, fv, mk, hh, alpha);
    return int1 ? nb : b;
  };
// colour of row yy: x0 is even, so cell a has colour (yy & 1), cell b the other
#define RELAX(A, B, DNA, DNB, UPA, UPB, FA, FB, MA, MB, YY, ODD)                                \
  do {                                                                                          \
    if ((((YY) + (ODD)) & 1) == 0) A = relax_a(A, B, DNA, UPA, FA, MA);                         \
    else B = relax_b(A, B, DNB, UPB, FB, MB);                                                   \
  } while (0)

  float pm1a, pm1b, p0a, p0b, p1a, p1b, p2a, p2b, p3a, p3b, pm2a, pm2b;
  float fm1a, fm1b, f0a, f0b, f1a, f1b, f2a, f2b;
  unsigned mm1a, mm1b, m0a, m0b, m1a, m1b, m2a, m2b;
  ldp(ya - 2, pm2a, pm2b);
  ldp(ya - 1, pm1a, pm1b);
  ldp(ya, p0a, p0b);
  ldp(ya + 1, p1a, p1b);
  ldp(ya + 2, p2a, p2b);
  ldp(ya + 3, p3a, p3b);
  ldfm(ya - 1, fm1a, fm1b, mm1a, mm1b);
  ldfm(ya, f0a, f0b, m0a, m0b);
  ldfm(ya + 1, f1a, f1b, m1a, m1b);
  ldfm(ya + 2, f2a, f2b, m2a, m2b);
  if (ya - 1 >= 1) RELAX(pm1a, pm1b, pm2a, pm2b, p0a, p0b, fm1a, fm1b, mm1a, mm1b, ya - 1, 1);
  RELAX(p0a, p0b, pm1a, pm1b, p1a, p1b, f0a, f0b, m0a, m0b, ya, 1);
  for (int y = ya; y <= yb; y++) {
    float p4a, p4b, f3a, f3b; // two rows of lead
    unsigned m3a, m3b;
    ldp(y + 4, p4a, p4b);
    ldfm(y + 3, f3a, f3b, m3a, m3b);
    if (y + 1 <= h - 2) RELAX(p1a, p1b, p0a, p0b, p2a, p2b, f1a, f1b, m1a, m1b, y + 1, 1);
    RELAX(p0a, p0b, pm1a, pm1b, p1a, p1b, f0a, f0b, m0a, m0b, y, 0); // row y final
    float oa = p0a, ob = p0b;
    if (ZG) { // setZeroGradientBC on the side columns (rows 1..h-2)
      float left = from_left(p0b);
      oa = (x0 == 0) ? p0b : ((x0 == w - 1) ? left : oa);
      ob = (x1 == w - 1) ? p0a : ob;
    }
    {
      float *row = dst + (size_t)y * w;
      if (AL) { // w even: an owned in-domain pair is whole and 8-byte aligned
        if (own && in0) *(float2 *)(row + x0) = make_float2(oa, ob);
      } else {
        if (own && in0) row[x0] = oa;
        if (own && in1) row[x1] = ob;
      }
    }
    if ((y == 1 || y == h - 2) && own) { // rows 0 / h-1: ZG copies, corners kept
      float *edge = dst + (y == 1 ? 0 : (size_t)(h - 1) * w);
      float ea = (ZG && int0) ? p0a : (y == 1 ? pm1a : p1a);
      float eb = (ZG && int1) ? p0b : (y == 1 ? pm1b : p1b);
      if (in0) edge[x0] = ea;
      if (in1) edge[x1] = eb;
    }
    pm1a = p0a; pm1b = p0b;
    p0a = p1a; p0b = p1b;
    p1a = p2a; p1b = p2b;
    p2a = p3a; p2b = p3b;
    p3a = p4a; p3b = p4b;
    f0a = f1a; f0b = f1b;
    f1a = f2a; f1b = f2b;
    f2a = f3a; f2b = f3b;
    m0a = m1a; m0b = m1b;
    m1a = m2a; m1b = m2b;
    m2a = m3a; m2b = m3b;
  }
#undef RELAX
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
