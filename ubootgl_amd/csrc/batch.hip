// batch.hip -- batched refinement solves (ubg_batch_*): many independent
// domains of uneven size, one launch per multigrid stage.
//
// Per refinement step and domain the arithmetic is exactly that of
// ubg_mg_ir_solve (refine.hip) around MG::solve (pressure_solver.cpp:201-248):
// nu_pre rbgs sweeps (pressure_solver.cpp:49-89), residual + restriction
// (:91-132), the coarse levels from the first one that fits in LDS (one
// workgroup per domain), prolongation + correction (:134-181), nu_post
// sweeps; then the fused float64 update x += e, r = f + A x, and the norm of
// r summed in the single-domain order.  Stages of all domains share one
// launch: a kernel walks a tile list (domain, tile origin) with a fixed grid
// (persistent loop), and reads per-domain buffer indices from a small table
// row of the stage, so domains with different level counts and ping-pong
// parities run in the same launch.  Domains whose solve has ended are
// skipped on the device; no host round trip is needed per stage.
#include "batch_engine.hpp"
#include "stream_device.hpp"
#include "tile_device.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace ubg {

// ===========================================================================
// kernels
// ===========================================================================
namespace {

constexpr int TPB = 256;

struct StageArgs {
  const BLev *lev;
  const BDom *dom;
  const int *off; // tile prefix over domain slots [ndom + 1], tiles per row [ndom], tile width [ndom]
  int ndom;
  int ntiles;
  int level;
  const uint8_t *ia; // per-domain index row a
  const uint8_t *ib; // per-domain index row b
  float alpha;
  int oty; // tile rows of the fused level passes (their widths are per domain)
  int zg;  // row-streamed level-0 passes apply setZeroGradientBC (projection solves)
};

__device__ __forceinline__ bool dom_active(const BDom &D) {
  return __builtin_amdgcn_readfirstlane(D.st->active) != 0;
}

// Tiles of a stage are split into equal contiguous chunks, one per block of
// the (persistent) grid, in XCD-grouped block order so that neighbouring
// chunks share an XCD's L2.  A block locates the domain slot of its first tile
// once (the last d with off[d] <= t; domains without tiles at this level
// have off[d] == off[d+1]) and then walks forward.
struct TileWalk {
  int t, t1; // current tile, end of the chunk
  int d, j;  // domain slot, tile index within the domain
  int nt, ntx, otx;
  __device__ __forceinline__ bool init(const StageArgs &a) {
    const int per = (a.ntiles + gridDim.x - 1) / gridDim.x;
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    t = b * per;
    t1 = min(t + per, a.ntiles);
    if (t >= t1) return false;
    int lo = 0, hi = a.ndom;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.off[mid] <= t) lo = mid;
      else hi = mid;
    }
    d = lo;
    j = t - a.off[d];
    nt = a.off[d + 1] - a.off[d];
    ntx = a.off[a.ndom + 1 + d];
    otx = a.off[2 * a.ndom + 1 + d];
    return true;
  }
  __device__ __forceinline__ bool next(const StageArgs &a) {
    if (++t >= t1) return false;
    if (++j == nt) {
      do {
        d++;
      } while (a.off[d + 1] == a.off[d]);
      j = 0;
      nt = a.off[d + 1] - a.off[d];
      ntx = a.off[a.ndom + 1 + d];
      otx = a.off[2 * a.ndom + 1 + d];
    }
    return true;
  }
  __device__ __forceinline__ int tx() const { return j % ntx; }
  __device__ __forceinline__ int ty() const { return j / ntx; }
};

// One red+black rbgs sweep of a TX x TY tile, ping-pong (read p[a], write
// p[a^1]).  p is staged with a 2-cell halo; red cells ((x+y) odd) of the tile
// and a 1-cell ring are updated in LDS from old black values, then black
// cells of the tile from the new red values -- the two-pass order of
// pressure_solver.cpp:65-71, bit for bit.
template <int TX, int TY>
__global__ void __launch_bounds__(TPB) kb_sweep(StageArgs a) {
  static_assert(TX % 2 == 0, "pairs need an even tile width");
  constexpr int PX = TX + 4, PY = TY + 4, RX = TX + 2, RY = TY + 2;
  __shared__ float sp[PY * PX];
  TileWalk tw;
  if (!tw.init(a)) return;
  do {
    const Tile T{tw.d, tw.tx() * TX, tw.ty() * TY, tw.j};
    const BDom &D = a.dom[T.d];
    if (!dom_active(D)) continue;
    const BLev &L = a.lev[T.d * BMAXL + a.level];
    const int s = a.ia[T.d];
    const float *__restrict__ src = L.p[s];
    float *__restrict__ dst = L.p[s ^ 1];
    const float *__restrict__ fv = L.f;
    const uint8_t *__restrict__ mk = L.mask;
    const int w = L.w, h = L.h;
    const float hh = L.hh;
    __syncthreads(); // the previous tile's readers are done with sp
    for (int i = threadIdx.x; i < PX * PY; i += TPB) {
      const int r = i / PX, c = i - r * PX;
      const int gx = T.x0 - 2 + c, gy = T.y0 - 2 + r;
      sp[i] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? src[(size_t)gy * w + gx] : 0.0f;
    }
    __syncthreads();
    // red: one red cell per horizontal pair (columns x0-1+2q, x0+2q)
    for (int i = threadIdx.x; i < RY * (RX / 2); i += TPB) {
      const int r = i / (RX / 2), q = i - r * (RX / 2);
      const int gy = T.y0 - 1 + r;
      const int c0 = T.x0 - 1 + 2 * q;
      const int gx = ((c0 + gy) & 1) ? c0 : c0 + 1;
      if (gx >= 1 && gx <= w - 2 && gy >= 1 && gy <= h - 2) {
        const size_t gi = (size_t)gy * w + gx;
        const int li = (r + 1) * PX + (gx - T.x0 + 2);
        sp[li] = smooth(sp[li], sp[li - 1], sp[li + 1], sp[li - PX], sp[li + PX], fv[gi], mk[gi],
                        hh, a.alpha);
      }
    }
    __syncthreads();
    // black: one black cell per pair (columns x0+2q, x0+2q+1); both cells of
    // the pair are stored
    for (int i = threadIdx.x; i < TY * (TX / 2); i += TPB) {
      const int r = i / (TX / 2), q = i - r * (TX / 2);
      const int gy = T.y0 + r;
      if (gy >= h) continue;
      const int c0 = T.x0 + 2 * q;
      const int li0 = (r + 2) * PX + 2 * q + 2;
      float v0 = sp[li0], v1 = sp[li0 + 1];
      const int bx = ((c0 + gy) & 1) ? c0 + 1 : c0; // the black one
      if (bx < w && bx >= 1 && bx <= w - 2 && gy >= 1 && gy <= h - 2) {
        const size_t gi = (size_t)gy * w + bx;
        const int li = li0 + (bx - c0);
        const float nb = smooth(sp[li], sp[li - 1], sp[li + 1], sp[li - PX], sp[li + PX], fv[gi],
                                mk[gi], hh, a.alpha);
        if (bx == c0) v0 = nb;
        else v1 = nb;
      }
      const size_t g0 = (size_t)gy * w + c0;
      if (c0 < w) dst[g0] = v0;
      if (c0 + 1 < w) dst[g0 + 1] = v1;
    }
  } while (tw.next(a));
}

// calculateResidualField + restrict (pressure_solver.cpp:91-132) for a tile
// of 32 x 8 interior coarse cells; the coarse solution is zeroed (:224).
// ia: fine p index, ib: coarse p index.
constexpr int RCX = 32, RCY = 8, RFX = 2 * RCX + 1, RFY = 2 * RCY + 1, RPX = RFX + 2,
              RPY = RFY + 2;
__global__ void __launch_bounds__(TPB) kb_rr(StageArgs a) {
  __shared__ float sp[RPY][RPX];
  __shared__ float sr[RFY][RFX + 1];
  TileWalk tw;
  if (!tw.init(a)) return;
  do {
    const Tile T{tw.d, 1 + tw.tx() * RCX, 1 + tw.ty() * RCY, tw.j};
    const BDom &D = a.dom[T.d];
    if (!dom_active(D)) continue;
    const BLev &F = a.lev[T.d * BMAXL + a.level];
    const BLev &C = a.lev[T.d * BMAXL + a.level + 1];
    const float *__restrict__ p = F.p[a.ia[T.d]];
    const int wf = F.w, hf = F.h, wc = C.w, hc = C.h;
    const int fx0 = 2 * T.x0 - 1, fy0 = 2 * T.y0 - 1; // residual region origin
    __syncthreads();
    for (int i = threadIdx.x; i < RPX * RPY; i += TPB) {
      const int r = i / RPX, c = i - r * RPX;
      const int gx = fx0 - 1 + c, gy = fy0 - 1 + r;
      (&sp[0][0])[i] = (gx >= 0 && gx < wf && gy >= 0 && gy < hf) ? p[(size_t)gy * wf + gx] : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < RFX * RFY; i += TPB) {
      const int r = i / RFX, c = i - r * RFX;
      const int gx = fx0 + c, gy = fy0 + r;
      float res = 0.0f;
      if (gx >= 1 && gx < wf - 1 && gy >= 1 && gy < hf - 1) {
        const size_t gi = (size_t)gy * wf + gx;
        res = residual_of(sp[r + 1][c + 1], sp[r + 1][c], sp[r + 1][c + 2], sp[r][c + 1],
                          sp[r + 2][c + 1], F.f[gi], F.mask[gi], F.ihsq);
      }
      sr[r][c] = res;
    }
    __syncthreads();
    const int cx = threadIdx.x % RCX, cy = threadIdx.x / RCX;
    const int X = T.x0 + cx, Y = T.y0 + cy;
    if (X < wc - 1 && Y < hc - 1) {
      const int r = 2 * cy + 1, c = 2 * cx + 1;
      auto R = [&](int dx, int dy) { return sr[r + dy][c + dx]; };
      float vv = 0.0f;
      vv = R(-1, -1) * 1.0f + R(0, -1) * 2.0f + R(1, -1) * 1.0f;
      vv += R(-1, 0) * 2.0f + R(0, 0) * 4.0f + R(1, 0) * 2.0f;
      vv += R(-1, 1) * 1.0f + R(0, 1) * 2.0f + R(1, 1) * 1.0f;
      const size_t ci = (size_t)Y * wc + X;
      C.f[ci] = vv * (1.0f / 16.0f);
      C.p[a.ib[T.d]][ci] = 0.0f;
    }
  } while (tw.next(a));
}

// prolongate + correct (pressure_solver.cpp:134-181) for a tile of 256 x 8
// interior fine cells (one column per thread), in place on p[ia]; the coarse
// solution is p[ib] of the next level.
constexpr int PCW = 256, PCH = 8;
__global__ void __launch_bounds__(TPB) kb_pc(StageArgs a) {
  TileWalk tw;
  if (!tw.init(a)) return;
  do {
    const Tile T{tw.d, 1 + tw.tx() * PCW, 1 + tw.ty() * PCH, tw.j};
    const BDom &D = a.dom[T.d];
    if (!dom_active(D)) continue;
    const BLev &F = a.lev[T.d * BMAXL + a.level];
    const BLev &C = a.lev[T.d * BMAXL + a.level + 1];
    const int w = F.w, h = F.h, wc = C.w;
    const int x = T.x0 + threadIdx.x;
    if (x >= w - 1) continue;
    float *__restrict__ p = F.p[a.ia[T.d]];
    const float *__restrict__ ec = C.p[a.ib[T.d]];
    const uint8_t *__restrict__ mc = C.mask;
    const int ax = x >> 1;
    const bool xe = !(x & 1);
    auto E = [&](int xx, int yy) { return ec[(size_t)yy * wc + xx]; };
    auto FC = [&](int xx, int yy) { return (int)(mc[(size_t)yy * wc + xx] & MK_SELF); };
    const int y1 = min(T.y0 + PCH, h - 1);
    for (int y = T.y0; y < y1; y++) {
      const size_t i = (size_t)y * w + x;
      const float Fv = (F.mask[i] & MK_SELF) ? 1.0f : 0.0f;
      const int by = y >> 1;
      const bool ye = !(y & 1);
      float e = 0.0f;
      if (xe && ye) {
        e = E(ax, by) * Fv;
      } else if (!xe && ye) {
        if (x <= w - 3) e = Fv * (E(ax, by) + E(ax + 1, by)) / eps_count(FC(ax, by) + FC(ax + 1, by));
      } else if (xe && !ye) {
        if (y <= h - 3) e = Fv * (E(ax, by) + E(ax, by + 1)) / eps_count(FC(ax, by) + FC(ax, by + 1));
      } else {
        if (x <= w - 3 && y <= h - 3)
          e = Fv * (E(ax, by) + E(ax + 1, by + 1) + E(ax + 1, by) + E(ax, by + 1)) /
              eps_count(FC(ax, by) + FC(ax + 1, by + 1) + FC(ax + 1, by) + FC(ax, by + 1));
      }
      p[i] += e;
    }
  } while (tw.next(a));
}

// The fused level passes of tile_device.hpp over the tiles of every member
// domain at one level (each domain's own even column split, oty rows, walked
// as the other stages): pre reads p[ia], writes p[ia^1], the coarse rhs and
// a zero coarse p[ib]; post reads p[ia] and the coarse solution p[ib],
// writes p[ia^1].  ZP (every level below the finest): p[ia] is the zero
// solution the restriction above wrote, and pre does not read it.
template <int NU, int TPB_T, bool ZP>
__global__ void __launch_bounds__(TPB_T) kb_tile_pre(StageArgs a) {
  extern __shared__ __align__(16) char smem[];
  TileWalk tw;
  if (!tw.init(a)) return;
  do {
    const BDom &D = a.dom[tw.d];
    if (!dom_active(D)) continue;
    const BLev &F = a.lev[tw.d * BMAXL + a.level];
    const BLev &C = a.lev[tw.d * BMAXL + a.level + 1];
    const int s = a.ia[tw.d];
    __syncthreads(); // the previous tile's LDS readers are done
    tile_pre_body<NU, TPB_T, ZP>(smem, tw.tx() * tw.otx, tw.ty() * a.oty, tw.otx, a.oty,
                                 F.p[s], F.p[s ^ 1], F.f, F.mask, F.w, F.h, F.hh, a.alpha,
                                 F.ihsq, C.f, C.p[a.ib[tw.d]], C.w, C.h);
  } while (tw.next(a));
}
template <int NU, int TPB_T>
__global__ void __launch_bounds__(TPB_T) kb_tile_post(StageArgs a) {
  extern __shared__ __align__(16) char smem[];
  TileWalk tw;
  if (!tw.init(a)) return;
  do {
    const BDom &D = a.dom[tw.d];
    if (!dom_active(D)) continue;
    const BLev &F = a.lev[tw.d * BMAXL + a.level];
    const BLev &C = a.lev[tw.d * BMAXL + a.level + 1];
    const int s = a.ia[tw.d];
    __syncthreads();
    tile_post_body<NU, TPB_T>(smem, tw.tx() * tw.otx, tw.ty() * a.oty, tw.otx, a.oty, F.p[s],
                              F.p[s ^ 1], F.f, F.mask, F.w, F.h, F.hh, a.alpha,
                              C.p[a.ib[tw.d]], C.mask, C.w, C.h);
  } while (tw.next(a));
}

// Row-streamed level passes (stream_device.hpp) over the units (strip x row
// block, one wave each) of every member domain whose level is streamed.  A
// wave finds its domain slot in the unit prefix (the last d with off[d] <=
// unit, as the tile walk does); the table's second and third rows hold the
// domain's strip count and row-block height.
struct StreamUnit {
  int d, unit, nstrips, R;
};
__device__ __forceinline__ bool stream_unit(const StageArgs &a, StreamUnit &u) {
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = xcd_tile(blockIdx.x, gridDim.x) * (TPB / 64) + wid;
  if (g >= a.ntiles) return false; // whole wave leaves together
  int lo = 0, hi = a.ndom;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.off[mid] <= g) lo = mid;
    else hi = mid;
  }
  u.d = lo;
  u.unit = g - a.off[lo];
  u.nstrips = a.off[a.ndom + 1 + lo];
  u.R = a.off[2 * a.ndom + 1 + lo];
  return true;
}

// one red+black sweep: reads p[ia], writes p[ia^1]
__global__ void __launch_bounds__(TPB) kb_stream_sweep(StageArgs a) {
  StreamUnit u;
  if (!stream_unit(a, u)) return;
  if (!dom_active(a.dom[u.d])) return;
  const BLev &F = a.lev[u.d * BMAXL + a.level];
  const int s = a.ia[u.d];
  const bool al = (F.w & 1) == 0;
  if (a.zg) {
    if (al)
      rbgs_pair_unit<true, true>(F.p[s], F.p[s ^ 1], F.f, F.mask, F.w, F.h, F.hh, a.alpha,
                                 u.nstrips, u.R, u.unit);
    else
      rbgs_pair_unit<true, false>(F.p[s], F.p[s ^ 1], F.f, F.mask, F.w, F.h, F.hh, a.alpha,
                                  u.nstrips, u.R, u.unit);
  } else if (al) {
    rbgs_pair_unit<false, true>(F.p[s], F.p[s ^ 1], F.f, F.mask, F.w, F.h, F.hh, a.alpha,
                                u.nstrips, u.R, u.unit);
  } else {
    rbgs_pair_unit<false, false>(F.p[s], F.p[s ^ 1], F.f, F.mask, F.w, F.h, F.hh, a.alpha,
                                 u.nstrips, u.R, u.unit);
  }
}

// MODE_RR: last pre-sweep + residual + restriction (coarse rhs, zero coarse
// p[ib]); MODE_PC: prolongation of the coarse p[ib] + correction + first
// post-sweep.  Both read p[ia] and write p[ia^1].
template <int MODE>
__global__ void __launch_bounds__(TPB) kb_stream_pass(StageArgs a) {
  StreamUnit u;
  if (!stream_unit(a, u)) return;
  if (!dom_active(a.dom[u.d])) return;
  const BLev &F = a.lev[u.d * BMAXL + a.level];
  const BLev &C = a.lev[u.d * BMAXL + a.level + 1];
  const int s = a.ia[u.d];
  StreamArgs A{};
  A.src = F.p[s];
  A.dst = F.p[s ^ 1];
  A.f = F.f;
  A.m = F.mask;
  A.w = F.w;
  A.h = F.h;
  A.hh = F.hh;
  A.alpha = a.alpha;
  A.ihsq = F.ihsq;
  A.crc = C.f;
  A.cec = C.p[a.ib[u.d]];
  A.cpin = C.p[a.ib[u.d]];
  A.cmask = C.mask;
  A.wc = C.w;
  A.hc = C.h;
  A.nstrips = u.nstrips;
  A.R = u.R;
  const bool al = (F.w & 1) == 0;
  if (a.zg) {
    if (al) mg_stream_unit<1, MODE, true, true>(A, u.unit);
    else mg_stream_unit<1, MODE, true, false>(A, u.unit);
  } else if (al) {
    mg_stream_unit<1, MODE, false, true>(A, u.unit);
  } else {
    mg_stream_unit<1, MODE, false, false>(A, u.unit);
  }
}

// the coarse levels of each domain in LDS, one workgroup per domain slot;
// ia: index of p at the domain's first LDS level
__global__ void __launch_bounds__(TPB_C) kb_coarse(StageArgs a, const CoarseArgs *cargs,
                                                   int nslots) {
  const int d = blockIdx.x;
  if (d >= nslots) return;
  const BDom &D = a.dom[d];
  if (!dom_active(D) || D.cfirst >= D.nlev) return;
  const BLev &L = a.lev[d * BMAXL + D.cfirst];
  coarse_vcycle_body(cargs[d], L.f, L.p[a.ia[d]]);
}

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// The refinement update of refine.hip's k_ir_update for 256 x 8 tiles of
// every domain: ia = e index of level 0; ib bit 0 = x index, bit 1 = add e.
__global__ void __launch_bounds__(B_UPD_W) kb_update(StageArgs a) {
  __shared__ double red[B_UPD_W / 64];
  TileWalk tw;
  if (!tw.init(a)) return;
  do {
    const Tile T{tw.d, tw.tx() * B_UPD_W, tw.ty() * B_UPD_H, tw.j};
    const BDom &D = a.dom[T.d];
    if (!dom_active(D)) continue;
    const BLev &L0 = a.lev[T.d * BMAXL];
    const int ie = a.ia[T.d], b = a.ib[T.d];
    const bool ACC = (b >> 1) & 1;
    const double *__restrict__ xa = D.x[b & 1];
    double *__restrict__ xb = D.x[(b & 1) ^ 1];
    const float *__restrict__ e = L0.p[ie];
    float *__restrict__ ez = L0.p[ie ^ 1];
    const float *__restrict__ f = D.rhs;
    const uint8_t *__restrict__ m = L0.mask;
    float *__restrict__ r32 = L0.f;
    const int w = D.w, h = D.h;
    const double ihsq = 1.0 / ((double)L0.hh * (double)L0.hh);
    const int x = T.x0 + threadIdx.x;
    const int y0 = T.y0, y1 = min(y0 + B_UPD_H, h);
    double acc = 0.0;
    if (x < w) {
      auto X = [&](size_t i) { return ACC ? xa[i] + (double)e[i] : xa[i]; };
      size_t c = (size_t)y0 * w + x;
      double xs = y0 > 0 ? X(c - w) : 0.0;
      double xc = X(c);
      for (int y = y0; y < y1; y++, c += w) {
        const double xn = y + 1 < h ? X(c + w) : 0.0;
        xb[c] = xc;
        ez[c] = 0.0f;
        double r = 0.0;
        const unsigned mk = m[c];
        if (x >= 1 && x <= w - 2 && y >= 1 && y <= h - 2 && (mk & MK_SELF)) {
          double s = 0.0;
          if (mk & MK_W) s += X(c - 1) - xc;
          if (mk & MK_E) s += X(c + 1) - xc;
          if (mk & MK_S) s += xs - xc;
          if (mk & MK_N) s += xn - xc;
          r = (double)f[c] + s * ihsq;
        }
        r32[c] = (float)r;
        acc += r * r;
        xs = xc;
        xc = xn;
      }
    }
    acc = wsum(acc);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int i = 0; i < B_UPD_W / 64; i++) s += red[i];
      D.partial[T.i] = s;
    }
  } while (tw.next(a));
}

// per-domain norm (the order of refine.hip's k_sum_partials) and the
// convergence rule of ubg_mg_ir_solve; ia = 1 for a domain's first update,
// ib = its x index after the update.  One 1024-thread block per domain slot.
__global__ void __launch_bounds__(1024) kb_norm(StageArgs a, int nslots) {
  const int d = blockIdx.x;
  if (d >= nslots) return;
  const BDom &D = a.dom[d];
  if (!dom_active(D)) return;
  const int n = D.gx * D.gy;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += D.partial[i];
  __shared__ double red[16];
  s = wsum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double n2 = 0.0;
  for (int i = 0; i < 16; i++) n2 += red[i];
  DState &S = *D.st;
  S.xcur = a.ib[d];
  if (a.ia[d]) { // first update: the initial residual
    const double r = sqrt(n2);
    S.r0 = S.ref > 0.0 ? S.ref : r;
    S.iters = 0;
    S.vcycles = 0;
    S.stall = 0;
    S.best = 1.0;
    S.converged = 0;
    if (!isfinite(r)) {
      S.nonfinite = 1;
      S.active = 0;
      return;
    }
    S.rel = r == 0.0 ? 0.0 : r / S.r0;
    if (r == 0.0 || S.rel <= S.tol) { // nothing to do (a warm start may already meet tol)
      S.converged = 1;
      S.active = 0;
    }
    return;
  }
  const double rel = sqrt(n2) / S.r0;
  S.iters += 1;
  S.vcycles += S.k;
  S.rel = rel;
  if (S.iters <= UBG_IR_HISTORY) S.hist[S.iters - 1] = rel;
  if (!isfinite(rel)) {
    S.nonfinite = 1;
    S.active = 0;
    return;
  }
  if (rel <= S.tol) {
    S.converged = 1;
    S.active = 0;
    return;
  }
  if (rel > 0.9 * S.best) {
    if (++S.stall >= 3) S.active = 0;
  } else {
    S.stall = 0;
  }
  S.best = fmin(S.best, rel);
  if (S.iters >= S.max_iter) S.active = 0;
}

// Level-0 copies of up to kCopyJobs domains in one launch (batched
// simulation steps): block (x, j) copies a stride of job j's p (and f when
// its source is set).  The jobs travel as the kernel's argument.
constexpr int kCopyJobs = 64;
struct CopyJob {
  const float *sp, *sf;
  float *dp, *df;
  int n;
};
struct CopyJobs {
  CopyJob j[kCopyJobs];
};
__global__ void __launch_bounds__(256) kb_copy_level0(CopyJobs jobs, int njobs) {
  if ((int)blockIdx.y >= njobs) return;
  const CopyJob &J = jobs.j[blockIdx.y];
  for (int i = blockIdx.x * 256 + threadIdx.x; i < J.n; i += gridDim.x * 256) {
    J.dp[i] = J.sp[i];
    if (J.sf) J.df[i] = J.sf[i];
  }
}

__global__ void __launch_bounds__(256) kb_count(const BDom *dom, int n, int *out,
                                                uint8_t *flags) {
  __shared__ int red[4];
  int c = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int a = dom[i].st->active ? 1 : 0;
    flags[i] = (uint8_t)a;
    c += a;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

// the compiled smoother tile shapes
struct Shape {
  int tx, ty;
};
constexpr Shape kShapes[] = {{32, 8}, {64, 4}, {64, 8}, {64, 16}, {128, 4}, {128, 8}, {256, 4}};

void launch_sweep(const Tuning &tn, int blocks, hipStream_t s, const StageArgs &a) {
#define UBG_SW(X, Y)                                                                            \
  if (tn.tx == X && tn.ty == Y) {                                                               \
    hipLaunchKernelGGL((kb_sweep<X, Y>), dim3(blocks), dim3(TPB), 0, s, a);                     \
    return;                                                                                     \
  }
  UBG_SW(32, 8)
  UBG_SW(64, 4)
  UBG_SW(64, 8)
  UBG_SW(64, 16)
  UBG_SW(128, 4)
  UBG_SW(128, 8)
  UBG_SW(256, 4)
#undef UBG_SW
}

bool shape_ok(int tx, int ty) {
  for (const Shape &s : kShapes)
    if (s.tx == tx && s.ty == ty) return true;
  return false;
}

size_t al(size_t n) { return align_up(n, 256); }

// rows of the fused-pass tiles the tuning may fix (0: chosen per level)
bool pass_rows_ok(int r) { return r == 0 || r == 8 || r == 16 || r == 32; }

void launch_tile_pass(bool pre, int nu, int tpb, int blocks, hipStream_t s, const StageArgs &a) {
  const size_t shm = pre ? tile_pre_lds(a.oty, nu) : tile_post_lds(a.oty, nu);
#define UBG_TP(N, T)                                                                            \
  if (nu == N && tpb == T) {                                                                    \
    if (pre && a.level > 0)                                                                     \
      hipLaunchKernelGGL((kb_tile_pre<N, T, true>), dim3(blocks), dim3(T), shm, s, a);         \
    else if (pre)                                                                               \
      hipLaunchKernelGGL((kb_tile_pre<N, T, false>), dim3(blocks), dim3(T), shm, s, a);        \
    else hipLaunchKernelGGL((kb_tile_post<N, T>), dim3(blocks), dim3(T), shm, s, a);           \
    return;                                                                                     \
  }
  UBG_TP(0, 512)
  UBG_TP(1, 512)
  UBG_TP(2, 512)
  UBG_TP(3, 512)
  UBG_TP(0, 1024)
  UBG_TP(1, 1024)
  UBG_TP(2, 1024)
  UBG_TP(3, 1024)
#undef UBG_TP
}

} // namespace

// ===========================================================================
// Arena
// ===========================================================================
int Arena::init(size_t bytes) {
  cap_ = al(bytes);
  UBG_HIP(hipMalloc((void **)&base_, cap_));
  free_.clear();
  live_.clear();
  free_[0] = cap_;
  used_ = 0;
  return UBG_OK;
}

void Arena::release() {
  if (base_) (void)hipFree(base_);
  base_ = nullptr;
  free_.clear();
  live_.clear();
  cap_ = used_ = 0;
}

char *Arena::alloc(size_t bytes) {
  bytes = al(bytes ? bytes : 1);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    size_t off = it->first, sz = it->second;
    free_.erase(it);
    if (sz > bytes) free_[off + bytes] = sz - bytes;
    live_[off] = bytes;
    used_ += bytes;
    return base_ + off;
  }
  return nullptr;
}

void Arena::free(char *p) {
  if (!p) return;
  size_t off = (size_t)(p - base_);
  auto it = live_.find(off);
  if (it == live_.end()) return;
  size_t sz = it->second;
  live_.erase(it);
  used_ -= sz;
  auto nx = free_.lower_bound(off);
  if (nx != free_.end() && off + sz == nx->first) {
    sz += nx->second;
    free_.erase(nx);
  }
  auto pv = free_.lower_bound(off);
  if (pv != free_.begin()) {
    --pv;
    if (pv->first + pv->second == off) {
      pv->second += sz;
      return;
    }
  }
  free_[off] = sz;
}

// ===========================================================================
// shapes, geometries, domains
// ===========================================================================
int batch_shape(int W, int H, BShape *s) {
  if (W < B_MIN_SIZE || H < B_MIN_SIZE)
    return fail(UBG_E_DOMAIN_TOO_SMALL, "batched domains need at least 16 x 16 cells");
  std::vector<int> ws, hs;
  level_sizes(W, H, ws, hs);
  if ((int)ws.size() > BMAXL) return fail(UBG_E_INVALID, "too many levels");
  *s = BShape{};
  s->nlev = (int)ws.size();
  for (int l = 0; l < s->nlev; l++) {
    s->w[l] = ws[l];
    s->h[l] = hs[l];
  }
  // the LDS V-cycle starts where the single-domain solver starts it (the
  // first level of at most kCoarseMaxCells cells: larger LDS levels cost more
  // than their tile passes), or at the first level whose sub-hierarchy fits
  // if that one does not
  s->cfirst = -1;
  int fit = -1;
  size_t fit_bytes = 0;
  for (int l = 1; l <= s->nlev - 2; l++) {
    size_t b = coarse_lds_layout(ws.data(), hs.data(), l, s->nlev, nullptr);
    if (b > LDS_BUDGET) continue;
    if (fit < 0) {
      fit = l;
      fit_bytes = b;
    }
    if ((size_t)ws[l] * hs[l] <= kCoarseMaxCells) {
      s->cfirst = l;
      s->lds_bytes = b;
      break;
    }
  }
  if (s->cfirst < 0 && fit >= 0) {
    s->cfirst = fit;
    s->lds_bytes = fit_bytes;
  }
  if (s->cfirst < 0) return fail(UBG_E_INVALID, "no coarse level fits in LDS");
  for (int l = 0; l < s->nlev; l++) s->geom_bytes += 2 * al((size_t)ws[l] * hs[l]);
  const size_t n0 = (size_t)W * H;
  const int gx = (W + B_UPD_W - 1) / B_UPD_W, gy = (H + B_UPD_H - 1) / B_UPD_H;
  size_t b = 2 * al(n0 * 8) + al(n0 * 4) + al((size_t)gx * gy * 8) + al(sizeof(DState));
  for (int l = 0; l <= s->cfirst; l++) b += 3 * al((size_t)ws[l] * hs[l] * 4);
  s->dom_bytes = b;
  return UBG_OK;
}

BatchEngine::~BatchEngine() {
  drop_graphs();
  if (d_lev_) (void)hipFree(d_lev_);
  if (d_dom_) (void)hipFree(d_dom_);
  if (d_cargs_) (void)hipFree(d_cargs_);
  if (d_off_) (void)hipFree(d_off_);
  if (d_idx_) (void)hipFree(d_idx_);
  if (d_count_) (void)hipFree(d_count_);
  if (d_active_) (void)hipFree(d_active_);
  if (h_active_) (void)hipHostFree(h_active_);
  if (count_host_) (void)hipHostFree(count_host_);
  if (h_idx_pinned_) (void)hipHostFree(h_idx_pinned_);
}

int BatchEngine::init(hipStream_t s, const ubg_params &prm, int vcycles_per_step) {
  stream_ = s;
  prm_ = prm;
  k_ = vcycles_per_step;
  UBG_HIP(hipMalloc((void **)&d_count_, sizeof(int)));
  UBG_HIP(hipHostMalloc((void **)&count_host_, sizeof(int), hipHostMallocDefault));
  *count_host_ = 0;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)kb_coarse, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)LDS_BUDGET);
    attr = true;
  }
  return UBG_OK;
}

int BatchEngine::geom_create(Arena &arena, int W, int H, const float *flag, BGeom **out,
                             bool need_ring) {
  BShape sh;
  UBG_TRY(batch_shape(W, H, &sh));
  std::vector<uint8_t> f8;
  UBG_TRY(validate_flag(flag, (size_t)W * H, f8));
  bool ring = true;
  for (int x = 0; x < W; x++)
    if (!f8[x] || !f8[(size_t)(H - 1) * W + x]) ring = false;
  for (int y = 0; y < H; y++)
    if (!f8[(size_t)y * W] || !f8[(size_t)y * W + W - 1]) ring = false;
  if (need_ring && !ring)
    return fail(UBG_E_FLAG_VALUE, "the outer ring of a solve domain must be fluid");
  auto *g = new BGeom();
  g->ring_fluid = ring;
  g->W = W;
  g->H = H;
  g->nlev = sh.nlev;
  g->cfirst = sh.cfirst;
  g->lds_bytes = sh.lds_bytes;
  g->bytes = sh.geom_bytes;
  g->mem = arena.alloc(g->bytes);
  if (!g->mem) {
    delete g;
    return fail(UBG_E_BUDGET, "device arena full (geometry)");
  }
  char *q = g->mem;
  for (int l = 0; l < g->nlev; l++) {
    g->w[l] = sh.w[l];
    g->h[l] = sh.h[l];
    size_t n = (size_t)g->w[l] * g->h[l];
    g->flag[l] = (uint8_t *)q;
    q += al(n);
    g->mask[l] = (uint8_t *)q;
    q += al(n);
  }
  UBG_HIP(hipMemcpyAsync(g->flag[0], f8.data(), f8.size(), hipMemcpyHostToDevice, stream_));
  launch_build_mask(g->flag[0], g->mask[0], W, H, stream_);
  for (int l = 1; l < g->nlev; l++) {
    launch_coarsen(g->flag[l - 1], g->w[l - 1], g->flag[l], g->w[l], g->h[l], stream_);
    launch_build_mask(g->flag[l], g->mask[l], g->w[l], g->h[l], stream_);
  }
  UBG_TRY(check_launch());
  UBG_HIP(hipStreamSynchronize(stream_)); // f8 is a host temporary
  g->refs = 1;
  *out = g;
  return UBG_OK;
}

void BatchEngine::geom_release(Arena &arena, BGeom *g) {
  if (!g || --g->refs > 0) return;
  arena.free(g->mem);
  delete g;
}

int BatchEngine::domain_create(Arena &arena, BGeom *g, float h, BDomain **out, hipStream_t s) {
  BShape sh;
  UBG_TRY(batch_shape(g->W, g->H, &sh));
  auto *d = new BDomain();
  d->g = g;
  d->h = h;
  d->bytes = sh.dom_bytes;
  d->mem = arena.alloc(d->bytes);
  if (!d->mem) {
    delete d;
    return fail(UBG_E_BUDGET, "device arena full (domain)");
  }
  const size_t n0 = (size_t)g->W * g->H;
  d->gx = (g->W + B_UPD_W - 1) / B_UPD_W;
  d->gy = (g->H + B_UPD_H - 1) / B_UPD_H;
  char *q = d->mem;
  d->x[0] = (double *)q;
  q += al(n0 * 8);
  d->x[1] = (double *)q;
  q += al(n0 * 8);
  d->rhs = (float *)q;
  q += al(n0 * 4);
  d->partial = (double *)q;
  q += al((size_t)d->gx * d->gy * 8);
  d->st = (DState *)q;
  q += al(sizeof(DState));
  for (int l = 0; l <= g->cfirst; l++) {
    size_t n = (size_t)g->w[l] * g->h[l];
    d->p[l][0] = (float *)q;
    q += al(n * 4);
    d->p[l][1] = (float *)q;
    q += al(n * 4);
    d->f[l] = (float *)q;
    q += al(n * 4);
  }
  // h of each level: h_c = h * (W_l - 1) / (W_{l+1} - 1) (pressure_solver.cpp:229)
  d->hh[0] = h;
  for (int l = 1; l < g->nlev; l++) d->hh[l] = d->hh[l - 1] * (g->w[l - 1] - 1.0f) / (g->w[l] - 1.0f);
  UBG_HIP(hipMemsetAsync(d->mem, 0, d->bytes, s ? s : stream_));
  g->refs++;
  *out = d;
  return UBG_OK;
}

void BatchEngine::domain_release(Arena &arena, BDomain *d) {
  if (!d) return;
  arena.free(d->mem);
  geom_release(arena, d->g);
  delete d;
}

int BatchEngine::domain_set_rhs(BDomain *d, const float *rhs) {
  const size_t n = (size_t)d->g->W * d->g->H;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(rhs[i])) return fail(UBG_E_NONFINITE, "rhs has a non-finite value");
  UBG_HIP(hipMemcpyAsync(d->rhs, rhs, n * 4, hipMemcpyHostToDevice, stream_));
  UBG_HIP(hipStreamSynchronize(stream_));
  return UBG_OK;
}

int BatchEngine::domain_set_guess(BDomain *d, const double *x) {
  const size_t n = (size_t)d->g->W * d->g->H;
  if (x) {
    for (size_t i = 0; i < n; i++)
      if (!std::isfinite(x[i])) return fail(UBG_E_NONFINITE, "guess has a non-finite value");
    UBG_HIP(hipMemcpyAsync(d->x[d->xcur], x, n * 8, hipMemcpyHostToDevice, stream_));
  } else {
    launch_fill_f64(d->x[d->xcur], n, 0.0, stream_);
    UBG_TRY(check_launch());
  }
  UBG_HIP(hipStreamSynchronize(stream_));
  return UBG_OK;
}

int BatchEngine::domain_start(BDomain *d, double tol, int max_iter, double ref, hipStream_t s) {
  if (!s) s = stream_;
  DState st{};
  st.active = 1;
  st.tol = tol;
  st.ref = ref;
  st.max_iter = max_iter;
  st.k = k_;
  st.best = 1.0;
  st.rel = 1.0;
  st.xcur = d->xcur;
  UBG_HIP(hipMemcpyAsync(d->st, &st, sizeof(st), hipMemcpyHostToDevice, s));
  UBG_HIP(hipStreamSynchronize(s));
  return UBG_OK;
}

int BatchEngine::domain_read_state(BDomain *d, DState *out, hipStream_t s) {
  if (!s) s = stream_;
  UBG_HIP(hipMemcpyAsync(out, d->st, sizeof(DState), hipMemcpyDeviceToHost, s));
  UBG_HIP(hipStreamSynchronize(s));
  return UBG_OK;
}

int BatchEngine::domain_get_solution(BDomain *d, double *x, hipStream_t s) {
  if (!s) s = stream_;
  const size_t n = (size_t)d->g->W * d->g->H;
  UBG_HIP(hipMemcpyAsync(x, d->x[d->xcur], n * 8, hipMemcpyDeviceToHost, s));
  UBG_HIP(hipStreamSynchronize(s));
  return UBG_OK;
}

int BatchEngine::domain_set_level0(BDomain *d, const float *p, const float *f,
                                   hipMemcpyKind kind, hipStream_t s) {
  if (!s) s = stream_;
  const size_t n = (size_t)d->g->W * d->g->H;
  float *p0 = d->p[0][d->cur[0]];
  if (p) UBG_HIP(hipMemcpyAsync(p0, p, n * 4, kind, s));
  else launch_fill_f32(p0, n, 0.0f, s);
  if (f) UBG_HIP(hipMemcpyAsync(d->f[0], f, n * 4, kind, s));
  UBG_TRY(check_launch());
  if (kind != hipMemcpyDeviceToDevice) UBG_HIP(hipStreamSynchronize(s));
  return UBG_OK;
}

int BatchEngine::domain_get_pressure(BDomain *d, float *p, hipMemcpyKind kind, hipStream_t s) {
  if (!s) s = stream_;
  const size_t n = (size_t)d->g->W * d->g->H;
  UBG_HIP(hipMemcpyAsync(p, d->p[0][d->cur[0]], n * 4, kind, s));
  if (kind != hipMemcpyDeviceToDevice) UBG_HIP(hipStreamSynchronize(s));
  return UBG_OK;
}

// ===========================================================================
// member tables, tiles, schedules
// ===========================================================================
void BatchEngine::drop_graphs() {
  for (auto &kv : graphs_) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.d_idx) (void)hipFree(kv.second.d_idx);
  }
  graphs_.clear();
}

int BatchEngine::set_members(const std::vector<BDomain *> &running,
                             const std::vector<BDomain *> &fresh) {
  if (running == run_ && fresh == fresh_ && d_lev_) return UBG_OK;
  run_ = running;
  fresh_ = fresh;
  all_ = running;
  all_.insert(all_.end(), fresh.begin(), fresh.end());
  drop_graphs();
  UBG_TRY(build_tables());
  UBG_TRY(build_tiles());
  return UBG_OK;
}

int BatchEngine::build_tables() {
  const size_t n = all_.size();
  if (n > tab_cap_) {
    if (d_lev_) (void)hipFree(d_lev_);
    if (d_dom_) (void)hipFree(d_dom_);
    if (d_cargs_) (void)hipFree(d_cargs_);
    if (d_active_) (void)hipFree(d_active_);
    if (h_active_) (void)hipHostFree(h_active_);
    d_lev_ = nullptr;
    d_dom_ = nullptr;
    d_cargs_ = nullptr;
    d_active_ = nullptr;
    h_active_ = nullptr;
    size_t cap = std::max<size_t>(n, 16);
    UBG_HIP(hipMalloc((void **)&d_lev_, cap * BMAXL * sizeof(BLev)));
    UBG_HIP(hipMalloc((void **)&d_dom_, cap * sizeof(BDom)));
    UBG_HIP(hipMalloc((void **)&d_cargs_, cap * sizeof(CoarseArgs)));
    UBG_HIP(hipMalloc((void **)&d_active_, cap));
    UBG_HIP(hipHostMalloc((void **)&h_active_, cap, hipHostMallocDefault));
    std::memset(h_active_, 0, cap);
    tab_cap_ = cap;
  }
  std::vector<BLev> lev(std::max<size_t>(n, 1) * BMAXL);
  std::vector<BDom> dom(std::max<size_t>(n, 1));
  std::vector<CoarseArgs> ca(std::max<size_t>(n, 1));
  coarse_lds_ = 0;
  lmax_ = 0;
  for (size_t s = 0; s < n; s++) {
    const BDomain &D = *all_[s];
    const BGeom &G = *D.g;
    for (int l = 0; l < G.nlev; l++) {
      BLev &L = lev[s * BMAXL + l];
      L.p[0] = l <= G.cfirst ? D.p[l][0] : nullptr;
      L.p[1] = l <= G.cfirst ? D.p[l][1] : nullptr;
      L.f = l <= G.cfirst ? D.f[l] : nullptr;
      L.mask = G.mask[l];
      L.w = G.w[l];
      L.h = G.h[l];
      L.hh = D.hh[l];
      L.ihsq = 1.0f / D.hh[l] / D.hh[l];
    }
    BDom &B = dom[s];
    B.st = D.st;
    B.x[0] = D.x[0];
    B.x[1] = D.x[1];
    B.rhs = D.rhs;
    B.partial = D.partial;
    B.w = G.W;
    B.h = G.H;
    B.gx = D.gx;
    B.gy = D.gy;
    B.nlev = G.nlev;
    B.cfirst = G.cfirst;
    CoarseArgs &a = ca[s];
    a = CoarseArgs{};
    coarse_lds_layout(G.w, G.h, G.cfirst, G.nlev, &a);
    a.n = G.nlev - 1 - G.cfirst;
    for (int k = 0; k < a.n; k++) {
      const int l = G.cfirst + k;
      a.w[k] = G.w[l];
      a.h[k] = G.h[l];
      a.mask[k] = G.mask[l];
      a.hh[k] = D.hh[l];
    }
    a.nu_pre = prm_.nu_pre;
    a.nu_post = prm_.nu_post;
    a.nu_coarse = prm_.nu_coarse;
    a.alpha = prm_.alpha;
    if (s < run_.size()) {
      coarse_lds_ = std::max(coarse_lds_, G.lds_bytes);
      lmax_ = std::max(lmax_, G.cfirst);
    }
  }
  if (n) {
    UBG_HIP(hipMemcpyAsync(d_lev_, lev.data(), n * BMAXL * sizeof(BLev), hipMemcpyHostToDevice,
                           stream_));
    UBG_HIP(hipMemcpyAsync(d_dom_, dom.data(), n * sizeof(BDom), hipMemcpyHostToDevice, stream_));
    UBG_HIP(hipMemcpyAsync(d_cargs_, ca.data(), n * sizeof(CoarseArgs), hipMemcpyHostToDevice,
                           stream_));
    UBG_HIP(hipStreamSynchronize(stream_));
  }
  return UBG_OK;
}

// Per stage kind and level: the prefix of tile counts over the member slots
// and each domain's tiles per row; kernels locate their tile from these, so
// a member change rebuilds O(domains x levels) integers, not tile lists.
// Row-streamed levels: at least kStreamMinW wide and kStreamMinCells large,
// with both smoothing counts >= 1 (the grid-transfer passes carry one sweep
// each) and at most 3 (the tiled mode); row blocks of kStreamR rows; used at
// a level only when the members' sweep units there reach kStreamMinUnits.
// Measured on gfx950 (256 uneven batched domains and the served mix): below
// about 4096 units a level is faster as LDS tiles; above, streaming its
// >= 512-wide, >= 2^18-cell domains gained ~3% on the served mix and was
// neutral on the batch.
constexpr int kStreamMinW = 512;
constexpr size_t kStreamMinCells = size_t(1) << 18;
constexpr long kStreamMinUnits = 4096;
constexpr int kStreamR = 16;
static_assert(kStreamR % 2 == 0, "streamed passes need an even row-block height");

bool BatchEngine::stream_eligible(const BGeom &G, int l) const {
  if (l < G.cfirst && tiled() && prm_.nu_pre >= 1 && prm_.nu_post >= 1 && l == 0 &&
      force_stream0_)
    return true;
  return l < G.cfirst && tiled() && prm_.nu_pre >= 1 && prm_.nu_post >= 1 &&
         G.w[l] >= kStreamMinW && (size_t)G.w[l] * G.h[l] >= kStreamMinCells;
}

int BatchEngine::build_tiles() {
  h_off_.clear();
  // per span: tile prefix over the domain slots [nd + 1], tiles per row [nd]
  // and the tile width [nd] (fused passes; 0 elsewhere)
  auto span = [&](Span &sp, int nd, auto tiles_of) {
    sp.off = h_off_.size();
    sp.nd = nd;
    h_off_.resize(sp.off + 3 * (size_t)nd + 1);
    int *pre = h_off_.data() + sp.off, *ntx = pre + nd + 1, *otx = ntx + nd;
    long tot = 0;
    for (int s = 0; s < nd; s++) {
      int nx = 0, ny = 0, ox = 0;
      tiles_of(s, nx, ny, ox);
      pre[s] = (int)tot;
      ntx[s] = nx > 0 ? nx : 1;
      otx[s] = ox;
      tot += (long)nx * ny;
    }
    pre[nd] = (int)tot;
    sp.n = (int)tot;
  };
  auto cdiv = [](int a, int b) { return a > 0 ? (a + b - 1) / b : 0; };
  const int tx = tune_.tx, ty = tune_.ty;
  const int nr = (int)run_.size();
  for (int l = 0; l < BMAXL; l++) {
    long units = 0;
    for (int s = 0; s < nr; s++) {
      const BGeom &G = *run_[s]->g;
      if (stream_eligible(G, l)) units += (long)cdiv(G.w[l], SW2) * cdiv(G.h[l] - 2, kStreamR);
    }
    stream_on_[l] = units >= kStreamMinUnits || (l == 0 && force_stream0_);
  }
  for (int l = 0; l < BMAXL; l++) {
    span(sw_[l], nr, [&](int s, int &nx, int &ny, int &) {
      const BGeom &G = *run_[s]->g;
      if (l >= G.cfirst) return;
      nx = cdiv(G.w[l], tx);
      ny = cdiv(G.h[l], ty);
    });
    span(rr_[l], nr, [&](int s, int &nx, int &ny, int &) {
      const BGeom &G = *run_[s]->g;
      if (l >= G.cfirst) return;
      nx = cdiv(G.w[l + 1] - 2, RCX);
      ny = cdiv(G.h[l + 1] - 2, RCY);
    });
    span(pc_[l], nr, [&](int s, int &nx, int &ny, int &) {
      const BGeom &G = *run_[s]->g;
      if (l >= G.cfirst) return;
      nx = cdiv(G.w[l] - 2, PCW);
      ny = cdiv(G.h[l] - 2, PCH);
    });
  }
  // Fused passes: each domain's columns split evenly (tile_cols); the rows
  // fixed by the tuning, or per level the fewest (even, 8 to 32) whose tiles
  // over all member domains fit in one residency wave of the launch's
  // target, with 512-thread workgroups for levels of at least kTileBigCells
  // cells in total and 1024-thread ones below.
  auto fused = [&](Span *out, int nu) {
    for (int l = 0; l < BMAXL; l++) {
      size_t cells = 0;
      for (int s = 0; s < nr; s++) {
        const BGeom &G = *run_[s]->g;
        if (l < G.cfirst && !streamed(G, l)) cells += (size_t)G.w[l] * G.h[l];
      }
      const int tpb = cells >= kTileBigCells ? 512 : 1024;
      int oty = tune_.prow;
      if (oty == 0) {
        auto tiles = [&](int r) {
          long t = 0;
          for (int s = 0; s < nr; s++) {
            const BGeom &G = *run_[s]->g;
            if (l >= G.cfirst || streamed(G, l)) continue;
            int ox, nx;
            tile_cols(G.w[l], nu, ox, nx);
            t += (long)nx * cdiv(G.h[l], r);
          }
          return t;
        };
        oty = 8;
        while (oty < 32 && tiles(oty) > tile_target(tpb)) oty += 2;
      }
      out[l].oty = oty;
      out[l].tpb = tpb;
      span(out[l], nr, [&](int s, int &nx, int &ny, int &ox) {
        const BGeom &G = *run_[s]->g;
        if (l >= G.cfirst || streamed(G, l)) return;
        tile_cols(G.w[l], nu, ox, nx);
        ny = cdiv(G.h[l], oty);
      });
    }
  };
  fused(tp_, prm_.nu_pre);
  fused(tq_, prm_.nu_post);
  // row-streamed levels: units = strips x row blocks (the table's third row
  // holds the row-block height)
  constexpr int own_pass = 2 * (64 - 2 * (1 + 2)); // columns per strip, one-sweep passes
  for (int l = 0; l < BMAXL; l++) {
    span(ss_[l], nr, [&](int s, int &nx, int &ny, int &r) {
      const BGeom &G = *run_[s]->g;
      if (!streamed(G, l)) return;
      nx = cdiv(G.w[l], SW2);
      ny = cdiv(G.h[l] - 2, kStreamR);
      r = kStreamR;
    });
    span(st_[l], nr, [&](int s, int &nx, int &ny, int &r) {
      const BGeom &G = *run_[s]->g;
      if (!streamed(G, l)) return;
      nx = cdiv(G.w[l], own_pass);
      ny = cdiv(G.h[l], kStreamR);
      r = kStreamR;
    });
  }
  span(upd_, (int)all_.size(), [&](int s, int &nx, int &ny, int &) {
    nx = all_[s]->gx;
    ny = all_[s]->gy;
  });
  if (h_off_.size() > off_cap_) {
    if (d_off_) (void)hipFree(d_off_);
    d_off_ = nullptr;
    off_cap_ = h_off_.size() + h_off_.size() / 4 + 64;
    UBG_HIP(hipMalloc((void **)&d_off_, off_cap_ * sizeof(int)));
  }
  UBG_HIP(hipMemcpyAsync(d_off_, h_off_.data(), h_off_.size() * sizeof(int), hipMemcpyHostToDevice,
                         stream_));
  UBG_HIP(hipStreamSynchronize(stream_));
  return UBG_OK;
}

enum { ST_SWEEP, ST_RR, ST_PC, ST_COARSE, ST_UPDATE, ST_NORM, ST_TPRE, ST_TPOST, ST_SSWEEP, ST_SRR, ST_SPC };

// One V-cycle of the running members in the tiled mode (the per-level
// passes of solveLevel, pressure_solver.cpp:201-248), simulated on (and
// advancing) the domains' host index state.
void BatchEngine::gen_vcycle_tiled(std::vector<StageRec> &st, std::vector<uint8_t> &idx) {
  const size_t n = all_.size();
  const size_t nr = run_.size();
  auto row = [&](int kind, int level) -> std::pair<uint8_t *, uint8_t *> {
    size_t off = idx.size();
    idx.resize(off + 2 * n, 0);
    st.push_back(StageRec{kind, level, off});
    return {idx.data() + off, idx.data() + off + n};
  };
  // per level: a streamed domain takes nu_pre - 1 sweeps and the fused
  // pre pass (each flipping the level's p), a tiled one the fused tile
  // pre pass (one flip); mirrored on the way up
  auto tile_at = [&](size_t d, int l) {
    return l < run_[d]->g->cfirst && !streamed(*run_[d]->g, l);
  };
  auto stream_at = [&](size_t d, int l) { return streamed(*run_[d]->g, l); };
  auto sweep_rows = [&](int l, int n) {
    for (int k = 0; k < n; k++) {
      auto r = row(ST_SSWEEP, l);
      for (size_t d = 0; d < nr; d++)
        if (stream_at(d, l)) {
          r.first[d] = (uint8_t)run_[d]->cur[l];
          run_[d]->cur[l] ^= 1;
        }
    }
  };
  auto pass_row = [&](int kind, int l, auto at) {
    auto r = row(kind, l);
    for (size_t d = 0; d < nr; d++)
      if (at(d, l)) {
        r.first[d] = (uint8_t)run_[d]->cur[l];
        r.second[d] = (uint8_t)run_[d]->cur[l + 1];
        run_[d]->cur[l] ^= 1;
      }
  };
  for (int l = 0; l < lmax_; l++) {
    if (st_[l].n) {
      sweep_rows(l, prm_.nu_pre - 1);
      pass_row(ST_SRR, l, stream_at);
    }
    if (tp_[l].n) pass_row(ST_TPRE, l, tile_at);
  }
  {
    auto r = row(ST_COARSE, 0);
    for (size_t d = 0; d < nr; d++) r.first[d] = (uint8_t)run_[d]->cur[run_[d]->g->cfirst];
  }
  for (int l = lmax_ - 1; l >= 0; l--) {
    if (st_[l].n) {
      pass_row(ST_SPC, l, stream_at);
      sweep_rows(l, prm_.nu_post - 1);
    }
    if (tq_[l].n) pass_row(ST_TPOST, l, tile_at);
  }
}

// The stage sequence of nsteps refinement steps and, per stage, the two
// per-domain index rows, simulated from (and advancing) the domains' host
// index state: a sweep reads p[cur] and writes p[cur^1]; the update reads
// e = p0[cur] and x[xcur], zeroes p0[cur^1], writes x[xcur^1].
void BatchEngine::gen_steps(int nsteps, std::vector<StageRec> &st, std::vector<uint8_t> &idx,
                            bool all_fresh) {
  const size_t n = all_.size();
  const size_t nr = all_fresh ? 0 : run_.size();
  auto row = [&](int kind, int level) -> std::pair<uint8_t *, uint8_t *> {
    size_t off = idx.size();
    idx.resize(off + 2 * n, 0);
    st.push_back(StageRec{kind, level, off});
    return {idx.data() + off, idx.data() + off + n};
  };
  for (int step = 0; step < nsteps; step++) {
    const bool with_fresh = step == 0 && (all_fresh || !fresh_.empty());
    for (int v = 0; v < k_ && nr && tiled(); v++) gen_vcycle_tiled(st, idx);
    for (int v = 0; v < k_ && nr && !tiled(); v++) {
      for (int l = 0; l < lmax_; l++) {
        if (!sw_[l].n && !rr_[l].n) continue;
        for (int s = 0; s < prm_.nu_pre; s++) {
          auto r = row(ST_SWEEP, l);
          for (size_t d = 0; d < nr; d++)
            if (l < run_[d]->g->cfirst) {
              r.first[d] = (uint8_t)run_[d]->cur[l];
              run_[d]->cur[l] ^= 1;
            }
        }
        auto r = row(ST_RR, l);
        for (size_t d = 0; d < nr; d++)
          if (l < run_[d]->g->cfirst) {
            r.first[d] = (uint8_t)run_[d]->cur[l];
            r.second[d] = (uint8_t)run_[d]->cur[l + 1];
          }
      }
      {
        auto r = row(ST_COARSE, 0);
        for (size_t d = 0; d < nr; d++) r.first[d] = (uint8_t)run_[d]->cur[run_[d]->g->cfirst];
      }
      for (int l = lmax_ - 1; l >= 0; l--) {
        if (!pc_[l].n && !sw_[l].n) continue;
        auto r = row(ST_PC, l);
        for (size_t d = 0; d < nr; d++)
          if (l < run_[d]->g->cfirst) {
            r.first[d] = (uint8_t)run_[d]->cur[l];
            r.second[d] = (uint8_t)run_[d]->cur[l + 1];
          }
        for (int s = 0; s < prm_.nu_post; s++) {
          auto q = row(ST_SWEEP, l);
          for (size_t d = 0; d < nr; d++)
            if (l < run_[d]->g->cfirst) {
              q.first[d] = (uint8_t)run_[d]->cur[l];
              run_[d]->cur[l] ^= 1;
            }
        }
      }
    }
    auto u = row(ST_UPDATE, 0);
    for (size_t d = 0; d < all_.size(); d++) {
      BDomain &D = *all_[d];
      const bool fresh = with_fresh && d >= nr;
      if (!fresh && d >= nr) continue; // fresh members join later steps only as running
      u.first[d] = (uint8_t)D.cur[0];
      u.second[d] = (uint8_t)(D.xcur | (fresh ? 0 : 2));
    }
    auto m = row(ST_NORM, 0);
    for (size_t d = 0; d < all_.size(); d++) {
      BDomain &D = *all_[d];
      const bool fresh = with_fresh && d >= nr;
      if (!fresh && d >= nr) continue;
      D.cur[0] ^= 1;
      D.xcur ^= 1;
      m.first[d] = fresh ? 1 : 0;
      m.second[d] = (uint8_t)D.xcur;
    }
  }
}

int BatchEngine::launch_stages(const std::vector<StageRec> &st, const uint8_t *d_idx) {
  const int n = (int)all_.size();
  auto grid = [&](int ntiles) {
    int b = tune_.blocks > 0 ? std::min(ntiles, tune_.blocks) : ntiles;
    return std::max(b, 1);
  };
  auto use = [&](StageArgs &a, const Span &sp) {
    a.off = d_off_ + sp.off;
    a.ndom = sp.nd;
    a.ntiles = sp.n;
  };
  for (const StageRec &r : st) {
    StageArgs a{};
    a.lev = d_lev_;
    a.dom = d_dom_;
    a.level = r.level;
    a.zg = zg0_ && r.level == 0;
    a.ia = d_idx + r.off;
    a.ib = d_idx + r.off + n;
    a.alpha = prm_.alpha;
    switch (r.kind) {
    case ST_SWEEP:
      use(a, sw_[r.level]);
      if (a.ntiles) launch_sweep(tune_, grid(a.ntiles), stream_, a);
      break;
    case ST_RR:
      use(a, rr_[r.level]);
      if (a.ntiles) hipLaunchKernelGGL(kb_rr, dim3(grid(a.ntiles)), dim3(TPB), 0, stream_, a);
      break;
    case ST_PC:
      use(a, pc_[r.level]);
      if (a.ntiles) hipLaunchKernelGGL(kb_pc, dim3(grid(a.ntiles)), dim3(TPB), 0, stream_, a);
      break;
    case ST_TPRE:
      use(a, tp_[r.level]);
      a.oty = tp_[r.level].oty;
      if (a.ntiles)
        launch_tile_pass(true, prm_.nu_pre, tp_[r.level].tpb, grid(a.ntiles), stream_, a);
      break;
    case ST_TPOST:
      use(a, tq_[r.level]);
      a.oty = tq_[r.level].oty;
      if (a.ntiles)
        launch_tile_pass(false, prm_.nu_post, tq_[r.level].tpb, grid(a.ntiles), stream_, a);
      break;
    case ST_SSWEEP:
      use(a, ss_[r.level]);
      if (a.ntiles)
        hipLaunchKernelGGL(kb_stream_sweep, dim3((a.ntiles + TPB / 64 - 1) / (TPB / 64)),
                           dim3(TPB), 0, stream_, a);
      break;
    case ST_SRR:
    case ST_SPC:
      use(a, st_[r.level]);
      if (a.ntiles) {
        const dim3 g((a.ntiles + TPB / 64 - 1) / (TPB / 64));
        if (r.kind == ST_SRR) hipLaunchKernelGGL(kb_stream_pass<MODE_RR>, g, dim3(TPB), 0, stream_, a);
        else hipLaunchKernelGGL(kb_stream_pass<MODE_PC>, g, dim3(TPB), 0, stream_, a);
      }
      break;
    case ST_COARSE:
      if (!run_.empty())
        hipLaunchKernelGGL(kb_coarse, dim3((unsigned)run_.size()), dim3(TPB_C), coarse_lds_,
                           stream_, a, (const CoarseArgs *)d_cargs_, (int)run_.size());
      break;
    case ST_UPDATE:
      use(a, upd_);
      if (a.ntiles) hipLaunchKernelGGL(kb_update, dim3(grid(a.ntiles)), dim3(B_UPD_W), 0, stream_, a);
      break;
    case ST_NORM:
      hipLaunchKernelGGL(kb_norm, dim3(n), dim3(1024), 0, stream_, a, n);
      break;
    }
  }
  hipLaunchKernelGGL(kb_count, dim3(1), dim3(256), 0, stream_, (const BDom *)d_dom_, n, d_count_,
                     d_active_);
  UBG_TRY(check_launch());
  UBG_HIP(hipMemcpyAsync(count_host_, d_count_, sizeof(int), hipMemcpyDeviceToHost, stream_));
  UBG_HIP(hipMemcpyAsync(h_active_, d_active_, (size_t)n, hipMemcpyDeviceToHost, stream_));
  return UBG_OK;
}

int BatchEngine::upload_idx(const std::vector<uint8_t> &idx, uint8_t **d_idx) {
  if (idx.size() > h_idx_cap_) {
    if (h_idx_pinned_) (void)hipHostFree(h_idx_pinned_);
    h_idx_pinned_ = nullptr;
    h_idx_cap_ = idx.size() + idx.size() / 2 + 256;
    UBG_HIP(hipHostMalloc((void **)&h_idx_pinned_, h_idx_cap_, hipHostMallocDefault));
  }
  if (idx.size() > idx_cap_) {
    if (d_idx_) (void)hipFree(d_idx_);
    d_idx_ = nullptr;
    idx_cap_ = idx.size() + idx.size() / 2 + 256;
    UBG_HIP(hipMalloc((void **)&d_idx_, idx_cap_));
  }
  // the pinned staging buffer may still feed the previous step's copy
  UBG_HIP(hipStreamSynchronize(stream_));
  std::memcpy(h_idx_pinned_, idx.data(), idx.size());
  UBG_HIP(hipMemcpyAsync(d_idx_, h_idx_pinned_, idx.size(), hipMemcpyHostToDevice, stream_));
  *d_idx = d_idx_;
  return UBG_OK;
}

int BatchEngine::enqueue_steps(int nsteps) {
  if (all_.empty()) {
    *count_host_ = 0;
    return UBG_OK;
  }
  if (!fresh_.empty() && nsteps != 1)
    return fail(UBG_E_STATE, "members with a pending first update take one step at a time");
  std::vector<StageRec> st;
  std::vector<uint8_t> idx;
  gen_steps(nsteps, st, idx);
  uint8_t *d_idx = nullptr;
  UBG_TRY(upload_idx(idx, &d_idx));
  UBG_TRY(launch_stages(st, d_idx));
  return UBG_OK;
}

int BatchEngine::enqueue_first_update() {
  if (all_.empty()) {
    *count_host_ = 0;
    return UBG_OK;
  }
  std::vector<StageRec> st;
  std::vector<uint8_t> idx;
  gen_steps(1, st, idx, true);
  uint8_t *d_idx = nullptr;
  UBG_TRY(upload_idx(idx, &d_idx));
  UBG_TRY(launch_stages(st, d_idx));
  return UBG_OK;
}

std::string BatchEngine::state_key() const {
  std::string k;
  k.reserve(all_.size() * (BMAXL + 1));
  for (const BDomain *d : all_) {
    for (int l = 0; l < BMAXL; l++) k.push_back((char)('0' + d->cur[l]));
    k.push_back((char)('0' + d->xcur));
  }
  return k;
}

int BatchEngine::run_steps_graph(int nsteps) {
  if (all_.empty()) {
    *count_host_ = 0;
    return UBG_OK;
  }
  if (!fresh_.empty()) return fail(UBG_E_STATE, "graph replay needs every member running");
  const std::string key = state_key() + "/" + std::to_string(nsteps);
  auto it = graphs_.find(key);
  if (it == graphs_.end()) {
    // generate on copies of the index state, capture, then drop the copies
    std::vector<std::pair<std::array<int, BMAXL>, int>> saved;
    for (BDomain *d : all_) {
      std::array<int, BMAXL> c;
      for (int l = 0; l < BMAXL; l++) c[l] = d->cur[l];
      saved.push_back({c, d->xcur});
    }
    std::vector<StageRec> st;
    std::vector<uint8_t> idx;
    gen_steps(nsteps, st, idx);
    for (size_t i = 0; i < all_.size(); i++) {
      for (int l = 0; l < BMAXL; l++) all_[i]->cur[l] = saved[i].first[l];
      all_[i]->xcur = saved[i].second;
    }
    Graph g;
    UBG_HIP(hipMalloc((void **)&g.d_idx, std::max<size_t>(idx.size(), 1)));
    UBG_HIP(hipMemcpy(g.d_idx, idx.data(), idx.size(), hipMemcpyHostToDevice));
    hipGraph_t graph = nullptr;
    UBG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    int rc = launch_stages(st, g.d_idx);
    hipError_t e = hipStreamEndCapture(stream_, &graph);
    if (rc != UBG_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipFree(g.d_idx);
      return rc != UBG_OK ? rc : fail(UBG_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      (void)hipFree(g.d_idx);
      return fail(UBG_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    it = graphs_.emplace(key, g).first;
  }
  UBG_HIP(hipGraphLaunch(it->second.exec, stream_));
  // advance the host index state as the schedule does
  std::vector<StageRec> st;
  std::vector<uint8_t> idx;
  gen_steps(nsteps, st, idx);
  return UBG_OK;
}

int BatchEngine::set_level0_streamed(bool on) {
  if (on == force_stream0_) return UBG_OK;
  force_stream0_ = on;
  drop_graphs();
  if (d_lev_) UBG_TRY(build_tiles());
  return UBG_OK;
}

int BatchEngine::enqueue_vcycles(int ncycles, bool zg) {
  if (!tiled() || prm_.nu_pre < 1 || prm_.nu_post < 1)
    return fail(UBG_E_INVALID, "projection batches need 1..3 pre- and post-smoothing sweeps");
  if (!fresh_.empty()) return fail(UBG_E_STATE, "projection batches take running members only");
  if (run_.empty() || ncycles <= 0) return UBG_OK;
  UBG_TRY(set_level0_streamed(zg));
  // one captured graph per index state (the V-cycles of all members, their
  // index rows in a device buffer of the graph's own), replayed
  const std::string key =
      state_key() + "/v" + std::to_string(ncycles) + (zg ? "z" : "n");
  auto it = graphs_.find(key);
  if (it == graphs_.end()) {
    std::vector<std::array<int, BMAXL>> saved;
    for (BDomain *d : all_) {
      std::array<int, BMAXL> c;
      for (int l = 0; l < BMAXL; l++) c[l] = d->cur[l];
      saved.push_back(c);
    }
    std::vector<StageRec> st;
    std::vector<uint8_t> idx;
    for (int v = 0; v < ncycles; v++) gen_vcycle_tiled(st, idx);
    for (size_t i = 0; i < all_.size(); i++)
      for (int l = 0; l < BMAXL; l++) all_[i]->cur[l] = saved[i][l];
    Graph g;
    UBG_HIP(hipMalloc((void **)&g.d_idx, std::max<size_t>(idx.size(), 1)));
    UBG_HIP(hipMemcpy(g.d_idx, idx.data(), idx.size(), hipMemcpyHostToDevice));
    hipGraph_t graph = nullptr;
    UBG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    zg0_ = zg;
    int rc = launch_stages(st, g.d_idx);
    zg0_ = false;
    hipError_t e = hipStreamEndCapture(stream_, &graph);
    if (rc != UBG_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipFree(g.d_idx);
      return rc != UBG_OK ? rc : fail(UBG_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      (void)hipFree(g.d_idx);
      return fail(UBG_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    it = graphs_.emplace(key, g).first;
  }
  UBG_HIP(hipGraphLaunch(it->second.exec, stream_));
  // advance the host index state as the schedule does
  std::vector<StageRec> st;
  std::vector<uint8_t> idx;
  for (int v = 0; v < ncycles; v++) gen_vcycle_tiled(st, idx);
  return UBG_OK;
}

std::vector<BDomain *> BatchEngine::active_members() const {
  std::vector<BDomain *> out;
  for (size_t i = 0; i < all_.size(); i++)
    if (h_active_ && h_active_[i]) out.push_back(all_[i]);
  return out;
}

int BatchEngine::sync() {
  UBG_HIP(hipStreamSynchronize(stream_));
  return UBG_OK;
}

void BatchEngine::set_tuning(const Tuning &t) {
  tune_ = t;
  drop_graphs();
  (void)build_tiles();
}

// nu_pre sweeps of every member level, `reps` times (scratch buffers only:
// the next update zeroes e before it is used)
int BatchEngine::launch_sweeps_only(int reps) {
  const int n = (int)all_.size();
  std::vector<uint8_t> idx(2 * (size_t)n, 0);
  uint8_t *d_idx = nullptr;
  UBG_TRY(upload_idx(idx, &d_idx));
  for (int r = 0; r < reps; r++)
    for (int l = 0; l < lmax_; l++) {
      if (!sw_[l].n) continue;
      StageArgs a{};
      a.lev = d_lev_;
      a.dom = d_dom_;
      a.level = l;
      a.ia = d_idx;
      a.ib = d_idx + n;
      a.alpha = prm_.alpha;
      a.off = d_off_ + sw_[l].off;
      a.ndom = sw_[l].nd;
      a.ntiles = sw_[l].n;
      int b = tune_.blocks > 0 ? std::min(a.ntiles, tune_.blocks) : a.ntiles;
      launch_sweep(tune_, std::max(b, 1), stream_, a);
    }
  UBG_TRY(check_launch());
  return UBG_OK;
}

// the fused pre passes of every member level, then the post passes, `reps`
// times (scratch buffers only, as launch_sweeps_only)
int BatchEngine::launch_passes_only(int reps) {
  const int n = (int)all_.size();
  std::vector<uint8_t> idx(2 * (size_t)n, 0);
  uint8_t *d_idx = nullptr;
  UBG_TRY(upload_idx(idx, &d_idx));
  auto pass = [&](bool pre, int l) {
    const Span &sp = pre ? tp_[l] : tq_[l];
    if (!sp.n) return;
    StageArgs a{};
    a.lev = d_lev_;
    a.dom = d_dom_;
    a.level = l;
    a.ia = d_idx;
    a.ib = d_idx + n;
    a.alpha = prm_.alpha;
    a.off = d_off_ + sp.off;
    a.ndom = sp.nd;
    a.ntiles = sp.n;
    a.oty = sp.oty;
    int b = tune_.blocks > 0 ? std::min(a.ntiles, tune_.blocks) : a.ntiles;
    launch_tile_pass(pre, pre ? prm_.nu_pre : prm_.nu_post, sp.tpb, std::max(b, 1), stream_, a);
  };
  for (int r = 0; r < reps; r++) {
    for (int l = 0; l < lmax_; l++) pass(true, l);
    for (int l = lmax_ - 1; l >= 0; l--) pass(false, l);
  }
  UBG_TRY(check_launch());
  return UBG_OK;
}

// Times the candidate shapes over the member levels and keeps the fastest:
// the fused passes' tile rows x block counts when the levels run as fused
// passes, else the smoother tile shapes x block counts.
int BatchEngine::autotune(double *best_ms) {
  const Tuning saved = tune_;
  Tuning best = saved;
  double best_t = 1e30;
  hipEvent_t e0, e1;
  UBG_HIP(hipEventCreate(&e0));
  UBG_HIP(hipEventCreate(&e1));
  const int block_opts[] = {0, 2048, 1024};
  std::vector<Tuning> cand;
  if (tiled()) {
    for (int r : {0, 8, 16, 32})
      for (int bo : block_opts) cand.push_back(Tuning{saved.tx, saved.ty, bo, r});
  } else {
    for (const Shape &sh : kShapes)
      for (int bo : block_opts) cand.push_back(Tuning{sh.tx, sh.ty, bo, saved.prow});
  }
  // the solve state must allow the sweeps: mark members active meanwhile
  std::vector<DState> st(all_.size());
  for (size_t i = 0; i < all_.size(); i++) UBG_TRY(domain_read_state(all_[i], &st[i]));
  for (size_t i = 0; i < all_.size(); i++) {
    DState s = st[i];
    s.active = 1;
    UBG_HIP(hipMemcpyAsync(all_[i]->st, &s, sizeof(s), hipMemcpyHostToDevice, stream_));
  }
  int rc = UBG_OK;
  for (const Tuning &c : cand) {
    tune_ = c;
    if ((rc = build_tiles()) != UBG_OK) break;
    auto run = [&](int reps) { return tiled() ? launch_passes_only(reps) : launch_sweeps_only(reps); };
    if ((rc = run(1)) != UBG_OK) break; // warm
    (void)hipEventRecord(e0, stream_);
    if ((rc = run(3)) != UBG_OK) break;
    (void)hipEventRecord(e1, stream_);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best_t) {
      best_t = ms;
      best = tune_;
    }
  }
  for (size_t i = 0; i < all_.size(); i++)
    (void)hipMemcpyAsync(all_[i]->st, &st[i], sizeof(DState), hipMemcpyHostToDevice, stream_);
  (void)hipStreamSynchronize(stream_);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  tune_ = rc == UBG_OK ? best : saved;
  drop_graphs();
  UBG_TRY(build_tiles());
  if (best_ms) *best_ms = best_t / 3.0;
  return rc;
}

} // namespace ubg

// ===========================================================================
// C ABI: ubg_batch
// ===========================================================================
struct ubg_batch {
  ubg::BatchEngine eng;
  ubg::Arena arena;
  hipStream_t stream = nullptr;
  std::vector<ubg::BDomain *> doms;
  std::vector<int> ws, hs;
  std::vector<bool> ready;
  std::recursive_mutex mu;
  // projection solves: every domain's solve state is active (refinement
  // solves end domains; the next projection re-activates them)
  bool proj_active = false;
  // batched simulation steps: the simulation and flag generation each
  // domain's geometry was built from, and the events that order the
  // simulations' streams around the batch's
  std::vector<const ubg_sim *> src;
  std::vector<unsigned> src_gen;
  std::vector<hipEvent_t> ev_ready;
  std::shared_ptr<ubg::ExtEvent> ev_done; // recorded after each batched step
  // the per-stage simulation tables of a batched step: two slots used in
  // turn (pinned staging and device copy), each free again once its event
  // has passed
  struct TabSlot {
    ubg::SimRec *host = nullptr, *dev = nullptr;
    size_t cap = 0;
    hipEvent_t used = nullptr;
  } slot[2];
  int next_slot = 0;
};

namespace {

int batch_check(ubg_batch *b, int d) {
  if (!b) return fail(UBG_E_INVALID, "null batch");
  if (d < 0 || d >= (int)b->ws.size()) return fail(UBG_E_INVALID, "domain index out of range");
  return UBG_OK;
}

std::string batch_key(const ubg_batch *b, const ubg_params &p, int k) {
  // FNV-1a over the sizes and the smoothing counts
  uint64_t hsh = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    for (int i = 0; i < 8; i++) {
      hsh ^= (v >> (8 * i)) & 0xff;
      hsh *= 1099511628211ull;
    }
  };
  for (size_t i = 0; i < b->ws.size(); i++) {
    mix((uint64_t)b->ws[i]);
    mix((uint64_t)b->hs[i]);
  }
  mix((uint64_t)p.nu_pre);
  mix((uint64_t)p.nu_post);
  mix((uint64_t)p.nu_coarse);
  mix((uint64_t)k);
  char buf[40];
  std::snprintf(buf, sizeof(buf), "%016llx", (unsigned long long)hsh);
  return buf;
}

} // namespace

extern "C" {

UBG_API int ubg_batch_bytes(int n, const int *widths, const int *heights, size_t *bytes) {
  if (n <= 0 || !widths || !heights || !bytes) return fail(UBG_E_INVALID, "bad arguments");
  size_t tot = 0;
  for (int i = 0; i < n; i++) {
    ubg::BShape s;
    UBG_TRY(ubg::batch_shape(widths[i], heights[i], &s));
    tot += s.geom_bytes + s.dom_bytes;
  }
  *bytes = tot;
  return UBG_OK;
}

UBG_API int ubg_batch_create(int n, const int *widths, const int *heights, const ubg_params *prm,
                             ubg_batch **out) {
  if (!out) return fail(UBG_E_INVALID, "null output pointer");
  int nd = 0;
  ubg_device_count(&nd);
  if (nd <= 0) return fail(UBG_E_NO_DEVICE, "no HIP device visible");
  size_t bytes = 0;
  UBG_TRY(ubg_batch_bytes(n, widths, heights, &bytes));
  ubg_params P;
  default_params(&P);
  if (prm) {
    UBG_TRY(check_params(prm));
    P = *prm;
  }
  auto *b = new ubg_batch();
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return fail(UBG_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  b->ws.assign(widths, widths + n);
  b->hs.assign(heights, heights + n);
  b->ready.assign(n, false);
  b->doms.assign(n, nullptr);
  int rc = b->arena.init(bytes + 4096);
  if (rc == UBG_OK) rc = b->eng.init(b->stream, P, 4);
  if (rc != UBG_OK) {
    ubg_batch_destroy(b);
    return rc;
  }
  *out = b;
  return UBG_OK;
}

UBG_API int ubg_batch_destroy(ubg_batch *b) {
  if (!b) return UBG_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto *d : b->doms)
    if (d) b->eng.domain_release(b->arena, d);
  b->doms.clear();
  b->arena.release();
  for (hipEvent_t e : b->ev_ready)
    if (e) (void)hipEventDestroy(e);
  for (auto &t : b->slot) {
    if (t.host) (void)hipHostFree(t.host);
    if (t.dev) (void)hipFree(t.dev);
    if (t.used) (void)hipEventDestroy(t.used);
  }
  hipStream_t s = b->stream;
  delete b; // the engine frees its tables first
  if (s) (void)hipStreamDestroy(s);
  return UBG_OK;
}

namespace {

int set_domain(ubg_batch *b, int d, const float *flag, const float *rhs, float h, bool need_ring) {
  UBG_TRY(batch_check(b, d));
  if (!flag || !rhs) return fail(UBG_E_INVALID, "null argument");
  if (!(h > 0.0f) || !std::isfinite(h)) return fail(UBG_E_INVALID, "h must be positive");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (b->doms[d]) {
    b->eng.invalidate_members();
    b->eng.domain_release(b->arena, b->doms[d]);
    b->doms[d] = nullptr;
  }
  ubg::BGeom *geo = nullptr;
  UBG_TRY(b->eng.geom_create(b->arena, b->ws[d], b->hs[d], flag, &geo, need_ring));
  ubg::BDomain *dom = nullptr;
  int rc = b->eng.domain_create(b->arena, geo, h, &dom);
  b->eng.geom_release(b->arena, geo); // the domain holds its reference
  if (rc) return rc;
  b->doms[d] = dom;
  b->proj_active = false;
  rc = b->eng.domain_set_rhs(dom, rhs);
  b->ready[d] = rc == UBG_OK;
  return rc;
}

} // namespace

UBG_API int ubg_batch_set_domain(ubg_batch *b, int d, const float *flag, const float *rhs, float h) {
  return set_domain(b, d, flag, rhs, h, true);
}

UBG_API int ubg_batch_set_projection_domain(ubg_batch *b, int d, const float *flag,
                                            const float *rhs, float h) {
  return set_domain(b, d, flag, rhs, h, false);
}

UBG_API int ubg_batch_set_rhs(ubg_batch *b, int d, const float *rhs) {
  UBG_TRY(batch_check(b, d));
  if (!rhs) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (!b->doms[d]) return fail(UBG_E_STATE, "domain not set");
  return b->eng.domain_set_rhs(b->doms[d], rhs);
}

UBG_API int ubg_batch_set_guess(ubg_batch *b, int d, const double *x) {
  UBG_TRY(batch_check(b, d));
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (!b->doms[d]) return fail(UBG_E_STATE, "domain not set");
  return b->eng.domain_set_guess(b->doms[d], x);
}

UBG_API int ubg_batch_set_pressure(ubg_batch *b, int d, const float *p) {
  UBG_TRY(batch_check(b, d));
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (!b->doms[d]) return fail(UBG_E_STATE, "domain not set");
  if (p) {
    const size_t n = (size_t)b->ws[d] * b->hs[d];
    for (size_t i = 0; i < n; i++)
      if (!std::isfinite(p[i])) return fail(UBG_E_NONFINITE, "pressure has a non-finite value");
  }
  return b->eng.domain_set_level0(b->doms[d], p, nullptr, hipMemcpyHostToDevice);
}

UBG_API int ubg_batch_get_pressure(ubg_batch *b, int d, float *p) {
  UBG_TRY(batch_check(b, d));
  if (!p) return fail(UBG_E_INVALID, "null output");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (!b->doms[d]) return fail(UBG_E_STATE, "domain not set");
  return b->eng.domain_get_pressure(b->doms[d], p, hipMemcpyDeviceToHost);
}

namespace {

// every domain a running member with an active solve state
int proj_members(ubg_batch *b) {
  if (!b->proj_active) {
    for (auto *d : b->doms) UBG_TRY(b->eng.domain_start(d, 0.5, 1));
    b->proj_active = true;
  }
  return b->eng.set_members(b->doms, {});
}

} // namespace

UBG_API int ubg_batch_vcycles(ubg_batch *b, int ncycles, int zero_gradient_bc) {
  if (!b) return fail(UBG_E_INVALID, "null batch");
  if (ncycles < 1 || ncycles > 1000) return fail(UBG_E_INVALID, "ncycles must be in 1..1000");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  for (size_t i = 0; i < b->doms.size(); i++)
    if (!b->ready[i]) return fail(UBG_E_STATE, "domain " + std::to_string(i) + " not set");
  // the level-0 rhs of the V-cycles is the domain's rhs
  for (auto *d : b->doms)
    UBG_HIP(hipMemcpyAsync(d->f[0], d->rhs, (size_t)d->g->W * d->g->H * 4,
                           hipMemcpyDeviceToDevice, b->stream));
  UBG_TRY(proj_members(b));
  UBG_TRY(b->eng.enqueue_vcycles(ncycles, zero_gradient_bc != 0));
  UBG_HIP(hipStreamSynchronize(b->stream));
  return check_launch();
}

namespace {

// the copies of `jobs`, kCopyJobs per launch
int launch_copies(const std::vector<ubg::CopyJob> &jobs, hipStream_t s) {
  for (size_t b = 0; b < jobs.size(); b += ubg::kCopyJobs) {
    ubg::CopyJobs cj{};
    const int nj = (int)std::min<size_t>(ubg::kCopyJobs, jobs.size() - b);
    int maxn = 0;
    for (int k = 0; k < nj; k++) {
      cj.j[k] = jobs[b + k];
      maxn = std::max(maxn, cj.j[k].n);
    }
    const int gx = std::max(1, std::min((maxn + 1023) / 1024, 512));
    hipLaunchKernelGGL(ubg::kb_copy_level0, dim3(gx, nj), dim3(256), 0, s, cj, nj);
  }
  return check_launch();
}

bool same_solver(const ubg_params &a, const ubg_params &b) {
  return a.nu_pre == b.nu_pre && a.nu_post == b.nu_post && a.nu_coarse == b.nu_coarse &&
         a.vcycles_per_step == b.vcycles_per_step && a.alpha == b.alpha;
}

// whether a simulation appears more than once in the list (one batched
// step would then advance its fields twice in the same launches)
bool repeats(ubg_sim *const *sims, int n) {
  std::vector<ubg_sim *> v(sims, sims + n);
  std::sort(v.begin(), v.end());
  return std::adjacent_find(v.begin(), v.end()) != v.end();
}

// (re)build domain d from simulation s's level-0 flag and h
int domain_from_sim(ubg_batch *b, int d, ubg_sim *s) {
  const ubg::SimLevel0 q = ubg::sim_level0(s);
  const size_t n = (size_t)q.W * q.H;
  std::vector<uint8_t> fl(n);
  UBG_HIP(hipStreamSynchronize(q.stream));
  UBG_HIP(hipMemcpy(fl.data(), q.flag, n, hipMemcpyDeviceToHost));
  std::vector<float> flag(n), zero(n, 0.0f);
  for (size_t i = 0; i < n; i++) flag[i] = fl[i] ? 1.0f : 0.0f;
  UBG_TRY(ubg_batch_set_projection_domain(b, d, flag.data(), zero.data(), q.h));
  b->src[d] = s;
  b->src_gen[d] = q.flag_gen;
  b->proj_active = false;
  return UBG_OK;
}

} // namespace

UBG_API int ubg_batch_create_for_sims(ubg_sim *const *sims, int n, ubg_batch **out) {
  if (!sims || !out || n < 1) return fail(UBG_E_INVALID, "null argument or n < 1");
  std::vector<int> ws(n), hs(n);
  ubg_params prm{};
  for (int i = 0; i < n; i++) {
    if (!sims[i]) return fail(UBG_E_INVALID, "null simulation");
    const ubg::SimLevel0 q = ubg::sim_level0(sims[i]);
    if (i == 0) prm = q.prm;
    else if (!same_solver(prm, q.prm))
      return fail(UBG_E_INVALID, "the simulations' multigrid parameters differ");
    ws[i] = q.W;
    hs[i] = q.H;
  }
  if (repeats(sims, n)) return fail(UBG_E_INVALID, "a simulation appears twice in the list");
  ubg_batch *b = nullptr;
  UBG_TRY(ubg_batch_create(n, ws.data(), hs.data(), &prm, &b));
  b->src.assign(n, nullptr);
  b->src_gen.assign(n, 0u);
  for (int i = 0; i < n; i++) {
    const int rc = domain_from_sim(b, i, sims[i]);
    if (rc != UBG_OK) {
      ubg_batch_destroy(b);
      return rc;
    }
  }
  *out = b;
  return UBG_OK;
}

UBG_API int ubg_sims_step(ubg_batch *b, ubg_sim *const *sims, int n, const float *dts) {
  if (!b || !sims || !dts) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (n != (int)b->doms.size() || b->src.size() != b->doms.size())
    return fail(UBG_E_INVALID, "the batch was not made for these simulations");
  const ubg_params &bp = b->eng.params();
  std::vector<ubg::SimLevel0> q(n);
  for (int i = 0; i < n; i++) {
    if (!sims[i]) return fail(UBG_E_INVALID, "null simulation");
    q[i] = ubg::sim_level0(sims[i]);
    if (q[i].W != b->ws[i] || q[i].H != b->hs[i] || !same_solver(bp, q[i].prm))
      return fail(UBG_E_INVALID, "simulation " + std::to_string(i) + " does not match its domain");
    if (!std::isfinite(dts[i])) return fail(UBG_E_INVALID, "dt is not finite");
  }
  if (repeats(sims, n)) return fail(UBG_E_INVALID, "a simulation appears twice in the list");
  // geometry of a domain whose simulation changed its flags (or is another one)
  for (int i = 0; i < n; i++)
    if (b->src[i] != sims[i] || b->src_gen[i] != q[i].flag_gen) {
      UBG_TRY(domain_from_sim(b, i, sims[i]));
      q[i] = ubg::sim_level0(sims[i]);
    }
  if (!b->ev_done) {
    auto e = std::make_shared<ubg::ExtEvent>();
    UBG_HIP(hipEventCreateWithFlags(&e->ev, hipEventDisableTiming));
    b->ev_done = e;
  }
  while ((int)b->ev_ready.size() < n) {
    hipEvent_t e = nullptr;
    UBG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->ev_ready.push_back(e);
  }
  // a free table slot (allocated before any simulation's state moves)
  auto &T = b->slot[b->next_slot];
  b->next_slot ^= 1;
  const size_t nt = (size_t)4 * n;
  if (!T.used) UBG_HIP(hipEventCreateWithFlags(&T.used, hipEventDisableTiming));
  UBG_HIP(hipEventSynchronize(T.used)); // the step two calls ago has read it
  if (T.cap < nt) {
    if (T.host) (void)hipHostFree(T.host);
    if (T.dev) (void)hipFree(T.dev);
    T.host = nullptr;
    T.dev = nullptr;
    T.cap = 0;
    UBG_HIP(hipHostMalloc((void **)&T.host, nt * sizeof(ubg::SimRec), hipHostMallocDefault));
    UBG_HIP(hipMalloc((void **)&T.dev, nt * sizeof(ubg::SimRec)));
    T.cap = nt;
  }
  // The per-cell stages of every simulation's step as records (their
  // applyAccumulatedVelocity / boundary / face-factor preludes, when due,
  // run on the simulations' own streams); the batch's stream then waits for
  // every simulation whose stream had work of its own since the last step.
  std::vector<ubg::SimRec> rec((size_t)4 * n);
  std::vector<char> sinks(n, 0);
  for (int i = 0; i < n; i++) {
    bool has = false;
    UBG_TRY(ubg::sim_batched_records(sims[i], dts[i], &rec[(size_t)4 * i], &has));
    sinks[i] = has;
  }
  for (int i = 0; i < n; i++)
    if (ubg::sim_take_own_work(sims[i])) {
      UBG_HIP(hipEventRecord(b->ev_ready[i], q[i].stream));
      UBG_HIP(hipStreamWaitEvent(b->stream, b->ev_ready[i], 0));
    }
  // stage tables: stage k of simulation i at [k n + i], rows numbered
  // through the simulations' interiors
  int rows = 0, maxw = 0;
  bool buf = true;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 4; k++) {
      ubg::SimRec r = rec[(size_t)4 * i + k];
      r.row0 = rows;
      T.host[(size_t)k * n + i] = r;
    }
    rows += q[i].H - 2;
    maxw = std::max(maxw, q[i].W);
    buf = buf && (size_t)q[i].W * q[i].H * 4 < ((size_t)1 << 31);
  }
  UBG_HIP(hipMemcpyAsync(T.dev, T.host, nt * sizeof(ubg::SimRec), hipMemcpyHostToDevice, b->stream));
  const ubg::SimRec *tab = T.dev;
  // diffuse (two iterations), advect, setVBCs + divergence of all
  // simulations, one launch each; sinks stamped in list order after the
  // divergence
  ubg::launch_sims_diffuse(tab, n, rows, maxw, b->stream);
  ubg::launch_sims_diffuse(tab + n, n, rows, maxw, b->stream);
  ubg::launch_sims_advect(tab + 2 * n, n, rows, maxw, buf, b->stream);
  ubg::launch_sims_vbcs_divergence(tab + 3 * n, n, rows, maxw, b->stream);
  UBG_TRY(check_launch());
  for (int i = 0; i < n; i++)
    if (sinks[i]) UBG_TRY(ubg::sim_apply_sinks_on(sims[i], b->stream));
  // the V-cycles of all of them (the level-0 pressure and rhs gathered in
  // one launch, the pressure scattered back in another)
  UBG_TRY(proj_members(b));
  std::vector<ubg::CopyJob> in(n), out(n);
  for (int i = 0; i < n; i++) {
    ubg::BDomain *d = b->doms[i];
    in[i] = ubg::CopyJob{q[i].p, q[i].f, d->p[0][d->cur[0]], d->f[0], q[i].W * q[i].H};
  }
  UBG_TRY(launch_copies(in, b->stream));
  UBG_TRY(b->eng.enqueue_vcycles(bp.vcycles_per_step, true));
  for (int i = 0; i < n; i++) {
    ubg::BDomain *d = b->doms[i];
    out[i] = ubg::CopyJob{d->p[0][d->cur[0]], nullptr, q[i].p, nullptr, q[i].W * q[i].H};
  }
  UBG_TRY(launch_copies(out, b->stream));
  // setPBC + correction + setVBCs + saveCurrent of all of them
  ubg::launch_sims_correct_tail(tab + 3 * n, n, rows, maxw, b->stream);
  UBG_TRY(check_launch());
  UBG_HIP(hipEventRecord(T.used, b->stream));
  // each simulation's next own work waits for this step
  UBG_HIP(hipEventRecord(b->ev_done->ev, b->stream));
  for (int i = 0; i < n; i++) ubg::sim_set_ext_event(sims[i], b->ev_done);
  return UBG_OK;
}

UBG_API int ubg_batch_solve(ubg_batch *b, const ubg_ir_params *prm, int *graph_replays) {
  if (!b) return fail(UBG_E_INVALID, "null batch");
  ubg_ir_params P;
  ubg_ir_params_default(&P);
  if (prm) P = *prm;
  if (!(P.tol > 0.0 && P.tol < 1.0)) return fail(UBG_E_INVALID, "tol must be in (0, 1)");
  if (P.vcycles_per_iter < 1 || P.vcycles_per_iter > 256 || P.max_iter < 1 || P.max_iter > 1000000)
    return fail(UBG_E_INVALID, "vcycles_per_iter or max_iter out of range");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  for (size_t i = 0; i < b->doms.size(); i++) {
    if (!b->ready[i]) return fail(UBG_E_STATE, "domain " + std::to_string(i) + " not set");
    if (!b->doms[i]->g->ring_fluid)
      return fail(UBG_E_FLAG_VALUE, "domain " + std::to_string(i) +
                                        " has a solid outer-ring cell: projection solves only");
  }
  b->eng.set_vcycles(P.vcycles_per_iter);
  b->proj_active = false;
  UBG_TRY(b->eng.set_level0_streamed(false));
  for (auto *d : b->doms) UBG_TRY(b->eng.domain_start(d, P.tol, P.max_iter));
  // first update of every domain (r0), then replayed two-step graphs
  UBG_TRY(b->eng.set_members(b->doms, {}));
  UBG_TRY(b->eng.enqueue_first_update());
  UBG_TRY(b->eng.sync());
  // replay two-step graphs; when domains finish, shrink the member set to
  // the active ones (new tables, tiles and graph) so finished domains cost
  // no dispatch
  int replays = 0;
  while (b->eng.active_count() > 0 && 2 * replays < P.max_iter + 2) {
    std::vector<ubg::BDomain *> act = b->eng.active_members();
    if (act.size() != b->eng.members()) UBG_TRY(b->eng.set_members(act, {}));
    UBG_TRY(b->eng.run_steps_graph(2));
    UBG_TRY(b->eng.sync());
    replays++;
  }
  if (graph_replays) *graph_replays = replays;
  int bad = 0;
  for (auto *d : b->doms) {
    ubg::DState s;
    UBG_TRY(b->eng.domain_read_state(d, &s));
    d->xcur = s.xcur;
    if (s.nonfinite) return fail(UBG_E_NONFINITE, "a domain's residual is not finite");
    if (!s.converged) bad++;
  }
  if (bad)
    return fail(UBG_E_NOT_CONVERGED, std::to_string(bad) + " domain(s) stopped short of tol");
  return UBG_OK;
}

UBG_API int ubg_batch_result(ubg_batch *b, int d, ubg_ir_result *res) {
  UBG_TRY(batch_check(b, d));
  if (!res) return fail(UBG_E_INVALID, "null result");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (!b->doms[d]) return fail(UBG_E_STATE, "domain not set");
  ubg::DState s;
  UBG_TRY(b->eng.domain_read_state(b->doms[d], &s));
  std::memset(res, 0, sizeof(*res));
  res->iterations = s.iters;
  res->vcycles = s.vcycles;
  res->converged = s.converged;
  res->r0_norm = s.r0;
  res->rel_residual = s.rel;
  for (int i = 0; i < std::min(s.iters, UBG_IR_HISTORY); i++) res->history[i] = s.hist[i];
  return UBG_OK;
}

UBG_API int ubg_batch_get_solution(ubg_batch *b, int d, double *x) {
  UBG_TRY(batch_check(b, d));
  if (!x) return fail(UBG_E_INVALID, "null destination");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  if (!b->doms[d]) return fail(UBG_E_STATE, "domain not set");
  return b->eng.domain_get_solution(b->doms[d], x);
}

UBG_API int ubg_batch_set_tuning(ubg_batch *b, const ubg_batch_tuning *t) {
  if (!b || !t) return fail(UBG_E_INVALID, "null argument");
  if (!shape_ok(t->tile_w, t->tile_h)) return fail(UBG_E_INVALID, "tile shape not compiled");
  if (t->blocks < 0) return fail(UBG_E_INVALID, "negative block count");
  const int rows = t->pass_rows;
  if (!pass_rows_ok(rows)) return fail(UBG_E_INVALID, "pass_rows must be 0, 8, 16 or 32");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  b->eng.set_tuning(ubg::Tuning{t->tile_w, t->tile_h, t->blocks, rows});
  return UBG_OK;
}

UBG_API int ubg_batch_get_tuning(ubg_batch *b, ubg_batch_tuning *t) {
  if (!b || !t) return fail(UBG_E_INVALID, "null argument");
  const ubg::Tuning &x = b->eng.tuning();
  t->tile_w = x.tx;
  t->tile_h = x.ty;
  t->blocks = x.blocks;
  t->pass_rows = x.prow;
  t->sweep_ms = 0.0;
  return UBG_OK;
}

UBG_API int ubg_batch_autotune(ubg_batch *b, const char *db_path, ubg_batch_tuning *out) {
  if (!b) return fail(UBG_E_INVALID, "null batch");
  std::lock_guard<std::recursive_mutex> g(b->mu);
  for (size_t i = 0; i < b->doms.size(); i++)
    if (!b->ready[i]) return fail(UBG_E_STATE, "domain " + std::to_string(i) + " not set");
  const std::string key = batch_key(b, b->eng.params(), b->eng.vcycles());
  if (db_path) {
    std::ifstream in(db_path);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string k;
      int tx, ty, bl, rows;
      double ms;
      if (ls >> k >> tx >> ty >> bl >> rows >> ms && k == key && shape_ok(tx, ty) && bl >= 0 &&
          pass_rows_ok(rows)) {
        b->eng.set_tuning(ubg::Tuning{tx, ty, bl, rows});
        if (out) *out = ubg_batch_tuning{tx, ty, bl, rows, ms};
        return UBG_OK;
      }
    }
  }
  UBG_TRY(b->eng.set_level0_streamed(false));
  UBG_TRY(b->eng.set_members(b->doms, {}));
  double ms = 0.0;
  UBG_TRY(b->eng.autotune(&ms));
  const ubg::Tuning &t = b->eng.tuning();
  if (out) *out = ubg_batch_tuning{t.tx, t.ty, t.blocks, t.prow, ms};
  if (db_path) {
    std::ofstream o(db_path, std::ios::app);
    if (!o) return fail(UBG_E_INVALID, std::string("cannot write tuning db ") + db_path);
    o << key << " " << t.tx << " " << t.ty << " " << t.blocks << " " << t.prow << " " << ms << "\n";
  }
  return UBG_OK;
}

} // extern "C"
