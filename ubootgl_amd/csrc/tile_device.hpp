// tile_device.hpp -- LDS-tiled multigrid level passes (MG::solveLevel,
// pressure_solver.cpp:201-248) as device functions over one tile, shared by
// the single-domain kernels (mg_tile.hip) and the batched solver (batch.hip):
//   pre : nu red/black sweeps, residual, full-weighting restriction
//   post: prolongation, correction, nu red/black sweeps
// A workgroup stages its tile plus a halo of 2 cells per sweep (and the
// residual's and restriction's reach) in LDS and recomputes the halo cells
// instead of exchanging them; a cell's value after k colour phases depends
// only on cells within k of it, so every cell a workgroup writes is the value
// the sweep-by-sweep schedule produces, bit for bit.  The sweeps read one
// buffer and write the other, so halos read by one workgroup are never
// overwritten by another.
#pragma once

#include "mg_device.hpp"

namespace ubg {

constexpr int PITCH = 130; // LDS row pitch; 64 lanes cover one row's cells of a colour
constexpr int CPITCH = 68; // coarse rows under a tile of at most 127 fine columns
constexpr int NU_TILE = 3; // largest smoothing count with a tiled pass
constexpr int KL = 8;      // global loads in flight per thread while staging

// Colour phase j (1-based) over the staged region: the cells of colour
// `want` ((x + y) & 1 == want; red = 1 first, pressure_solver.cpp:35-47) in
// the domain's interior that can still become exact: after phase j-1 only
// cells at distance >= j-1 from a halo edge of the region hold the
// sweep-by-sweep values, so phase j updates those at distance >= j (an edge
// on the domain boundary is no halo edge: nothing outside it is missing).
// F holds (f * h) * h, or with RAWF f itself, scaled here by the same two
// products (hh is h).
template <int WAVES, bool RAWF = false>
__device__ __forceinline__ void tile_phase(float *P, const float *F, const uint8_t *M, int LW,
                                           int LH, int gx0, int gy0, int w, int h, int want,
                                           int j, float alpha, float hh = 0.0f) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // row loop in SALU
  const int xlo = gx0 == 0 ? 1 : j, ylo = gy0 == 0 ? 1 : j;
  const int xhi = (gx0 + LW == w) ? LW - 2 : LW - 1 - j;
  const int yhi = (gy0 + LH == h) ? LH - 2 : LH - 1 - j;
  // a wave's rows step by WAVES (even), so its column of the colour is the
  // same in every row
  static_assert(WAVES % 2 == 0, "a wave's rows must share the colour's column parity");
  const int ly0 = ylo + wave;
  const int lx = xlo + (((gx0 + xlo + gy0 + ly0) ^ want) & 1) + 2 * lane;
  if (lx <= xhi) {
    for (int ly = ly0; ly <= yhi; ly += WAVES) {
      const int i = ly * PITCH + lx;
      const float fhh = RAWF ? F[i] * hh * hh : F[i];
      P[i] = smooth_fhh(P[i], P[i - 1], P[i + 1], P[i - PITCH], P[i + PITCH], fhh, M[i], alpha);
    }
  }
  __syncthreads();
}

// Stage an LW x LH region of p, (f * hh) * hh and the mask (origin gx0,
// gy0) into LDS with pitch PITCH.  Each round issues KL loads of each array
// per thread before the first LDS store, so a round costs one memory latency.
// ZP: p is zero on entry and is not read (the region's p words are set to
// +0, the value the restriction above stored).  RAWF: F holds f unscaled.
template <int TPB_M, bool ZP = false, bool RAWF = false>
__device__ __forceinline__ void stage_fine(float *P, float *F, uint8_t *M, const float *p,
                                           const float *f, const uint8_t *m, int w, int gx0,
                                           int gy0, int LW, int LH, float hh) {
  const int n = LW * LH;
  // element j = threadIdx.x + TPB_M * k of the region, as (row, column)
  // stepped by (dq, dr) = divmod(TPB_M, LW) instead of divided per element
  const int dq = TPB_M / LW, dr = TPB_M - dq * LW;
  int ly = (int)threadIdx.x / LW, lx = (int)threadIdx.x - ly * LW;
  for (int base = 0; base < n; base += KL * TPB_M) {
    float vp[KL], vf[KL];
    uint8_t vm[KL];
    int li[KL];
#pragma unroll
    for (int k = 0; k < KL; k++) {
      li[k] = -1;
      if (ly < LH) {
        const size_t g = (size_t)(gy0 + ly) * w + gx0 + lx;
        vp[k] = ZP ? 0.0f : p[g];
        vf[k] = f[g];
        vm[k] = m[g];
        li[k] = ly * PITCH + lx;
      }
      lx += dr;
      ly += dq;
      if (lx >= LW) {
        lx -= LW;
        ly++;
      }
    }
#pragma unroll
    for (int k = 0; k < KL; k++)
      if (li[k] >= 0) {
        P[li[k]] = vp[k];
        F[li[k]] = RAWF ? vf[k] : vf[k] * hh * hh;
        M[li[k]] = vm[k];
      }
  }
}

// LDS bytes of the passes for tiles of oty rows
inline size_t tile_pre_lds(int oty, int nu) { return (size_t)PITCH * (oty + 3 + 4 * nu) * 9; }
inline size_t tile_post_lds(int oty, int nu) {
  const int lh = oty + 4 * nu;
  return (size_t)PITCH * lh * 9 + (size_t)CPITCH * (lh / 2 + 2) * 5;
}
// widest even tile whose staged rows fit the LDS pitch
inline int tile_otx_max(int nu) { return (PITCH - 3 - 4 * nu) & ~1; }
// columns of a w-wide level's tiles: as few as the pitch allows, of even,
// evenly split width otx (ntx of them)
inline void tile_cols(int w, int nu, int &otx, int &ntx) {
  const int otx_max = tile_otx_max(nu);
  ntx = (w + otx_max - 1) / otx_max;
  otx = ((w + ntx - 1) / ntx + 1) & ~1;
  ntx = (w + otx - 1) / otx;
}
// Tile workgroups have 1024 threads: 16 waves share a tile's colour phases,
// so a phase walks few rows per wave.  Two fit a CU, so a launch aims for
// kTileTarget tiles, one residency wave (a second, partial wave costs a
// whole tile lifetime).  Rows per tile: the fewest (even, 8 to kTileRowsMax)
// whose tiles fit that target, within 64 KiB of LDS.  Measured on gfx950,
// 2048x1024 with 38 rows per 1024-thread tile (486 tiles) against 20 rows
// per 512-thread tile (936 tiles): pre 16.0 instead of 17.8 us, post 16.0
// instead of 17.2 us (fewer halo rows, fewer rows per wave per phase).
constexpr int TILE_TPB = 1024;
constexpr int kTileTarget = 512;
constexpr int kTileRowsMax = 48;
inline bool tile_lds_ok(int oty, int nu) {
  return tile_pre_lds(oty, nu) <= 65536 && tile_post_lds(oty, nu) <= 65536;
}
// The batched solver's tiles (many domains' tiles walked by each workgroup)
// keep 512-thread workgroups for levels of at least kTileBigCells cells in
// all and 1024-thread ones below, of 8 to 32 rows, aiming for
// tile_target(tpb) workgroups: measured on gfx950, the single-domain rule
// above made the 256-domain batch 12% slower and the served mix 9%.
constexpr size_t kTileBigCells = size_t(1) << 18;
inline int tile_target(int tpb) { return tpb == 512 ? 1024 : 512; }

// Pre-smoothing + calculateResidualField + restrict (pressure_solver.cpp:
// 213-227): reads p, writes the smoothed level to p_out (every cell of the
// tile), the coarse rhs fc and a zero coarse solution ec (interior cells;
// ec null: left as it is, for a coarse pass that does not read it).
// Tile origin (X0, Y0) of an otx x oty partition (otx, oty even); smem
// holds tile_pre_lds(oty, NU) bytes.  ZP: p is zero on entry (a level below
// the finest, whose rhs and zero solution the restriction above wrote) and
// is not read.
template <int NU, int TPB_M, bool ZP = false>
__device__ __forceinline__ void
tile_pre_body(char *smem, int X0, int Y0, int otx, int oty, const float *__restrict__ p,
              float *__restrict__ p_out, const float *__restrict__ f,
              const uint8_t *__restrict__ m, int w, int h, float hh, float alpha, float ihsq,
              float *__restrict__ fc, float *__restrict__ ec, int wc, int hc) {
  const int lh_max = oty + 3 + 4 * NU;
  float *P = (float *)smem;
  float *F = P + PITCH * lh_max;
  uint8_t *M = (uint8_t *)(F + PITCH * lh_max);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // the restriction reads residuals at [X0-1, X0+otx-1], which read p at
  // [X0-2, X0+otx]; each sweep widens that by 2
  const int gx0 = max(X0 - 2 - 2 * NU, 0), gy0 = max(Y0 - 2 - 2 * NU, 0);
  const int gx1 = min(X0 + otx + 2 * NU, w - 1), gy1 = min(Y0 + oty + 2 * NU, h - 1);
  const int LW = gx1 - gx0 + 1, LH = gy1 - gy0 + 1;
  constexpr int WAVES = TPB_M / 64;
  // F holds f itself: the sweeps form (f * h) * h where they use it and the
  // residual reads f from LDS (rows ry0..ry1, columns rx0..rx1 <= 127)
  // instead of loading it a second time
  const int rx0 = max(X0 - 1, 1), rx1 = min(X0 + otx - 1, w - 2);
  const int ry0 = max(Y0 - 1, 1), ry1 = min(Y0 + oty - 1, h - 2);
  stage_fine<TPB_M, ZP, true>(P, F, M, p, f, m, w, gx0, gy0, LW, LH, hh);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NU; s++) {
    tile_phase<WAVES, true>(P, F, M, LW, LH, gx0, gy0, w, h, 1, 2 * s + 1, alpha, hh);
    tile_phase<WAVES, true>(P, F, M, LW, LH, gx0, gy0, w, h, 0, 2 * s + 2, alpha, hh);
  }
  // the tile's own cells (boundary cells unchanged) ...
  const int ox1 = min(X0 + otx, w), oy1 = min(Y0 + oty, h);
  for (int y = Y0 + wave; y < oy1; y += TPB_M / 64)
    for (int x = X0 + lane; x < ox1; x += 64)
      p_out[(size_t)y * w + x] = P[(y - gy0) * PITCH + (x - gx0)];
  // ... and the residual in place of F where the restriction reads it
  for (int y = ry0 + wave; y <= ry1; y += WAVES)
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int x = rx0 + lane + 64 * c;
      if (x <= rx1) {
        const int i = (y - gy0) * PITCH + (x - gx0);
        F[i] = residual_of(P[i], P[i - 1], P[i + 1], P[i - PITCH], P[i + PITCH], F[i], M[i],
                           ihsq);
      }
    }
  __syncthreads();
  // restrict (pressure_solver.cpp:118-132) for the coarse cells whose centre
  // 2xc lies in the tile
  const int cx0 = max(X0 / 2, 1), cx1 = min((X0 + otx) / 2, wc - 1);
  const int cy0 = max(Y0 / 2, 1), cy1 = min((Y0 + oty) / 2, hc - 1);
  for (int yc = cy0 + wave; yc < cy1; yc += TPB_M / 64)
    for (int xc = cx0 + lane; xc < cx1; xc += 64) {
      const int c = (2 * yc - gy0) * PITCH + (2 * xc - gx0);
      auto R = [&](int dx, int dy) { return F[c + dy * PITCH + dx]; };
      float v = 0.0f;
      v = R(-1, -1) * 1.0f + R(0, -1) * 2.0f + R(1, -1) * 1.0f;
      v += R(-1, 0) * 2.0f + R(0, 0) * 4.0f + R(1, 0) * 2.0f;
      v += R(-1, 1) * 1.0f + R(0, 1) * 2.0f + R(1, 1) * 1.0f;
      const size_t ci = (size_t)yc * wc + xc;
      fc[ci] = v * (1.0f / 16.0f);
      if (ec) ec[ci] = 0.0f;
    }
}

// flag sum + the double literal 0.0001, stored as float (pressure_solver.cpp:149)
__device__ __forceinline__ float eps_flag_sum(int s) { return (float)((double)s + 0.0001); }

// prolongate + correct (pressure_solver.cpp:134-181) + post-smoothing: reads
// p and the coarse solution ec, writes p_out (every cell of the tile).
// smem holds tile_post_lds(oty, NU) bytes.
template <int NU, int TPB_M>
__device__ __forceinline__ void
tile_post_body(char *smem, int X0, int Y0, int otx, int oty, const float *__restrict__ p,
               float *__restrict__ p_out, const float *__restrict__ f,
               const uint8_t *__restrict__ m, int w, int h, float hh, float alpha,
               const float *__restrict__ ec, const uint8_t *__restrict__ mc, int wc, int hc) {
  const int lh_max = oty + 4 * NU;
  const int ch_max = lh_max / 2 + 2;
  float *P = (float *)smem;
  float *F = P + PITCH * lh_max;
  float *E = F + PITCH * lh_max;          // coarse solution, pitch CPITCH
  uint8_t *M = (uint8_t *)(E + CPITCH * ch_max);
  uint8_t *FC = M + PITCH * lh_max;       // coarse flags, pitch CPITCH
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gx0 = max(X0 - 2 * NU, 0), gy0 = max(Y0 - 2 * NU, 0);
  const int gx1 = min(X0 + otx - 1 + 2 * NU, w - 1), gy1 = min(Y0 + oty - 1 + 2 * NU, h - 1);
  const int LW = gx1 - gx0 + 1, LH = gy1 - gy0 + 1;
  const int ex0 = gx0 >> 1, ey0 = gy0 >> 1;
  const int ex1 = min((gx1 >> 1) + 1, wc - 1), ey1 = min((gy1 >> 1) + 1, hc - 1);
  {
    // coarse values, two per thread per round: the first round's loads are
    // issued before the fine region's and stored after it, so the two
    // memory latencies overlap (further rounds, for regions of more than
    // 2 TPB_M coarse cells, follow one after the other)
    const int CW = ex1 - ex0 + 1, CH = ey1 - ey0 + 1, nc = CW * CH;
    // (row, column) of element threadIdx.x + TPB_M * k, stepped as in stage_fine
    const int CWs = CW > 0 ? CW : 1; // CW < 1: nc < 1, nothing is staged
    const int dq = TPB_M / CWs, dr = TPB_M - dq * CWs;
    int cy = (int)threadIdx.x / CWs, cx = (int)threadIdx.x - cy * CWs;
    float ve[2];
    uint8_t vc[2];
    int li[2];
    auto load_round = [&] {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        li[k] = -1;
        if (cy < CH) {
          const size_t g = (size_t)(ey0 + cy) * wc + ex0 + cx;
          ve[k] = ec[g];
          vc[k] = mc[g] & MK_SELF;
          li[k] = cy * CPITCH + cx;
        }
        cx += dr;
        cy += dq;
        if (cx >= CWs) {
          cx -= CWs;
          cy++;
        }
      }
    };
    auto store_round = [&] {
#pragma unroll
      for (int k = 0; k < 2; k++)
        if (li[k] >= 0) {
          E[li[k]] = ve[k];
          FC[li[k]] = vc[k];
        }
    };
    load_round();
    stage_fine<TPB_M>(P, F, M, p, f, m, w, gx0, gy0, LW, LH, hh);
    store_round();
    for (int base = 2 * TPB_M; base < nc; base += 2 * TPB_M) {
      load_round();
      store_round();
    }
  }
  __syncthreads();
  // e for every interior cell of the region (no neighbour dependence), by
  // 2x2 blocks of fine cells over one coarse cell (a, b): even-even, odd-even,
  // even-odd, odd-odd, each by its parity rule
  {
    const int xlo = max(gx0, 1), xhi = min(gx1, w - 2);
    const int ylo = max(gy0, 1), yhi = min(gy1, h - 2);
    const int a0 = xlo >> 1, na = (xhi >> 1) - a0 + 1;
    const int b0 = ylo >> 1, nb = (yhi >> 1) - b0 + 1;
    auto Ev = [&](int xx, int yy) { return E[(yy - ey0) * CPITCH + (xx - ex0)]; };
    auto Fv = [&](int xx, int yy) { return (int)FC[(yy - ey0) * CPITCH + (xx - ex0)]; };
    // block (a0 + ja, b0 + jb) of index j = jb * na + ja, j = threadIdx.x +
    // TPB_M * k, stepped by divmod(TPB_M, na)
    const int nas = na > 0 ? na : 1;
    const int dq = TPB_M / nas, dr = TPB_M - dq * nas;
    int jb = na > 0 ? (int)threadIdx.x / nas : nb, ja = (int)threadIdx.x - jb * nas;
    while (jb < nb) {
      const int a = a0 + ja, b = b0 + jb;
      ja += dr;
      jb += dq;
      if (ja >= na) {
        ja -= na;
        jb++;
      }
      const float e00 = Ev(a, b);
      const int f00 = Fv(a, b);
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
          const int x = 2 * a + dx, y = 2 * b + dy;
          if (x < xlo || x > xhi || y < ylo || y > yhi) continue;
          const int i = (y - gy0) * PITCH + (x - gx0);
          const float Fs = (M[i] & MK_SELF) ? 1.0f : 0.0f;
          float e = 0.0f;
          if (dx == 0 && dy == 0) {
            e = e00 * Fs;
          } else if (dy == 0) {
            if (x <= w - 3) e = Fs * (e00 + Ev(a + 1, b)) / eps_flag_sum(f00 + Fv(a + 1, b));
          } else if (dx == 0) {
            if (y <= h - 3) e = Fs * (e00 + Ev(a, b + 1)) / eps_flag_sum(f00 + Fv(a, b + 1));
          } else {
            if (x <= w - 3 && y <= h - 3)
              e = Fs * (e00 + Ev(a + 1, b + 1) + Ev(a + 1, b) + Ev(a, b + 1)) /
                  eps_flag_sum(f00 + Fv(a + 1, b + 1) + Fv(a + 1, b) + Fv(a, b + 1));
          }
          P[i] += e;
        }
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NU; s++) {
    tile_phase<TPB_M / 64>(P, F, M, LW, LH, gx0, gy0, w, h, 1, 2 * s + 1, alpha);
    tile_phase<TPB_M / 64>(P, F, M, LW, LH, gx0, gy0, w, h, 0, 2 * s + 2, alpha);
  }
  const int ox1 = min(X0 + otx, w), oy1 = min(Y0 + oty, h);
  for (int y = Y0 + wave; y < oy1; y += TPB_M / 64)
    for (int x = X0 + lane; x < ox1; x += 64)
      p_out[(size_t)y * w + x] = P[(y - gy0) * PITCH + (x - gx0)];
}

} // namespace ubg
