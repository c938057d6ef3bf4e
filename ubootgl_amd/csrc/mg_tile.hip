// mg_tile.hip -- LDS-tiled level passes of the single-domain multigrid for
// the levels below the streaming size: one launch per pass, one workgroup
// per tile (tile bodies in tile_device.hpp).
#include "tile_device.hpp"

#include <cstdlib>

namespace ubg {
namespace {

// tile (tx, ty) of an otx x oty partition, blocks in XCD-grouped order
__device__ __forceinline__ void tile_origin(int ntx, int &X0, int &Y0, int otx, int oty) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  X0 = (b % ntx) * otx;
  Y0 = (b / ntx) * oty;
}

template <int NU, int TPB_M, bool ZP>
__global__ void __launch_bounds__(TPB_M)
    k_tile_pre(const float *__restrict__ p, float *__restrict__ p_out, const float *__restrict__ f,
               const uint8_t *__restrict__ m, int w, int h, float hh, float alpha, float ihsq,
               float *__restrict__ fc, float *__restrict__ ec, int wc, int hc, int otx, int oty,
               int ntx) {
  extern __shared__ __align__(16) char smem[];
  int X0, Y0;
  tile_origin(ntx, X0, Y0, otx, oty);
  tile_pre_body<NU, TPB_M, ZP>(smem, X0, Y0, otx, oty, p, p_out, f, m, w, h, hh, alpha, ihsq, fc,
                               ec, wc, hc);
}

template <int NU, int TPB_M>
__global__ void __launch_bounds__(TPB_M)
    k_tile_post(const float *__restrict__ p, float *__restrict__ p_out,
                const float *__restrict__ f, const uint8_t *__restrict__ m, int w, int h,
                float hh, float alpha, const float *__restrict__ ec,
                const uint8_t *__restrict__ mc, int wc, int hc, int otx, int oty, int ntx) {
  extern __shared__ __align__(16) char smem[];
  int X0, Y0;
  tile_origin(ntx, X0, Y0, otx, oty);
  tile_post_body<NU, TPB_M>(smem, X0, Y0, otx, oty, p, p_out, f, m, w, h, hh, alpha, ec, mc, wc,
                            hc);
}

struct TileGrid {
  int otx, oty, ntx, nty;
};

// Tile shape for a level: columns as wide as the LDS pitch allows, split
// evenly; rows: the fewest (even, 8 to kTileRowsMax) whose tile count still
// fits in one residency wave of kTileTarget workgroups, within 64 KiB of LDS.
// UBG_TILE_ROWS (an even number 8..kTileRowsMax whose LDS fits) overrides
// the rule: a test switch, so every row height the rule can pick is checked
// on one level.
TileGrid tile_grid(int w, int h, int nu) {
  TileGrid t;
  tile_cols(w, nu, t.otx, t.ntx);
  int oty = 8;
  while (oty < kTileRowsMax && (long)t.ntx * ((h + oty - 1) / oty) > kTileTarget &&
         tile_lds_ok(oty + 2, nu))
    oty += 2;
  if (const char *e = std::getenv("UBG_TILE_ROWS")) {
    const int r = std::atoi(e);
    if (r >= 8 && r <= kTileRowsMax && r % 2 == 0 && tile_lds_ok(r, nu)) oty = r;
  }
  t.oty = oty;
  t.nty = (h + oty - 1) / oty;
  return t;
}

template <int NU, int TPB>
void pre_launch(Level &L, Level &C, float h, float alpha, bool zero_p, bool zero_coarse,
                hipStream_t s) {
  const TileGrid t = tile_grid(L.w, L.h, NU);
  const size_t shm = tile_pre_lds(t.oty, NU);
  const float ihsq = 1.0f / h / h;
  hipLaunchKernelGGL((zero_p ? k_tile_pre<NU, TPB, true> : k_tile_pre<NU, TPB, false>),
                     dim3(t.ntx * t.nty), dim3(TPB), shm, s, L.p, L.p_alt, L.f, L.mask, L.w, L.h, h,
                     alpha, ihsq, C.f, zero_coarse ? C.p : nullptr, C.w, C.h, t.otx, t.oty,
                     t.ntx);
}

template <int NU, int TPB>
void post_launch(Level &L, const Level &C, float h, float alpha, hipStream_t s) {
  const TileGrid t = tile_grid(L.w, L.h, NU);
  const size_t shm = tile_post_lds(t.oty, NU);
  hipLaunchKernelGGL((k_tile_post<NU, TPB>), dim3(t.ntx * t.nty), dim3(TPB), shm, s, L.p,
                     L.p_alt, L.f, L.mask, L.w, L.h, h, alpha, C.p, C.mask, C.w, C.h, t.otx,
                     t.oty, t.ntx);
}

template <int NU>
void pre_nu(Level &L, Level &C, float h, float alpha, bool zero_p, bool zero_coarse,
            hipStream_t s) {
  pre_launch<NU, TILE_TPB>(L, C, h, alpha, zero_p, zero_coarse, s);
}

template <int NU>
void post_nu(Level &L, const Level &C, float h, float alpha, hipStream_t s) {
  post_launch<NU, TILE_TPB>(L, C, h, alpha, s);
}

void swap_p(Level &L) {
  float *t = L.p;
  L.p = L.p_alt;
  L.p_alt = t;
}

} // namespace

bool tile_pass_ok(int nu) { return nu >= 0 && nu <= NU_TILE; }

void launch_tile_pre(Level &L, Level &C, int nu, float h, float alpha, bool zero_p,
                     bool zero_coarse, hipStream_t s) {
  switch (nu) {
  case 0: pre_nu<0>(L, C, h, alpha, zero_p, zero_coarse, s); break;
  case 1: pre_nu<1>(L, C, h, alpha, zero_p, zero_coarse, s); break;
  case 2: pre_nu<2>(L, C, h, alpha, zero_p, zero_coarse, s); break;
  default: pre_nu<3>(L, C, h, alpha, zero_p, zero_coarse, s); break;
  }
  swap_p(L);
}

void launch_tile_post(Level &L, const Level &C, int nu, float h, float alpha, hipStream_t s) {
  switch (nu) {
  case 0: post_nu<0>(L, C, h, alpha, s); break;
  case 1: post_nu<1>(L, C, h, alpha, s); break;
  case 2: post_nu<2>(L, C, h, alpha, s); break;
  default: post_nu<3>(L, C, h, alpha, s); break;
  }
  swap_p(L);
}

} // namespace ubg
