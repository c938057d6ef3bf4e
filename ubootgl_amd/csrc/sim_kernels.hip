// sim_kernels.hip -- fluid-step kernels for gfx950 (simulation.cpp of the
// reference): boundary conditions, accumulation, masked Jacobi diffusion,
// semi-Lagrangian RK2 advection with Catmull-Rom sampling, divergence,
// velocity correction, bilinear velocity sampling and tracer advection.
// Memory-bound gathers and stencils; one thread per face / cell / point.
#include "mg_device.hpp"

#include <cstdlib>

namespace ubg {

namespace {

constexpr int TPB = 256;
inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// simulation.cpp:50-78
__device__ __forceinline__ float vbc_par(int bc, float a, float b) {
  if (bc == BC_INFLOW) return b;
  if (bc == BC_OUTFLOW || bc == BC_OUTFLOW_ZERO_PRESSURE) return fmaxf(a, 0.0f);
  return 0.0f;
}
__device__ __forceinline__ float vbc_per(int bc, float a, float b) {
  if (bc == BC_INFLOW) return b;
  if (bc == BC_OUTFLOW || bc == BC_OUTFLOW_ZERO_PRESSURE) return fmaxf(a, 0.0f);
  return -a;
}

// setVBCs (simulation.cpp:80-102) in one launch.  The reference runs four
// sequential loops: vx W/E columns (all rows), vx S/N rows (all columns), vy
// W/E columns, vy S/N rows; BC values are written to both buffers.  A row loop
// reads the column loop's result only at the corners' neighbours, and a
// column value at rows 0/H-1 is overwritten by the row loop.  So every
// boundary cell is a function of interior values and old boundary values,
// except the corners: each corner group -- the corner plus its two boundary
// neighbours -- is computed by one thread from old values in the sequential
// order; all other boundary cells by one thread each.  No thread reads a
// cell another thread writes.
struct BCGrid {
  float *f, *b;
  int w, h; // grid size
  int pw, pe, ps, pn; // BC kinds: west, east, south, north
  bool xpar;          // W/E use VBCPar (vx) or VBCPer (vy)
};
__device__ __forceinline__ float bc_we(const BCGrid &g, int kind, float a, float bb) {
  return g.xpar ? vbc_par(kind, a, bb) : vbc_per(kind, a, bb);
}
__device__ __forceinline__ float bc_sn(const BCGrid &g, int kind, float a, float bb) {
  return g.xpar ? vbc_per(kind, a, bb) : vbc_par(kind, a, bb);
}
// the front is written only when its bits change (an INFLOW side keeps its
// value), so a reader of an unchanged front cell never meets a write
__device__ __forceinline__ void put(const BCGrid &g, int x, int y, float v) {
  size_t i = (size_t)y * g.w + x;
  if (__float_as_int(g.f[i]) != __float_as_int(v)) g.f[i] = v;
  g.b[i] = v;
}
__device__ void bc_thread(const BCGrid &g, int t) {
  const int w = g.w, h = g.h;
  auto F = [&](int x, int y) { return g.f[(size_t)y * w + x]; };
  if (t >= 2 && t <= h - 3) { // columns, rows away from the corners
    float a = bc_we(g, g.pw, F(1, t), F(0, t));
    float b = bc_we(g, g.pe, F(w - 2, t), F(w - 1, t));
    put(g, 0, t, a);
    put(g, w - 1, t, b);
  }
  if (t >= 2 && t <= w - 3) { // rows, columns away from the corners
    float a = bc_sn(g, g.ps, F(t, 1), F(t, 0));
    float b = bc_sn(g, g.pn, F(t, h - 2), F(t, h - 1));
    put(g, t, 0, a);
    put(g, t, h - 1, b);
  }
  if (t < 4) { // corner groups: SW, SE, NW, NE
    const int cx = (t & 1) ? w - 1 : 0, ix = (t & 1) ? w - 2 : 1;  // boundary / interior column
    const int cy = (t & 2) ? h - 1 : 0, iy = (t & 2) ? h - 2 : 1;  // boundary / interior row
    const int kx = (t & 1) ? g.pe : g.pw, ky = (t & 2) ? g.pn : g.ps;
    float side = bc_we(g, kx, F(ix, iy), F(cx, iy));      // (cx, iy) after the column loop
    float cornA = bc_we(g, kx, F(ix, cy), F(cx, cy));     // (cx, cy) after the column loop
    float edge = bc_sn(g, ky, F(ix, iy), F(ix, cy));      // (ix, cy) after the row loop
    float corn = bc_sn(g, ky, side, cornA);               // (cx, cy) after the row loop
    put(g, cx, iy, side);
    put(g, ix, cy, edge);
    put(g, cx, cy, corn);
  }
}
__global__ void k_set_vbcs(BCGrid gx, BCGrid gy) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  bc_thread(gx, t);
  bc_thread(gy, t);
}
__global__ void k_set_vbcs_one(BCGrid g) { bc_thread(g, blockIdx.x * blockDim.x + threadIdx.x); }

// The same setVBCs writes, issued by the thread that produced interior cell
// (x, y) of grid g (g.f is the front buffer; v its new value there): every
// boundary cell is written by the thread owning the interior cell its value
// is computed from -- the side neighbour, or (x, y) = (ix, iy) for a corner
// group.  The boundary's old values are read from g.f before any write; no
// other thread reads or writes those cells.  cur (optional) receives the
// same values (saveCurrentVelocityFields after the setVBCs).
__device__ __forceinline__ void bc_duty(const BCGrid &g, int x, int y, float v, float *cur) {
  const int w = g.w, h = g.h;
  auto Fo = [&](int xx, int yy) { return g.f[(size_t)yy * w + xx]; };
  auto putc = [&](int xx, int yy, float val) {
    put(g, xx, yy, val);
    if (cur) cur[(size_t)yy * w + xx] = val;
  };
  if (y >= 2 && y <= h - 3) {
    if (x == 1) putc(0, y, bc_we(g, g.pw, v, Fo(0, y)));
    if (x == w - 2) putc(w - 1, y, bc_we(g, g.pe, v, Fo(w - 1, y)));
  }
  if (x >= 2 && x <= w - 3) {
    if (y == 1) putc(x, 0, bc_sn(g, g.ps, v, Fo(x, 0)));
    if (y == h - 2) putc(x, h - 1, bc_sn(g, g.pn, v, Fo(x, h - 1)));
  }
  if ((x == 1 || x == w - 2) && (y == 1 || y == h - 2)) {
    const bool e = x == w - 2, n = y == h - 2;
    const int cx = e ? w - 1 : 0, cy = n ? h - 1 : 0;
    const int kx = e ? g.pe : g.pw, ky = n ? g.pn : g.ps;
    const float side = bc_we(g, kx, v, Fo(cx, y));
    const float cornA = bc_we(g, kx, Fo(x, cy), Fo(cx, cy));
    const float edge = bc_sn(g, ky, v, Fo(x, cy));
    const float corn = bc_sn(g, ky, side, cornA);
    putc(cx, y, side);
    putc(x, cy, edge);
    putc(cx, cy, corn);
  }
}

// setPBC (simulation.cpp:36-45) in one launch: W/E columns over all rows,
// then S/N rows over all columns; only OUTFLOW_ZERO_PRESSURE negates.  Corner
// groups as in setVBCs.
__device__ __forceinline__ float pbc(int bc, float a) {
  return bc == BC_OUTFLOW_ZERO_PRESSURE ? -a : a;
}
__global__ void k_set_pbc(float *__restrict__ p, int w, int h, int bw, int be, int bn, int bs) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  auto P = [&](int x, int y) -> float & { return p[(size_t)y * w + x]; };
  if (t >= 2 && t <= h - 3) {
    float a = pbc(bw, P(1, t)), b = pbc(be, P(w - 2, t));
    P(0, t) = a;
    P(w - 1, t) = b;
  }
  if (t >= 2 && t <= w - 3) {
    float a = pbc(bs, P(t, 1)), b = pbc(bn, P(t, h - 2));
    P(t, 0) = a;
    P(t, h - 1) = b;
  }
  if (t < 4) {
    const int cx = (t & 1) ? w - 1 : 0, ix = (t & 1) ? w - 2 : 1;
    const int cy = (t & 2) ? h - 1 : 0, iy = (t & 2) ? h - 2 : 1;
    const int kx = (t & 1) ? be : bw, ky = (t & 2) ? bn : bs;
    float side = pbc(kx, P(ix, iy));
    float edge = pbc(ky, P(ix, iy));
    float corn = pbc(ky, side);
    P(cx, iy) = side;
    P(ix, cy) = edge;
    P(cx, cy) = corn;
  }
}

// applyAccumulatedVelocity (simulation.cpp:376-396), interior faces only.
__global__ void k_accumulate(float *__restrict__ xf, float *__restrict__ xa,
                             float *__restrict__ yf, float *__restrict__ ya, int W, int H) {
  int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int y = blockIdx.y + 1;
  const int Wx = W - 1, Hy = H - 1;
  if (x < Wx - 1 && y < H - 1) {
    size_t i = (size_t)y * Wx + x;
    xf[i] += xa[i];
    xa[i] = 0.0f;
  }
  if (x < W - 1 && y < Hy - 1) {
    size_t i = (size_t)y * W + x;
    yf[i] += ya[i];
    ya[i] = 0.0f;
  }
}


__device__ __forceinline__ float fbit(unsigned m, int k) { return ((m >> k) & 1u) ? 1.0f : 0.0f; }

__global__ void k_face_masks(const uint8_t *__restrict__ fl, uint16_t *__restrict__ fm, int W,
                             int H) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W) return;
  auto F = [&](int xx, int yy) -> unsigned {
    return (xx >= 0 && xx < W && yy >= 0 && yy < H) ? fl[(size_t)yy * W + xx] : 0u;
  };
  unsigned m = 0;
  if (x <= W - 2) { // vx face (x, y)
    m |= (F(x, y) & F(x + 1, y)) << 0;
    m |= (F(x + 1, y) & F(x + 2, y)) << 1;
    m |= (F(x, y) & F(x - 1, y)) << 2;
    m |= (F(x, y + 1) & F(x - 1, y + 1)) << 3;
    m |= (F(x, y - 1) & F(x - 1, y - 1)) << 4;
  }
  if (y <= H - 2) { // vy face (x, y)
    m |= (F(x, y) & F(x, y + 1)) << 8;
    m |= (F(x, y) & F(x, y - 1)) << 9;
    m |= (F(x, y + 1) & F(x, y + 2)) << 10;
    m |= (F(x + 1, y) & F(x + 1, y + 1)) << 11;
    m |= (F(x - 1, y) & F(x - 1, y + 1)) << 12;
  }
  fm[(size_t)y * W + x] = (uint16_t)m;
}

// One iteration of both diffuse loops (simulation.cpp:114-129 and 140-155)
// with the setVBCs that follows each of them: vx and vy do not read each
// other, and setVBCs on one component only reads that component.  Reads the
// front buffers sx / sy, writes the back buffers dx / dy (the fronts after
// the swap) and the boundary cells of both buffers.  The vy boundary must be
// current for vy's interior when the first iteration starts (the reference
// refreshes it in the setVBCs after the first vx iteration).  The flag
// products come from the face factors: s * F1 * F2 evaluated left to right
// equals s * (F1 AND F2) for 0/1 flags, signed zeros included.
__device__ __forceinline__ void diffuse_cell(int x, int y, const float *__restrict__ sx,
                                             float *__restrict__ dx, const float *__restrict__ sy,
                                             float *__restrict__ dy,
                                             const uint16_t *__restrict__ fm, int W, int H,
                                             float a, const BCGrid &gx, const BCGrid &gy) {
  const int Wx = W - 1, Hy = H - 1;
  if (x >= W - 1) return;
  const unsigned m = fm[(size_t)y * W + x];
  float vxn = 0.0f, vyn = 0.0f;
  if (x < Wx - 1) {
    const float *s = sx;
    size_t i = (size_t)y * Wx + x;
    float vc = s[i];
    float val = 0.0f;
    val += s[i + 1] * fbit(m, 1);
    val += s[i - 1] * fbit(m, 2);
    float fvn = fbit(m, 3);
    val += s[i + Wx] * fvn + (1.0f - fvn) * -vc;
    float fvs = fbit(m, 4);
    val += s[i - Wx] * fvs + (1.0f - fvs) * -vc;
    vxn = fbit(m, 0) * (vc + a * val) / (1.0f + 4.0f * a);
    dx[i] = vxn;
  }
  if (y < Hy - 1) {
    const float *s = sy;
    size_t i = (size_t)y * W + x;
    float vc = s[i];
    float val = 0.0f;
    val += s[i - W] * fbit(m, 9);
    val += s[i + W] * fbit(m, 10);
    float fve = fbit(m, 11);
    val += s[i + 1] * fve + (1.0f - fve) * -vc;
    float fvw = fbit(m, 12);
    val += s[i - 1] * fvw + (1.0f - fvw) * -vc;
    vyn = fbit(m, 8) * (vc + a * val) / (1.0f + 4.0f * a);
    dy[i] = vyn;
  }
  if (x == 1 || x >= W - 3 || y == 1 || y >= H - 3) { // boundary duties
    if (x < Wx - 1) bc_duty(gx, x, y, vxn, nullptr);
    if (y < Hy - 1) bc_duty(gy, x, y, vyn, nullptr);
  }
}
__global__ void k_diffuse_both(const float *__restrict__ sx, float *__restrict__ dx,
                               const float *__restrict__ sy, float *__restrict__ dy,
                               const uint16_t *__restrict__ fm, int W, int H, float a,
                               BCGrid gx, BCGrid gy) {
  diffuse_cell(blockIdx.x * blockDim.x + threadIdx.x + 1, blockIdx.y + 1, sx, dx, sy, dy, fm, W, H,
               a, gx, gy);
}

// CubicHermite (interpolators.hpp:78-85), Catmull-Rom, evaluated as
//   a = -A/2 + (3B)/2 - (3C)/2 + D/2,  b = A - (5B)/2 + 2C - D/2,
//   c = -A/2 + C/2,  a t t t + b t t + c t + B,
// each operation rounded as written, left to right.  A halving or a
// doubling is exact (a power-of-two scaling of a normal number), and fma(x,
// k, s) rounds x*k + s once, which for an exact product x*k is the rounding
// of the written sum: so each "(...)/2 +" and "2C +" term is one fma with the
// written result.  (Only a product that leaves the normal range -- a
// nonzero input below 2^-125 or above 2^126 in magnitude -- could round
// differently; velocities never come near either.)  The polynomial's
// products are inexact and stay separate.  20 instead of 26 operations.
__device__ __forceinline__ float cubic_hermite(float t, float A, float B, float C, float D) {
  const float mA = A * -0.5f;                       // -A / 2
  float a = __builtin_fmaf(3.0f * B, 0.5f, mA);     // + (3B) / 2
  a = __builtin_fmaf(3.0f * C, -0.5f, a);           // - (3C) / 2
  a = __builtin_fmaf(D, 0.5f, a);                   // + D / 2
  float b = __builtin_fmaf(5.0f * B, -0.5f, A);     // A - (5B) / 2
  b = __builtin_fmaf(C, 2.0f, b);                   // + 2C
  b = __builtin_fmaf(D, -0.5f, b);                  // - D / 2
  const float c = __builtin_fmaf(C, 0.5f, mA);      // -A / 2 + C / 2
  return a * t * t * t + b * t * t + c * t + B;
}

// bicubicSample for one lane (interpolators.hpp:92-206): clamp to [3, n-3],
// truncate, Catmull-Rom down each of four columns, then across.  BUF (grids
// under 2 GiB): the four 4-float rows are buffer loads, a 32-bit byte offset
// for the first row and the row pitch as the (scalar) offset of the others,
// instead of 64-bit address arithmetic per row.  The clamp keeps the 4 x 4
// stencil (columns icx-1 .. icx+2, rows icy-1 .. icy+2) inside the grid for
// gw, gh >= 6 (icx <= gw-3); a simulation is at least 8 x 8 cells, so its
// velocity grids are at least 7 x 8.  The descriptor's third word is the gfx9
// raw-buffer format.
struct Taps {
  float r[4][4];
  float stx, sty;
};
template <bool BUF>
__device__ __forceinline__ Taps bicubic_taps(const float *__restrict__ g, int gw, int gh,
                                             float cx, float cy) {
  Taps t;
  cx = fmaxf(fminf(cx, gw - 3.0f), 3.0f);
  cy = fmaxf(fminf(cy, gh - 3.0f), 3.0f);
  int icx = (int)cx, icy = (int)cy;
  t.stx = cx - truncf(cx);
  t.sty = cy - truncf(cy);
  float (&r)[4][4] = t.r;
  if (BUF) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)g, (short)0, gw * gh * 4, 0x00020000);
    const int o = ((icy - 1) * gw + (icx - 1)) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, o, k * gw * 4, 0);
#pragma unroll
      for (int j = 0; j < 4; j++) r[k][j] = __uint_as_float(v[j]);
    }
  } else {
    const float *b0 = g + (ptrdiff_t)(icy - 1) * gw + (icx - 1);
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int j = 0; j < 4; j++) r[k][j] = b0[(ptrdiff_t)k * gw + j];
  }
  return t;
}
__device__ __forceinline__ float bicubic_eval(const Taps &t) {
  const float(&r)[4][4] = t.r;
  float c0 = cubic_hermite(t.sty, r[0][0], r[1][0], r[2][0], r[3][0]);
  float c1 = cubic_hermite(t.sty, r[0][1], r[1][1], r[2][1], r[3][1]);
  float c2 = cubic_hermite(t.sty, r[0][2], r[1][2], r[2][2], r[3][2]);
  float c3 = cubic_hermite(t.sty, r[0][3], r[1][3], r[2][3], r[3][3]);
  return cubic_hermite(t.stx, c0, c1, c2, c3);
}
template <bool BUF>
__device__ __forceinline__ float bicubic(const float *__restrict__ g, int gw, int gh, float cx,
                                         float cy) {
  return bicubic_eval(bicubic_taps<BUF>(g, gw, gh, cx, cy));
}

// One launch for both sweeps (advect, simulation.cpp:241-354): a thread
// advects the vx face and the vy face of its cell.  The two faces are
// computed side by side in straight-line code and only their stores are
// predicated, so each of the three dependent gather rounds (the face's own
// velocity and its neighbours, the two midpoint samples, the end sample)
// issues the loads of both faces together: a wave waits three times per
// cell instead of six.  Lanes outside a face's range, and skipped chunks,
// compute on in-bounds operands and store nothing.  Measured on gfx950 at
// 4096x2048: 124.5 us against 128 us with the faces one after the other;
// the same code held to 71 or 100 VGPRs (7 or 5 waves per SIMD instead of
// the compiler's 66) takes 136 or 140 us.
//
// The chunk of lanes x0 .. x0+7 (x0 = 1 + 8k) runs if any of its lanes has
// both flags of the skip test set.  A chunk is 8 consecutive threads of one
// wave, 8-aligned in it, and either all or none of them lie in a face's
// range (nck is a count of chunks), so one ballot of the lanes' own tests
// answers it for every chunk of the wave.
template <bool BUF>
__device__ __forceinline__ void advect_pair(int lane, int y, const float *__restrict__ xf,
                                            const float *__restrict__ yf,
                                            float *__restrict__ xb, float *__restrict__ yb,
                                            const uint8_t *__restrict__ fl, int W, int H,
                                            int nck_x, int nck_y, float hdt, float fdt) {
  // lane: 0-based (x = lane + 1), lane & 63 == threadIdx.x & 63
  const int Wx = W - 1, Hy = H - 1;
  const bool in_x = lane < 8 * nck_x, in_y = lane < 8 * nck_y;
  // every lane of either range has x <= W - 2, so x - 1 .. x + 1 and rows
  // y - 1 .. y + 1 are inside the flag and velocity grids; other lanes load
  // at x = 1
  const int x = (in_x || in_y) ? lane + 1 : 1;
  const uint8_t *frow = fl + (size_t)y * W;
  const unsigned f0 = frow[x], fw = frow[x - 1], fe = frow[x + 1];
  const unsigned fs = frow[x - W], fn = frow[x + W];
  const unsigned long long vx_votes = __ballot(in_x && (fw & f0) != 0);
  const unsigned long long vy_votes = __ballot(in_y && (f0 & fs) != 0);
  const int sh = threadIdx.x & 56;
  const bool run_x = in_x && ((vx_votes >> sh) & 0xffull) != 0;
  const bool run_y = in_y && ((vy_votes >> sh) & 0xffull) != 0;
  if (!run_x && !run_y) return;
  const size_t ry = (size_t)y, rs = (size_t)(y - 1);
  // the faces' own velocities and the four-face averages of the other
  // component (simulation.cpp:265-273, 315-323)
  const float xc = xf[x + ry * Wx], xcs = xf[x + rs * Wx];
  const float xe = xf[x + 1 + ry * Wx], xes = xf[x + 1 + rs * Wx];
  const float yc = yf[x + ry * W], ycs = yf[x + rs * W];
  const float ye = yf[x + 1 + ry * W], yes = yf[x + 1 + rs * W];
  // vx face (x, y): position (x + 0.5, y)
  const float ax = (float)x + 0.5f, ay = (float)y;
  const float a_vy1 = ((yc + ycs) + (ye + yes)) * 0.25f;
  const float a_midx = ax - xc * hdt, a_midy = ay - a_vy1 * hdt;
  // vy face (x, y): position (x, y + 0.5)
  const float bx = (float)x, by = (float)y + 0.5f;
  const float b_vx1 = ((xc + xcs) + (xe + xes)) * 0.25f;
  const float b_midx = bx - b_vx1 * hdt, b_midy = by - yc * hdt;
  // the four midpoint samples' gathers are issued before any of them is
  // evaluated, then the two end samples'
  const Taps t1 = bicubic_taps<BUF>(xf, Wx, H, a_midx - 0.5f, a_midy);
  const Taps t2 = bicubic_taps<BUF>(yf, W, Hy, a_midx, a_midy - 0.5f);
  const Taps t3 = bicubic_taps<BUF>(xf, Wx, H, b_midx - 0.5f, b_midy);
  const Taps t4 = bicubic_taps<BUF>(yf, W, Hy, b_midx, b_midy - 0.5f);
  const float a_vx2 = bicubic_eval(t1), a_vy2 = bicubic_eval(t2);
  const float b_vx2 = bicubic_eval(t3), b_vy2 = bicubic_eval(t4);
  const float a_endx = ax - a_vx2 * fdt, a_endy = ay - a_vy2 * fdt;
  const float b_endx = bx - b_vx2 * fdt, b_endy = by - b_vy2 * fdt;
  const Taps e1 = bicubic_taps<BUF>(xf, Wx, H, a_endx - 0.5f, a_endy);
  const Taps e2 = bicubic_taps<BUF>(yf, W, Hy, b_endx, b_endy - 0.5f);
  float va = bicubic_eval(e1) * (float)f0 * (float)fe;
  float vb = bicubic_eval(e2) * (float)f0 * (float)fn;
  // both values exist here, so neither face's work is sunk into its store's
  // branch (which would serialize the faces again); no instruction is emitted
  asm volatile("" : "+v"(va), "+v"(vb));
  if (run_x) xb[(size_t)Wx * y + x] = va;
  if (run_y) yb[(size_t)W * y + x] = vb;
}
template <bool BUF>
__global__ void k_advect(const float *__restrict__ xf, const float *__restrict__ yf,
                         float *__restrict__ xb, float *__restrict__ yb,
                         const uint8_t *__restrict__ fl, int W, int H, int nck_x, int nck_y,
                         float hdt, float fdt) {
  advect_pair<BUF>(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y + 1, xf, yf, xb, yb, fl, W, H,
                   nck_x, nck_y, hdt, fdt);
}

// Collocated velocity for rendering (interp_shader.cs:14-33, dispatched by
// VelocityTextures::updateFromStaggered, velocity_textures.cpp:63-93): pixel
// (gx, gy) of the (2W-1) x (2H-1) image samples vx at normalized
// (gx / (2W-2), (gy+1) / 2H) and vy at ((gx+1) / 2W, gy / (2H-2)) with
// texture(): linear filtering between texel centres and GL_REPEAT wrapping
// (the default sampler state).  In texel space these are the points
// ((gx-1)/2, gy/2) of vx and (gx/2, (gy-1)/2) of vy, so every weight is 0 or
// 1/2; the filter is evaluated as lerp(lerp(t00, t10, wx), lerp(t01, t11,
// wx), wy) written with (1-w) and w.  mag = length(vx, vy).
__device__ __forceinline__ float texel_linear(const float *__restrict__ g, int gw, int gh,
                                              int twx, int twy) {
  // twx, twy: texel-space coordinate times 2 (exact), wrapped with REPEAT
  const int i0 = twx >> 1, j0 = twy >> 1; // floor for negatives too
  const float wx = (twx & 1) ? 0.5f : 0.0f, wy = (twy & 1) ? 0.5f : 0.0f;
  auto wrap = [](int i, int n) { return ((i % n) + n) % n; };
  const int xa = wrap(i0, gw), xb = wrap(i0 + 1, gw);
  const int ya = wrap(j0, gh), yb = wrap(j0 + 1, gh);
  const float t00 = g[(size_t)ya * gw + xa], t10 = g[(size_t)ya * gw + xb];
  const float t01 = g[(size_t)yb * gw + xa], t11 = g[(size_t)yb * gw + xb];
  const float r0 = (1.0f - wx) * t00 + wx * t10;
  const float r1 = (1.0f - wx) * t01 + wx * t11;
  return (1.0f - wy) * r0 + wy * r1;
}

__global__ void k_collocate(const float *__restrict__ vx, const float *__restrict__ vy,
                            float *__restrict__ vxy, float *__restrict__ mag, int W, int H) {
  const int gx = blockIdx.x * blockDim.x + threadIdx.x;
  const int gy = blockIdx.y;
  const int ow = 2 * W - 1;
  if (gx >= ow) return;
  const float a = texel_linear(vx, W - 1, H, gx - 1, gy);
  const float b = texel_linear(vy, W, H - 1, gx, gy - 1);
  const size_t o = (size_t)gy * ow + gx;
  vxy[2 * o] = a;
  vxy[2 * o + 1] = b;
  mag[o] = sqrtf(a * a + b * b);
}

// UbootGlApp::shiftMap's simulation part (ubootgl_app.cpp:248-296).  One
// workgroup per row stages the row in LDS, so the sequential left-to-right
// copies (each reading a value not yet overwritten) become: new(x-1) =
// old(x).  field 0: vx (W-1 wide, both buffers from the front one, x-1 in
// [1, W-3]); 1: vy (W wide, x-1 in [1, W-2]); 2: p (x-1 in [0, W-2]).
__global__ void k_shift_rows(float *__restrict__ vxf, float *__restrict__ vxb,
                             float *__restrict__ vyf, float *__restrict__ vyb,
                             float *__restrict__ p, int W, int H) {
  extern __shared__ float row[];
  const int field = blockIdx.y, y = blockIdx.x;
  const int w = field == 0 ? W - 1 : W;
  const int rows = field == 1 ? H - 1 : H;
  if (y >= rows) return;
  float *f = field == 0 ? vxf : (field == 1 ? vyf : p);
  float *b = field == 0 ? vxb : (field == 1 ? vyb : nullptr);
  float *fr = f + (size_t)y * w;
  for (int x = threadIdx.x; x < w; x += blockDim.x) row[x] = fr[x];
  __syncthreads();
  const int lo = field == 2 ? 0 : 1, hi = w - 2; // targets x-1
  for (int t = lo + threadIdx.x; t <= hi; t += blockDim.x) {
    fr[t] = row[t + 1];
    if (b) b[(size_t)y * w + t] = row[t + 1];
  }
}

// setGrids(c, flag(c)) for every cell (simulation.hpp:82-98): the flag, and
// for a solid cell zero velocities on its four faces and its pressure (front
// buffers).  Only zeros are written, so the order of the cells is immaterial.
__global__ void k_apply_flags(const uint8_t *__restrict__ nf, uint8_t *__restrict__ fl,
                              float *__restrict__ vxf, float *__restrict__ vyf,
                              float *__restrict__ p, int W, int H) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  const size_t c = (size_t)y * W + x;
  const uint8_t v = nf[c];
  fl[c] = v;
  if (v) return;
  if (x < W - 1) vxf[(size_t)y * (W - 1) + x] = 0.0f;
  if (x > 0) vxf[(size_t)y * (W - 1) + x - 1] = 0.0f;
  if (y < H - 1) vyf[c] = 0.0f;
  if (y > 0) vyf[c - W] = 0.0f;
  p[c] = 0.0f;
}

// inlet velocity after a map shift (ubootgl_app.cpp:283-291): inletArea =
// 1 + sum of flag(0, y) over the H-1 rows of vy (integers, exact in float);
// vx(0, y) = vx.b(0, y) = 0.07 H / inletArea * flag(0, y).
__global__ void k_inlet(const uint8_t *__restrict__ fl, float *__restrict__ vxf,
                        float *__restrict__ vxb, int W, int H) {
  __shared__ int red[4];
  int c = 0;
  for (int y = threadIdx.x; y < H - 1; y += 256) c += fl[(size_t)y * W];
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  const float area = 1.0f + (float)(red[0] + red[1] + red[2] + red[3]);
  const float vin = 0.07f * (float)H / area;
  for (int y = threadIdx.x; y < H; y += 256) {
    const float v = vin * (float)fl[(size_t)y * W];
    vxf[(size_t)y * (W - 1)] = v;
    vxb[(size_t)y * (W - 1)] = v;
  }
}

// Terrain's flags (terrain.cpp:34-45): full resolution = the image's red
// channel / 255 with rows flipped; the simulation flag (Terrain::subSample,
// terrain.cpp:3-11) = the mean of a scale x scale block, summed in row-major
// order, above 0.99.
__global__ void k_terrain_full(const uint8_t *__restrict__ rgba, int iw, int ih,
                               float *__restrict__ full) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= iw) return;
  full[(size_t)y * iw + x] = (float)rgba[4 * ((size_t)(ih - y - 1) * iw + x)] / 255.0f;
}

__global__ void k_terrain_subsample(const float *__restrict__ full, int iw, int scale,
                                    float *__restrict__ sim, int sw) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= sw) return;
  float sum = 0.0f;
  for (int sy = 0; sy < scale; sy++)
    for (int sx = 0; sx < scale; sx++) sum += full[(size_t)(y * scale + sy) * iw + x * scale + sx];
  sim[(size_t)y * sw + x] = (sum / (float)scale / (float)scale) > 0.99f ? 1.0f : 0.0f;
}

// project's divergence (simulation.cpp:166-171), interior, not flag-masked.
__global__ void k_divergence(const float *__restrict__ xf, const float *__restrict__ yf,
                             float *__restrict__ f, int W, int H, float ih) {
  int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int y = blockIdx.y + 1;
  if (x >= W - 1) return;
  const int Wx = W - 1;
  size_t ix = (size_t)y * Wx + x;
  size_t iy = (size_t)y * W + x;
  f[iy] = -ih * (xf[ix] - xf[ix - 1] + yf[iy] - yf[iy - W]);
}

// sink stamps (simulation.cpp:173-182): the cells of the 3x3 patches and
// their final values, settled on the host in sink order (one write per cell)
__global__ void k_stamp_sinks(float *__restrict__ f, const int *__restrict__ cell,
                              const float *__restrict__ val, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  f[cell[t]] = val[t];
}

// project's velocity correction (simulation.cpp:196-207)
__global__ void k_correct_velocity(float *__restrict__ xf, float *__restrict__ yf,
                                   const float *__restrict__ p, const uint8_t *__restrict__ fl,
                                   int W, int H, float ih) {
  int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int y = blockIdx.y + 1;
  if (x >= W - 1) return;
  const int Wx = W - 1;
  size_t c = (size_t)y * W + x;
  float Fc = (float)fl[c];
  float pc = p[c];
  if (x < W - 2)
    xf[(size_t)y * Wx + x] -= Fc * (float)fl[c + 1] * ih * (p[c + 1] - pc);
  if (y < H - 2)
    yf[(size_t)y * W + x] -= Fc * (float)fl[c + W] * ih * (p[c + W] - pc);
}

// setPBC's writes (simulation.cpp:36-45, corner groups as k_set_pbc) issued
// by the thread of interior cell (x, y); p's interior is only read.
__device__ __forceinline__ void pbc_duty(float *__restrict__ p, int w, int h, int x, int y,
                                         int bw, int be, int bn, int bs) {
  auto P = [&](int xx, int yy) -> float & { return p[(size_t)yy * w + xx]; };
  if (y >= 2 && y <= h - 3) {
    if (x == 1) P(0, y) = pbc(bw, P(1, y));
    if (x == w - 2) P(w - 1, y) = pbc(be, P(w - 2, y));
  }
  if (x >= 2 && x <= w - 3) {
    if (y == 1) P(x, 0) = pbc(bs, P(x, 1));
    if (y == h - 2) P(x, h - 1) = pbc(bn, P(x, h - 2));
  }
  if ((x == 1 || x == w - 2) && (y == 1 || y == h - 2)) {
    const bool e = x == w - 2, n = y == h - 2;
    const int cx = e ? w - 1 : 0, cy = n ? h - 1 : 0;
    const int kx = e ? be : bw, ky = n ? bn : bs;
    const float side = pbc(kx, P(x, y));
    const float edge = pbc(ky, P(x, y));
    const float corn = pbc(ky, side);
    P(cx, y) = side;
    P(x, cy) = edge;
    P(cx, cy) = corn;
  }
}

// The end of project() and step(): setPBC, the velocity correction
// (simulation.cpp:196-207), setVBCs and saveCurrentVelocityFields in one
// pass.  The correction reads p at interior cells only, so setPBC's writes
// do not reach it; every vx / vy face is either corrected here or a boundary
// face, so the copy into vx_cur / vy_cur is complete (cur may be null).  A
// thread issues all its loads before its first store: 12% less time on
// gfx950 than loading each face's operands beside its own store.
__device__ __forceinline__ void correct_tail_cell(int x, int y, float *__restrict__ xf,
                                                  float *__restrict__ yf, float *__restrict__ p,
                                                  const uint16_t *__restrict__ fm, int W, int H,
                                                  float ih, const BCGrid &gx, const BCGrid &gy,
                                                  float *__restrict__ xcur,
                                                  float *__restrict__ ycur, int bw, int be,
                                                  int bn, int bs) {
  if (x >= W - 1) return;
  const int Wx = W - 1;
  const bool hx = x < W - 2, hy = y < H - 2; // the thread owns a vx / vy face
  const size_t c = (size_t)y * W + x;
  const unsigned m = fm[c]; // F(x)F(x+1) (bit 0) and F(y)F(y+1) (bit 8)
  const float pc = p[c];
  float pe = 0.0f, pn = 0.0f, ux = 0.0f, uy = 0.0f;
  if (hx) {
    pe = p[c + 1];
    ux = xf[(size_t)y * Wx + x];
  }
  if (hy) {
    pn = p[c + W];
    uy = yf[c];
  }
  float vx = 0.0f, vy = 0.0f;
  if (hx) {
    const size_t i = (size_t)y * Wx + x;
    vx = ux - fbit(m, 0) * ih * (pe - pc);
    xf[i] = vx;
    if (xcur) xcur[i] = vx;
  }
  if (hy) {
    vy = uy - fbit(m, 8) * ih * (pn - pc);
    yf[c] = vy;
    if (ycur) ycur[c] = vy;
  }
  if (x == 1 || x >= W - 3 || y == 1 || y >= H - 3) { // boundary duties
    if (hx) bc_duty(gx, x, y, vx, xcur);
    if (hy) bc_duty(gy, x, y, vy, ycur);
    pbc_duty(p, W, H, x, y, bw, be, bn, bs);
  }
}
__global__ void k_correct_tail(float *__restrict__ xf, float *__restrict__ yf,
                               float *__restrict__ p, const uint16_t *__restrict__ fm, int W,
                               int H, float ih, BCGrid gx, BCGrid gy, float *__restrict__ xcur,
                               float *__restrict__ ycur, int bw, int be, int bn, int bs) {
  correct_tail_cell(blockIdx.x * blockDim.x + threadIdx.x + 1, blockIdx.y + 1, xf, yf, p, fm, W,
                    H, ih, gx, gy, xcur, ycur, bw, be, bn, bs);
}

// setVBCs after advect (simulation.cpp:369) fused with project's divergence
// (:166-171).  Thread (x, y) owns the vx face (x, y) and the vy face (x, y):
// it issues their boundary writes and computes f at cell (x, y) from the
// face values after setVBCs -- a boundary face it reads is the side / edge
// value of the rule applied to its interior neighbour (never a corner).
__device__ __forceinline__ void vbcs_divergence_cell(int x, int y,
                                                     const float *__restrict__ xf,
                                                     const float *__restrict__ yf,
                                                     float *__restrict__ f, int W, int H,
                                                     float ih, const BCGrid &gx,
                                                     const BCGrid &gy) {
  if (x >= W - 1) return;
  const int Wx = W - 1, Hy = H - 1;
  // vx faces x - 1 and x of row y, vy faces y - 1 and y of column x (all
  // exist for x <= W - 2, y <= H - 2), then their values after setVBCs: a
  // boundary face among them gets the rule applied to the other face, its
  // interior neighbour.  The west / south boundary faces are written by this
  // thread (after its reads); the east / north ones by the thread of their
  // interior neighbour, so they are read only for INFLOW, whose rule keeps
  // the old value and whose front write is skipped (put).
  const bool east = x == Wx - 1, north = y == Hy - 1;
  const float xw = xf[(size_t)y * Wx + x - 1];
  const float xe = (!east || gx.pe == BC_INFLOW) ? xf[(size_t)y * Wx + x] : 0.0f;
  const float ys = yf[(size_t)(y - 1) * W + x];
  const float yn = (!north || gy.pn == BC_INFLOW) ? yf[(size_t)y * W + x] : 0.0f;
  const float vx_w = (x == 1) ? bc_we(gx, gx.pw, xe, xw) : xw;
  const float vx_e = east ? bc_we(gx, gx.pe, xw, xe) : xe;
  const float vy_s = (y == 1) ? bc_sn(gy, gy.ps, yn, ys) : ys;
  const float vy_n = north ? bc_sn(gy, gy.pn, ys, yn) : yn;
  // same association as the reference's expression: ((e - w) + n) - s
  f[(size_t)y * W + x] = -ih * (vx_e - vx_w + vy_n - vy_s);
  if (x == 1 || x >= W - 3 || y == 1 || y >= H - 3) { // boundary duties
    if (x <= Wx - 2) bc_duty(gx, x, y, xe, nullptr);
    if (y <= Hy - 2) bc_duty(gy, x, y, yn, nullptr);
  }
}
__global__ void k_vbcs_divergence(const float *__restrict__ xf, const float *__restrict__ yf,
                                  float *__restrict__ f, int W, int H, float ih, BCGrid gx,
                                  BCGrid gy) {
  vbcs_divergence_cell(blockIdx.x * blockDim.x + threadIdx.x + 1, blockIdx.y + 1, xf, yf, f, W, H,
                       ih, gx, gy);
}

// scalar bilinearSample (interpolators.hpp:11-27), clamp to [0, n-1.1]
__device__ __forceinline__ float bilinear(const float *__restrict__ g, int gw, int gh, float cx,
                                          float cy) {
  cx = fminf(fmaxf(cx, 0.0f), gw - 1.1f);
  cy = fminf(fmaxf(cy, 0.0f), gh - 1.1f);
  int ix = (int)cx, iy = (int)cy;
  float sx = cx - floorf(cx), sy = cy - floorf(cy);
  const float *r0 = g + (size_t)iy * gw + ix;
  float v1 = r0[0], v2 = r0[1], v3 = r0[gw], v4 = r0[gw + 1];
  float vm1 = v1 + (v2 - v1) * sx;
  float vm2 = v3 + (v4 - v3) * sx;
  return vm1 + (vm2 - vm1) * sy;
}

// Simulation::bilinearVel (advect_floating_items.cpp:11-14)
__device__ __forceinline__ float2 bilinear_vel(const float *__restrict__ xf,
                                               const float *__restrict__ yf, int W, int H,
                                               float cx, float cy) {
  return make_float2(bilinear(xf, W - 1, H, cx - 0.5f, cy), bilinear(yf, W, H - 1, cx, cy - 0.5f));
}

__global__ void k_bilinear_vel(const float *__restrict__ xf, const float *__restrict__ yf, int W,
                               int H, const float *__restrict__ c, float *__restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float2 v = bilinear_vel(xf, yf, W, H, c[2 * i], c[2 * i + 1]);
  out[2 * i] = v.x;
  out[2 * i + 1] = v.y;
}

// advect_tracer_points.cs:21-39
__device__ __forceinline__ uint32_t wang_hash(uint32_t seed) {
  seed = (seed ^ 61u) ^ (seed >> 16);
  seed *= 9u;
  seed = seed ^ (seed >> 4);
  seed *= 0x27d4eb2du;
  seed = seed ^ (seed >> 15);
  return seed;
}
__device__ __forceinline__ float randf(uint32_t &s) {
  s = 1664525u * s + 1013904223u;
  return (float)(s % 100000u) / 100000.0f;
}

// psampleFlagLinear (simulation.cpp:398-411); mix(x,y,a) = x*(1-a) + y*a.
// The float-to-int conversion of glm::ivec2(vec2) is x86's truncating
// conversion, which gives INT_MIN for NaN and out-of-range values; the
// clamp then picks column / row 0 for them, as it does on the reference.
__device__ __forceinline__ int trunc_i32_x86(float v) {
  return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000u;
}
__device__ __forceinline__ float flag_linear(const uint8_t *__restrict__ F, int W, int H,
                                             float pwidth, float px, float py) {
  float cx = px / (pwidth / W) - 0.5f, cy = py / (pwidth / W) - 0.5f;
  int ix = trunc_i32_x86(cx), iy = trunc_i32_x86(cy);
  ix = max(0, min(W - 2, ix));
  iy = max(0, min(H - 2, iy));
  float sx = cx - floorf(cx), sy = cy - floorf(cy);
  const uint8_t *r0 = F + (size_t)iy * W + ix;
  float p00 = r0[0], p10 = r0[1], p01 = r0[W], p11 = r0[W + 1];
  float m0 = p00 * (1.0f - sx) + p10 * sx;
  float m1 = p01 * (1.0f - sx) + p11 * sx;
  return m0 * (1.0f - sy) + m1 * sy;
}

// One tracer step (advect_tracer_points.cs:42-82 with bilinearVel sampling):
// RK2 midpoint; a point outside [0,pdim] or in a solid (linear flag < 0.6)
// freezes and ages 0.1; past age 2*3.141 it respawns from the Wang hash and
// the LCG, otherwise it ages 0.02.
__global__ void k_tracers(const float *__restrict__ xf, const float *__restrict__ yf,
                          const uint8_t *__restrict__ F, int W, int H, float h, float pwidth,
                          float dt, uint32_t seed, float2 *__restrict__ xy,
                          float *__restrict__ age, const int *__restrict__ perm, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x; // storage slot
  if (j >= n) return;
  const int i = perm ? perm[j] : j; // tracer id (seeds the respawn hash)
  const float pdx = pwidth, pdy = pwidth * H / W;
  float2 P = xy[j];
  float2 v1 = bilinear_vel(xf, yf, W, H, P.x / h, P.y / h);
  float mx = P.x + v1.x * dt * 0.5f, my = P.y + v1.y * dt * 0.5f;
  float2 v2 = bilinear_vel(xf, yf, W, H, mx / h, my / h);
  float nx = P.x + v2.x * dt, ny = P.y + v2.y * dt;
  float a = age[j];
  if (P.x < 0.0f || P.y < 0.0f || P.x > pdx || P.y > pdy ||
      flag_linear(F, W, H, pwidth, P.x, P.y) < 0.6f) {
    nx = P.x;
    ny = P.y;
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
  xy[j] = make_float2(nx, ny);
  age[j] = a;
}

// psampleFlagNearest / psampleFlagLinear / psampleFlagNormal
// (simulation.cpp:398-425) at n physical positions pc = (x0, y0, x1, ...).
// Nearest reads the flag at the truncated grid position pc / (pwidth / W);
// a position whose cell lies outside the grid (the reference indexes out of
// bounds there) reads as solid, 0.  Normal writes two values per point from
// the four linear samples at pc -+ (h, h), summed left to right.
__global__ void k_sample_flag(const uint8_t *__restrict__ F, int W, int H, float pwidth, float h,
                              int mode, const float *__restrict__ pc, float *__restrict__ out,
                              int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pc[2 * i], py = pc[2 * i + 1];
  if (mode == 0) {
    const float cell = pwidth / W;
    const float gx = px / cell, gy = py / cell;
    const bool in = gx > -1.0f && gx < (float)W && gy > -1.0f && gy < (float)H;
    out[i] = in ? (float)F[(size_t)(int)gy * W + (int)gx] : 0.0f;
  } else if (mode == 1) {
    out[i] = flag_linear(F, W, H, pwidth, px, py);
  } else {
    const float p01 = flag_linear(F, W, H, pwidth, px - h, py + h);
    const float p11 = flag_linear(F, W, H, pwidth, px + h, py + h);
    const float p00 = flag_linear(F, W, H, pwidth, px - h, py - h);
    const float p10 = flag_linear(F, W, H, pwidth, px + h, py - h);
    out[2 * i] = p11 + p10 - p01 - p00;
    out[2 * i + 1] = p01 + p11 - p00 - p10;
  }
}

// getDT's scan (simulation.cpp:226-238): max over interior cells x in
// [1, W-2], y in [1, H-2] of max(vx(x, y), vy(x, y)), one partial per
// block (a maximum does not depend on the order it is taken in).
__global__ void __launch_bounds__(256) k_max_vel(const float *__restrict__ xf,
                                                 const float *__restrict__ yf, int W, int H,
                                                 float *__restrict__ part) {
  __shared__ float red[4];
  const int iw = W - 2;
  const size_t n = (size_t)iw * (H - 2);
  float m = -INFINITY;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const int x = 1 + (int)(k % iw), y = 1 + (int)(k / iw);
    m = fmaxf(m, fmaxf(xf[(size_t)y * (W - 1) + x], yf[(size_t)y * W + x]));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// centerP's sum (simulation.cpp:210-217) over interior cells, in float64
// per block (a fixed order, so the result does not vary between calls)
__global__ void __launch_bounds__(256) k_sum_interior(const float *__restrict__ p, int W, int H,
                                                      double *__restrict__ part) {
  __shared__ double red[4];
  const int iw = W - 2;
  const size_t n = (size_t)iw * (H - 2);
  double s = 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    s += (double)p[(size_t)(1 + k / iw) * W + 1 + k % iw];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}
// centerP's update (simulation.cpp:218-223): p -= v on interior cells
__global__ void k_sub_interior(float *__restrict__ p, int W, int H, float v) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x + 1, y = blockIdx.y + 1;
  if (x >= W - 1) return;
  p[(size_t)y * W + x] -= v;
}

// gs (pressure_solver.cpp:26-33): a lexicographic Gauss-Seidel sweep.  The
// cell (x, y) reads the new values of (x-1, y) and (x, y-1) and the old ones
// of (x+1, y) and (x, y+1), so the cells of one anti-diagonal x + y = d
// depend only on diagonal d-1 (already updated) and d+1 (not yet): one launch
// per diagonal, in increasing d, gives the sequential sweep's values.
__global__ void k_gs_diagonal(float *__restrict__ p, const float *__restrict__ f,
                              const uint8_t *__restrict__ mask, int W, int H, float h, float alpha,
                              int d) {
  const int y0 = max(1, d - (W - 2));
  const int y = y0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (y > min(H - 2, d - 1)) return;
  const int x = d - y;
  const size_t c = (size_t)y * W + x;
  p[c] = smooth(p[c], p[c - 1], p[c + 1], p[c - W], p[c + W], f[c], mask[c], h, alpha);
}

// ---- batched steps of several simulations ---------------------------------
__device__ __forceinline__ const SimRec &sim_of_row(const SimRec *tab, int n, int row) {
  int lo = 0, hi = n; // the last simulation whose row0 <= row
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tab[mid].row0 <= row) lo = mid;
    else hi = mid;
  }
  return tab[lo];
}

// blockIdx.x: interior row in the step's numbering; blockIdx.y: column chunk
__global__ void kb_sims_diffuse(const SimRec *__restrict__ tab, int n) {
  const SimRec &S = sim_of_row(tab, n, blockIdx.x);
  const BCGrid gx{S.vxb, S.vxf, S.W - 1, S.H, S.west, S.east, S.south, S.north, true};
  const BCGrid gy{S.vyb, S.vyf, S.W, S.H - 1, S.west, S.east, S.south, S.north, false};
  diffuse_cell(blockIdx.y * blockDim.x + threadIdx.x + 1, blockIdx.x - S.row0 + 1, S.vxf, S.vxb,
               S.vyf, S.vyb, S.fmask, S.W, S.H, S.a, gx, gy);
}

template <bool BUF> __global__ void kb_sims_advect(const SimRec *__restrict__ tab, int n) {
  const SimRec &S = sim_of_row(tab, n, blockIdx.x);
  const int Wx = S.W - 1;
  const int nck_x = Wx > 9 ? (Wx - 2) / 8 : 0, nck_y = S.W > 9 ? (S.W - 2) / 8 : 0;
  const int lanes = 8 * (nck_x > nck_y ? nck_x : nck_y);
  const int lane = blockIdx.y * blockDim.x + threadIdx.x;
  if ((int)(blockIdx.y * blockDim.x) >= lanes) return; // whole block past this row's lanes
  advect_pair<BUF>(lane, blockIdx.x - S.row0 + 1, S.vxf, S.vyf, S.vxb, S.vyb, S.flag, S.W, S.H,
                   nck_x, nck_y, S.hdt, S.fdt);
}

__global__ void kb_sims_vbcs_divergence(const SimRec *__restrict__ tab, int n) {
  const SimRec &S = sim_of_row(tab, n, blockIdx.x);
  const BCGrid gx{S.vxf, S.vxb, S.W - 1, S.H, S.west, S.east, S.south, S.north, true};
  const BCGrid gy{S.vyf, S.vyb, S.W, S.H - 1, S.west, S.east, S.south, S.north, false};
  vbcs_divergence_cell(blockIdx.y * blockDim.x + threadIdx.x + 1, blockIdx.x - S.row0 + 1, S.vxf,
                       S.vyf, S.f, S.W, S.H, S.ih, gx, gy);
}

__global__ void kb_sims_correct_tail(const SimRec *__restrict__ tab, int n) {
  const SimRec &S = sim_of_row(tab, n, blockIdx.x);
  const BCGrid gx{S.vxf, S.vxb, S.W - 1, S.H, S.west, S.east, S.south, S.north, true};
  const BCGrid gy{S.vyf, S.vyb, S.W, S.H - 1, S.west, S.east, S.south, S.north, false};
  correct_tail_cell(blockIdx.y * blockDim.x + threadIdx.x + 1, blockIdx.x - S.row0 + 1, S.vxf,
                    S.vyf, S.p, S.fmask, S.W, S.H, S.ih, gx, gy, S.xcur, S.ycur, S.west, S.east,
                    S.north, S.south);
}

} // namespace

// ---------------------------------------------------------------------------
void launch_set_vbcs(const VelView &v, const BCs &bc, hipStream_t s) {
  BCGrid gx{v.vxf, v.vxb, v.W - 1, v.H, bc.west, bc.east, bc.south, bc.north, true};
  BCGrid gy{v.vyf, v.vyb, v.W, v.H - 1, bc.west, bc.east, bc.south, bc.north, false};
  int n = v.W > v.H ? v.W : v.H;
  hipLaunchKernelGGL(k_set_vbcs, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, gx, gy);
}
void launch_set_vbcs_vy(const VelView &v, const BCs &bc, hipStream_t s) {
  BCGrid gy{v.vyf, v.vyb, v.W, v.H - 1, bc.west, bc.east, bc.south, bc.north, false};
  int n = v.W > v.H ? v.W : v.H;
  hipLaunchKernelGGL(k_set_vbcs_one, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, gy);
}
// one iteration of both diffuse loops with their setVBCs; the back buffers
// become the fronts (the caller swaps both)
void launch_face_masks(const uint8_t *flag, uint16_t *fmask, int W, int H, hipStream_t s) {
  hipLaunchKernelGGL(k_face_masks, dim3(ceil_div(W, TPB), H), dim3(TPB), 0, s, flag, fmask, W, H);
}
void launch_diffuse_both(const VelView &v, const uint16_t *fmask, float a, const BCs &bc,
                         hipStream_t s) {
  BCGrid gx{v.vxb, v.vxf, v.W - 1, v.H, bc.west, bc.east, bc.south, bc.north, true};
  BCGrid gy{v.vyb, v.vyf, v.W, v.H - 1, bc.west, bc.east, bc.south, bc.north, false};
  hipLaunchKernelGGL(k_diffuse_both, dim3(ceil_div(v.W - 2, TPB), v.H - 2), dim3(TPB), 0, s, v.vxf,
                     v.vxb, v.vyf, v.vyb, fmask, v.W, v.H, a, gx, gy);
}
void launch_vbcs_divergence(const VelView &v, const BCs &bc, float *f, float ih, hipStream_t s) {
  BCGrid gx{v.vxf, v.vxb, v.W - 1, v.H, bc.west, bc.east, bc.south, bc.north, true};
  BCGrid gy{v.vyf, v.vyb, v.W, v.H - 1, bc.west, bc.east, bc.south, bc.north, false};
  hipLaunchKernelGGL(k_vbcs_divergence, dim3(ceil_div(v.W - 2, TPB), v.H - 2), dim3(TPB), 0, s,
                     v.vxf, v.vyf, f, v.W, v.H, ih, gx, gy);
}
void launch_correct_tail(const VelView &v, float *p, const uint16_t *fmask, float ih,
                         const BCs &bc, float *xcur, float *ycur, hipStream_t s) {
  BCGrid gx{v.vxf, v.vxb, v.W - 1, v.H, bc.west, bc.east, bc.south, bc.north, true};
  BCGrid gy{v.vyf, v.vyb, v.W, v.H - 1, bc.west, bc.east, bc.south, bc.north, false};
  hipLaunchKernelGGL(k_correct_tail, dim3(ceil_div(v.W - 2, TPB), v.H - 2), dim3(TPB), 0, s, v.vxf,
                     v.vyf, p, fmask, v.W, v.H, ih, gx, gy, xcur, ycur, bc.west, bc.east, bc.north,
                     bc.south);
}
void launch_set_pbc(float *p, int W, int H, const BCs &bc, hipStream_t s) {
  int n = W > H ? W : H;
  hipLaunchKernelGGL(k_set_pbc, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, p, W, H, bc.west, bc.east,
                     bc.north, bc.south);
}
void launch_accumulate(const VelView &v, float *xa, float *ya, hipStream_t s) {
  hipLaunchKernelGGL(k_accumulate, dim3(ceil_div(v.W - 2, TPB), v.H - 2), dim3(TPB), 0, s, v.vxf,
                     xa, v.vyf, ya, v.W, v.H);
}
void launch_advect(const VelView &v, const uint8_t *flag, float hdt, float fdt, hipStream_t s) {
  const int Wx = v.W - 1;
  int nck_x = Wx > 9 ? (Wx - 2) / 8 : 0;
  int nck_y = v.W > 9 ? (v.W - 2) / 8 : 0;
  int lanes = 8 * (nck_x > nck_y ? nck_x : nck_y);
  if (lanes == 0 || v.H < 3) return;
  // buffer-load gathers while every byte offset of a velocity grid fits in
  // 31 bits; UBG_ADVECT_FLAT_LOADS=1 forces the 64-bit flat loads of larger
  // grids (a test switch: both paths give the same values)
  const char *flat = std::getenv("UBG_ADVECT_FLAT_LOADS");
  const bool buf = (size_t)v.W * v.H * 4 < ((size_t)1 << 31) && !(flat && flat[0] == '1');
  launch_k(buf ? k_advect<true> : k_advect<false>, dim3(ceil_div(lanes, TPB), v.H - 2),
           dim3(TPB), 0, s, v.vxf, v.vyf, v.vxb, v.vyb, flag, v.W, v.H, nck_x, nck_y, hdt, fdt);
}
void launch_divergence(const VelView &v, float *f, float ih, hipStream_t s) {
  hipLaunchKernelGGL(k_divergence, dim3(ceil_div(v.W - 2, TPB), v.H - 2), dim3(TPB), 0, s, v.vxf,
                     v.vyf, f, v.W, v.H, ih);
}
void launch_stamp_sinks(float *f, const int *cell, const float *val, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_stamp_sinks, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, f, cell, val, n);
}
void launch_correct_velocity(const VelView &v, const float *p, const uint8_t *flag, float ih,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_correct_velocity, dim3(ceil_div(v.W - 2, TPB), v.H - 2), dim3(TPB), 0, s,
                     v.vxf, v.vyf, p, flag, v.W, v.H, ih);
}
void launch_shift_map(const VelView &v, float *p, const uint8_t *new_flag, uint8_t *flag,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_shift_rows, dim3(v.H, 3), dim3(TPB), sizeof(float) * v.W, s, v.vxf, v.vxb,
                     v.vyf, v.vyb, p, v.W, v.H);
  hipLaunchKernelGGL(k_apply_flags, dim3(ceil_div(v.W, TPB), v.H), dim3(TPB), 0, s, new_flag, flag,
                     v.vxf, v.vyf, p, v.W, v.H);
  hipLaunchKernelGGL(k_inlet, dim3(1), dim3(256), 0, s, flag, v.vxf, v.vxb, v.W, v.H);
}
void launch_terrain_flags(const uint8_t *rgba, int iw, int ih, int scale, float *full, float *sim,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_terrain_full, dim3(ceil_div(iw, TPB), ih), dim3(TPB), 0, s, rgba, iw, ih,
                     full);
  const int sw = iw / scale, sh = ih / scale;
  if (sw <= 0 || sh <= 0) return;
  hipLaunchKernelGGL(k_terrain_subsample, dim3(ceil_div(sw, TPB), sh), dim3(TPB), 0, s, full, iw,
                     scale, sim, sw);
}
void launch_collocate(const float *vx, const float *vy, float *vxy, float *mag, int W, int H,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_collocate, dim3(ceil_div(2 * W - 1, TPB), 2 * H - 1), dim3(TPB), 0, s, vx,
                     vy, vxy, mag, W, H);
}
void launch_bilinear_vel(const VelView &v, const float *c, float *out, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_bilinear_vel, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, v.vxf, v.vyf, v.W,
                     v.H, c, out, n);
}
void launch_sample_flag(const uint8_t *flag, int W, int H, float pwidth, float h, int mode,
                        const float *pc, float *out, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sample_flag, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, flag, W, H, pwidth, h,
                     mode, pc, out, n);
}
void launch_max_vel(const VelView &v, float *part, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_max_vel, dim3(nblocks), dim3(256), 0, s, v.vxf, v.vyf, v.W, v.H, part);
}
void launch_sum_interior(const float *p, int W, int H, double *part, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_interior, dim3(nblocks), dim3(256), 0, s, p, W, H, part);
}
void launch_sub_interior(float *p, int W, int H, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_sub_interior, dim3(ceil_div(W - 2, TPB), H - 2), dim3(TPB), 0, s, p, W, H,
                     v);
}
void launch_gs(float *p, const float *f, const uint8_t *mask, int W, int H, float h, float alpha,
               hipStream_t s) {
  for (int d = 2; d <= (W - 2) + (H - 2); d++) {
    const int y0 = d - (W - 2) > 1 ? d - (W - 2) : 1;
    const int y1 = (H - 2) < d - 1 ? (H - 2) : d - 1;
    hipLaunchKernelGGL(k_gs_diagonal, dim3(ceil_div(y1 - y0 + 1, TPB)), dim3(TPB), 0, s, p, f, mask,
                       W, H, h, alpha, d);
  }
}
void launch_sims_diffuse(const SimRec *tab, int n, int rows, int max_w, hipStream_t s) {
  if (n <= 0 || rows <= 0) return;
  hipLaunchKernelGGL(kb_sims_diffuse, dim3(rows, ceil_div(max_w - 2, TPB)), dim3(TPB), 0, s, tab,
                     n);
}
void launch_sims_advect(const SimRec *tab, int n, int rows, int max_w, bool buf, hipStream_t s) {
  if (n <= 0 || rows <= 0) return;
  const int lanes = 8 * ((max_w - 2) / 8); // a bound on every simulation's chunked lanes
  if (lanes <= 0) return;
  hipLaunchKernelGGL(buf ? kb_sims_advect<true> : kb_sims_advect<false>,
                     dim3(rows, ceil_div(lanes, TPB)), dim3(TPB), 0, s, tab, n);
}
void launch_sims_vbcs_divergence(const SimRec *tab, int n, int rows, int max_w, hipStream_t s) {
  if (n <= 0 || rows <= 0) return;
  hipLaunchKernelGGL(kb_sims_vbcs_divergence, dim3(rows, ceil_div(max_w - 2, TPB)), dim3(TPB), 0,
                     s, tab, n);
}
void launch_sims_correct_tail(const SimRec *tab, int n, int rows, int max_w, hipStream_t s) {
  if (n <= 0 || rows <= 0) return;
  hipLaunchKernelGGL(kb_sims_correct_tail, dim3(rows, ceil_div(max_w - 2, TPB)), dim3(TPB), 0, s,
                     tab, n);
}
void launch_tracers(const VelView &v, const uint8_t *flag, float h, float pwidth, float dt,
                    uint32_t seed, float *xy, float *age, const int *perm, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_tracers, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, v.vxf, v.vyf, flag, v.W,
                     v.H, h, pwidth, dt, seed, (float2 *)xy, age, perm, n);
}

} // namespace ubg
