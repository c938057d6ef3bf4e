// mg_kernels.hip -- multigrid pressure-solver kernels for gfx950.
//
// Follows pressure_solver.cpp / pressure_solver.hpp of the reference:
//   smoothingKernel :10-24, rbgs :49-89 (two-pass red/black semantics),
//   calculateResidualField :91-116, restrict :118-132, prolongate :134-172,
//   correct :174-181, setZeroGradientBC :183-192, updateFields hpp:34-57.
// All kernels are memory-bound stencils (no MFMA); each thread owns one cell
// and one wave64 covers 64 consecutive cells of a row, so every load and store
// of a wave is one contiguous 256-byte segment.
#include "ubg_internal.hpp"

namespace ubg {

namespace {

constexpr int TPB = 256;

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

__global__ void k_fill_f32(float *__restrict__ d, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = v;
}
__global__ void k_fill_f64(double *__restrict__ d, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = v;
}
__global__ void k_f32_from_u8(const uint8_t *__restrict__ s, float *__restrict__ d, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}

// Sum of squares, accumulated in double: block reduction through LDS, one
// atomic per block.
__global__ void k_sumsq(const float *__restrict__ s, size_t n, double *acc) {
  __shared__ double red[TPB];
  double v = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double x = s[i];
    v += x * x;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = TPB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(acc, red[0]);
}

// mask = own flag (bit 0) + the four neighbours' flags (bits 1-4) + their
// count (bits 5-7); neighbours outside the grid count as 0.
__global__ void k_build_mask(const uint8_t *__restrict__ fl, uint8_t *__restrict__ m, int w,
                             int h) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= w) return;
  size_t i = (size_t)y * w + x;
  uint8_t b = fl[i] ? MK_SELF : 0;
  if (x > 0 && fl[i - 1]) b |= MK_W;
  if (x < w - 1 && fl[i + 1]) b |= MK_E;
  if (y > 0 && fl[i - w]) b |= MK_S;
  if (y < h - 1 && fl[i + w]) b |= MK_N;
  unsigned cnt = ((b >> 1) & 1) + ((b >> 2) & 1) + ((b >> 3) & 1) + ((b >> 4) & 1);
  m[i] = (uint8_t)(b | (cnt << 5));
}

// pressure_solver.hpp:43-56: coarse flag is 1 iff sum(w * f)/16 > 0.2 over the
// full-weighting stencil; for 0/1 fine flags that is the integer rule
// weighted count >= 4.  The coarse boundary is 1 (fill(1.0), :41).
__global__ void k_coarsen(const uint8_t *__restrict__ ff, int wf, uint8_t *__restrict__ fc,
                          int wc, int hc) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= wc) return;
  uint8_t out = 1;
  if (x >= 1 && x < wc - 1 && y >= 1 && y < hc - 1) {
    const uint8_t *r0 = ff + (size_t)(2 * y - 1) * wf + 2 * x;
    const uint8_t *r1 = r0 + wf;
    const uint8_t *r2 = r1 + wf;
    int v = (r0[-1] + 2 * r0[0] + r0[1]) + 2 * (r1[-1] + 2 * r1[0] + r1[1]) +
            (r2[-1] + 2 * r2[0] + r2[1]);
    out = v >= 4 ? 1 : 0;
  }
  fc[(size_t)y * wc + x] = out;
}

// setZeroGradientBC (pressure_solver.cpp:183-192): edges copy the first
// interior cell; corners untouched.  No read depends on another write.
__global__ void k_zero_gradient(float *__restrict__ p, int w, int h) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 1 && t < h - 1) {
    size_t r = (size_t)t * w;
    p[r] = p[r + 1];
    p[r + w - 1] = p[r + w - 2];
  }
  if (t >= 1 && t < w - 1) {
    p[t] = p[(size_t)w + t];
    p[(size_t)(h - 1) * w + t] = p[(size_t)(h - 2) * w + t];
  }
}

// calculateResidualField for one interior cell (pressure_solver.cpp:100-108):
// r = F*(f + (sum_n(F_n p_n + (1-F_n) p) - 4p) / h^2), neighbours summed in
// the order x-1, x+1, y-1, y+1.
__device__ __forceinline__ float residual_cell(const float *__restrict__ p,
                                               const float *__restrict__ f,
                                               const uint8_t *__restrict__ m, size_t i, int w,
                                               float ihsq) {
  uint8_t mk = m[i];
  float pc = p[i];
  float val = 0.0f;
  val += (mk & MK_W) ? p[i - 1] : pc;
  val += (mk & MK_E) ? p[i + 1] : pc;
  val += (mk & MK_S) ? p[i - w] : pc;
  val += (mk & MK_N) ? p[i + w] : pc;
  val += -4.0f * pc;
  val *= ihsq;
  return (mk & MK_SELF) ? (f[i] + val) : 0.0f;
}

__global__ void k_residual(const float *__restrict__ p, const float *__restrict__ f,
                           const uint8_t *__restrict__ m, float *__restrict__ r, int w,
                           float ihsq) {
  int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int y = blockIdx.y + 1;
  if (x >= w - 1) return;
  size_t i = (size_t)y * w + x;
  r[i] = residual_cell(p, f, m, i, w, ihsq);
}

// Fused calculateResidualField + restrict (pressure_solver.cpp:217-227): the
// fine residual never reaches HBM.  Each thread owns one coarse interior cell,
// computes the nine fine residuals under the full-weighting stencil and
// writes the coarse right-hand side; the coarse solution starts at zero
// (ec.fill(0.0), :224).
__global__ void k_residual_restrict(const float *__restrict__ p, const float *__restrict__ f,
                                    const uint8_t *__restrict__ m, int wf, float ihsq,
                                    float *__restrict__ rc, float *__restrict__ ec, int wc) {
  int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int y = blockIdx.y + 1;
  if (x >= wc - 1) return;
  size_t c = (size_t)(2 * y) * wf + 2 * x;
  auto R = [&](int dx, int dy) { return residual_cell(p, f, m, c + (ptrdiff_t)dy * wf + dx, wf, ihsq); };
  float v = 0.0f;
  v = R(-1, -1) * 1.0f + R(0, -1) * 2.0f + R(1, -1) * 1.0f;
  v += R(-1, 0) * 2.0f + R(0, 0) * 4.0f + R(1, 0) * 2.0f;
  v += R(-1, 1) * 1.0f + R(0, 1) * 2.0f + R(1, 1) * 1.0f;
  size_t ci = (size_t)y * wc + x;
  rc[ci] = v * (1.0f / 16.0f);
  ec[ci] = 0.0f;
}

// LDS-tiled variant for large levels: a workgroup owns 32x8 coarse cells,
// computes the 65x17 fine residuals under them once into LDS (p staged with
// a one-cell halo, all loads issued before the first LDS store), then
// restricts.  Same arithmetic as k_residual_restrict.
constexpr int RCX = 32, RCY = 8;
constexpr int RFX = 2 * RCX + 1, RFY = 2 * RCY + 1;  // residual region
constexpr int RPX = RFX + 2, RPY = RFY + 2;          // p region
__global__ void __launch_bounds__(256)
    k_residual_restrict_tile(const float *__restrict__ p, const float *__restrict__ f,
                             const uint8_t *__restrict__ m, int wf, int hf, float ihsq,
                             float *__restrict__ rc, float *__restrict__ ec, int wc, int hc,
                             int ntx) {
  __shared__ float sp[RPY][RPX];
  __shared__ float sr[RFY][RFX + 1];
  const int tid = threadIdx.x;
  const int X0 = 1 + (blockIdx.x % ntx) * RCX, Y0 = 1 + (blockIdx.x / ntx) * RCY;
  const int fx0 = 2 * X0 - 1, fy0 = 2 * Y0 - 1; // residual region origin
  constexpr int NP = RPX * RPY, NPI = (NP + 255) / 256;
  float v[NPI];
#pragma unroll
  for (int k = 0; k < NPI; k++) {
    int li = tid + 256 * k;
    int r = li / RPX, c = li - r * RPX;
    int gx = fx0 - 1 + c, gy = fy0 - 1 + r;
    v[k] = (li < NP && gx >= 0 && gx < wf && gy >= 0 && gy < hf) ? p[(size_t)gy * wf + gx] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < NPI; k++) {
    int li = tid + 256 * k;
    if (li < NP) (&sp[0][0])[li] = v[k];
  }
  __syncthreads();
  for (int li = tid; li < RFX * RFY; li += 256) {
    int r = li / RFX, c = li - r * RFX;
    int gx = fx0 + c, gy = fy0 + r;
    float res = 0.0f;
    if (gx >= 1 && gx < wf - 1 && gy >= 1 && gy < hf - 1) {
      size_t gi = (size_t)gy * wf + gx;
      unsigned mk = m[gi];
      float pc = sp[r + 1][c + 1];
      float val = 0.0f;
      val += (mk & MK_W) ? sp[r + 1][c] : pc;
      val += (mk & MK_E) ? sp[r + 1][c + 2] : pc;
      val += (mk & MK_S) ? sp[r][c + 1] : pc;
      val += (mk & MK_N) ? sp[r + 2][c + 1] : pc;
      val += -4.0f * pc;
      val *= ihsq;
      res = (mk & MK_SELF) ? (f[gi] + val) : 0.0f;
    }
    sr[r][c] = res;
  }
  __syncthreads();
  const int cx = tid % RCX, cy = tid / RCX; // 256 threads = 32 x 8 coarse cells
  const int X = X0 + cx, Y = Y0 + cy;
  if (X < wc - 1 && Y < hc - 1) {
    const int r = 2 * cy + 1, c = 2 * cx + 1; // fine (2X, 2Y) in the residual region
    auto R = [&](int dx, int dy) { return sr[r + dy][c + dx]; };
    float vv = 0.0f;
    vv = R(-1, -1) * 1.0f + R(0, -1) * 2.0f + R(1, -1) * 1.0f;
    vv += R(-1, 0) * 2.0f + R(0, 0) * 4.0f + R(1, 0) * 2.0f;
    vv += R(-1, 1) * 1.0f + R(0, 1) * 2.0f + R(1, 1) * 1.0f;
    size_t ci = (size_t)Y * wc + X;
    rc[ci] = vv * (1.0f / 16.0f);
    ec[ci] = 0.0f;
  }
}

// flag sum + the double literal 0.0001, stored as float (pressure_solver.cpp:149)
__device__ __forceinline__ float eps_sum(int s) { return (float)((double)s + 0.0001); }

// Fused prolongate + correct (pressure_solver.cpp:134-181): e is computed per
// fine interior cell by its parity class and added to p; cells the
// reference's loops do not reach (last odd column/row of an odd-sized grid)
// keep e = 0.
__global__ void k_prolong_correct(float *__restrict__ p, const uint8_t *__restrict__ m, int w,
                                  int h, const float *__restrict__ ec,
                                  const uint8_t *__restrict__ fc, int wc) {
  int x = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int y = blockIdx.y + 1;
  if (x >= w - 1) return;
  size_t i = (size_t)y * w + x;
  float F = (m[i] & MK_SELF) ? 1.0f : 0.0f;
  int a = x >> 1, b = y >> 1;
  bool xe = !(x & 1), ye = !(y & 1);
  auto E = [&](int xx, int yy) { return ec[(size_t)yy * wc + xx]; };
  auto FC = [&](int xx, int yy) { return (int)fc[(size_t)yy * wc + xx]; };
  float e = 0.0f;
  if (xe && ye) {
    e = E(a, b) * F;
  } else if (!xe && ye) {
    if (x <= w - 3) e = F * (E(a, b) + E(a + 1, b)) / eps_sum(FC(a, b) + FC(a + 1, b));
  } else if (xe && !ye) {
    if (y <= h - 3) e = F * (E(a, b) + E(a, b + 1)) / eps_sum(FC(a, b) + FC(a, b + 1));
  } else {
    if (x <= w - 3 && y <= h - 3)
      e = F * (E(a, b) + E(a + 1, b + 1) + E(a + 1, b) + E(a, b + 1)) /
          eps_sum(FC(a, b) + FC(a + 1, b + 1) + FC(a + 1, b) + FC(a, b + 1));
  }
  p[i] += e;
}

} // namespace

// ---------------------------------------------------------------------------
void launch_fill_f32(float *dst, size_t n, float v, hipStream_t s) {
  if (!n) return;
  unsigned g = ceil_div(n, TPB);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fill_f32, dim3(g), dim3(TPB), 0, s, dst, n, v);
}
void launch_fill_f64(double *dst, size_t n, double v, hipStream_t s) {
  if (!n) return;
  unsigned g = ceil_div(n, TPB);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fill_f64, dim3(g), dim3(TPB), 0, s, dst, n, v);
}
void launch_f32_from_u8(const uint8_t *src, float *dst, size_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_f32_from_u8, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, src, dst, n);
}
void launch_sumsq(const float *src, size_t n, double *acc, hipStream_t s) {
  if (!n) return;
  unsigned g = ceil_div(n, TPB);
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_sumsq, dim3(g), dim3(TPB), 0, s, src, n, acc);
}
void launch_build_mask(const uint8_t *flag, uint8_t *mask, int w, int h, hipStream_t s) {
  hipLaunchKernelGGL(k_build_mask, dim3(ceil_div(w, TPB), h), dim3(TPB), 0, s, flag, mask, w, h);
}
void launch_coarsen(const uint8_t *ff, int wf, uint8_t *fc, int wc, int hc, hipStream_t s) {
  hipLaunchKernelGGL(k_coarsen, dim3(ceil_div(wc, TPB), hc), dim3(TPB), 0, s, ff, wf, fc, wc, hc);
}
void launch_zero_gradient(float *p, int w, int h, hipStream_t s) {
  int n = w > h ? w : h;
  hipLaunchKernelGGL(k_zero_gradient, dim3(ceil_div(n, TPB)), dim3(TPB), 0, s, p, w, h);
}
void launch_residual(const Level &L, float h, hipStream_t s) {
  float ihsq = 1.0f / h / h;
  hipLaunchKernelGGL(k_residual, dim3(ceil_div(L.w - 2, TPB), L.h - 2), dim3(TPB), 0, s, L.p,
                     L.f, L.mask, L.r, L.w, ihsq);
}
void launch_residual_restrict(const Level &F, const Level &C, float h, hipStream_t s) {
  if (C.w < 3 || C.h < 3) return;
  float ihsq = 1.0f / h / h;
  if ((size_t)C.w * C.h >= 16384) {
    int ntx = (C.w - 2 + RCX - 1) / RCX, nty = (C.h - 2 + RCY - 1) / RCY;
    hipLaunchKernelGGL(k_residual_restrict_tile, dim3(ntx * nty), dim3(256), 0, s, F.p, F.f,
                       F.mask, F.w, F.h, ihsq, C.f, C.p, C.w, C.h, ntx);
  } else {
    hipLaunchKernelGGL(k_residual_restrict, dim3(ceil_div(C.w - 2, TPB), C.h - 2), dim3(TPB), 0,
                       s, F.p, F.f, F.mask, F.w, ihsq, C.f, C.p, C.w);
  }
}
void launch_prolong_correct(const Level &F, const Level &C, hipStream_t s) {
  hipLaunchKernelGGL(k_prolong_correct, dim3(ceil_div(F.w - 2, TPB), F.h - 2), dim3(TPB), 0, s,
                     F.p, F.mask, F.w, F.h, C.p, C.flag, C.w);
}

} // namespace ubg
