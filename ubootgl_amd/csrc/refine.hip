// refine.hip -- FP64 iterative refinement around FP32 V-cycles.
//
// No counterpart in the reference; it wraps MG::solve (pressure_solver.hpp:
// 59-62) for a standalone Poisson problem in the form of mgtest
// (mgtest/mgtest.cpp:18-53: all-fluid outer ring holding Dirichlet values,
// no zero-gradient BC).  The fine-grid operator is calculateResidualField's
// (pressure_solver.cpp:91-116) evaluated in float64:
//
//   r(c) = F(c) * (f(c) + sum_n F(n) (x(n) - x(c)) / h^2)   interior cells
//   r(c) = 0                                               ring and solids
//
// One refinement step: k FP32 V-cycles on A e = (float)r from e = 0, then
// x += e and the new r in float64, fused in one pass (k_ir_update) that also
// zeroes the next step's e.  ||r||_2 is reduced in a fixed order (per-block
// partial sums, then one block), so a solve is deterministic run to run.
//
// HBM bytes per fine cell of the fused pass: x64 read 8 + e32 read 4 +
// f32 read 4 + mask 1 + x64 write 8 + r32 write 4 + e32 zero 4 = 33.
#include "capi_common.hpp"

struct IRState {
  char *base = nullptr;
  double *x[2] = {nullptr, nullptr};
  int cur = 0;
  float *f = nullptr;
  double *partial = nullptr;
  double *acc = nullptr;
  int npartial = 0;
};

namespace ubg {
namespace {

constexpr int IR_TPB = 256; // cells of one row per block
constexpr int IR_ROWS = 8;  // rows per block

__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// x_b = x_a + e (ACC) or x_a, r32 = (float) r(x_b), e_zero = 0, and one
// partial sum of r^2 per block.  Neighbour values x_a(n) + e(n) are formed
// from the inputs, so no thread reads a value another thread writes.
template <bool ACC>
__global__ void __launch_bounds__(IR_TPB)
k_ir_update(const double *__restrict__ xa, double *__restrict__ xb, const float *__restrict__ e,
            float *__restrict__ ez, const float *__restrict__ f, const uint8_t *__restrict__ m,
            float *__restrict__ r32, int w, int h, double ihsq, double *__restrict__ partial) {
  const int x = blockIdx.x * IR_TPB + threadIdx.x;
  const int y0 = blockIdx.y * IR_ROWS;
  const int y1 = min(y0 + IR_ROWS, h);
  double acc = 0.0;
  if (x < w) {
    auto X = [&](size_t i) { return ACC ? xa[i] + (double)e[i] : xa[i]; };
    size_t c = (size_t)y0 * w + x;
    double xs = y0 > 0 ? X(c - w) : 0.0; // row y-1
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
  __shared__ double red[IR_TPB / 64];
  acc = warp_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < IR_TPB / 64; i++) s += red[i];
    partial[blockIdx.y * gridDim.x + blockIdx.x] = s;
  }
}

// out = sum of n partials, one block, fixed order
__global__ void __launch_bounds__(1024) k_sum_partials(const double *__restrict__ p, int n,
                                                         double *__restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += p[i];
  __shared__ double red[16];
  s = warp_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < 16; i++) t += red[i];
    out[0] = t;
  }
}

dim3 ir_grid(int w, int h) { return dim3((w + IR_TPB - 1) / IR_TPB, (h + IR_ROWS - 1) / IR_ROWS); }

int ir_ensure(ubg_mg *m) {
  if (m->ir) return UBG_OK;
  const size_t n = (size_t)m->W * m->H;
  const dim3 g = ir_grid(m->W, m->H);
  auto *s = new IRState();
  s->npartial = (int)(g.x * g.y);
  const size_t bytes = 2 * align_up(n * 8) + align_up(n * 4) + align_up((size_t)s->npartial * 8) +
                       align_up(8);
  hipError_t e = hipMalloc((void **)&s->base, bytes);
  if (e != hipSuccess) {
    delete s;
    return fail(UBG_E_OOM, std::string("refinement buffers: ") + hipGetErrorString(e));
  }
  char *q = s->base;
  s->x[0] = (double *)q;
  q += align_up(n * 8);
  s->x[1] = (double *)q;
  q += align_up(n * 8);
  s->f = (float *)q;
  q += align_up(n * 4);
  s->partial = (double *)q;
  q += align_up((size_t)s->npartial * 8);
  s->acc = (double *)q;
  m->ir = s;
  UBG_HIP(hipMemsetAsync(s->base, 0, bytes, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

// one fused update pass; returns ||r||_2^2 of the new x (synchronises)
int ir_update(ubg_mg *m, bool accumulate, double ihsq, double *norm2) {
  IRState &s = *m->ir;
  Level &L = m->lv[0];
  const dim3 g = ir_grid(L.w, L.h);
  const double *xa = s.x[s.cur];
  double *xb = s.x[s.cur ^ 1];
  if (accumulate)
    hipLaunchKernelGGL(k_ir_update<true>, g, dim3(IR_TPB), 0, m->stream, xa, xb, L.p, L.p_alt,
                       s.f, L.mask, L.f, L.w, L.h, ihsq, s.partial);
  else
    hipLaunchKernelGGL(k_ir_update<false>, g, dim3(IR_TPB), 0, m->stream, xa, xb, L.p, L.p_alt,
                       s.f, L.mask, L.f, L.w, L.h, ihsq, s.partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, m->stream, s.partial, s.npartial,
                     s.acc);
  UBG_TRY(check_launch());
  s.cur ^= 1;
  std::swap(L.p, L.p_alt); // e := the zeroed buffer
  UBG_HIP(hipMemcpyAsync(norm2, s.acc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

} // namespace

void ir_release(ubg_mg *m) {
  if (!m || !m->ir) return;
  if (m->ir->base) (void)hipFree(m->ir->base);
  delete m->ir;
  m->ir = nullptr;
}

} // namespace ubg

extern "C" {

UBG_API void ubg_ir_params_default(ubg_ir_params *out) {
  if (!out) return;
  out->tol = 1e-10;
  out->vcycles_per_iter = 4;
  out->max_iter = 200;
  out->ref_norm = 0.0;
}

UBG_API int ubg_mg_ir_set_rhs(ubg_mg *m, const float *f) {
  if (!m || !f) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  UBG_TRY(ir_ensure(m));
  const size_t n = (size_t)m->W * m->H;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(f[i])) return fail(UBG_E_NONFINITE, "rhs has a non-finite value");
  UBG_HIP(hipMemcpyAsync(m->ir->f, f, n * 4, hipMemcpyHostToDevice, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

UBG_API int ubg_mg_ir_set_guess(ubg_mg *m, const double *x) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  UBG_TRY(ir_ensure(m));
  const size_t n = (size_t)m->W * m->H;
  double *d = m->ir->x[m->ir->cur];
  if (x) {
    for (size_t i = 0; i < n; i++)
      if (!std::isfinite(x[i])) return fail(UBG_E_NONFINITE, "guess has a non-finite value");
    UBG_HIP(hipMemcpyAsync(d, x, n * 8, hipMemcpyHostToDevice, m->stream));
  } else {
    launch_fill_f64(d, n, 0.0, m->stream);
    UBG_TRY(check_launch());
  }
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

UBG_API int ubg_mg_ir_solve(ubg_mg *m, float h, const ubg_ir_params *prm, ubg_ir_result *res) {
  if (!m || !res) return fail(UBG_E_INVALID, "null argument");
  ubg_ir_params P;
  ubg_ir_params_default(&P);
  if (prm) P = *prm;
  if (!(P.tol > 0.0 && P.tol < 1.0)) return fail(UBG_E_INVALID, "tol must be in (0, 1)");
  if (P.vcycles_per_iter < 1 || P.vcycles_per_iter > 256 || P.max_iter < 1 || P.max_iter > 1000000)
    return fail(UBG_E_INVALID, "vcycles_per_iter or max_iter out of range");
  if (!(h > 0.0f) || !std::isfinite(h)) return fail(UBG_E_INVALID, "h must be positive");
  if (!(P.ref_norm >= 0.0) || !std::isfinite(P.ref_norm))
    return fail(UBG_E_INVALID, "ref_norm must be finite and >= 0");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  UBG_TRY(ir_ensure(m));
  std::memset(res, 0, sizeof(*res));
  const double hd = (double)h;
  const double ihsq = 1.0 / (hd * hd);
  double n2 = 0.0;
  UBG_TRY(ir_update(m, false, ihsq, &n2));
  if (!std::isfinite(n2)) return fail(UBG_E_NONFINITE, "initial residual is not finite");
  const double r0 = std::sqrt(n2);
  res->r0_norm = r0;
  if (r0 == 0.0) {
    res->converged = 1;
    return UBG_OK;
  }
  // relative to ref_norm when given (the batched solver's rule: a warm start
  // that already meets tol takes no step)
  const double denom = P.ref_norm > 0.0 ? P.ref_norm : r0;
  double best = 1.0, rel = r0 / denom;
  res->rel_residual = rel;
  if (rel <= P.tol) {
    res->converged = 1;
    return UBG_OK;
  }
  int stall = 0;
  for (int it = 1; it <= P.max_iter; it++) {
    for (int v = 0; v < P.vcycles_per_iter; v++) m->solve_level(0, h, false);
    UBG_TRY(check_launch());
    UBG_TRY(ir_update(m, true, ihsq, &n2));
    rel = std::sqrt(n2) / denom;
    res->iterations = it;
    res->vcycles += P.vcycles_per_iter;
    res->rel_residual = rel;
    if (it <= UBG_IR_HISTORY) res->history[it - 1] = rel;
    if (!std::isfinite(rel)) return fail(UBG_E_NONFINITE, "residual norm is not finite");
    if (rel <= P.tol) {
      res->converged = 1;
      return UBG_OK;
    }
    // stop once three steps in a row fail to gain 10%: the FP32 correction
    // no longer resolves what is left
    if (rel > 0.9 * best) {
      if (++stall >= 3) break;
    } else {
      stall = 0;
    }
    best = std::min(best, rel);
  }
  return fail(UBG_E_NOT_CONVERGED, "refinement stopped at relative residual " +
                                       std::to_string(rel));
}

UBG_API int ubg_mg_ir_get_solution(ubg_mg *m, double *x) {
  if (!m || !x) return fail(UBG_E_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  if (!m->ir) return fail(UBG_E_STATE, "no refinement solve has run");
  const size_t n = (size_t)m->W * m->H;
  UBG_HIP(hipMemcpyAsync(x, m->ir->x[m->ir->cur], n * 8, hipMemcpyDeviceToHost, m->stream));
  UBG_HIP(hipStreamSynchronize(m->stream));
  return UBG_OK;
}

UBG_API int ubg_mg_ir_device_ptrs(ubg_mg *m, double **x, float **f) {
  if (!m) return fail(UBG_E_INVALID, "null mg");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  m->enter();
  UBG_TRY(ir_ensure(m));
  if (x) *x = m->ir->x[m->ir->cur];
  if (f) *f = m->ir->f;
  return UBG_OK;
}

} // extern "C"
