// components.hip -- connected components of the fluid cells and the
// per-component mean removal of a right-hand side (ubg_flag_components,
// ubg_remove_component_means).
//
// With solid walls every fluid component of the flag is a pure-Neumann
// problem of its own: its rhs is compatible only when it sums to zero over
// the component.  The labelling is integer work and exact:
//   * components under 4-neighbour connectivity (the stencil's:
//     calculateResidualField, pressure_solver.cpp:100-108, couples a cell to
//     x-1, x+1, y-1, y+1 only);
//   * a lock-free union-find over the cells: every link joins a cell to its
//     right and lower fluid neighbours, and a root is always hooked below a
//     smaller index with an atomic minimum, so the surviving root of a
//     component is its smallest linear index whatever the order the links
//     ran in (parent indices only decrease, so no cycle can form);
//   * components are numbered 0..n-1 in raster order of their first cell.
// The mean removal sums f over each component as 64-bit fixed point
// (f * 2^s rounded to an integer, s chosen from max |f| and the cell count
// so no sum can overflow): integer addition is associative, so the sums and
// the result do not depend on the order the atomics ran in.
#include "capi_common.hpp"

namespace ubg {
namespace {

constexpr int TPB = 256;
constexpr int CC_BLOCK = 1024;           // cells per numbering block
constexpr int CC_MAX_STEPS = 1 << 26;    // bound on a find or unite loop

__device__ __forceinline__ int ld(const int *L, int i) {
  return __hip_atomic_load(L + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root of x: follow parents until a fixed point (parents only decrease)
__device__ __forceinline__ int find_root(const int *L, int x, int *err) {
  for (int k = 0; k < CC_MAX_STEPS; k++) {
    const int p = ld(L, x);
    if (p == x) return x;
    x = p;
  }
  atomicOr(err, 1);
  return x;
}

__device__ __forceinline__ void unite(int *L, int a, int b, int *err) {
  for (int k = 0; k < CC_MAX_STEPS; k++) {
    a = find_root(L, a, err);
    b = find_root(L, b, err);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    // hook root b below a; if b stopped being a root meanwhile, retry from
    // what it points to now
    const int old = atomicMin(L + b, a);
    if (old == b) return;
    b = old;
  }
  atomicOr(err, 1);
}

__global__ void k_cc_init(const uint8_t *__restrict__ fl, int *__restrict__ L, int n) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) L[i] = fl[i] ? i : -1;
}

__global__ void k_cc_link(const uint8_t *__restrict__ fl, int *L, int W, int H, int *err) {
  const int x = blockIdx.x * TPB + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  const int i = y * W + x;
  if (!fl[i]) return;
  if (x + 1 < W && fl[i + 1]) unite(L, i, i + 1, err);
  if (y + 1 < H && fl[i + W]) unite(L, i, i + W, err);
}

__global__ void k_cc_flatten(int *L, int n, int *err) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n && ld(L, i) >= 0) L[i] = find_root(L, i, err);
}

// roots per block of CC_BLOCK cells
__global__ void __launch_bounds__(CC_BLOCK) k_cc_count(const int *__restrict__ L, int n,
                                                        int *__restrict__ cnt) {
  __shared__ int wsum[CC_BLOCK / 64];
  const int i = blockIdx.x * CC_BLOCK + threadIdx.x;
  const bool root = i < n && L[i] == i;
  const unsigned long long b = __ballot(root);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < CC_BLOCK / 64; w++) s += wsum[w];
    cnt[blockIdx.x] = s;
  }
}

// exclusive prefix of the per-block counts, one workgroup; off[nb] = total
__global__ void __launch_bounds__(1024) k_cc_scan(const int *__restrict__ cnt, int nb,
                                                  int *__restrict__ off) {
  __shared__ int part[1024];
  const int per = (nb + 1023) / 1024;
  const int lo = threadIdx.x * per, hi = min(lo + per, nb);
  int s = 0;
  for (int k = lo; k < hi; k++) s += cnt[k];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int t = 0; t < 1024; t++) {
      const int v = part[t];
      part[t] = run;
      run += v;
    }
    off[nb] = run;
  }
  __syncthreads();
  int run = part[threadIdx.x];
  for (int k = lo; k < hi; k++) {
    off[k] = run;
    run += cnt[k];
  }
}

// number of each root: its rank among the roots in raster order
__global__ void __launch_bounds__(CC_BLOCK) k_cc_number(const int *__restrict__ L, int n,
                                                         const int *__restrict__ off,
                                                         int *__restrict__ id) {
  __shared__ int wsum[CC_BLOCK / 64];
  const int i = blockIdx.x * CC_BLOCK + threadIdx.x;
  const bool root = i < n && L[i] == i;
  const unsigned long long b = __ballot(root);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = __popcll(b);
  __syncthreads();
  int base = off[blockIdx.x];
  for (int k = 0; k < w; k++) base += wsum[k];
  if (root) id[i] = base + __popcll(b & ((1ull << lane) - 1ull));
}

__global__ void k_cc_label(const int *__restrict__ L, const int *__restrict__ id, int n,
                           int *__restrict__ out) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) out[i] = L[i] >= 0 ? id[L[i]] : -1;
}

// max |f| over fluid cells as float bits (non-negative floats order as
// their bit patterns)
__global__ void k_absmax(const float *__restrict__ f, const int *__restrict__ lab, int n,
                         unsigned *mx) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  unsigned v = 0u;
  if (i < n && lab[i] >= 0) v = __float_as_uint(fabsf(f[i]));
  for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o, 64));
  if ((threadIdx.x & 63) == 0 && v) atomicMax(mx, v);
}

// fixed-point sums and counts per component; a wave whose active cells all
// lie in one component adds its total once
__global__ void k_comp_sums(const float *__restrict__ f, const int *__restrict__ lab, int n,
                            double scale, unsigned long long *sums, int *counts) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  const int c = i < n ? lab[i] : -1;
  const long long q = c >= 0 ? __double2ll_rn((double)f[i] * scale) : 0ll;
  const int c0 = __shfl(c, __builtin_amdgcn_readfirstlane(__ffsll(__ballot(c >= 0)) - 1), 64);
  const bool uniform = __ballot(c >= 0 && c != c0) == 0ull && c0 >= 0;
  if (uniform) {
    long long s = q;
    int k = c >= 0 ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) {
      s += __shfl_xor(s, o, 64);
      k += __shfl_xor(k, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(sums + c0, (unsigned long long)s);
      atomicAdd(counts + c0, k);
    }
  } else if (c >= 0) {
    atomicAdd(sums + c, (unsigned long long)q);
    atomicAdd(counts + c, 1);
  }
}

__global__ void k_sub_means(float *__restrict__ f, const int *__restrict__ lab, int n,
                            const unsigned long long *__restrict__ sums,
                            const int *__restrict__ counts, double inv_scale) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const int c = lab[i];
  if (c < 0) return;
  const double mean = (double)(long long)sums[c] * inv_scale / (double)counts[c];
  f[i] = (float)((double)f[i] - mean);
}

int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// labels of the flag (device), component count to *ncomp
int label_device(const uint8_t *d_fl, int W, int H, int *d_lab, int *ncomp, hipStream_t s) {
  const int n = W * H;
  const int nb = cdiv(n, CC_BLOCK);
  int *L = nullptr, *id = nullptr, *cnt = nullptr, *off = nullptr, *err = nullptr;
  size_t bytes = align_up((size_t)n * 4) * 2 + align_up((size_t)nb * 4) + align_up(((size_t)nb + 1) * 4) + 256;
  char *mem = nullptr;
  UBG_HIP(hipMalloc((void **)&mem, bytes));
  char *q = mem;
  L = (int *)q;
  q += align_up((size_t)n * 4);
  id = (int *)q;
  q += align_up((size_t)n * 4);
  cnt = (int *)q;
  q += align_up((size_t)nb * 4);
  off = (int *)q;
  q += align_up(((size_t)nb + 1) * 4);
  err = (int *)q;
  int rc = UBG_OK, herr = 0, total = 0;
  launch_fill_f32((float *)err, 1, 0.0f, s); // 0.0f is the all-zero bit pattern
  hipLaunchKernelGGL(k_cc_init, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, d_fl, L, n);
  hipLaunchKernelGGL(k_cc_link, dim3(cdiv(W, TPB), H), dim3(TPB), 0, s, d_fl, L, W, H, err);
  hipLaunchKernelGGL(k_cc_flatten, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, L, n, err);
  hipLaunchKernelGGL(k_cc_count, dim3(nb), dim3(CC_BLOCK), 0, s, L, n, cnt);
  hipLaunchKernelGGL(k_cc_scan, dim3(1), dim3(1024), 0, s, cnt, nb, off);
  hipLaunchKernelGGL(k_cc_number, dim3(nb), dim3(CC_BLOCK), 0, s, L, n, off, id);
  hipLaunchKernelGGL(k_cc_label, dim3(cdiv(n, TPB)), dim3(TPB), 0, s, L, id, n, d_lab);
  rc = check_launch();
  if (rc == UBG_OK && hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = fail(UBG_E_HIP, "component count copy");
  if (rc == UBG_OK && hipMemcpyAsync(&total, off + nb, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = fail(UBG_E_HIP, "component count copy");
  if (hipStreamSynchronize(s) != hipSuccess && rc == UBG_OK) rc = fail(UBG_E_HIP, "labelling");
  (void)hipFree(mem);
  if (rc != UBG_OK) return rc;
  if (herr) return fail(UBG_E_HIP, "component labelling did not converge");
  *ncomp = total;
  return UBG_OK;
}

struct Scratch {
  hipStream_t s = nullptr;
  uint8_t *fl = nullptr;
  int *lab = nullptr;
  ~Scratch() {
    if (fl) (void)hipFree(fl);
    if (lab) (void)hipFree(lab);
    if (s) (void)hipStreamDestroy(s);
  }
};

int setup(const float *flag, int W, int H, Scratch &sc) {
  if (!flag) return fail(UBG_E_INVALID, "null flag");
  if (W < 1 || H < 1 || (long)W * H >= (1l << 31) / 2)
    return fail(UBG_E_INVALID, "grid size out of range");
  std::vector<uint8_t> f8;
  UBG_TRY(validate_flag(flag, (size_t)W * H, f8));
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0)
    return fail(UBG_E_NO_DEVICE, "no HIP device visible");
  UBG_HIP(hipStreamCreateWithFlags(&sc.s, hipStreamNonBlocking));
  UBG_HIP(hipMalloc((void **)&sc.fl, f8.size()));
  UBG_HIP(hipMalloc((void **)&sc.lab, f8.size() * 4));
  UBG_HIP(hipMemcpyAsync(sc.fl, f8.data(), f8.size(), hipMemcpyHostToDevice, sc.s));
  UBG_HIP(hipStreamSynchronize(sc.s));
  return UBG_OK;
}

} // namespace
} // namespace ubg

using namespace ubg;

extern "C" {

UBG_API int ubg_flag_components(const float *flag, int width, int height, int32_t *labels,
                                int *ncomp) {
  if (!labels || !ncomp) return fail(UBG_E_INVALID, "null output");
  Scratch sc;
  UBG_TRY(setup(flag, width, height, sc));
  UBG_TRY(label_device(sc.fl, width, height, sc.lab, ncomp, sc.s));
  UBG_HIP(hipMemcpy(labels, sc.lab, (size_t)width * height * 4, hipMemcpyDeviceToHost));
  return UBG_OK;
}

UBG_API int ubg_remove_component_means(const float *flag, int width, int height, float *f,
                                       int *ncomp) {
  if (!f) return fail(UBG_E_INVALID, "null rhs");
  if (width < 1 || height < 1 || (long)width * height >= (1l << 31) / 2)
    return fail(UBG_E_INVALID, "grid size out of range");
  const size_t n = (size_t)width * height;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(f[i])) return fail(UBG_E_NONFINITE, "rhs has a non-finite value");
  Scratch sc;
  UBG_TRY(setup(flag, width, height, sc));
  int nc = 0;
  UBG_TRY(label_device(sc.fl, width, height, sc.lab, &nc, sc.s));
  if (ncomp) *ncomp = nc;
  if (nc == 0) return UBG_OK;
  char *mem = nullptr;
  const size_t bytes = align_up(n * 4) + align_up((size_t)nc * 8) + align_up((size_t)nc * 4) + 256;
  UBG_HIP(hipMalloc((void **)&mem, bytes));
  float *df = (float *)mem;
  auto *sums = (unsigned long long *)(mem + align_up(n * 4));
  int *counts = (int *)(mem + align_up(n * 4) + align_up((size_t)nc * 8));
  auto *mx = (unsigned *)(mem + align_up(n * 4) + align_up((size_t)nc * 8) + align_up((size_t)nc * 4));
  int rc = UBG_OK;
  unsigned hmx = 0;
  auto done = [&](int r) {
    (void)hipStreamSynchronize(sc.s);
    (void)hipFree(mem);
    return r;
  };
  if (hipMemcpyAsync(df, f, n * 4, hipMemcpyHostToDevice, sc.s) != hipSuccess)
    return done(fail(UBG_E_HIP, "rhs upload"));
  launch_fill_f32((float *)sums, (size_t)nc * 2, 0.0f, sc.s);
  launch_fill_f32((float *)counts, (size_t)nc, 0.0f, sc.s);
  launch_fill_f32((float *)mx, 1, 0.0f, sc.s);
  const int g = (int)((n + TPB - 1) / TPB);
  hipLaunchKernelGGL(k_absmax, dim3(g), dim3(TPB), 0, sc.s, df, sc.lab, (int)n, mx);
  if ((rc = check_launch()) != UBG_OK) return done(rc);
  if (hipMemcpyAsync(&hmx, mx, 4, hipMemcpyDeviceToHost, sc.s) != hipSuccess ||
      hipStreamSynchronize(sc.s) != hipSuccess)
    return done(fail(UBG_E_HIP, "max |f| copy"));
  const float amax = __builtin_bit_cast(float, hmx);
  if (amax == 0.0f) return done(UBG_OK); // every component sums to 0 already
  // 2^s scaling with max|f| 2^s n < 2^62: no component sum overflows
  int e = 0;
  (void)std::frexp((double)amax, &e); // amax < 2^e
  int en = 0;
  (void)std::frexp((double)n, &en);   // n < 2^en
  const int sh = 62 - e - en;
  const double scale = std::ldexp(1.0, sh), inv = std::ldexp(1.0, -sh);
  hipLaunchKernelGGL(k_comp_sums, dim3(g), dim3(TPB), 0, sc.s, df, sc.lab, (int)n, scale, sums,
                     counts);
  hipLaunchKernelGGL(k_sub_means, dim3(g), dim3(TPB), 0, sc.s, df, sc.lab, (int)n, sums, counts,
                     inv);
  if ((rc = check_launch()) != UBG_OK) return done(rc);
  if (hipMemcpyAsync(f, df, n * 4, hipMemcpyDeviceToHost, sc.s) != hipSuccess)
    return done(fail(UBG_E_HIP, "rhs download"));
  return done(UBG_OK);
}

} // extern "C"
