// noalloc.hip -- ubg_mg_solve_device_noalloc: MG::solve (pressure_solver.hpp:
// 59-62) on caller-owned device fields, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).  Every
// buffer beyond the caller's p, f and flag -- the level-0 second solution
// buffer, flags and neighbour masks, and every coarse level -- is carved out
// of one file-scope device array, `scratch`.  The call makes no hipMalloc,
// hipFree, memset or synchronisation; it only enqueues kernels (and one
// device-to-device copy) on the caller's stream.
//
// Each GPU holds its own copy of `scratch` until the process ends, at its own
// address: the address is looked up with hipGetSymbolAddress once per device
// (first call with that device current) and only that address is passed to
// kernels.  All calls on one GPU share the array and every call clears the
// part it uses first, so a call's result does not depend on earlier calls;
// calls that overlap on one GPU (other streams, other host threads) race on
// it and must be ordered by the caller.
#include "capi_common.hpp"

namespace ubg {
namespace {

__device__ __align__(256) char scratch[UBG_NOALLOC_SCRATCH_BYTES];

// flags as 0/1 bytes (any nonzero byte is fluid)
__global__ void k_flag_01(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] != 0 ? 1 : 0;
}

struct DeviceScratch {
  char *base = nullptr; // device address of `scratch` on this device
  ubg_mg mg;            // host-side level table, reused by every call
  std::mutex mu;        // serialises the host side of calls on this device
};
constexpr int kMaxDevices = 64;
DeviceScratch g_dev[kMaxDevices];
std::mutex g_init_mu;

// the device's scratch address and level table, set up on its first call
int device_scratch(DeviceScratch **out) {
  int dev = 0;
  UBG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(UBG_E_INVALID, "device index out of range");
  std::lock_guard<std::mutex> g(g_init_mu);
  DeviceScratch &d = g_dev[dev];
  if (!d.base) {
    void *ptr = nullptr;
    UBG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    d.base = (char *)ptr;
    d.mg.lv.reserve(32); // level sizes halve: at most 31 levels for int sizes
    default_params(&d.mg.prm);
  }
  *out = &d;
  return UBG_OK;
}

} // namespace
} // namespace ubg

extern "C" UBG_API int ubg_mg_solve_device_noalloc(float *p, const float *f, const uint8_t *flag,
                                                   int width, int height, float h,
                                                   int zero_gradient_bc, int ncycles,
                                                   const ubg_params *prm, void *stream) {
  if (!p || !f || !flag) return fail(UBG_E_INVALID, "null field pointer");
  if (width < 8 || height < 8)
    return fail(UBG_E_DOMAIN_TOO_SMALL, "MG needs at least 8 cells in each dimension");
  if (ncycles < 0) return fail(UBG_E_INVALID, "negative cycle count");
  if (prm) UBG_TRY(check_params(prm));
  // level sizes without a heap allocation: at most 31 halvings of an int
  int ws_a[32], hs_a[32], nl = 0;
  for (int cw = width, ch = height; cw > 3 && ch > 3 && nl < 32; cw /= 2, ch /= 2) {
    ws_a[nl] = cw;
    hs_a[nl] = ch;
    nl++;
  }
  size_t need = 0;
  for (int l = 0; l < nl; l++) {
    const size_t n = (size_t)ws_a[l] * hs_a[l];
    need += l == 0 ? align_up(n * 4) + 2 * align_up(n) : 3 * align_up(n * 4) + 2 * align_up(n);
  }
  if (need > (size_t)UBG_NOALLOC_SCRATCH_BYTES)
    return fail(UBG_E_BUDGET, "a " + std::to_string(width) + " x " + std::to_string(height) +
                                  " solve needs " + std::to_string(need) +
                                  " bytes of scratch; the static array holds " +
                                  std::to_string((size_t)UBG_NOALLOC_SCRATCH_BYTES));
  DeviceScratch *d = nullptr;
  UBG_TRY(device_scratch(&d));
  std::lock_guard<std::mutex> g(d->mu);
  ubg_mg &m = d->mg;
  hipStream_t s = (hipStream_t)stream;
  m.W = width;
  m.H = height;
  m.levels = nl;
  m.stream = s;
  if (prm) m.prm = *prm;
  else default_params(&m.prm);
  m.lv.resize(nl); // within the reserved capacity: no reallocation
  char *base = d->base;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base + off;
    off += align_up(bytes);
    return q;
  };
  for (int l = 0; l < nl; l++) {
    Level &L = m.lv[l];
    L = Level{};
    L.w = ws_a[l];
    L.h = hs_a[l];
    const size_t n = (size_t)L.w * L.h;
    if (l == 0) {
      L.p = p;
      L.f = const_cast<float *>(f); // the V-cycle never writes the level-0 rhs
      L.p_alt = (float *)take(n * 4);
    } else {
      L.p = (float *)take(n * 4);
      L.p_alt = (float *)take(n * 4);
      L.f = (float *)take(n * 4);
    }
    L.flag = (uint8_t *)take(n);
    L.mask = (uint8_t *)take(n);
  }
  // a fresh hierarchy: zero everything this call uses (a fill kernel, no
  // memset node under capture), then the flags and the coarse classification
  launch_fill_f32((float *)base, off / 4, 0.0f, s);
  const size_t n0 = (size_t)width * height;
  hipLaunchKernelGGL(k_flag_01, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, s, flag,
                     m.lv[0].flag, n0);
  m.classify();
  for (int c = 0; c < ncycles; c++) m.solve_level(0, h, zero_gradient_bc != 0);
  // the sweeps ping-pong between p and p_alt: the result may sit in scratch
  if (m.lv[0].p != p)
    UBG_HIP(hipMemcpyAsync(p, m.lv[0].p, n0 * 4, hipMemcpyDeviceToDevice, s));
  m.lv[0].p = p;
  return check_launch();
}
