// mg_stream.hip -- row-pipelined multigrid level kernels for gfx950.
//
// One launch runs NS (0 or 1) rbgs sweeps (pressure_solver.cpp:49-89, each
// followed by setZeroGradientBC at level 0, :183-192) over a whole level and,
// depending on MODE, fuses what the V-cycle does next to them
// (solveLevel, pressure_solver.cpp:201-248):
//   MODE_RR     NS pre-smoothing sweeps + calculateResidualField + restrict
//               (:212-227): writes p and the coarse rhs (ec of the coarse
//               level is zeroed, :224)
//   MODE_PC     prolongate + correct (+ZG) (:231-236) + NS post-smoothing
//               sweeps (:238-243)
// so the sweep next to the grid transfer costs no pass of its own.  The
// pipeline is written for any NS, but every extra stage adds VALU work and
// halo lanes: on gfx950 a deeper pipeline is instruction-bound and loses to
// separate single-sweep launches, so only NS <= 1 is instantiated.
//
// Layout: a wave owns a strip of columns; lane l holds the horizontal cell
// pair (x0, x0+1), x0 even, so each row of a pair has one red and one black
// cell.  The wave streams down the rows.  Input row t (p from memory, or
// p + prolongated correction) enters sweep stage 1; stage k holds a window of
// four rows of its input and, when row j+2 arrives, updates the red cell of
// row j+1 (from old black values) and then the black cell of row j (from the
// new red values) -- exactly the two-pass order -- and emits row j, which is
// final for that sweep.  Stage k+1 consumes stage k's rows two rows behind.
// The residual is computed one row behind the last stage and restricted to a
// coarse row whenever three residual rows are complete.  The horizontal
// neighbour outside a pair comes from the adjacent lane by a DPP lane shift.
// Each stage invalidates one halo lane per side, so a wave owns the
// 64 - 2(NS+2) middle lanes; rows outside a wave's range are recomputed by
// the neighbouring wave from the same inputs.  Reads come from `src`, writes
// go to `dst` (ping-pong), so no wave sees another wave's new values, and the
// result is bit-identical to the sequential sweeps.
#include "stream_device.hpp"

namespace ubg {

namespace {

constexpr int TPB = 256;

// one wave per unit (strip x row block) of the level; body in stream_device.hpp.
// The smoothing updates take the mask's constants from an LDS table (8 KiB
// per workgroup) instead of decoding the byte: these passes are bound by
// VALU issue, and the decode is about 15 of the ~38 instructions of an
// update.
template <int NS, int MODE, bool ZG, bool AL>
__global__ void __launch_bounds__(TPB) k_mg_stream(const StreamArgs A) {
  __shared__ uint4 tab[kRuleEntries];
  fill_rule_table(tab);
  __syncthreads();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int unit = xcd_tile(blockIdx.x, gridDim.x) * (TPB / 64) + wid;
  if (unit >= A.nunits) return; // whole wave leaves together
  mg_stream_unit<NS, MODE, ZG, AL, true>(A, unit, tab);
}

template <int NS, int MODE>
void launch_ns(const StreamArgs &a, bool zg, bool al, int nblocks, hipStream_t s) {
  if (zg) {
    if (al) launch_k((k_mg_stream<NS, MODE, true, true>), dim3(nblocks), dim3(TPB), 0, s, a);
    else launch_k((k_mg_stream<NS, MODE, true, false>), dim3(nblocks), dim3(TPB), 0, s, a);
  } else {
    if (al) launch_k((k_mg_stream<NS, MODE, false, true>), dim3(nblocks), dim3(TPB), 0, s, a);
    else launch_k((k_mg_stream<NS, MODE, false, false>), dim3(nblocks), dim3(TPB), 0, s, a);
  }
}

template <int MODE>
void launch_mode(int ns, const StreamArgs &a, bool zg, bool al, int nb, hipStream_t s) {
  if (ns == 0) launch_ns<0, MODE>(a, zg, al, nb, s);
  else launch_ns<1, MODE>(a, zg, al, nb, s);
}

void launch_stream(int mode, int ns, Level &L, const Level &C, float h, float alpha, bool zg,
                   bool zero_coarse, hipStream_t s) {
  StreamArgs a{};
  a.src = L.p;
  a.dst = L.p_alt;
  a.f = L.f;
  a.m = L.mask;
  a.w = L.w;
  a.h = L.h;
  a.hh = h;
  a.alpha = alpha;
  a.ihsq = 1.0f / h / h;
  a.crc = C.f;
  a.cec = zero_coarse ? C.p : nullptr;
  a.cpin = C.p;
  a.cmask = C.mask;
  a.wc = C.w;
  a.hc = C.h;
  const int own = 64 - 2 * (ns + 2);
  a.nstrips = (L.w + 2 * own - 1) / (2 * own);
  // rows per wave: nstrips*h/4608 (about 4600 waves, 4.5 per SIMD), clamped
  // to [16, 64], even.  Each row block runs up to 5NS+4 extra pipeline rows
  // (halo rows, re-read from L2), so shorter blocks cost VALU work; longer
  // ones leave too few waves in flight to cover the serial row loop's load
  // latency.  Measured on gfx950 at 4096x2048: R=16 beats R=12 and R=14
  // (~0.5% of the step), R=18 (~1.5%), R=20 (~3%), R=2 (2.2x slower) and
  // R=32 (1.5x slower).
  int R = (int)((long)a.nstrips * L.h / 4608);
  R = R < 16 ? 16 : (R > 64 ? 64 : R);
  R &= ~1;
  a.R = R;
  const int nrb = (L.h + R - 1) / R;
  a.nunits = a.nstrips * nrb;
  const int nb = (a.nunits + (TPB / 64) - 1) / (TPB / 64);
  const bool al = (L.w % 2) == 0;
  if (mode == MODE_RR) launch_mode<MODE_RR>(ns, a, zg, al, nb, s);
  else launch_mode<MODE_PC>(ns, a, zg, al, nb, s);
  float *t = L.p;
  L.p = L.p_alt;
  L.p_alt = t;
}

} // namespace

void launch_presmooth_restrict(Level &L, Level &C, int ns, float h, float alpha, bool zg,
                               bool zero_coarse, hipStream_t s) {
  launch_stream(MODE_RR, ns, L, C, h, alpha, zg, zero_coarse, s);
}

void launch_prolong_postsmooth(Level &L, const Level &C, int ns, float h, float alpha, bool zg,
                               hipStream_t s) {
  launch_stream(MODE_PC, ns, L, C, h, alpha, zg, false, s);
}

} // namespace ubg
