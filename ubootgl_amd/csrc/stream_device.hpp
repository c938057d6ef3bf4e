// stream_device.hpp -- row-pipelined multigrid level passes for gfx950, as
// per-wave-unit device functions shared by the single-domain kernels
// (mg_fast.hip, mg_stream.hip) and the batched solver (batch.hip).
//
// mg_stream_unit: a launch over every unit of a level runs NS (0 or 1) rbgs
// sweeps (pressure_solver.cpp:49-89, each followed by setZeroGradientBC at
// level 0, :183-192) over the whole level and, depending on MODE, fuses what
// the V-cycle does next to them (solveLevel, pressure_solver.cpp:201-248):
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
#pragma once

#include "mg_device.hpp"

namespace ubg {

enum { MODE_RR = 1, MODE_PC = 2 };
constexpr int SW2 = 124; // output columns per wave of the stand-alone streaming sweep

struct StreamArgs {
  const float *src; // level p, input
  float *dst;       // level p, output
  const float *f;
  const uint8_t *m;
  int w, h;
  float hh, alpha, ihsq;
  float *crc;            // MODE_RR: coarse rhs (rcs[l+1])
  float *cec;            // MODE_RR: coarse solution, zeroed (ecs[l+1]); null: left
                         // as it is (the coarse level's next pass does not read it)
  const float *cpin;     // MODE_PC: coarse solution (ecs[l+1])
  const uint8_t *cmask;  // MODE_PC: coarse flag mask
  int wc, hc;
  int nstrips, R, nunits; // R: rows per block, even (mg_stream_unit needs it)
};

struct P2 {
  float a, b;
};
struct M2 {
  unsigned a, b;
};

// The smoothing constants of every mask byte, for passes that look them up
// in LDS instead of decoding the byte at each update: entry 2 mk holds the
// four neighbour selectors (0 / -1 for the bits W, E, S, N), entry 2 mk + 1
// RN(1/s) and -s for the neighbour count s, the quotient's select (0 when
// s == 0, else -1) and F, as bits.  Filled by the whole workgroup.
constexpr int kRuleEntries = 512;
__device__ __forceinline__ void fill_rule_table(uint4 *tab) {
  for (unsigned mk = threadIdx.x; mk < 256u; mk += blockDim.x) {
    const unsigned s = mk >> 5;
    const float pow2 = __int_as_float((int)((127u - (s >> 1)) << 23));
    const float inv = (s == 3u) ? (1.0f / 3.0f) : pow2;
    tab[2 * mk] = make_uint4(0u - ((mk >> 1) & 1u), 0u - ((mk >> 2) & 1u), 0u - ((mk >> 3) & 1u),
                             0u - ((mk >> 4) & 1u));
    tab[2 * mk + 1] = make_uint4(__float_as_uint(inv), __float_as_uint(-(float)s),
                                 s == 0u ? 0u : ~0u, __float_as_uint((float)(mk & MK_SELF)));
  }
}
// smooth_fhh with the mask's constants from the table: the same operations
// in the same order (keep_if, div_count, the zero-count case, F), so the
// same value.
__device__ __forceinline__ float smooth_tab(float pc, float pw, float pe, float ps, float pn,
                                            float fhh, const uint4 *tab, unsigned mk,
                                            float alpha) {
  const uint4 nb = tab[2 * mk], dc = tab[2 * mk + 1];
  float val = 0.0f;
  val += __uint_as_float(__float_as_uint(pw) & nb.x);
  val += __uint_as_float(__float_as_uint(pe) & nb.y);
  val += __uint_as_float(__float_as_uint(ps) & nb.z);
  val += __uint_as_float(__float_as_uint(pn) & nb.w);
  val += fhh;
  const float inv = __uint_as_float(dc.x);
  const float q = val * inv;
  const float r = fmaf(__uint_as_float(dc.y), q, val);
  const float q2 = fmaf(r, inv, q);
  val = __uint_as_float(__float_as_uint(q2) & dc.z);
  return __uint_as_float(dc.w) * (alpha * val + (1.0f - alpha) * pc);
}

// TAB: the smoothing constants come from `tab` (fill_rule_table, in LDS).
template <int NS, int MODE, bool ZG, bool AL, bool TAB = false>
__device__ __forceinline__ void mg_stream_unit(const StreamArgs &A, int unit,
                                               const uint4 *tab = nullptr) {
  constexpr int HALO = NS + 2;
  constexpr int OWN = 64 - 2 * HALO;
  // stage k (0-based) consumes row t-3k at iteration t (stage k-1's output
  // of the previous iteration, a pipeline register, so the stages of one
  // iteration are independent) and updates rows t-3k-2 and t-3k-1; the
  // residual consumes the last stage's output of the previous iteration.
  constexpr int Q = 3 * NS + 2; // f/mask rows t, t-1, ..., t-Q+1
  constexpr int NW = NS > 0 ? NS : 1;
  // rows produced by a sweep or by the correction get setZeroGradientBC; raw
  // input rows from memory do not (the first sweep reads p as it is)
  constexpr bool OUT_ZG = ZG && (NS > 0 || MODE == MODE_PC);
  const int lane = threadIdx.x & 63;
  const int w = A.w, h = A.h;
  const int s = unit % A.nstrips, r = unit / A.nstrips;
  const int x0 = 2 * OWN * s - 2 * HALO + 2 * lane, x1 = x0 + 1;
  const bool in0 = x0 >= 0 && x0 < w, in1 = x1 >= 0 && x1 < w;
  const bool int0 = x0 >= 1 && x0 <= w - 2, int1 = x1 >= 1 && x1 <= w - 2;
  const bool own = lane >= HALO && lane < HALO + OWN;
  const int FA = A.R * r, FB = min(FA + A.R - 1, h - 1); // owned fine rows
  const int LNS = FA - (MODE == MODE_RR ? 2 : 0);
  const int UNS = FB + (MODE == MODE_RR ? 1 : 0);
  // first input row: a sweep's final row j depends on rows j-2..j+2 of its
  // input.  Under ZG, row h-1 is a copy of row h-2 (of the last sweep's
  // output, or of the corrected input), so a block that owns only row h-1
  // (odd h) needs row h-2 as well: two more rows (L0 stays even)
  const int L0 =
      LNS - 2 * NS - ((ZG && (NS > 0 || MODE == MODE_PC) && FA == h - 1) ? 2 : 0);
  const int TEND = UNS + 3 * NS + 1; // last iteration

  // clamped (always valid) load addresses; clamped values only ever feed
  // cells that are not updated or not kept
  const int c0 = min(max(x0, 0), w - 1), c1 = min(max(x1, 0), w - 1);
  const int cp = min(max(x0, 0), w - 2);
  auto rowoff = [&](int yy) { return (size_t)min(max(yy, 0), h - 1) * w; };
  auto ld2 = [&](const float *base, int yy) -> P2 {
    const float *row = base + rowoff(yy);
    if (AL) {
      float2 v = *(const float2 *)(row + cp);
      return P2{v.x, v.y};
    }
    return P2{row[c0], row[c1]};
  };
  // a row's mask pair as loaded (AL: both bytes in one 16-bit value); it is
  // split into the two cells' masks only when the row enters the queue, a
  // trip after its load was issued (splitting at the load would wait for it)
  auto ldm2 = [&](int yy) -> M2 {
    const uint8_t *row = A.m + rowoff(yy);
    if (AL) return M2{(unsigned)*(const uint16_t *)(row + cp), 0u};
    return M2{row[c0], row[c1]};
  };
  auto relax_a = [&](P2 v, float dn, float up, float fv, unsigned mk) {
    float na;
    if constexpr (TAB)
      na = smooth_tab(v.a, from_left(v.b), v.b, dn, up, fv * A.hh * A.hh, tab, mk, A.alpha);
    else
      na = smooth(v.a, from_left(v.b), v.b, dn, up, fv, mk, A.hh, A.alpha);
    return int0 ? na : v.a;
  };
  auto relax_b = [&](P2 v, float dn, float up, float fv, unsigned mk) {
    float nb;
    if constexpr (TAB)
      nb = smooth_tab(v.b, v.a, from_right(v.a), dn, up, fv * A.hh * A.hh, tab, mk, A.alpha);
    else
      nb = smooth(v.b, v.a, from_right(v.a), dn, up, fv, mk, A.hh, A.alpha);
    return int1 ? nb : v.b;
  };
  // side-column setZeroGradientBC of an interior row
  auto zg_cols = [&](P2 v) {
    float left = from_left(v.b);
    P2 o = v;
    if (x0 == 0) o.a = v.b;
    if (x0 == w - 1) o.a = left;
    if (x1 == w - 1) o.b = v.a;
    return o;
  };
  // rows 0 / h-1 under ZG copy the interior cells of rows 1 / h-2
  auto zg_row = [&](P2 edge, P2 inner) {
    return P2{int0 ? inner.a : edge.a, int1 ? inner.b : edge.b};
  };

  P2 W[NW][4];
#pragma unroll
  for (int k = 0; k < NW; k++)
#pragma unroll
    for (int i = 0; i < 4; i++) W[k][i] = P2{0.0f, 0.0f};
  P2 FQ[Q];
  unsigned MQa[Q], MQb[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    FQ[q] = P2{0.0f, 0.0f};
    MQa[q] = MQb[q] = 0u;
  }
  P2 prev_in{0.0f, 0.0f};                      // MODE_PC: previous corrected row
  P2 out1{0.0f, 0.0f}, out2{0.0f, 0.0f};       // last two emitted final rows
  P2 res1{0.0f, 0.0f}, res2{0.0f, 0.0f};       // last two residual rows

  // each stage's output row of the previous iteration: stage k emitted row
  // (t - 1) - 3k - 2 there, so stage k's input at iteration t is row t - 3k
  // (before the pipeline is full these are zero rows; no update or store
  // that uses them is kept)
  P2 E[NW];
#pragma unroll
  for (int k = 0; k < NW; k++) E[k] = P2{0.0f, 0.0f};

  // Two rows per trip: L0 is even (R is even), so each half of a trip
  // has a constant row parity and the register queues shift by renaming
  // inside the trip instead of by moves every row.  The input rows (p, f,
  // mask) of a trip are loaded one trip ahead: a trip first issues the loads
  // of the next trip's two rows and hands them over at its end, so the wait
  // for them falls a whole trip after they were issued.
  P2 px[2] = {ld2(A.src, L0), ld2(A.src, L0 + 1)};
  P2 pf[2] = {ld2(A.f, L0), ld2(A.f, L0 + 1)};
  M2 pm[2] = {ldm2(L0), ldm2(L0 + 1)};
  // MODE_PC: the coarse solution and coarse flags under this lane's pair
  // (columns X, X1) of one coarse row.  A trip's two fine rows t2, t2+1
  // prolongate from coarse rows Y = t2/2 and Y+1 (clamped into the level;
  // rows outside it only feed cells that are not corrected), and the next
  // trip shares row Y+1, so a trip loads one coarse row, for the next trip,
  // and the rows rotate through cr0, cr1.
  struct CRow {
    float e, e1;
    int g, g1;
  };
  const int cX = max(0, min(x0 >> 1, A.wc - 1)), cX1 = min(cX + 1, A.wc - 1);
  auto ldc = [&](int Y) -> CRow {
    if (MODE != MODE_PC) return CRow{0.0f, 0.0f, 0, 0};
    const int Yc = min(max(Y, 0), A.hc - 1);
    const float *e = A.cpin + (size_t)Yc * A.wc;
    const uint8_t *g = A.cmask + (size_t)Yc * A.wc;
    return CRow{e[cX], e[cX1], g[cX] & 1, g[cX1] & 1};
  };
  CRow cr0 = ldc(L0 >> 1), cr1 = ldc((L0 >> 1) + 1);
  for (int t2 = L0; t2 <= TEND; t2 += 2) {
    const P2 lx0 = ld2(A.src, t2 + 2), lx1 = ld2(A.src, t2 + 3);
    const P2 lf0 = ld2(A.f, t2 + 2), lf1 = ld2(A.f, t2 + 3);
    const M2 lm0 = ldm2(t2 + 2), lm1 = ldm2(t2 + 3);
    const CRow crn = ldc((t2 >> 1) + 2);
#pragma unroll
    for (int par = 0; par < 2; par++) {
      const int t = t2 + par;
      if (par == 1 && t > TEND) break;
      // ---- input row t and the f / mask queue (FQ[q] = row t - q)
#pragma unroll
      for (int q = Q - 1; q > 0; q--) {
        FQ[q] = FQ[q - 1];
        MQa[q] = MQa[q - 1];
        MQb[q] = MQb[q - 1];
      }
      FQ[0] = pf[par];
      MQa[0] = AL ? pm[par].a & 0xffu : pm[par].a;
      MQb[0] = AL ? pm[par].a >> 8 : pm[par].b;
      P2 x = px[par];
      if (MODE == MODE_PC) {
        // prolongate + correct (pressure_solver.cpp:134-181) on interior cells
        if (t >= 1 && t <= h - 2) {
          // coarse row Y = t/2 is cr0, row min(Y + 1, hc - 1) is cr1
          const float Fa = (MQa[0] & MK_SELF) ? 1.0f : 0.0f, Fb = (MQb[0] & MK_SELF) ? 1.0f : 0.0f;
          float ea = 0.0f, eb = 0.0f;
          if (par == 0) { // t even
            ea = cr0.e * Fa;
            if (x1 <= w - 3) eb = Fb * (cr0.e + cr0.e1) / eps_count(cr0.g + cr0.g1);
          } else if (t <= h - 3) {
            ea = Fa * (cr0.e + cr1.e) / eps_count(cr0.g + cr1.g);
            if (x1 <= w - 3)
              eb = Fb * (cr0.e + cr1.e1 + cr0.e1 + cr1.e) / eps_count(cr0.g + cr1.g1 + cr0.g1 + cr1.g);
          }
          x.a = int0 ? x.a + ea : x.a;
          x.b = int1 ? x.b + eb : x.b;
          if (ZG) x = zg_cols(x);
        } else if (ZG && t == h - 1) {
          x = zg_row(x, prev_in);
        }
        prev_in = x;
      }
      // ---- sweep stages; stage k's input is stage k-1's output of the
      // previous iteration (stage 0: the new input row)
      P2 nx[NW];
#pragma unroll
      for (int k = 0; k < NS; k++) {
        const P2 in = k == 0 ? x : E[k - 1];
        const int jin = t - 3 * k;
        W[k][0] = W[k][1];
        W[k][1] = W[k][2];
        W[k][2] = W[k][3];
        W[k][3] = in;
        if (ZG && (k > 0 || MODE == MODE_PC) && jin == 1)
          W[k][2] = zg_row(W[k][2], in); // row 0 of a produced row set takes row 1's interior
        const int j = jin - 2;
        const int qj = 3 * k + 2, qj1 = 3 * k + 1; // queue slots of rows j and j+1
        const int jpar = (par + k) & 1;             // parity of j = t - 3k - 2
        if (j + 1 >= 1 && j + 1 <= h - 2) {         // red cells of row j+1
          if (jpar == 0) W[k][2].a = relax_a(W[k][2], W[k][1].a, W[k][3].a, FQ[qj1].a, MQa[qj1]);
          else W[k][2].b = relax_b(W[k][2], W[k][1].b, W[k][3].b, FQ[qj1].b, MQb[qj1]);
        }
        if (j >= 1 && j <= h - 2) { // black cells of row j: final
          if (jpar == 0) W[k][1].a = relax_a(W[k][1], W[k][0].a, W[k][2].a, FQ[qj].a, MQa[qj]);
          else W[k][1].b = relax_b(W[k][1], W[k][0].b, W[k][2].b, FQ[qj].b, MQb[qj]);
        }
        P2 e = W[k][1];
        if (ZG && j >= 1 && j <= h - 2) e = zg_cols(e);
        if (ZG && j == h - 1) e = zg_row(e, W[k][0]);
        nx[k] = e;
      }
      // ---- the final row of this iteration: last stage's output of the
      // previous iteration (or the input itself when there is no sweep)
      const P2 fin = NS > 0 ? E[NW - 1] : x;
      const int jo = NS > 0 ? t - 3 * NS : t;
#pragma unroll
      for (int k = 0; k < NS; k++) E[k] = nx[k];
      x = fin;

      // ---- x is final row jo of the last sweep
      if (OUT_ZG && jo == 1) out1 = zg_row(out1, x); // row 0 takes row 1's interior
      if (own) {
        // write owned rows; row 0 is written once row 1 is known
        if (jo >= FA && jo <= FB && jo >= 1) {
          float *row = A.dst + (size_t)jo * w;
          if (AL) { // w even: an owned in-domain pair is whole and 8-byte aligned
            if (in0) *(float2 *)(row + x0) = make_float2(x.a, x.b);
          } else {
            if (in0) row[x0] = x.a;
            if (in1) row[x1] = x.b;
          }
        }
        if (jo == 1 && FA == 0) {
          if (in0) A.dst[x0] = out1.a;
          if (in1) A.dst[x1] = out1.b;
        }
      }
      if (MODE == MODE_RR) {
        // residual of row i = jo - 1 (rows i-1, i, i+1 = out2, out1, x)
        const int i = jo - 1;
        const int qi = NS > 0 ? 3 * NS + 1 : 1;
        P2 rr{0.0f, 0.0f};
        if (i >= 1 && i <= h - 2) {
          rr.a = residual_of(out1.a, from_left(out1.b), out1.b, out2.a, x.a, FQ[qi].a, MQa[qi], A.ihsq);
          rr.b = residual_of(out1.b, out1.a, from_right(out1.a), out2.b, x.b, FQ[qi].b, MQb[qi], A.ihsq);
        }
        // restrict (pressure_solver.cpp:118-132) when rows 2Y-1, 2Y, 2Y+1 are done
        if (((par + NS + 1) & 1) && i >= 1) { // i = t - 3NS - 1 odd
          const int Y = (i - 1) >> 1, X = x0 >> 1;
          const float l2 = from_left(res2.b), l1 = from_left(res1.b), l0 = from_left(rr.b);
          float v = 0.0f;
          v = l2 * 1.0f + res2.a * 2.0f + res2.b * 1.0f;
          v += l1 * 2.0f + res1.a * 4.0f + res1.b * 2.0f;
          v += l0 * 1.0f + rr.a * 2.0f + rr.b * 1.0f;
          if (own && Y >= 1 && Y <= A.hc - 2 && X >= 1 && X <= A.wc - 2 && 2 * Y >= FA &&
              2 * Y <= FA + A.R - 1) {
            size_t ci = (size_t)Y * A.wc + X;
            A.crc[ci] = v * (1.0f / 16.0f);
            if (A.cec) A.cec[ci] = 0.0f;
          }
        }
        res2 = res1;
        res1 = rr;
      }
      out2 = out1;
      out1 = x;
    }
    px[0] = lx0;
    px[1] = lx1;
    pf[0] = lf0;
    pf[1] = lf1;
    pm[0] = lm0;
    pm[1] = lm1;
    cr0 = cr1;
    cr1 = crn;
  }
}

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
__device__ __forceinline__ void rbgs_pair_unit(const float *__restrict__ src,
                                               float *__restrict__ dst,
                                               const float *__restrict__ f,
                                               const uint8_t *__restrict__ m, int w, int h,
                                               float hh, float alpha, int nstrips, int R,
                                               int unit) {
  const int lane = threadIdx.x & 63;
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
  // f pair and mask pair of a row; the mask pair as loaded (AL: both bytes
  // in one 16-bit value), split by `split` one row later, so the split does
  // not wait for the load issued in the same iteration
  auto ldfm = [&](int yy, float &fa, float &fb, M2 &mq) {
    const size_t o = rowoff(yy);
    if (AL) {
      float2 v = *(const float2 *)(f + o + cp);
      fa = v.x;
      fb = v.y;
      mq = M2{(unsigned)*(const uint16_t *)(m + o + cp), 0u};
    } else {
      fa = f[o + c0];
      fb = f[o + c1];
      mq = M2{m[o + c0], m[o + c1]};
    }
  };
  auto split = [&](M2 q, unsigned &ma, unsigned &mb) {
    ma = AL ? q.a & 0xffu : q.a;
    mb = AL ? q.a >> 8 : q.b;
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
  unsigned mm1a, mm1b, m0a, m0b, m1a, m1b;
  M2 mq;
  ldp(ya - 2, pm2a, pm2b);
  ldp(ya - 1, pm1a, pm1b);
  ldp(ya, p0a, p0b);
  ldp(ya + 1, p1a, p1b);
  ldp(ya + 2, p2a, p2b);
  ldp(ya + 3, p3a, p3b);
  ldfm(ya - 1, fm1a, fm1b, mq);
  split(mq, mm1a, mm1b);
  ldfm(ya, f0a, f0b, mq);
  split(mq, m0a, m0b);
  ldfm(ya + 1, f1a, f1b, mq);
  split(mq, m1a, m1b);
  M2 m2; // row y + 2 (raw)
  ldfm(ya + 2, f2a, f2b, m2);
  if (ya - 1 >= 1) RELAX(pm1a, pm1b, pm2a, pm2b, p0a, p0b, fm1a, fm1b, mm1a, mm1b, ya - 1, 1);
  RELAX(p0a, p0b, pm1a, pm1b, p1a, p1b, f0a, f0b, m0a, m0b, ya, 1);
  for (int y = ya; y <= yb; y++) {
    float p4a, p4b, f3a, f3b; // two rows of lead
    M2 m3;
    ldp(y + 4, p4a, p4b);
    ldfm(y + 3, f3a, f3b, m3);
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
    split(m2, m1a, m1b);
    m2 = m3;
  }
#undef RELAX
}

} // namespace ubg
