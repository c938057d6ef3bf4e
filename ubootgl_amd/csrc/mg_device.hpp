// mg_device.hpp -- device helpers shared by the multigrid kernels: the
// smoothing point update, the exact division by the neighbour count, lane
// shifts and the XCD-aware block order.
#pragma once

#include "ubg_internal.hpp"

namespace ubg {

// v / s for the neighbour count s in {0..4} without the IEEE division
// sequence: q = v * RN(1/s), then one FMA correction q + (v - s*q) * RN(1/s).
// By Markstein's theorem this is the correctly rounded quotient; checked
// exhaustively against v / (float)s for all 2^32 float v (the only
// difference is the sign of a zero quotient of v = -0, which the smoother
// never produces: its sum starts at +0).  s = 0 gives +0, as the reference's
// "if (sum == 0) val = 0" does.
__device__ __forceinline__ float div_count(float v, unsigned s) {
  // RN(1/s) without branches: 2^-(s>>1) for s = 1, 2, 4 from the exponent
  // bits, one select for s = 3
  const float pow2 = __int_as_float((int)((127u - (s >> 1)) << 23));
  const float inv = (s == 3u) ? (1.0f / 3.0f) : pow2;
  const float q = v * inv;
  const float r = fmaf(-(float)s, q, v);
  const float q2 = fmaf(r, inv, q);
  return (s == 0u) ? 0.0f : q2;
}

// select v if bit K of the mask is set, else +0 (bitwise, no compare): the
// bit sign-extended to 0 / -1 by one v_bfe_i32 (written out because the
// compiler turns the extract back into an and / compare / select triple)
template <int K>
__device__ __forceinline__ float keep_if(unsigned mk, float v) {
  int sel;
  asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(sel) : "v"(mk), "i"(K));
  return __int_as_float(__float_as_int(v) & sel);
}

// v if bit K of the mask is set, else c (a bitwise select on the bit
// sign-extended to 0 / -1)
template <int K>
__device__ __forceinline__ float pick_if(unsigned mk, float v, float c) {
  int sel;
  asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(sel) : "v"(mk), "i"(K));
  return __int_as_float((__float_as_int(v) & sel) | (__float_as_int(c) & ~sel));
}

// smoothingKernel (pressure_solver.cpp:10-24) from neighbour values and mask;
// the neighbour-flag count is stored in mask bits 5-7.
// ... with the source term fhh = (f * h) * h already formed
__device__ __forceinline__ float smooth_fhh(float pc, float pw, float pe, float ps, float pn,
                                            float fhh, unsigned mk, float alpha) {
  float val = 0.0f;
  val += keep_if<1>(mk, pw);
  val += keep_if<2>(mk, pe);
  val += keep_if<3>(mk, ps);
  val += keep_if<4>(mk, pn);
  val += fhh;
  val = div_count(val, mk >> 5);
  const float F = (float)(mk & MK_SELF); // 0 or 1 (MK_SELF is bit 0)
  return F * (alpha * val + (1.0f - alpha) * pc);
}
__device__ __forceinline__ float smooth(float pc, float pw, float pe, float ps, float pn,
                                        float fv, unsigned mk, float h, float alpha) {
  return smooth_fhh(pc, pw, pe, ps, pn, fv * h * h, mk, alpha);
}

// calculateResidualField for one cell (pressure_solver.cpp:100-108):
// r = F*(f + (sum_n(F_n p_n + (1-F_n) p) - 4p)/h^2), neighbours in the order
// x-1, x+1, y-1, y+1.
// The selects are bitwise on the mask bit sign-extended to 0 / -1 (two
// instructions each instead of an and / compare / select triple); the
// values are those of the conditional expressions.
__device__ __forceinline__ float residual_of(float pc, float pw, float pe, float ps, float pn,
                                             float fv, unsigned mk, float ihsq) {
  float val = 0.0f;
  val += pick_if<1>(mk, pw, pc); // (mk & MK_W) ? pw : pc
  val += pick_if<2>(mk, pe, pc);
  val += pick_if<3>(mk, ps, pc);
  val += pick_if<4>(mk, pn, pc);
  val += -4.0f * pc;
  val *= ihsq;
  return keep_if<0>(mk, fv + val); // (mk & MK_SELF) ? fv + val : +0
}

// Whole-wave lane shifts by one (DPP wave_shr:1 / wave_shl:1, a VALU move):
// lane i receives lane i-1 (from_left) or lane i+1 (from_right); the edge lane
// keeps its own value.  Lane layout verified on gfx950.
__device__ __forceinline__ float from_left(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138,
                                                    0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130,
                                                    0xf, 0xf, false));
}

// Blocks are dispatched round-robin over the 8 XCDs; hand each XCD a
// contiguous range of work units so halo re-reads hit its own L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  int q = nb >> 3, rem = nb & 7, k = b & 7;
  return k * q + min(k, rem) + (b >> 3);
}

// flag count + the double literal 0.0001, stored as float (pressure_solver.cpp:149)
__device__ __forceinline__ float eps_count(int s) {
  return s == 0 ? (float)(0.0 + 0.0001)
                : (s == 1 ? (float)(1.0 + 0.0001)
                          : (s == 2 ? (float)(2.0 + 0.0001)
                                    : (s == 3 ? (float)(3.0 + 0.0001) : (float)(4.0 + 0.0001))));
}

} // namespace ubg
