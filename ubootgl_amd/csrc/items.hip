// items.hip -- floating items coupled to the fluid: Simulation::
// advectFloatingItems and advectFloatingItemsSimple
// (advect_floating_items.cpp:16-274), one thread per item.
//
// An item reads the fluid (bilinearVel of the front vx/vy, the flag samplers,
// for the simple kind also vx/vy/p at its cell) and scatters its reaction
// into vx_accum / vy_accum, which the next step's applyAccumulatedVelocity
// consumes.  Items do not read the accumulation fields, so they are
// independent of each other and run in parallel; the scatter uses float
// atomics, so only the summation order into vx_accum / vy_accum differs from
// the reference's item-by-item loop.  Operation order per item follows the
// reference (glm's vec2 arithmetic, left to right; double where the
// reference's literals promote to double).  sin / cos are the Cephes
// single-precision method (S. L. Moshier, Cephes Math Library: Cody-Waite
// reduction by pi/4 in three parts, degree-7/8 polynomials), written out so
// the result is the same on any IEEE machine.
#include "ubg_internal.hpp"
#include "../../include/ubg_c.h"

namespace ubg {
namespace {

struct V2 {
  float x, y;
};
__device__ __forceinline__ V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ V2 operator*(V2 a, V2 b) { return {a.x * b.x, a.y * b.y}; }
__device__ __forceinline__ V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ V2 operator*(float s, V2 a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ V2 operator/(V2 a, float s) { return {a.x / s, a.y / s}; }
__device__ __forceinline__ float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float length(V2 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V2 normalize(V2 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ V2 reflect(V2 i, V2 n) { return i - n * dot(n, i) * 2.0f; }
__device__ __forceinline__ V2 vabs(V2 a) { return {fabsf(a.x), fabsf(a.y)}; }
// glm::min / glm::max / glm::clamp: comparison based
__device__ __forceinline__ float gmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

// Cephes sinf / cosf
constexpr float FOPI = 1.27323954473516f;
constexpr float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
__device__ __forceinline__ float sin_poly(float x, float z) {
  return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
__device__ __forceinline__ float cos_poly(float z) {
  return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z -
         0.5f * z + 1.0f;
}
__device__ __forceinline__ void reduce(float ax, int &j, float &r) {
  j = (int)(FOPI * ax);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  r = ((ax - y * DP1) - y * DP2) - y * DP3;
}
__device__ __forceinline__ float c_sin(float x) {
  bool neg = x < 0.0f;
  int j;
  float r;
  reduce(fabsf(x), j, r);
  if (j > 3) {
    neg = !neg;
    j -= 4;
  }
  const float z = r * r;
  const float v = (j == 1 || j == 2) ? cos_poly(z) : sin_poly(r, z);
  return neg ? -v : v;
}
__device__ __forceinline__ float c_cos(float x) {
  int j;
  float r;
  reduce(fabsf(x), j, r);
  bool neg = false;
  if (j > 3) {
    j -= 4;
    neg = !neg;
  }
  if (j > 1) neg = !neg;
  const float z = r * r;
  const float v = (j == 1 || j == 2) ? sin_poly(r, z) : cos_poly(z);
  return neg ? -v : v;
}
// glm::rotate(vec2, angle)
__device__ __forceinline__ V2 rotate(V2 v, float a) {
  const float c = c_cos(a), s = c_sin(a);
  return {v.x * c - v.y * s, v.x * s + v.y * c};
}

struct Fluid {
  const float *vx, *vy, *p; // front vx, vy; pressure
  const uint8_t *flag;
  float *ax, *ay; // vx_accum, vy_accum
  int W, H;
  float h, pwidth;
};

// scalar bilinearSample (interpolators.hpp:11-27)
__device__ __forceinline__ float bilinear(const float *g, int gw, int gh, float cx, float cy) {
  cx = gclamp(cx, 0.0f, gw - 1.1f);
  cy = gclamp(cy, 0.0f, gh - 1.1f);
  const int ix = (int)cx, iy = (int)cy;
  const float sx = cx - floorf(cx), sy = cy - floorf(cy);
  const float *r0 = g + (size_t)iy * gw + ix;
  const float v1 = r0[0], v2 = r0[1], v3 = r0[gw], v4 = r0[gw + 1];
  const float vm1 = v1 + (v2 - v1) * sx;
  const float vm2 = v3 + (v4 - v3) * sx;
  return vm1 + (vm2 - vm1) * sy;
}
// Simulation::bilinearVel (advect_floating_items.cpp:11-14), grid space
__device__ __forceinline__ V2 bilinear_vel(const Fluid &F, V2 c) {
  return {bilinear(F.vx, F.W - 1, F.H, c.x - 0.5f, c.y), bilinear(F.vy, F.W, F.H - 1, c.x, c.y - 0.5f)};
}
// bilinearScatter (interpolators.hpp:29-40)
__device__ __forceinline__ void scatter(float *g, int gw, int gh, V2 c, float v) {
  c.x = gclamp(c.x, 0.0f, gw - 1.1f);
  c.y = gclamp(c.y, 0.0f, gh - 1.1f);
  const int ix = (int)c.x, iy = (int)c.y;
  const float sx = c.x - floorf(c.x), sy = c.y - floorf(c.y);
  const float ax = 1.0f - sx, ay = 1.0f - sy;
  float *r0 = g + (size_t)iy * gw + ix;
  atomicAdd(r0 + gw + 1, sx * sy * v);
  atomicAdd(r0 + gw, ax * sy * v);
  atomicAdd(r0 + 1, sx * ay * v);
  atomicAdd(r0, ax * ay * v);
}
// psampleFlagLinear (simulation.cpp:398-411); glm::mix(a, b, t) = a(1-t) + bt
__device__ __forceinline__ float flag_linear(const Fluid &F, V2 pc) {
  const float cell = F.pwidth / (float)F.W;
  const float cx = pc.x / cell - 0.5f, cy = pc.y / cell - 0.5f;
  const int ix = min(max((int)cx, 0), F.W - 2), iy = min(max((int)cy, 0), F.H - 2);
  const float sx = cx - floorf(cx), sy = cy - floorf(cy);
  auto fl = [&](int x, int y) { return (float)F.flag[(size_t)y * F.W + x]; };
  const float p01 = fl(ix, iy + 1), p11 = fl(ix + 1, iy + 1), p00 = fl(ix, iy), p10 = fl(ix + 1, iy);
  const float a = p00 * (1.0f - sx) + p10 * sx;
  const float b = p01 * (1.0f - sx) + p11 * sx;
  return a * (1.0f - sy) + b * sy;
}
// psampleFlagNormal (simulation.cpp:413-420)
__device__ __forceinline__ V2 flag_normal(const Fluid &F, V2 pc) {
  const float h = F.h;
  const float p01 = flag_linear(F, {pc.x - h, pc.y + h});
  const float p11 = flag_linear(F, {pc.x + h, pc.y + h});
  const float p00 = flag_linear(F, {pc.x - h, pc.y - h});
  const float p10 = flag_linear(F, {pc.x + h, pc.y - h});
  return {p11 + p10 - p01 - p00, p01 + p11 - p00 - p10};
}

struct Item {
  V2 size, pos;
  float rotation, mass;
  V2 vel, force;
  float ang_vel, ang_force;
  int bump;
};

__device__ __forceinline__ Item load_item(const ubg_item &s) {
  return Item{{s.size[0], s.size[1]}, {s.pos[0], s.pos[1]}, s.rotation, s.mass,
              {s.vel[0], s.vel[1]},   {s.force[0], s.force[1]}, s.ang_vel, s.ang_force,
              s.bump_count};
}
__device__ __forceinline__ void store_item(ubg_item &d, const Item &it) {
  d.size[0] = it.size.x;
  d.size[1] = it.size.y;
  d.pos[0] = it.pos.x;
  d.pos[1] = it.pos.y;
  d.rotation = it.rotation;
  d.mass = it.mass;
  d.vel[0] = it.vel.x;
  d.vel[1] = it.vel.y;
  d.force[0] = it.force.x;
  d.force[1] = it.force.y;
  d.ang_vel = it.ang_vel;
  d.ang_force = it.ang_force;
  d.bump_count = it.bump;
}

// sub-steps of a frame (advect_floating_items.cpp:24-26, :173-175):
// min(15, max(1, max|v| * dt / h * 2.5)) with the 2.5 in double; abs is taken
// as the float overload
__device__ __forceinline__ int iteration_steps(V2 vel, float dt, float h) {
  const double s = (double)(gmax(fabsf(vel.x), fabsf(vel.y)) * dt / h) * 2.5;
  return (int)fmin(15.0, fmax(1.0, s));
}
// rotation update (:35-36, :184-185): fmod in double with 2 pi
__device__ __forceinline__ float advance_rotation(float rot, float sub_dt, float ang_vel) {
  const double two_pi = 2.0 * 3.14159265358979323846;
  return (float)fmod((double)(rot + sub_dt * ang_vel) + two_pi, two_pi);
}
__device__ __forceinline__ bool outside(const Fluid &F, V2 pos) {
  const float cell = F.pwidth / (float)F.W;
  const V2 g = pos / cell;
  // written so that a non-finite position counts as outside: the reference
  // would index the grid at (int)NaN for such an item
  return !(g.x < F.W - 2 && g.x > 1.0f && g.y < F.H - 2 && g.y > 1.0f);
}

// Simulation::advectFloatingItems (advect_floating_items.cpp:16-138)
__global__ void k_items(ubg_item *__restrict__ items, int n, Fluid F, float game_dt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  Item it = load_item(items[t]);
  const float h = F.h;
  const int steps = iteration_steps(it.vel, game_dt, h);
  const float sub_dt = game_dt / (float)steps;
  const V2 samples[5] = {{1.0f, 1.0f}, {-1.0f, 1.0f}, {1.0f, -1.0f}, {-1.0f, -1.0f}, {0.0f, 0.0f}};
  const V2 surf[4] = {{-1.0f, 0.0f}, {1.0f, 0.0f}, {0.0f, -1.0f}, {0.0f, 1.0f}};
  for (int step = 0; step < steps; step++) {
    const V2 before = it.pos;
    it.pos = it.pos + sub_dt * it.vel;
    it.rotation = advance_rotation(it.rotation, sub_dt, it.ang_vel);
    if (outside(F, it.pos)) continue;
    it.force = it.force + V2{0.0f, -0.5f} * it.mass;
    // terrain collision (:47-74)
    for (int k = 0; k < 5; k++) {
      const V2 sp = rotate(samples[k] * 0.5f * it.size, it.rotation);
      if (flag_linear(F, it.pos + sp) < 0.5f) {
        V2 nrm = normalize(flag_normal(F, 0.5f * (before + it.pos) + sp));
        if (flag_linear(F, before + sp) > 0.5f) it.pos = before;
        if (length(nrm) > 0.0f) {
          nrm = nrm / length(nrm);
          if (dot(it.vel, nrm) < 0.0f) it.vel = reflect(it.vel, nrm) * 0.7f;
          it.vel = it.vel + nrm * 0.07f * flag_linear(F, it.pos + sp);
          if (dot(it.force, nrm) < 0.0f) it.force = it.force + 1.1f * dot(it.force, nrm) * nrm;
        }
        it.bump++;
      }
    }
    const V2 external = V2{0.0f, -0.5f} * it.mass + it.force;
    V2 central = {0.0f, 0.0f};
    const float ang_force = it.ang_force;
    const float side[4] = {it.size.y, it.size.y, it.size.x, it.size.x};
    // surface pressure points (:82-120)
    for (int i = 0; i < 4; i++) {
      const int nsp = (int)fmaxf(2.0f, side[i] / h);
      for (int m = 0; m < nsp; m++) {
        const V2 sp = surf[i] + vabs(V2{surf[i].y, surf[i].x}) * (1.0f - m * 2.0f / (float)(nsp - 1));
        V2 tsp = sp * it.size * 0.5f;
        tsp = rotate(tsp, it.rotation);
        tsp = tsp + it.pos;
        const V2 dv = bilinear_vel(F, tsp / h) -
                      (it.vel + 3.141f * rotate(sp * it.size * 0.5f, it.rotation + 0.5f * 3.141f) *
                                    it.ang_vel);
        const V2 nrm = rotate(normalize(sp), it.rotation);
        const V2 force = nrm * gmin(0.0f, dot(dv, rotate(surf[i], it.rotation)));
        central = central + force * 400000.0f * (0.003f + side[i]) * side[i] / (float)nsp;
        if (tsp.x / h < 1.0f || tsp.x / h > (F.W - 1) - 2.0f || tsp.y / h < 1.0f ||
            tsp.y / h > F.H - 2.0f)
          continue;
        const V2 dvec = force * (0.003f + side[i]) * side[i] / (float)nsp * sub_dt * 18000000.0f;
        scatter(F.ax, F.W - 1, F.H, tsp / h - V2{0.5f, 0.0f}, -dvec.x);
        scatter(F.ay, F.W, F.H - 1, tsp / h - V2{0.0f, 0.5f}, -dvec.y);
      }
    }
    central = central + 1000.0f * (it.size.x + it.size.y) * (bilinear_vel(F, it.pos / h) - it.vel);
    it.vel = it.vel + sub_dt * (external + central) / it.mass;
    const float ang_mass = it.size.x * it.size.y * it.mass * (1.0f / 12.0f);
    it.ang_vel += sub_dt * ang_force / ang_mass;
    it.ang_vel = (float)((double)it.ang_vel * 0.98);
  }
  it.ang_force = 0.0f;
  it.force = {0.0f, 0.0f};
  store_item(items[t], it);
}

// Simulation::advectFloatingItemsSimple (advect_floating_items.cpp:140-274).
// The 100 position bins (:144-151) are snapshots taken before any item
// moves: `snap` holds the input positions; an item visits the items of its
// bin in input order.
__global__ void k_items_simple(ubg_item *__restrict__ items, const float *__restrict__ snap, int n,
                               Fluid F, float game_dt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  Item it = load_item(items[t]);
  const float h = F.h;
  const int bin = ((int)(it.pos.x * 100)) % 100;
  V2 rep = {0.0f, 0.0f};
  int contacts = 0;
  for (int o = 0; o < n; o++) {
    const V2 p2 = {snap[2 * o], snap[2 * o + 1]};
    if (((int)(p2.x * 100)) % 100 != bin) continue;
    const V2 dis = it.pos - p2;
    if (dis.x * dis.x + dis.y * dis.y < (it.size.x * it.size.y * 0.4f)) {
      const float len = gmax(0.1f * it.size.x, length(dis));
      rep = rep + 0.0001f * (dis / len / len);
      contacts++;
    }
  }
  rep = rep / (float)max(1, contacts);
  it.force = it.force + rep * 10.0f;
  const int steps = iteration_steps(it.vel, game_dt, h);
  const float sub_dt = game_dt / (float)steps;
  for (int step = 0; step < steps; step++) {
    it.pos = it.pos + sub_dt * it.vel;
    it.rotation = advance_rotation(it.rotation, sub_dt, it.ang_vel);
    const float cell = F.pwidth / (float)F.W;
    const V2 grid_pos = it.pos / cell;
    if (outside(F, it.pos)) continue;
    // terrain collision (:196-220)
    if (flag_linear(F, it.pos) < 1.0f) {
      V2 nrm = flag_normal(F, it.pos);
      if (length(nrm) > 0.0f) {
        nrm = nrm / length(nrm);
        if (dot(it.vel, nrm) < 0.0) it.vel = reflect(it.vel, nrm) * 0.7f;
        if (dot(it.force, nrm) < 0.0f) it.force = reflect(it.force, nrm) * 0.7f;
        const V2 lat = rotate(nrm, 0.5f * (float)3.14159265358979323846);
        const float lat_vel = dot(lat, it.vel);
        const float rot_vel = it.ang_vel * (it.size.x + it.size.y) * 0.5f;
        const float lat_diff = lat_vel - rot_vel;
        it.force = it.force + lat_diff * lat * 1.0f;
      } else {
        it.vel = {0.0f, 0.0f};
      }
      it.bump++;
    }
    V2 external = V2{0.0f, -0.5f} * it.mass + it.force;
    const V2 dvel = bilinear_vel(F, it.pos / h) - it.vel;
    external = external + 2000.0f * (it.size.x + it.size.y) * dvel;
    // (the reference's y test reads "pos.y / h < 1"; kept)
    if (it.pos.x / h > 1.0f && it.pos.x / h < (F.W - 1) - 2.0f && it.pos.y / h < 1.0f &&
        it.pos.y / h < F.H - 2.0f) {
      const V2 dvec = dvel * it.size.x * it.size.y;
      scatter(F.ax, F.W - 1, F.H, it.pos / h - V2{0.5f, 0.0f}, -dvec.x);
      scatter(F.ay, F.W, F.H - 1, it.pos / h - V2{0.0f, 0.5f}, -dvec.y);
    }
    it.vel = it.vel + sub_dt * external / it.mass;
    const int gx = (int)grid_pos.x, gy = (int)grid_pos.y;
    auto VX = [&](int x, int y) { return F.vx[(size_t)y * (F.W - 1) + x]; };
    auto VY = [&](int x, int y) { return F.vy[(size_t)y * F.W + x]; };
    auto P = [&](int x, int y) { return F.p[(size_t)y * F.W + x]; };
    const float fluid_ang = -((VX(gx, gy) - VX(gx, gy - 1)) - (VY(gx, gy) - VY(gx - 1, gy))) / h / 2;
    const float ang_mass = it.size.x * it.size.y * it.mass * (1.0f / 12.0f);
    const float k = gmin(1.0f, sub_dt / ang_mass * 0.0005f * (it.size.x + it.size.y) / 4.0f);
    it.ang_vel += k * (fluid_ang - it.ang_vel);
    it.ang_vel += sub_dt / ang_mass * it.ang_force;
    atomicAdd(&F.ax[(size_t)gy * (F.W - 1) + gx], -(it.ang_vel * k * 0.01f * P(gx, gy)));
    atomicAdd(&F.ax[(size_t)(gy - 1) * (F.W - 1) + gx], it.ang_vel * k * 0.01f * P(gx, gy - 1));
    atomicAdd(&F.ay[(size_t)gy * F.W + gx], it.ang_vel * k * 0.01f * P(gx, gy));
    atomicAdd(&F.ay[(size_t)gy * F.W + gx - 1], -(it.ang_vel * k * 0.01f * P(gx - 1, gy)));
  }
  it.ang_force = 0.0f;
  it.force = {0.0f, 0.0f};
  store_item(items[t], it);
}

} // namespace

void launch_items(ubg_item *items, const float *snap, int n, const float *vx, const float *vy,
                  const float *p, const uint8_t *flag, float *ax, float *ay, int W, int H, float h,
                  float pwidth, float game_dt, bool simple, hipStream_t s) {
  if (n <= 0) return;
  const Fluid F{vx, vy, p, flag, ax, ay, W, H, h, pwidth};
  if (simple)
    hipLaunchKernelGGL(k_items_simple, dim3((n + 63) / 64), dim3(64), 0, s, items, snap, n, F,
                       game_dt);
  else
    hipLaunchKernelGGL(k_items, dim3((n + 63) / 64), dim3(64), 0, s, items, n, F, game_dt);
}

} // namespace ubg
