// ubg_internal.hpp -- device data layout and kernel launchers of libubg.so.
//
// HBM layout (one GPU):
//   * every field is a dense row-major fp32 array, index y*w + x, exactly the
//     host layout of Single2DGrid (db2dgrid.hpp:12-50), so host copies are
//     plain memcpys and flat-offset quirks of the reference carry over;
//   * flags are one byte per cell: `flag` holds the 0/1 value, `mask` holds the
//     cell's own flag in bit 0, its four neighbours' flags in bits 1-4 and
//     their count in bits 5-7, so the smoother, the residual and the
//     prolongation read one byte per cell;
//   * an MG hierarchy is one allocation carved into levels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <cstdint>

struct ubg_item;

namespace ubg {

// Kernel timing: an event pair set here is attached to the next launch made
// through launch_k, on the dispatch itself (hipExtLaunchKernelGGL), so the
// elapsed time is the kernel's own begin-to-end -- the interval a kernel
// trace reports.  Used from the thread that owns the handle being timed.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
inline LaunchEvents g_launch_ev;

template <typename F, typename... Args>
inline void launch_k(F kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s,
                     Args... args) {
  if (g_launch_ev.start) {
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, g_launch_ev.start, g_launch_ev.stop, 0,
                          args...);
    g_launch_ev = LaunchEvents{};
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
  }
}

enum : uint8_t {
  MK_SELF = 1u, // F(x, y)
  MK_W = 2u,    // F(x-1, y)
  MK_E = 4u,    // F(x+1, y)
  MK_S = 8u,    // F(x, y-1)
  MK_N = 16u,   // F(x, y+1)
};

// Boundary-condition kinds of Simulation::BC (simulation.hpp:69)
enum BCKind : int { BC_INFLOW = 0, BC_OUTFLOW = 1, BC_OUTFLOW_ZERO_PRESSURE = 2, BC_NOSLIP = 3 };

struct Level {
  int w = 0, h = 0;
  float *p = nullptr;      // solution: level 0 pressure, coarse levels ecs
  float *p_alt = nullptr;  // second buffer: sweeps read p, write p_alt, then swap
  float *f = nullptr;      // right-hand side: level 0 f, coarse levels rcs
  float *r = nullptr;      // residual scratch (stage-wise residual only)
  uint8_t *flag = nullptr; // 0/1 flag (level 0 input, coarse flagcs)
  uint8_t *mask = nullptr; // own + neighbour flag bits
};

struct BCs {
  int west = BC_INFLOW, east = BC_OUTFLOW_ZERO_PRESSURE, north = BC_NOSLIP, south = BC_NOSLIP;
};

// Staggered velocity state of one domain: vx is (W-1) x H, vy is W x (H-1).
struct VelView {
  int W, H;
  float *vxf, *vxb, *vyf, *vyb;
};

// ---- generic -------------------------------------------------------------
void launch_fill_f32(float *dst, size_t n, float v, hipStream_t s);
void launch_fill_f64(double *dst, size_t n, double v, hipStream_t s);
void launch_f32_from_u8(const uint8_t *src, float *dst, size_t n, hipStream_t s);
void launch_sumsq(const float *src, size_t n, double *acc, hipStream_t s);

// ---- multigrid (pressure_solver.hpp / pressure_solver.cpp) ---------------
void launch_build_mask(const uint8_t *flag, uint8_t *mask, int w, int h, hipStream_t s);
void launch_coarsen(const uint8_t *ff, int wf, uint8_t *fc, int wc, int hc, hipStream_t s);
// one red+black sweep fused in one launch (streaming waves, ping-pong: reads
// L.p, writes L.p_alt, then swaps the two pointers); zero_gradient applies
// setZeroGradientBC to the result.
void launch_rbgs_fused(Level &L, float h, float alpha, bool zero_gradient, hipStream_t s);
void launch_zero_gradient(float *p, int w, int h, hipStream_t s);
// Row-pipelined level passes (ping-pong: read L.p, write L.p_alt, swap):
//   ns (0 or 1) sweeps, then residual + restriction into C
//   (C.f = restricted residual, C.p = 0 if zero_coarse, else C.p is left as
//   it is for a coarse pass that does not read it);
void launch_presmooth_restrict(Level &L, Level &C, int ns, float h, float alpha, bool zg,
                               bool zero_coarse, hipStream_t s);
//   prolongation of C.p + correction (+ZG), then ns (0 or 1) sweeps.
void launch_prolong_postsmooth(Level &L, const Level &C, int ns, float h, float alpha, bool zg,
                               hipStream_t s);
// LDS-tiled level passes (mg_tile.hip) for levels below the streaming size:
// nu sweeps + residual + restriction, and prolongation + correction + nu
// sweeps, one launch each; both swap L.p / L.p_alt.  No zero-gradient BC.
// zero_p: L.p is zero on entry (every level below the finest: the
// restriction above zeroed it), so the pre pass does not read it;
// zero_coarse: the pre pass zeroes C.p (else C.p is left as it is, for a
// coarse pass that does not read it).
bool tile_pass_ok(int nu);
void launch_tile_pre(Level &L, Level &C, int nu, float h, float alpha, bool zero_p,
                     bool zero_coarse, hipStream_t s);
void launch_tile_post(Level &L, const Level &C, int nu, float h, float alpha, hipStream_t s);
void launch_residual(const Level &L, float h, hipStream_t s);
void launch_residual_restrict(const Level &F, const Level &C, float h, hipStream_t s);
void launch_prolong_correct(const Level &F, const Level &C, hipStream_t s);
// single-workgroup V-cycle of every level from `first` down, all in LDS
bool coarse_fits_lds(const Level *lv, int first, int levels);
void launch_coarse_vcycle(const Level *lv, int first, int levels, float h, int nu_pre,
                          int nu_post, int nu_coarse, float alpha, hipStream_t s);

// ---- fluid step (simulation.cpp) -----------------------------------------
void launch_set_vbcs(const VelView &v, const BCs &bc, hipStream_t s);
void launch_set_pbc(float *p, int W, int H, const BCs &bc, hipStream_t s);
// setVBCs of vy alone (vx untouched)
void launch_set_vbcs_vy(const VelView &v, const BCs &bc, hipStream_t s);
// fused passes of the step (sim_kernels.hip); each gives the same values as
// the reference's sequence of loops it replaces
// Face factors from the 0/1 flag F (one uint16 per cell (x, y)): vx face
// bits 0-4 = F(x)F(x+1), F(x+1)F(x+2), F(x)F(x-1), F(x,y+1)F(x-1,y+1),
// F(x,y-1)F(x-1,y-1); vy face bits 8-12 = F(y)F(y+1), F(y)F(y-1),
// F(y+1)F(y+2), F(x+1,y)F(x+1,y+1), F(x-1,y)F(x-1,y+1) (zero where a face
// or a neighbour is missing).  A 0/1 float product of flags equals the AND.
void launch_face_masks(const uint8_t *flag, uint16_t *fmask, int W, int H, hipStream_t s);
void launch_diffuse_both(const VelView &v, const uint16_t *fmask, float a, const BCs &bc,
                         hipStream_t s);
void launch_vbcs_divergence(const VelView &v, const BCs &bc, float *f, float ih, hipStream_t s);
void launch_correct_tail(const VelView &v, float *p, const uint16_t *fmask, float ih,
                         const BCs &bc, float *xcur, float *ycur, hipStream_t s);
void launch_accumulate(const VelView &v, float *vx_acc, float *vy_acc, hipStream_t s);
void launch_advect(const VelView &v, const uint8_t *flag, float hdt, float fdt, hipStream_t s);

// One simulation in a batched step of several (ubg_sims_step): the buffers
// and constants its per-cell passes read at one stage, and the first row of
// its interior in the step's row numbering (row0 = sum of H - 2 over the
// simulations before it).  A launch covers every simulation's interior
// rows, one block row per interior row, found by binary search on row0.
struct SimRec {
  float *vxf, *vxb, *vyf, *vyb; // front / back roles at this stage
  const uint16_t *fmask;
  const uint8_t *flag;
  float *f, *p, *xcur, *ycur; // xcur / ycur may be null
  int W, H;
  int west, east, north, south; // boundary-condition kinds
  float a;                      // diffuse: dt mu (W - 1) / pwidth
  float hdt, fdt, ih;           // advect: dt / 2h, dt / h; 1 / h
  int row0;
};
// the stages of Simulation::step after applyAccumulatedVelocity for n
// simulations: one diffuse iteration, advect, setVBCs + divergence,
// setPBC + correction + setVBCs + saveCurrent; rows = total interior rows
void launch_sims_diffuse(const SimRec *tab, int n, int rows, int max_w, hipStream_t s);
void launch_sims_advect(const SimRec *tab, int n, int rows, int max_w, bool buf, hipStream_t s);
void launch_sims_vbcs_divergence(const SimRec *tab, int n, int rows, int max_w, hipStream_t s);
void launch_sims_correct_tail(const SimRec *tab, int n, int rows, int max_w, hipStream_t s);
void launch_divergence(const VelView &v, float *f, float ih, hipStream_t s);
// f[cell[i]] = val[i] for n distinct cells
void launch_stamp_sinks(float *f, const int *cell, const float *val, int n, hipStream_t s);
void launch_correct_velocity(const VelView &v, const float *p, const uint8_t *flag, float ih,
                             hipStream_t s);
void launch_bilinear_vel(const VelView &v, const float *c, float *out, int n, hipStream_t s);
// the simulation part of a map shift: fields shifted one cell left, new
// flags with setGrids zeroing, inlet velocity (front and back vx)
void launch_shift_map(const VelView &v, float *p, const uint8_t *new_flag, uint8_t *flag,
                      hipStream_t s);
// terrain flags from an RGBA image's red channel: full resolution (iw x ih)
// and the simulation flag ((iw/scale) x (ih/scale))
void launch_terrain_flags(const uint8_t *rgba, int iw, int ih, int scale, float *full, float *sim,
                          hipStream_t s);
// floating items (one thread each), scatter into ax / ay; snap: 2n input
// positions (bins of the simple kind)
void launch_items(ubg_item *items, const float *snap, int n, const float *vx, const float *vy,
                  const float *p, const uint8_t *flag, float *ax, float *ay, int W, int H, float h,
                  float pwidth, float game_dt, bool simple, hipStream_t s);
// (2W-1) x (2H-1) collocated (vx, vy) pairs and magnitudes from vx, vy
void launch_collocate(const float *vx, const float *vy, float *vxy, float *mag, int W, int H,
                      hipStream_t s);
// psampleFlagNearest (mode 0), psampleFlagLinear (1), psampleFlagNormal (2)
// at n physical positions (x, y pairs); out holds n values, 2n for mode 2
void launch_sample_flag(const uint8_t *flag, int W, int H, float pwidth, float h, int mode,
                        const float *pc, float *out, int n, hipStream_t s);
// getDT's maximum of max(vx, vy) over interior cells: one partial per block
void launch_max_vel(const VelView &v, float *part, int nblocks, hipStream_t s);
// centerP: float64 partial sums of the interior cells, then p -= v there
void launch_sum_interior(const float *p, int W, int H, double *part, int nblocks, hipStream_t s);
void launch_sub_interior(float *p, int W, int H, float v, hipStream_t s);
// one lexicographic Gauss-Seidel sweep (gs), one launch per anti-diagonal
void launch_gs(float *p, const float *f, const uint8_t *mask, int W, int H, float h, float alpha,
               hipStream_t s);
void launch_tracers(const VelView &v, const uint8_t *flag, float h, float pwidth, float dt,
                    uint32_t seed, float *xy, float *age, const int *perm, int n,
                    hipStream_t s);

} // namespace ubg
