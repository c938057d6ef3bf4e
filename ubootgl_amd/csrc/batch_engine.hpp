// batch_engine.hpp -- internal to libubg.so: the engine that advances many
// independent refinement solves together (batched and served solves).
//
// Layout: every domain owns, from a device arena, the float64 solution x
// (two buffers), its float32 rhs, the FP32 V-cycle buffers of the levels down
// to the first one solved in LDS (p twice, f) and its solve state; the flag
// and mask pyramid (a "geometry") can be shared by domains of the same shape.
// The engine keeps device tables of the current member domains and, per
// step, a row of per-domain buffer indices for every stage, so that one
// launch of a stage serves every domain whatever its level count.
#pragma once

#include "capi_common.hpp"
#include "coarse_device.hpp"

#include <map>
#include <memory>

namespace ubg {

constexpr int BMAXL = 16;        // levels per domain at most
constexpr int B_UPD_W = 256;     // update tile: 256 x 8 cells, as the
constexpr int B_UPD_H = 8;       //   single-domain refinement's blocks
constexpr int B_MIN_SIZE = 16;

// solve state of one domain (device; written by the norm kernel)
struct DState {
  double r0, rel, best, tol;
  double ref; // > 0: residuals are relative to ref, not to the first one
  int active, converged, iters, vcycles, stall, max_iter, xcur, k;
  int nonfinite, pad;
  double hist[UBG_IR_HISTORY];
};

// device table rows
struct BLev {
  float *p[2];
  float *f;
  const uint8_t *mask;
  int w, h;
  float hh, ihsq;
};
struct BDom {
  DState *st;
  double *x[2];
  const float *rhs;
  double *partial;
  int w, h, gx, gy, nlev, cfirst;
};
struct Tile {
  int d, x0, y0, i;
};

// first-fit device sub-allocator over one hipMalloc (256-byte granules)
class Arena {
public:
  int init(size_t bytes);
  void release();
  char *alloc(size_t bytes); // nullptr if it does not fit
  void free(char *p);
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  char *base() const { return base_; }

private:
  char *base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
  std::map<size_t, size_t> free_; // offset -> size
  std::map<size_t, size_t> live_; // offset -> size
};

// flag/mask pyramid of one domain shape (device, read-only once built)
struct BGeom {
  int W = 0, H = 0, nlev = 0, cfirst = 0;
  int w[BMAXL] = {}, h[BMAXL] = {};
  uint8_t *flag[BMAXL] = {}, *mask[BMAXL] = {};
  char *mem = nullptr;
  size_t bytes = 0;
  size_t lds_bytes = 0;
  int refs = 0;
  bool ring_fluid = false; // every outer-ring cell is fluid
};

struct BDomain {
  BGeom *g = nullptr;
  float h = 0.0f;
  float hh[BMAXL] = {};
  float *p[BMAXL][2] = {};
  float *f[BMAXL] = {};
  double *x[2] = {};
  float *rhs = nullptr;
  double *partial = nullptr;
  DState *st = nullptr;
  int gx = 0, gy = 0;
  char *mem = nullptr;
  size_t bytes = 0;
  // host copy of the buffer indices (which p of each level, which x is current)
  int cur[BMAXL] = {};
  int xcur = 0;
};

struct Tuning {
  int tx = 64, ty = 8, blocks = 0; // smoother tile (sweep stages), workgroups per launch
  int prow = 0;                    // rows of the fused-pass tiles (0: chosen per level)
};

// Level sizes, first LDS level and byte counts of a W x H domain.
struct BShape {
  int nlev = 0, cfirst = 0;
  int w[BMAXL] = {}, h[BMAXL] = {};
  size_t lds_bytes = 0, geom_bytes = 0, dom_bytes = 0;
};
int batch_shape(int W, int H, BShape *s);

class BatchEngine {
public:
  ~BatchEngine();
  int init(hipStream_t s, const ubg_params &prm, int vcycles_per_step);
  hipStream_t stream() const { return stream_; }
  const ubg_params &params() const { return prm_; }
  int vcycles() const { return k_; }
  void set_vcycles(int k) {
    if (k != k_) drop_graphs();
    k_ = k;
  }

  // geometry / domain lifetime (memory from `arena`)
  // need_ring: the outer ring must be fluid (refinement solves hold it as
  // Dirichlet data); projection domains take any flag
  int geom_create(Arena &arena, int W, int H, const float *flag_host, BGeom **out,
                  bool need_ring = true);
  void geom_release(Arena &arena, BGeom *g);
  // The domain calls take an optional stream (default: the engine's own).
  // A caller that keeps a step in flight on the engine's stream may use
  // another stream for domains that are not members of that step.
  int domain_create(Arena &arena, BGeom *g, float h, BDomain **out, hipStream_t s = nullptr);
  void domain_release(Arena &arena, BDomain *d);
  int domain_set_rhs(BDomain *d, const float *rhs_host);
  int domain_set_guess(BDomain *d, const double *x_host);
  // reset a domain's solve state: active, tol/max_iter, fresh (next update
  // computes r0 and does not add e)
  int domain_start(BDomain *d, double tol, int max_iter, double ref = 0.0,
                   hipStream_t s = nullptr);
  int domain_read_state(BDomain *d, DState *out, hipStream_t s = nullptr);
  int domain_get_solution(BDomain *d, double *x_host, hipStream_t s = nullptr);
  // level-0 pressure and rhs of the plain V-cycles (projection solves); kind
  // is a hipMemcpyKind (host or device arrays); p == nullptr clears p
  int domain_set_level0(BDomain *d, const float *p, const float *f, hipMemcpyKind kind,
                        hipStream_t s = nullptr);
  int domain_get_pressure(BDomain *d, float *p, hipMemcpyKind kind, hipStream_t s = nullptr);

  // the member set of the following steps: `running` domains take V-cycles
  // and an accumulating update; `fresh` ones only their first update
  int set_members(const std::vector<BDomain *> &running, const std::vector<BDomain *> &fresh);
  // forget the member set (a member domain was replaced: a new domain may
  // sit at the old one's address with other buffers), so the next
  // set_members rebuilds the tables
  void invalidate_members() {
    run_.clear();
    fresh_.clear();
    all_.clear();
    drop_graphs();
  }
  // enqueue `nsteps` refinement steps of the members; the active-domain
  // count lands in *count_host (pinned) when the stream reaches it
  int enqueue_steps(int nsteps);
  // the first update (r0, no V-cycles) of every member, as if all were fresh
  int enqueue_first_update();
  // eager steps, or one captured graph per index state, replayed
  int run_steps_graph(int nsteps);
  // ncycles plain FP32 V-cycles (MG::solve, pressure_solver.cpp:250-257) of
  // every running member on its level-0 pressure and rhs; zg: level 0 takes
  // setZeroGradientBC after every sweep and after the correction (the
  // simulation's projection).  Level 0 then runs as row-streamed passes for
  // every domain (they carry the boundary copy); needs 1 <= nu_pre, nu_post
  // <= 3.
  int enqueue_vcycles(int ncycles, bool zg);
  // level 0 as row-streamed passes for every member domain (projection
  // solves) or by the size rule (refinement solves); rebuilds the tile maps
  // when it changes
  int set_level0_streamed(bool on);
  int active_count() const { return *count_host_; }
  // members whose solve was still active at the last completed step
  std::vector<BDomain *> active_members() const;
  size_t members() const { return all_.size(); }
  int sync();

  void set_tuning(const Tuning &t);
  const Tuning &tuning() const { return tune_; }
  // time each compiled smoother tile shape and block count over the member
  // levels; keeps the fastest
  int autotune(double *best_ms);

private:
  struct StageRec {
    int kind, level;
    size_t off;
  };
  int build_tables();
  int build_tiles();
  void gen_steps(int nsteps, std::vector<StageRec> &st, std::vector<uint8_t> &idx,
                 bool all_fresh = false);
  void gen_vcycle_tiled(std::vector<StageRec> &st, std::vector<uint8_t> &idx);
  int launch_stages(const std::vector<StageRec> &st, const uint8_t *d_idx);
  int upload_idx(const std::vector<uint8_t> &idx, uint8_t **d_idx);
  int launch_sweeps_only(int reps);
  std::string state_key() const;

  hipStream_t stream_ = nullptr;
  ubg_params prm_{};
  int k_ = 4;
  Tuning tune_{};
  std::vector<BDomain *> run_, fresh_, all_;
  int lmax_ = 0;
  // device tables
  BLev *d_lev_ = nullptr;
  BDom *d_dom_ = nullptr;
  CoarseArgs *d_cargs_ = nullptr;
  size_t tab_cap_ = 0;
  size_t coarse_lds_ = 0;
  // tile maps: per level sweep / rr / pc (running members), update (all):
  // prefix of tile counts over slots and tiles per row, in one int array
  std::vector<int> h_off_;
  int *d_off_ = nullptr;
  size_t off_cap_ = 0;
  struct Span {
    size_t off = 0; // into the int array
    int n = 0;      // tiles
    int nd = 0;     // domain slots covered
    int oty = 0;    // fused passes: tile rows
    int tpb = 512;  //   and workgroup size
  };
  Span sw_[BMAXL], rr_[BMAXL], pc_[BMAXL], upd_;
  Span tp_[BMAXL], tq_[BMAXL]; // fused level passes (pre, post)
  Span ss_[BMAXL], st_[BMAXL]; // row-streamed levels: sweeps, pre / post passes
  bool stream_on_[BMAXL] = {};  // level l has enough streamed units over the members
  bool force_stream0_ = false;  // level 0 streamed for every domain (projection solves)
  bool zg0_ = false;            // the stages being launched apply ZG at level 0
  // levels run as fused tile passes (smoothing counts with a tiled pass)
  bool tiled() const { return tile_pass_ok(prm_.nu_pre) && tile_pass_ok(prm_.nu_post); }
  // level l of a domain of this shape may run as row-streamed passes (wide
  // and large levels) ...
  bool stream_eligible(const BGeom &G, int l) const;
  // ... and does, when the member domains' streamed units at that level are
  // enough to fill the GPU (else the LDS tiles cost less per cell)
  bool streamed(const BGeom &G, int l) const {
    return l < BMAXL && stream_on_[l] && stream_eligible(G, l);
  }
  int launch_passes_only(int reps);
  // stage index rows
  uint8_t *d_idx_ = nullptr;
  size_t idx_cap_ = 0;
  uint8_t *h_idx_pinned_ = nullptr;
  size_t h_idx_cap_ = 0;
  int *d_count_ = nullptr;
  int *count_host_ = nullptr;
  uint8_t *d_active_ = nullptr; // per slot, written with the count
  uint8_t *h_active_ = nullptr; // pinned copy
  // graphs by index state (batched solves with a fixed member set)
  struct Graph {
    hipGraphExec_t exec = nullptr;
    uint8_t *d_idx = nullptr;
  };
  std::map<std::string, Graph> graphs_;
  void drop_graphs();
};

} // namespace ubg
