// pressure_solver.hpp -- drop-in MG (pressure_solver.hpp:3-76 of the
// reference) backed by libubg.so.  Same class name, constructors and
// signatures; the hierarchy, coarse flags and V-cycle live on the GPU.
//
//   MG()                    default (no hierarchy)
//   MG(w, h)                hierarchy with all-fluid coarse flags      (hpp:16-31)
//   MG(flag)                + updateFields(flag)                       (hpp:32)
//   updateFields(flag)      coarse classification on the device        (hpp:34-57)
//   solve(p, f, flag, h, zg) one V-cycle; p, f, flag are host grids;
//                           p is uploaded, solved and downloaded       (hpp:59-62)
// Free functions gs(), rbgs() and calculateResidualField() (hpp:3-11) run on
// the GPU through a per-thread cached MG of the grid's size.
//
// Errors: the reference has none; here a failing call throws
// ubg_error (status code + ubg_last_error() text).  There is no class named
// PressureSolver in the reference; the alias is provided for callers that
// use that name.
#pragma once

#include "../ubg_c.h"
#include "db2dgrid.hpp"

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>

class ubg_error : public std::runtime_error {
public:
  ubg_error(int code, const std::string &what) : std::runtime_error(what), code(code) {}
  int code;
};

inline void ubg_check(int rc) {
  if (rc != UBG_OK) throw ubg_error(rc, std::string("libubg: ") + ubg_last_error());
}

class MG {
public:
  MG() {}
  MG(int width, int height) : width(width), height(height) {
    ubg_check(ubg_mg_create(width, height, &h_));
    ubg_check(ubg_mg_num_levels(h_, &levels));
  }
  explicit MG(Single2DGrid &flag) : MG(flag.width, flag.height) { updateFields(flag); }
  // borrowed handle (Simulation::mg)
  static MG borrow(ubg_mg *h, int width, int height) {
    MG m;
    m.h_ = h;
    m.owned_ = false;
    m.width = width;
    m.height = height;
    ubg_check(ubg_mg_num_levels(h, &m.levels));
    return m;
  }
  MG(const MG &) = delete;
  MG &operator=(const MG &) = delete;
  MG(MG &&o) noexcept { *this = std::move(o); }
  MG &operator=(MG &&o) noexcept {
    if (this != &o) {
      release();
      h_ = o.h_;
      owned_ = o.owned_;
      levels = o.levels;
      width = o.width;
      height = o.height;
      o.h_ = nullptr;
    }
    return *this;
  }
  ~MG() { release(); }

  void updateFields(Single2DGrid &flag) { ubg_check(ubg_mg_update_fields(h_, flag.data())); }

  void solve(Single2DGrid &p, Single2DGrid &f, Single2DGrid &flag, float h,
             bool zeroGradientBC = false) {
    ubg_check(ubg_mg_solve_host(h_, p.data(), f.data(), flag.data(), h, zeroGradientBC ? 1 : 0));
  }

  void setParams(const ubg_params &prm) { ubg_check(ubg_mg_set_params(h_, &prm)); }
  ubg_mg *handle() { return h_; }

  int levels = 0;
  int width = 0, height = 0;

private:
  void release() {
    if (h_ && owned_) ubg_mg_destroy(h_);
    h_ = nullptr;
  }
  ubg_mg *h_ = nullptr;
  bool owned_ = true;
};

using PressureSolver = MG;

namespace ubg_detail {
inline MG &cached_mg(int w, int h) {
  thread_local std::map<std::pair<int, int>, std::unique_ptr<MG>> cache;
  auto &m = cache[{w, h}];
  if (!m) m.reset(new MG(w, h));
  return *m;
}
} // namespace ubg_detail

// gs(p, f, flag, h, alpha): one lexicographic Gauss-Seidel sweep
// (pressure_solver.cpp:26-33)
inline void gs(Single2DGrid &p, Single2DGrid &f, Single2DGrid &flag, float h, float alpha) {
  MG &m = ubg_detail::cached_mg(p.width, p.height);
  ubg_check(ubg_mg_update_fields(m.handle(), flag.data()));
  ubg_check(ubg_mg_set_level_field(m.handle(), 0, UBG_LEVEL_P, p.data()));
  ubg_check(ubg_mg_set_level_field(m.handle(), 0, UBG_LEVEL_F, f.data()));
  ubg_check(ubg_mg_gs(m.handle(), h, alpha, 1));
  ubg_check(ubg_mg_get_level_field(m.handle(), 0, UBG_LEVEL_P, p.data()));
}

// rbgs(p, f, flag, h, alpha): one red+black sweep (pressure_solver.cpp:49-89)
inline void rbgs(Single2DGrid &p, Single2DGrid &f, Single2DGrid &flag, float h, float alpha) {
  MG &m = ubg_detail::cached_mg(p.width, p.height);
  ubg_check(ubg_mg_update_fields(m.handle(), flag.data()));
  ubg_check(ubg_mg_set_level_field(m.handle(), 0, UBG_LEVEL_P, p.data()));
  ubg_check(ubg_mg_set_level_field(m.handle(), 0, UBG_LEVEL_F, f.data()));
  ubg_check(ubg_mg_rbgs(m.handle(), h, alpha, 1));
  ubg_check(ubg_mg_get_level_field(m.handle(), 0, UBG_LEVEL_P, p.data()));
}

// calculateResidualField (pressure_solver.cpp:91-116): fills r, returns ||r||_2
inline float calculateResidualField(Single2DGrid &p, Single2DGrid &f, Single2DGrid &flag,
                                    Single2DGrid &r, float h) {
  MG &m = ubg_detail::cached_mg(p.width, p.height);
  ubg_check(ubg_mg_update_fields(m.handle(), flag.data()));
  ubg_check(ubg_mg_set_level_field(m.handle(), 0, UBG_LEVEL_P, p.data()));
  ubg_check(ubg_mg_set_level_field(m.handle(), 0, UBG_LEVEL_F, f.data()));
  double n = 0.0;
  ubg_check(ubg_mg_residual(m.handle(), h, r.data(), &n));
  return (float)n;
}
