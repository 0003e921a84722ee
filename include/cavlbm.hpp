// cavlbm.hpp — C++ host mirror of the reference's uniform `CavitySolver`
// (bbrangeo/cfd_cavity src/uniform/solver.rs:35-1020, post.rs) over the C-ABI
// of include/cavlbm.h.  RAII ownership like the Rust struct; the reference's
// assert! panics become std::invalid_argument, runtime failures
// std::runtime_error.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cavlbm.h"

namespace cavlbm {

class CavitySolver {
 public:
  // CavitySolver::new (solver.rs:108); nu_override < 0 means None.
  CavitySolver(size_t nx, size_t ny, double ra, double pr, double cs, double pr_t,
               double utarget, uint64_t seed, size_t jobs, double alpha_mult,
               double nu_start_mult, double nu_override = -1.0,
               int storage = CAVLBM_STORAGE_FP64, int geq_order = 2, int device = 0) {
    cavlbm_params p;
    cavlbm_params_default(&p);
    p.nx = nx;
    p.ny = ny;
    p.ra = ra;
    p.pr = pr;
    p.cs = cs;
    p.pr_t = pr_t;
    p.utarget = utarget;
    p.seed = seed;
    p.jobs = jobs;
    p.alpha_mult = alpha_mult;
    p.nu_start_mult = nu_start_mult;
    p.has_nu_override = nu_override >= 0.0 ? 1 : 0;
    p.nu_override = nu_override >= 0.0 ? nu_override : 0.0;
    p.storage = storage;
    p.geq_order = geq_order;
    p.device = device;
    const int rc = cavlbm_create(&p, &s_);
    if (rc == CAVLBM_E_PARAM) throw std::invalid_argument(cavlbm_last_error(nullptr));
    if (rc) throw std::runtime_error(cavlbm_last_error(nullptr));
    nx_ = nx;
    ny_ = ny;
  }
  ~CavitySolver() { cavlbm_destroy(s_); }
  CavitySolver(const CavitySolver&) = delete;
  CavitySolver& operator=(const CavitySolver&) = delete;
  CavitySolver(CavitySolver&& o) noexcept : s_(std::exchange(o.s_, nullptr)), nx_(o.nx_), ny_(o.ny_) {}

  size_t nx() const { return nx_; }
  size_t ny() const { return ny_; }
  size_t idx(size_t ix, size_t iy) const { return iy * nx_ + ix; }

  void step() { check(cavlbm_step(s_)); }                      // solver.rs:279
  void run(uint64_t n) { check(cavlbm_run(s_, n)); }
  void init_from_fields(const std::vector<double>& rho, const std::vector<double>& ux,
                        const std::vector<double>& uy, const std::vector<double>& t) {
    const size_t n = nx_ * ny_;
    if (rho.size() != n || ux.size() != n || uy.size() != n || t.size() != n)
      throw std::invalid_argument("init_from_fields: fields of size nx*ny expected");
    check(cavlbm_init_from_fields(s_, rho.data(), ux.data(), uy.data(), t.data(), n));
  }
  std::vector<double> field(int which) {
    std::vector<double> v(nx_ * ny_);
    check(cavlbm_get_field(s_, which, v.data(), v.size()));
    return v;
  }
  std::vector<double> rho() { return field(CAVLBM_FIELD_RHO); }
  std::vector<double> ux() { return field(CAVLBM_FIELD_UX); }
  std::vector<double> uy() { return field(CAVLBM_FIELD_UY); }
  std::vector<double> temperature() { return field(CAVLBM_FIELD_T); }
  std::vector<double> nu_t() { return field(CAVLBM_FIELD_NU_T); }
  cavlbm_diag diagnostics() {
    cavlbm_diag d;
    check(cavlbm_diagnostics(s_, &d));
    return d;
  }
  cavlbm_derived derived() const {
    cavlbm_derived d;
    cavlbm_derived_params(s_, &d);
    return d;
  }
  double total_mass() { return diagnostics().total_mass; }        // solver.rs:1012
  double total_f_mass() { return diagnostics().total_f_mass; }    // solver.rs:1018
  uint64_t time_step() const {
    uint64_t t = 0;
    cavlbm_time_step(s_, &t);
    return t;
  }
  std::pair<double, double> max_uy_on_row(size_t iy) {            // post.rs:137
    double w = 0.0, x = 0.0;
    check(cavlbm_max_uy_on_row(s_, iy, &w, &x));
    return {w, x};
  }
  void set_dynamic_sgs(bool on, double a = 0.5, double m = -0.1, double e = 2.0) {
    check(cavlbm_set_dynamic_sgs(s_, on ? 1 : 0, a, m, e));   // solver.rs:215
  }
  // cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max (solver.rs:95-102)
  std::vector<double> dynamic_stats() {
    std::vector<double> v(4);
    check(cavlbm_dynamic_stats(s_, v.data()));
    return v;
  }
  // TimeAverager (post.rs:11-57), accumulated on the device
  void average_accumulate() { check(cavlbm_average_accumulate(s_)); }
  uint64_t average_count() const {
    uint64_t n = 0;
    cavlbm_average_count(s_, &n);
    return n;
  }
  void average_means(std::vector<double>& ux, std::vector<double>& uy, std::vector<double>& t,
                     std::vector<double>& nut, std::vector<double>& uu) {
    const size_t n = nx_ * ny_;
    for (auto* v : {&ux, &uy, &t, &nut, &uu}) v->resize(n);
    check(cavlbm_average_means(s_, ux.data(), uy.data(), t.data(), nut.data(), uu.data(), n));
  }
  cavlbm_solver* raw() { return s_; }

 private:
  void check(int rc) const {
    if (rc) throw std::runtime_error(cavlbm_last_error(s_));
  }
  cavlbm_solver* s_ = nullptr;
  size_t nx_ = 0, ny_ = 0;
};

// Octree CavitySolver (src/octree/solver.rs:118-1206) over cavlbm_octree_*.
class OctreeCavitySolver {
 public:
  // CavitySolver::new (solver.rs:340-353); nu_override < 0 means None.
  OctreeCavitySolver(size_t nx0, size_t ny0, size_t bl_width, double ra, double pr,
                     double utarget, uint64_t seed, size_t jobs, bool odd_exact, int levels,
                     double nu_start_mult, double nu_override = -1.0, int geq_order = 2,
                     int device = 0) {
    cavlbm_octree_params p;
    cavlbm_octree_params_default(&p);
    p.nx0 = nx0;
    p.ny0 = ny0;
    p.bl_width = bl_width;
    p.ra = ra;
    p.pr = pr;
    p.utarget = utarget;
    p.seed = seed;
    p.jobs = jobs;
    p.odd_exact = odd_exact ? 1 : 0;
    p.levels = levels;
    p.nu_start_mult = nu_start_mult;
    p.has_nu_override = nu_override >= 0.0 ? 1 : 0;
    p.nu_override = nu_override >= 0.0 ? nu_override : 0.0;
    p.geq_order = geq_order;
    p.device = device;
    const int rc = cavlbm_octree_create(&p, &o_);
    if (rc == CAVLBM_E_PARAM) throw std::invalid_argument(cavlbm_octree_last_error(nullptr));
    if (rc) throw std::runtime_error(cavlbm_octree_last_error(nullptr));
    check(cavlbm_octree_derived_params(o_, &d_));
    for (int g = 0; g < d_.n_grids; ++g) {
      uint64_t nx = 0, ny = 0;
      check(cavlbm_octree_grid_dims(o_, g, &nx, &ny));
      dims_.push_back({(size_t)nx, (size_t)ny});
    }
  }
  ~OctreeCavitySolver() { cavlbm_octree_destroy(o_); }
  OctreeCavitySolver(const OctreeCavitySolver&) = delete;
  OctreeCavitySolver& operator=(const OctreeCavitySolver&) = delete;

  void step_cavity() { check(cavlbm_octree_step(o_)); }          // solver.rs:647
  void run(uint64_t n) { check(cavlbm_octree_run(o_, n)); }
  const cavlbm_octree_derived& derived() const { return d_; }
  const std::vector<std::pair<size_t, size_t>>& dims() const { return dims_; }
  int n_grids() const { return d_.n_grids; }
  cavlbm_octree_diag diagnostics() {
    cavlbm_octree_diag d;
    check(cavlbm_octree_diagnostics(o_, &d));
    return d;
  }
  std::vector<double> field(int grid, int which) {
    std::vector<double> v(dims_[grid].first * dims_[grid].second);
    check(cavlbm_octree_get_field(o_, grid, which, v.data(), v.size()));
    return v;
  }
  // level_fields (solver.rs:1048-1066), concatenated in grid order
  void level_fields(std::vector<double>& rho, std::vector<double>& ux, std::vector<double>& uy,
                    std::vector<double>& t) {
    rho.clear();
    ux.clear();
    uy.clear();
    t.clear();
    for (int g = 0; g < n_grids(); ++g) {
      for (auto pr : {std::make_pair(&rho, CAVLBM_FIELD_RHO), std::make_pair(&ux, CAVLBM_FIELD_UX),
                      std::make_pair(&uy, CAVLBM_FIELD_UY), std::make_pair(&t, CAVLBM_FIELD_T)}) {
        const auto v = field(g, pr.second);
        pr.first->insert(pr.first->end(), v.begin(), v.end());
      }
    }
  }
  void set_from_snapshot(const std::vector<double>& rho, const std::vector<double>& ux,
                         const std::vector<double>& uy, const std::vector<double>& t) {
    check(cavlbm_octree_set_from_snapshot(o_, rho.data(), ux.data(), uy.data(), t.data(),
                                          rho.size()));
  }
  void set_dynamic_sgs(bool on, double a = 0.5, double m = -0.1, double e = 2.0) {
    check(cavlbm_octree_set_dynamic_sgs(o_, on ? 1 : 0, a, m, e));
  }
  // cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max, nu_t_max
  std::vector<double> dynamic_stats() {
    std::vector<double> v(5);
    check(cavlbm_octree_dynamic_stats(o_, v.data()));
    return v;
  }
  cavlbm_octree* raw() { return o_; }

 private:
  void check(int rc) const {
    if (rc) throw std::runtime_error(cavlbm_octree_last_error(o_));
  }
  cavlbm_octree* o_ = nullptr;
  cavlbm_octree_derived d_{};
  std::vector<std::pair<size_t, size_t>> dims_;
};

}  // namespace cavlbm
