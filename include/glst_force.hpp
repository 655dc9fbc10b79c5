// glst_force<CT>: the reference's force object (include/glst_force.hcu:24-128)
// as a header-only C++ facade over the C ABI of glst_hip.h.  Method names,
// arguments and meaning follow the reference; host failures throw
// std::runtime_error with the library's message, as the reference throws.
//
//   glst_force<double> f(natom, tol, Lx, Ly, Lz, rcut);   // plain FP64 path
//   glst_force<float>  g(natom, tol, Lx, Ly, Lz, rcut);   // mixed FP32/FP64
//   f.calc_ener_force(d_rx, d_ry, d_rz, d_qc);            // device FP64 SoA
//   f.get_ef(fx, fy, fz, en);                              // original order
//
// One GPU holds the whole system, so the per-device vectors returned by
// fx()/fy()/fz()/en() have one element (a view of the engine's cell-sorted
// accumulators).
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "glst_hip.h"
#include "hip_container.hpp"

template <typename CT> class glst_force {
  static_assert(std::is_same<CT, float>::value ||
                    std::is_same<CT, double>::value,
                "CT must be float or double");

 public:
  glst_force() = default;
  glst_force(unsigned int natom, double tol, double box_dim_x,
             double box_dim_y, double box_dim_z, double rcut) {
    init(natom, tol, box_dim_x, box_dim_y, box_dim_z, rcut);
  }
  glst_force(unsigned int natom, double tol, double box_dim_x,
             double box_dim_y, double box_dim_z, unsigned int ncell_x,
             unsigned int ncell_y, unsigned int ncell_z) {
    glst_handle_t h = nullptr;
    check_global(glst_create_ncell(ct(), natom, tol, box_dim_x, box_dim_y,
                                   box_dim_z, ncell_x, ncell_y, ncell_z,
                                   nullptr, &h));
    adopt(h, natom);
  }
  glst_force(const glst_force &) = delete;
  glst_force &operator=(const glst_force &) = delete;
  ~glst_force() {
    if (h_) glst_destroy(h_);
  }

  void init(unsigned int natom, double tol, double box_dim_x,
            double box_dim_y, double box_dim_z, double rcut) {
    glst_handle_t h = nullptr;
    check_global(glst_create(ct(), natom, tol, box_dim_x, box_dim_y,
                             box_dim_z, rcut, nullptr, &h));
    if (h_) glst_destroy(h_);
    adopt(h, natom);
  }

  // ---- energy and force calculation ---------------------------------------
  void assign_atoms(const double *d_rx, const double *d_ry,
                    const double *d_rz, const double *d_qc) {
    check(glst_assign_atoms(h_, d_rx, d_ry, d_rz, d_qc));
  }
  void calc_sf() { check(glst_calc_sf(h_)); }
  void sum_rmt_sf() { check(glst_sum_rmt_sf(h_)); }
  void calc_lr_ef() { check(glst_calc_lr_ef(h_)); }
  void calc_sr_ef() { check(glst_calc_sr_ef(h_)); }
  void comm_ef() { check(glst_comm_ef(h_)); }
  void calc_ener_force(const double *d_rx, const double *d_ry,
                       const double *d_rz, const double *d_qc) {
    check(glst_calc_ener_force(h_, d_rx, d_ry, d_rz, d_qc));
  }

  // ---- results ----------------------------------------------------------------
  // Energies and forces in original atom order, host and device copies.
  void get_ef(hip_container<double> &fx, hip_container<double> &fy,
              hip_container<double> &fz, hip_container<double> &en) {
    fx.resize(natom_);
    fy.resize(natom_);
    fz.resize(natom_);
    en.resize(natom_);
    check(glst_get_ef(h_, fx.d_array().data(), fy.d_array().data(),
                      fz.d_array().data(), en.d_array().data()));
    fx.transfer_to_host();
    fy.transfer_to_host();
    fz.transfer_to_host();
    en.transfer_to_host();
  }

  std::vector<hip_container<double>> &fx() { return fx_; }
  std::vector<hip_container<double>> &fy() { return fy_; }
  std::vector<hip_container<double>> &fz() { return fz_; }
  std::vector<hip_container<double>> &en() { return en_; }
  const std::vector<hip_container<double>> &fx() const { return fx_; }
  const std::vector<hip_container<double>> &fy() const { return fy_; }
  const std::vector<hip_container<double>> &fz() const { return fz_; }
  const std::vector<hip_container<double>> &en() const { return en_; }

  // ---- setup inspection (beyond the reference API) ---------------------------
  glst_info_t info() const {
    glst_info_t i;
    check(glst_get_info(h_, &i));
    return i;
  }
  glst_group_info_t group_info(unsigned int g) const {
    glst_group_info_t i;
    check(glst_get_group_info(h_, g, &i));
    return i;
  }
  glst_handle_t handle() const { return h_; }

 private:
  static int ct() {
    return std::is_same<CT, float>::value ? GLST_CT_FLOAT : GLST_CT_DOUBLE;
  }
  static void check_global(glst_status_t s) {
    if (s != GLST_OK) throw std::runtime_error(glst_global_last_error());
  }
  void check(glst_status_t s) const {
    if (s != GLST_OK)
      throw std::runtime_error(h_ ? glst_last_error(h_)
                                  : "glst_force: object not initialised");
  }
  void adopt(glst_handle_t h, unsigned int natom) {
    h_ = h;
    natom_ = natom;
    double *p[4];
    check(glst_sorted_ef_ptrs(h_, &p[0], &p[1], &p[2], &p[3]));
    fx_.clear();
    fy_.clear();
    fz_.clear();
    en_.clear();
    fx_.push_back(hip_container<double>::view(p[0], natom));
    fy_.push_back(hip_container<double>::view(p[1], natom));
    fz_.push_back(hip_container<double>::view(p[2], natom));
    en_.push_back(hip_container<double>::view(p[3], natom));
  }

  glst_handle_t h_ = nullptr;
  unsigned int natom_ = 0;
  std::vector<hip_container<double>> fx_, fy_, fz_, en_;
};
