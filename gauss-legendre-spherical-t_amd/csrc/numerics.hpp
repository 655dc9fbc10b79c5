// Host numerics that fix the shapes of the GLST force step: grid dimensions,
// alpha groups, Gauss-Legendre radial rules, the half-sphere spherical rule and
// the cubature node table.  Pure C++ (no HIP), all arithmetic in FP64 except the
// grid derivation, which runs in the calculation type CT as the reference does.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace glst {

// Newton iteration for erfc^{-1}(y) from x = 0, stopping when
// |1 - erf(x) - y| < tol (src/glst_force.cu:1616-1640).  Throws
// std::runtime_error when it does not converge.
double erfc_inv(double y, double tol);

struct AlphaGroups {
  std::vector<uint32_t> width;  // shell width (cells) of each group
  std::vector<double> rmin, rmax, alpha, zcut;
  std::vector<uint32_t> r_in, r_out;  // Chebyshev shell (r_in, r_out]
};

// init_alpha_groups (src/glst_force.cu:1642-1724).  cd* are the cell
// dimensions as the reference holds them (CT values widened to double).
AlphaGroups alpha_groups(uint32_t nx, uint32_t ny, uint32_t nz, double cdx,
                         double cdy, double cdz, double tol);

// Legendre P_n(x) and P_n'(x) by the three-term recurrence
// (src/cubature.cu:161-193).
void legendre(double x, uint32_t n, double &p, double &dp);

// Non-negative roots of P_n in ascending order (root 0 first for odd n) and
// the Gauss-Legendre weights 2 / ((1 - x^2) P_n'(x)^2) (src/cubature.cu:
// 195-351).  Roots come from Newton's method started at Tricomi's asymptotic
// estimate.
void gauss_legendre_nonneg(uint32_t n, std::vector<double> &x,
                           std::vector<double> &w);

struct GLRule {
  uint32_t order = 0;  // n (even)
  std::vector<double> x, w;  // the n/2 positive roots and weights
};

// get_gauss_legendre (src/cubature.cu:353-402): smallest even n in [2, 324]
// with |sum_i w_i e^{-z^2 x_i^2} sin(2 a d z x_i)/x_i - (pi/2) erf(a d)|
// < (pi/2) d tol.
GLRule select_gauss_legendre(double tol, double dist, double alpha,
                             double zcut);

struct SphRule {
  uint32_t m = 0;       // degree index: exact for spherical polynomials of
                        // degree <= 2m+1
  std::vector<std::array<double, 3>> s;  // half-sphere unit vectors
  std::vector<double> w;                 // weights, sum to 1 over the half set
};

// Antipodally reduced product rule of degree 2m+1: m+1 Gauss-Legendre nodes
// in cos(theta) times 2m+2 equispaced azimuths, keeping one point of every
// antipodal pair (upper hemisphere, half of the equator ring when m+1 is
// odd) with the pair's combined weight.
SphRule product_rule(uint32_t m);

// Womersley symmetric spherical design of degree t = 2m+1 read from `dir`
// (files ssTTT.NNNNN, 3 x 25-character columns; the first N/2 points, equal
// weights 1/(N/2)), as src/cubature.cu:422-459 reads them.
SphRule design_rule(uint32_t m, const std::string &dir);
// The same for any degree t (get_tdesign_props + read_tdesign): among the
// files of degree t the one whose name sorts first; names not of the form
// ...TTT.NNNNN are skipped.
SphRule design_file(uint32_t t, const std::string &dir);

// get_spherical_tdesign (src/cubature.cu:461-523): for each axis j the first
// degree index m >= m0 (starting at len(x)/2, carried across axes) whose
// radial x spherical rule reproduces erf(a d)/d along that axis within tol.
// `dir` empty selects product_rule, otherwise design_rule.
SphRule select_spherical(const GLRule &gl, double tol, double dist,
                         double alpha, double zcut, const std::string &dir);

struct NodeTable {
  std::vector<double> kx, ky, kz, w;
  std::vector<uint32_t> group;
  std::vector<uint32_t> first, count;  // per group
  std::vector<GLRule> gl;
  std::vector<uint32_t> sph_m, sph_count;
  // per group: radial magnitudes kappa_i = 2 a z x_i and the half-sphere
  // directions s_m, so node (i, m) has k = kappa_i s_m
  std::vector<std::vector<double>> kappa;
  std::vector<std::vector<std::array<double, 3>>> dirs;
};

// cubature<CT>::initialize (src/cubature.cu:99-159): nodes ordered
// group -> positive GL root (ascending) -> direction, with
// k = 2 a z x_i s_m and W = 4 a z w_i e^{-z^2 x_i^2} omega_m / pi.
NodeTable build_nodes(double tol, const AlphaGroups &g, const std::string &dir);

// Grid derivation of init() in precision CT (src/glst_force.cu:223-241).
template <typename CT>
void grid_dims(double lx, double ly, double lz, double rcut, uint32_t &nx,
               uint32_t &ny, uint32_t &nz);

}  // namespace glst
