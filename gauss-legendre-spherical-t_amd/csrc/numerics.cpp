// Host numerics for the GLST engine (see numerics.hpp).
#include "numerics.hpp"

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <dirent.h>
#include <fstream>
#include <limits>
#include <stdexcept>

namespace glst {

static constexpr double kPi = 3.14159265358979323846;

double erfc_inv(double y, double tol) {
  // Newton on f(x) = 1 - erf(x) - y from x = 0 with f'(x) = -2/sqrt(pi)
  // e^{-x^2} (src/glst_force.cu:1616-1640).  erfc^{-1} is only defined on
  // (0, 2); outside it the iteration cannot converge, so say so up front.
  if (!(y > 0.0 && y < 2.0))
    throw std::runtime_error(
        "FATAL ERROR: erfc_inv(const double, const double): argument " +
        std::to_string(y) + " outside (0, 2)");
  constexpr unsigned long long kMaxIt = 100000000ULL;
  const double two_over_sqrt_pi = 2.0 / std::sqrt(kPi);
  double x = 0.0;
  double f = std::numeric_limits<double>::max();
  for (unsigned long long it = 0; it < kMaxIt; ++it) {
    f = 1.0 - std::erf(x) - y;
    if (std::fabs(f) < tol) break;
    const double slope = -std::exp(-x * x) * two_over_sqrt_pi;
    x -= f / slope;
  }
  if (std::fabs(f) >= tol)
    throw std::runtime_error(
        "FATAL ERROR: erfc_inv(const double, const double): Inverse of erfc "
        "not found after " + std::to_string(kMaxIt) + " iterations");
  return x;
}

AlphaGroups alpha_groups(uint32_t nx, uint32_t ny, uint32_t nz, double cdx,
                         double cdy, double cdz, double tol) {
  AlphaGroups g;
  const double lmin = std::min(cdx, std::min(cdy, cdz));
  const double lmax = std::max(cdx, std::max(cdy, cdz));
  int remaining = static_cast<int>(std::max(nx, std::max(ny, nz))) - 2;
  int width = 1;
  int covered = 0;
  const double eps = std::numeric_limits<double>::epsilon();
  while (remaining >= 0) {
    g.width.push_back(static_cast<uint32_t>(width));
    // Closest and farthest distance a pair of this shell can have.
    const double rmin = lmin * static_cast<double>(covered + 1);
    const double rmax =
        std::sqrt(3.0) * (rmin + lmax * static_cast<double>(width + 1));
    // alpha: erfc(alpha r)/r = tol at r = rmin; zcut: the Gaussian tail of
    // the radial integral is below tol.
    const double alpha = erfc_inv(tol * rmin, eps) / rmin;
    const double zcut = erfc_inv(0.5 * std::sqrt(kPi) / alpha * tol, eps);
    g.rmin.push_back(rmin);
    g.rmax.push_back(rmax);
    g.alpha.push_back(alpha);
    g.zcut.push_back(zcut);
    remaining -= width;
    covered += width;
    if (remaining < 1) break;
    // Widths double; the tail is split so no group is tiny.
    const int doubled = 2 * width;
    const int next_two = 6 * width;
    if (remaining <= doubled)
      width = remaining;
    else if (doubled < remaining && remaining < next_two)
      width = remaining / 2;
    else
      width = doubled;
  }
  const size_t ng = g.width.size();
  g.r_in.assign(std::max<size_t>(ng, 1), 0);
  g.r_out.assign(std::max<size_t>(ng, 1), 0);
  g.r_in[0] = 1;
  for (size_t i = 0; i < ng; ++i) {
    if (i > 0) g.r_in[i] = g.r_out[i - 1];
    g.r_out[i] = g.r_in[i] + g.width[i];
  }
  g.r_in.resize(ng);
  g.r_out.resize(ng);
  return g;
}

void legendre(double x, uint32_t n, double &p, double &dp) {
  if (n == 0) {
    p = 1.0;
    dp = 0.0;
    return;
  }
  if (n == 1) {
    p = x;
    dp = 1.0;
    return;
  }
  // (i+1) P_{i+1} = (2i+1) x P_i - i P_{i-1};  P_{i+1}' = (i+1) P_i + x P_i'
  double pm1 = 1.0, pi = x, dpi = 1.0, pn = 0.0, dpn = 0.0;
  for (uint32_t i = 1; i < n; ++i) {
    const double ip1 = static_cast<double>(i) + 1.0;
    const double tip1 = 2.0 * static_cast<double>(i) + 1.0;
    pn = (tip1 * x * pi - static_cast<double>(i) * pm1) / ip1;
    dpn = ip1 * pi + x * dpi;
    pm1 = pi;
    pi = pn;
    dpi = dpn;
  }
  p = pn;
  dp = dpn;
}

void gauss_legendre_nonneg(uint32_t n, std::vector<double> &x,
                           std::vector<double> &w) {
  x.clear();
  w.clear();
  if (n == 0) throw std::runtime_error("No roots for n == 0");
  std::vector<double> pos;
  const uint32_t npos = n / 2;
  const double dn = static_cast<double>(n);
  for (uint32_t k = 1; k <= npos; ++k) {
    // Tricomi: x_k ~ (1 - 1/(8n^2) + 1/(8n^3)) cos(pi (4k-1)/(4n+2))
    const double th = kPi * (4.0 * k - 1.0) / (4.0 * dn + 2.0);
    double r = (1.0 - 1.0 / (8.0 * dn * dn) + 1.0 / (8.0 * dn * dn * dn)) *
               std::cos(th);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(r, n, p, dp);
      const double dx = p / dp;
      r -= dx;
      if (std::fabs(dx) <= 4.0 * DBL_EPSILON * std::fabs(r)) break;
    }
    pos.push_back(r);
  }
  std::reverse(pos.begin(), pos.end());
  if (n % 2 == 1) x.push_back(0.0);
  x.insert(x.end(), pos.begin(), pos.end());
  for (double r : x) {
    double p, dp;
    legendre(r, n, p, dp);
    w.push_back(2.0 / ((1.0 - r * r) * dp * dp));
  }
}

GLRule select_gauss_legendre(double tol, double dist, double alpha,
                             double zcut) {
  constexpr uint32_t kMin = 2, kMax = 324;
  // int_0^1 e^{-z^2 u^2} sin(2 a d z u)/u du -> (pi/2) erf(a d)
  const double target = 0.5 * kPi * std::erf(alpha * dist);
  const double bound = 0.5 * kPi * dist * tol;
  for (uint32_t n = kMin; n <= kMax; n += 2) {
    GLRule r;
    r.order = n;
    gauss_legendre_nonneg(n, r.x, r.w);
    double s = 0.0;
    for (size_t i = 0; i < r.x.size(); ++i) {
      const double v = std::exp(-zcut * zcut * r.x[i] * r.x[i]) *
                       std::sin(2.0 * alpha * dist * zcut * r.x[i]) / r.x[i];
      s += r.w[i] * v;
    }
    if (std::fabs(s - target) < bound) return r;
  }
  throw std::runtime_error(
      "Suitable Gauss-Legendre quadrature could not be found");
}

SphRule product_rule(uint32_t m) {
  SphRule r;
  r.m = m;
  const uint32_t nu = m + 1;      // GL nodes in u = cos(theta): degree 2m+1
  const uint32_t nphi = 2 * m + 2;  // azimuths: trigonometric degree 2m+1
  std::vector<double> u, wu;
  gauss_legendre_nonneg(nu, u, wu);
  for (size_t i = 0; i < u.size(); ++i) {
    const bool equator = (u[i] == 0.0);
    // Full rule weight per point is wu/(2 nphi) (sums to 1 over the sphere);
    // merging an antipodal pair doubles it.
    const double wpt = wu[i] / static_cast<double>(nphi);
    const double st = std::sqrt((1.0 - u[i]) * (1.0 + u[i]));
    const uint32_t jmax = equator ? nphi / 2 : nphi;
    for (uint32_t j = 0; j < jmax; ++j) {
      const double phi = 2.0 * kPi * static_cast<double>(j) /
                         static_cast<double>(nphi);
      r.s.push_back({st * std::cos(phi), st * std::sin(phi), u[i]});
      r.w.push_back(wpt);
    }
  }
  return r;
}

static bool parse_design_name(const std::string &name, uint32_t &t,
                              uint32_t &npts) {
  // ...TTT.NNNNN
  if (name.size() < 9) return false;
  const std::string tail = name.substr(name.size() - 9);
  if (tail[3] != '.') return false;
  for (int i = 0; i < 9; ++i)
    if (i != 3 && (tail[i] < '0' || tail[i] > '9')) return false;
  t = static_cast<uint32_t>(std::stoul(tail.substr(0, 3)));
  npts = static_cast<uint32_t>(std::stoul(tail.substr(4, 5)));
  return true;
}

SphRule design_rule(uint32_t m, const std::string &dir) {
  SphRule r = design_file(2 * m + 1, dir);
  r.m = m;
  return r;
}

SphRule design_file(uint32_t t, const std::string &dir) {
  DIR *d = opendir(dir.c_str());
  if (!d)
    throw std::runtime_error("Failed to open spherical design directory \"" +
                             dir + "\"");
  std::string best;
  uint32_t best_n = 0;
  while (dirent *e = readdir(d)) {
    const std::string name = e->d_name;
    uint32_t tt, nn;
    if (!parse_design_name(name, tt, nn) || tt != t) continue;
    if (best.empty() || name < best) {
      best = name;
      best_n = nn;
    }
  }
  closedir(d);
  if (best.empty())
    throw std::runtime_error("No spherical design file of degree " +
                             std::to_string(t) + " in directory \"" + dir +
                             "\"");
  const std::string path = dir + "/" + best;
  std::ifstream fs(path);
  if (!fs.is_open())
    throw std::runtime_error("Failed to open file \"" + path + "\"");
  SphRule r;
  r.m = t / 2;
  const uint32_t half = best_n / 2;  // upper hemisphere of a symmetric design
  for (uint32_t i = 0; i < half; ++i) {
    std::string line;
    if (!std::getline(fs, line) || line.size() < 75)
      throw std::runtime_error("Malformed spherical design file \"" + path +
                               "\"");
    r.s.push_back({std::stod(line.substr(0, 25)), std::stod(line.substr(25, 25)),
                   std::stod(line.substr(50, 25))});
  }
  r.w.assign(half, 1.0 / static_cast<double>(half));
  return r;
}

SphRule select_spherical(const GLRule &gl, double tol, double dist,
                         double alpha, double zcut, const std::string &dir) {
  constexpr uint32_t kMaxM = 162;
  const double z2 = zcut * zcut;
  const double target = std::erf(alpha * dist) / dist;
  const bool files = !dir.empty();
  uint32_t m0 = static_cast<uint32_t>(gl.x.size()) / 2;
  int best_m = -1;
  double best_err = std::numeric_limits<double>::max();
  SphRule rule;
  for (int axis = 0; axis < 3; ++axis) {
    for (uint32_t m = m0; m < kMaxM; ++m) {
      rule = files ? design_rule(m, dir) : product_rule(m);
      double s = 0.0;
      for (size_t i = 0; i < gl.x.size(); ++i) {
        const double coef = 2.0 * alpha * zcut * dist * gl.x[i];
        double v = 0.0;
        if (files) {
          for (size_t k = 0; k < rule.s.size(); ++k)
            v += std::cos(coef * rule.s[k][axis]);
        } else {
          for (size_t k = 0; k < rule.s.size(); ++k)
            v += rule.w[k] * std::cos(coef * rule.s[k][axis]);
        }
        s += gl.w[i] * std::exp(-z2 * gl.x[i] * gl.x[i]) * v;
      }
      if (files)
        s *= 4.0 * alpha * zcut /
             (kPi * static_cast<double>(rule.s.size()));
      else
        s *= 4.0 * alpha * zcut / kPi;
      const double err = std::fabs(s - target);
      if (err < best_err) {
        best_err = err;
        best_m = static_cast<int>(m);
      }
      if (err < tol) {
        m0 = m;
        best_m = -2;
        break;
      }
    }
  }
  if (best_m > 0)
    throw std::runtime_error("Suitable t-design could not be found");
  return rule;
}

NodeTable build_nodes(double tol, const AlphaGroups &g,
                      const std::string &dir) {
  NodeTable t;
  const size_t ng = g.width.size();
  for (size_t grp = 0; grp < ng; ++grp) {
    GLRule gl = select_gauss_legendre(tol, g.rmax[grp], g.alpha[grp],
                                      g.zcut[grp]);
    SphRule sph = select_spherical(gl, tol, g.rmax[grp], g.alpha[grp],
                                   g.zcut[grp], dir);
    t.first.push_back(static_cast<uint32_t>(t.kx.size()));
    t.count.push_back(static_cast<uint32_t>(gl.x.size() * sph.s.size()));
    const double a = g.alpha[grp], z = g.zcut[grp];
    for (size_t i = 0; i < gl.x.size(); ++i) {
      const double gauss = std::exp(-z * z * gl.x[i] * gl.x[i]);
      for (size_t k = 0; k < sph.s.size(); ++k) {
        t.kx.push_back(2.0 * a * z * gl.x[i] * sph.s[k][0]);
        t.ky.push_back(2.0 * a * z * gl.x[i] * sph.s[k][1]);
        t.kz.push_back(2.0 * a * z * gl.x[i] * sph.s[k][2]);
        t.w.push_back(4.0 * a * z * gl.w[i] * gauss * sph.w[k] / kPi);
        t.group.push_back(static_cast<uint32_t>(grp));
      }
    }
    std::vector<double> kap;
    for (size_t i = 0; i < gl.x.size(); ++i) kap.push_back(2.0 * a * z * gl.x[i]);
    t.kappa.push_back(std::move(kap));
    t.dirs.push_back(sph.s);
    t.gl.push_back(std::move(gl));
    t.sph_m.push_back(sph.m);
    t.sph_count.push_back(static_cast<uint32_t>(sph.s.size()));
  }
  return t;
}

template <typename CT>
void grid_dims(double lx, double ly, double lz, double rcut, uint32_t &nx,
               uint32_t &ny, uint32_t &nz) {
  const CT cell = static_cast<CT>(rcut);
  auto one = [&](double l) {
    const CT L = static_cast<CT>(l);
    uint32_t n = static_cast<uint32_t>(L / cell);
    if (static_cast<CT>(n) * cell < L) ++n;
    return n;
  };
  nx = one(lx);
  ny = one(ly);
  nz = one(lz);
}

template void grid_dims<float>(double, double, double, double, uint32_t &,
                               uint32_t &, uint32_t &);
template void grid_dims<double>(double, double, double, double, uint32_t &,
                                uint32_t &, uint32_t &);

}  // namespace glst
