/*
 * glst_oracle.c — CPU restatement of the reference GLST force path.
 *
 * TEST INFRASTRUCTURE ONLY (see glst_oracle.h): the checker for tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing in the
 * product loads this file.
 *
 * Citations are to the reference sources (src/glst_force.cu,
 * src/cubature.cu, utils/coulomb.cu).
 */
#include "glst_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static const double ORC_PI = 3.14159265358979323846;

void orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

int orc_get_threads(void) { return omp_get_max_threads(); }

/* ---- erfc^{-1}: src/glst_force.cu:1616-1640 ---------------------------- */
int orc_erfc_inv(double y, double tol, double *out) {
  /* Newton from 0 converges only for y in (0, 2); outside it the reference
   * would iterate 1e8 times.  Report failure directly. */
  if (!(y > 0.0 && y < 2.0)) return 3;
  const double rdum = 2.0 / sqrt(ORC_PI);
  double x = 0.0;
  double err = DBL_MAX;
  for (unsigned long long i = 0; i < 100000000ULL; i++) {
    err = 1.0 - erf(x) - y;
    if (fabs(err) < tol) break;
    const double deriv = -exp(-x * x) * rdum;
    x -= err / deriv;
  }
  if (fabs(err) >= tol) return 3;
  *out = x;
  return 0;
}

/* ---- alpha groups: src/glst_force.cu:1642-1724 -------------------------- */
int orc_alpha_groups(uint32_t nx, uint32_t ny, uint32_t nz, double cdx,
                     double cdy, double cdz, double tol, uint32_t cap,
                     uint32_t *ngroup, uint32_t *width_out, double *rmin_out,
                     double *rmax_out, double *alpha_out, double *zcut_out,
                     uint32_t *r_in, uint32_t *r_out) {
  int width = 1, tot_width = 0;
  int remain = (int)nx;
  if ((int)ny > remain) remain = (int)ny;
  if ((int)nz > remain) remain = (int)nz;
  remain -= 2;
  double lmin = cdx, lmax = cdx;
  if (cdy < lmin) lmin = cdy;
  if (cdz < lmin) lmin = cdz;
  if (cdy > lmax) lmax = cdy;
  if (cdz > lmax) lmax = cdz;
  uint32_t ng = 0;
  uint32_t widths[64];
  while (remain >= 0) {
    if (ng >= 64) return 1;
    widths[ng] = (uint32_t)width;
    const double rmin0 = lmin * (double)(tot_width + 1);
    const double rmax0 = sqrt(3.0) * (rmin0 + lmax * (double)(width + 1));
    double a = 0.0, z = 0.0;
    if (orc_erfc_inv(tol * rmin0, DBL_EPSILON, &a)) return 3;
    a /= rmin0;
    if (orc_erfc_inv(0.5 * sqrt(ORC_PI) / a * tol, DBL_EPSILON, &z)) return 3;
    if (ng < cap) {
      if (rmin_out) rmin_out[ng] = rmin0;
      if (rmax_out) rmax_out[ng] = rmax0;
      if (alpha_out) alpha_out[ng] = a;
      if (zcut_out) zcut_out[ng] = z;
    }
    ng++;
    remain -= width;
    tot_width += width;
    if (remain < 1) break;
    const int w1 = 2 * width, w2 = 6 * width;
    if (remain <= w1)
      width = remain;
    else if (w1 < remain && remain < w2)
      width = remain / 2;
    else
      width = w1;
  }
  *ngroup = ng;
  uint32_t rin = 1;
  for (uint32_t g = 0; g < ng; g++) {
    if (g < cap) {
      if (width_out) width_out[g] = widths[g];
      if (r_in) r_in[g] = rin;
      if (r_out) r_out[g] = rin + widths[g];
    }
    rin += widths[g];
  }
  return 0;
}

/* ---- Legendre / Gauss-Legendre: src/cubature.cu:161-351 ----------------- */
void orc_legendre(double x, uint32_t n, double *p, double *dpdx) {
  if (n == 0) {
    *p = 1.0;
    *dpdx = 0.0;
    return;
  }
  if (n == 1) {
    *p = x;
    *dpdx = 1.0;
    return;
  }
  double p0 = 1.0, p1 = x, p2 = 0.0, d1 = 1.0, d2 = 0.0;
  for (uint32_t i = 1; i < n; i++) {
    const double a1 = (double)i + 1.0;
    const double a2 = 2.0 * (double)i + 1.0;
    p2 = (a2 * x * p1 - (double)i * p0) / a1;
    d2 = a1 * p1 + x * d1;
    p0 = p1;
    p1 = p2;
    d1 = d2;
  }
  *p = p2;
  *dpdx = d2;
}

/* eval_legendre_root: at most 10 Newton steps (src/cubature.cu:216-238) */
static double ref_newton(double x0, uint32_t n) {
  double p = 1.0, dp = 0.0, dx1 = 0.0, dx2 = -1.0, x = x0;
  for (int i = 0; i < 10; i++) {
    orc_legendre(x, n, &p, &dp);
    dx2 = dx1;
    dx1 = p / dp;
    if (isnan(dx1)) dx1 = 0.0;
    x -= dx1;
    if (fabs(dx1) < DBL_EPSILON) break;
    if (fabs(dx1 - dx2) < DBL_EPSILON) break;
  }
  return x;
}

/* get_xgl: scan for the extremum, then for the closest approach to zero,
 * then Newton (src/cubature.cu:240-351). */
static int ref_roots(uint32_t n, double *xs, uint32_t *count) {
  if (n == 0) return 1;
  if (n == 1) {
    xs[0] = 0.0;
    *count = 1;
    return 0;
  }
  if (n == 2) {
    xs[0] = 1.0 / sqrt(3.0);
    *count = 1;
    return 0;
  }
  const int even = (n % 2 == 0);
  const uint32_t m = even ? n / 2 : (n + 1) / 2;
  double dx = 0.1 / (double)(m + 1);
  double x = dx;
  double p0 = 0, p1 = 0, d0 = 0, d1 = 0;
  uint32_t k = 0;
  if (!even) {
    xs[k++] = 0.0;
    for (int i = 0; i < 10; i++) {
      orc_legendre(x, n, &p0, &d0);
      x += dx;
      orc_legendre(x, n, &p1, &d1);
      if (d0 * d1 < 0.0) break;
    }
  }
  for (int i = 0; i < 10; i++) {
    orc_legendre(x, n, &p0, &d0);
    x += dx;
    orc_legendre(x, n, &p1, &d1);
    if (fabs(p0) < fabs(p1)) {
      x -= dx;
      break;
    }
  }
  double r0 = ref_newton(x, n);
  xs[k++] = r0;
  uint32_t found = even ? 1 : 2;
  x = r0;
  dx = even ? 0.2 * r0 : 0.1 * r0;
  while (found < m) {
    for (int i = 0; i < 10; i++) {
      orc_legendre(x, n, &p0, &d0);
      x += dx;
      orc_legendre(x, n, &p1, &d1);
      if (d0 * d1 < 0.0) break;
    }
    for (int i = 0; i < 10; i++) {
      orc_legendre(x, n, &p0, &d0);
      x += dx;
      orc_legendre(x, n, &p1, &d1);
      if (fabs(p0) < fabs(p1)) {
        x -= dx;
        break;
      }
    }
    const double r1 = ref_newton(x, n);
    xs[k++] = r1;
    x = r1;
    dx = 0.1 * (r1 - r0);
    r0 = r1;
    found++;
  }
  *count = k;
  return 0;
}

/* Odd orders (used only by the spherical product rule; the reference calls
 * get_xgl with even orders only, and its scan does not reach the first
 * extremum for small odd n, e.g. n = 3): Newton from
 * x0 = cos(pi (k - 1/4) / (n + 1/2)). */
static int odd_roots(uint32_t n, double *xs, uint32_t *count) {
  const uint32_t m = (n - 1) / 2;
  xs[0] = 0.0;
  for (uint32_t k = 1; k <= m; k++) {
    double x = cos(ORC_PI * ((double)k - 0.25) / ((double)n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p, dp;
      orc_legendre(x, n, &p, &dp);
      const double dx = p / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    xs[m + 1 - k] = x; /* ascending */
  }
  *count = m + 1;
  return 0;
}

int orc_gauss_legendre(uint32_t n, double *x, double *w, uint32_t *count) {
  double buf[400];
  uint32_t c = 0;
  if (n == 0 || n > 790) return 1;
  if (n % 2 == 1 && n > 1) {
    if (odd_roots(n, buf, &c)) return 1;
  } else if (ref_roots(n, buf, &c)) {
    return 1;
  }
  for (uint32_t i = 0; i < c; i++) {
    double p, dp;
    orc_legendre(buf[i], n, &p, &dp);
    if (x) x[i] = buf[i];
    if (w) w[i] = 2.0 / ((1.0 - buf[i] * buf[i]) * dp * dp);
  }
  *count = c;
  return 0;
}

/* get_gauss_legendre: src/cubature.cu:353-402 */
int orc_select_gl(double tol, double dist, double alpha, double zcut,
                  uint32_t *order) {
  const double erf0 = 0.5 * ORC_PI * erf(alpha * dist);
  const double tol0 = 0.5 * ORC_PI * dist * tol;
  double x[200], w[200];
  for (uint32_t n = 2; n <= 324; n += 2) {
    uint32_t c = 0;
    if (orc_gauss_legendre(n, x, w, &c)) return 3;
    double s0 = 0.0;
    for (uint32_t i = 0; i < c; i++) {
      const double val = exp(-zcut * zcut * x[i] * x[i]) *
                         sin(2.0 * alpha * dist * zcut * x[i]) / x[i];
      s0 += w[i] * val;
    }
    if (fabs(s0 - erf0) < tol0) {
      *order = n;
      return 0;
    }
  }
  return 3;
}

/* ---- spherical rule ------------------------------------------------------ */
int orc_sph_rule(uint32_t m, double *xyz, double *w, uint32_t *count) {
  const uint32_t nu = m + 1, nphi = 2 * m + 2;
  double u[200], wu[200];
  uint32_t cu = 0;
  if (nu > 390) return 1;
  if (orc_gauss_legendre(nu, u, wu, &cu)) return 1;
  uint32_t k = 0;
  for (uint32_t i = 0; i < cu; i++) {
    const int equator = (u[i] == 0.0);
    const uint32_t nj = equator ? nphi / 2 : nphi;
    const double st = sqrt((1.0 - u[i]) * (1.0 + u[i]));
    for (uint32_t j = 0; j < nj; j++) {
      const double phi = 2.0 * ORC_PI * (double)j / (double)nphi;
      if (xyz) {
        xyz[3 * k] = st * cos(phi);
        xyz[3 * k + 1] = st * sin(phi);
        xyz[3 * k + 2] = u[i];
      }
      if (w) w[k] = wu[i] / (double)nphi;
      k++;
    }
  }
  *count = k;
  return 0;
}

static uint32_t sph_count(uint32_t m) {
  uint32_t c = 0;
  orc_sph_rule(m, NULL, NULL, &c);
  return c;
}

/* get_spherical_tdesign: src/cubature.cu:461-523 (product rule weights) */
int orc_select_sph(uint32_t gl_order, double tol, double dist, double alpha,
                   double zcut, uint32_t *m_out) {
  double x[200], w[200];
  uint32_t c = 0;
  if (orc_gauss_legendre(gl_order, x, w, &c)) return 3;
  const double zcut2 = zcut * zcut;
  const double erf0 = erf(alpha * dist) / dist;
  uint32_t n0 = c / 2;
  int nmin = -1;
  double dmin = DBL_MAX;
  uint32_t last = n0;
  for (int j = 0; j < 3; j++) {
    for (uint32_t n = n0; n < 162; n++) {
      const uint32_t M = sph_count(n);
      double *s = (double *)malloc(sizeof(double) * 3 * M);
      double *om = (double *)malloc(sizeof(double) * M);
      uint32_t cc = 0;
      orc_sph_rule(n, s, om, &cc);
      last = n;
      double s0 = 0.0;
      for (uint32_t i = 0; i < c; i++) {
        const double coef = 2.0 * alpha * zcut * dist * x[i];
        double val = 0.0;
        for (uint32_t k = 0; k < cc; k++) val += om[k] * cos(coef * s[3 * k + j]);
        s0 += w[i] * exp(-zcut2 * x[i] * x[i]) * val;
      }
      free(s);
      free(om);
      s0 *= 4.0 * alpha * zcut / ORC_PI;
      const double err = fabs(s0 - erf0);
      if (err < dmin) {
        dmin = err;
        nmin = (int)n;
      }
      if (err < tol) {
        n0 = n;
        nmin = -2;
        break;
      }
    }
  }
  if (nmin > 0) return 3;
  *m_out = last;
  return 0;
}

/* ---- node table: src/cubature.cu:99-159 ---------------------------------- */
int orc_build_nodes(double tol, uint32_t nx, uint32_t ny, uint32_t nz,
                    double cell, size_t cap, uint32_t *K, double *kx,
                    double *ky, double *kz, double *wk, uint32_t *group) {
  uint32_t ng = 0;
  double rmax[64], alpha[64], zcut[64];
  int st = orc_alpha_groups(nx, ny, nz, cell, cell, cell, tol, 64, &ng, NULL,
                            NULL, rmax, alpha, zcut, NULL, NULL);
  if (st) return st;
  size_t k = 0;
  for (uint32_t g = 0; g < ng; g++) {
    uint32_t order = 0, m = 0, c = 0, M = 0;
    if (orc_select_gl(tol, rmax[g], alpha[g], zcut[g], &order)) return 3;
    if (orc_select_sph(order, tol, rmax[g], alpha[g], zcut[g], &m)) return 3;
    double x[200], w[200];
    orc_gauss_legendre(order, x, w, &c);
    M = sph_count(m);
    double *s = (double *)malloc(sizeof(double) * 3 * M);
    double *om = (double *)malloc(sizeof(double) * M);
    orc_sph_rule(m, s, om, &M);
    const double a = alpha[g], z = zcut[g];
    for (uint32_t i = 0; i < c; i++) {
      const double gauss = exp(-z * z * x[i] * x[i]);
      for (uint32_t d = 0; d < M; d++) {
        if (k < cap) {
          if (kx) kx[k] = 2.0 * a * z * x[i] * s[3 * d];
          if (ky) ky[k] = 2.0 * a * z * x[i] * s[3 * d + 1];
          if (kz) kz[k] = 2.0 * a * z * x[i] * s[3 * d + 2];
          if (wk) wk[k] = 4.0 * a * z * w[i] * gauss * om[d] / ORC_PI;
          if (group) group[k] = g;
        }
        k++;
      }
    }
    free(s);
    free(om);
  }
  *K = (uint32_t)k;
  return 0;
}

/* ---- grid and binning: src/glst_force.cu:223-241, 356-534 ----------------- */
int orc_grid_dims(int ct, double lx, double ly, double lz, double rcut,
                  uint32_t *nx, uint32_t *ny, uint32_t *nz) {
  const double L[3] = {lx, ly, lz};
  uint32_t *out[3] = {nx, ny, nz};
  if (!(lx > 0 && ly > 0 && lz > 0 && rcut > 0)) return 1;
  for (int d = 0; d < 3; d++) {
    uint32_t n;
    if (ct == 1) {
      const float c = (float)rcut, l = (float)L[d];
      n = (uint32_t)(l / c);
      if ((float)n * c < l) n++;
    } else {
      n = (uint32_t)(L[d] / rcut);
      if ((double)n * rcut < L[d]) n++;
    }
    *out[d] = n;
  }
  return 0;
}

/* (int) conversion with the device's saturating semantics (NaN -> 0), then
 * the clamp of src/glst_force.cu:372-377. */
static uint32_t cell_index(double v, uint32_t n) {
  int c;
  if (isnan(v))
    c = 0;
  else if (v >= 2147483647.0)
    c = 2147483647;
  else if (v <= -2147483648.0)
    c = (-2147483647 - 1);
  else
    c = (int)v;
  if (c >= (int)n) c = (int)n - 1;
  if (c < 0) c = 0;
  return (uint32_t)c;
}

int orc_bin(int ct, const double *rx, const double *ry, const double *rz,
            uint32_t natom, uint32_t nx, uint32_t ny, uint32_t nz,
            double rcut, uint32_t *atom_cell, uint32_t *sorted_idx,
            uint32_t *cell_point, uint32_t *cell_count) {
  const size_t ncell = (size_t)nx * ny * nz;
  const float inv_f = 1.0f / (float)rcut;
  const double inv_d = 1.0 / rcut;
  for (uint32_t i = 0; i < natom; i++) {
    uint32_t c3[3];
    const double r3[3] = {rx[i], ry[i], rz[i]};
    const uint32_t n3[3] = {nx, ny, nz};
    for (int d = 0; d < 3; d++) {
      double v;
      if (ct == 1) {
        const float f = (float)r3[d] * inv_f;
        v = (double)f;
      } else {
        v = r3[d] * inv_d;
      }
      c3[d] = cell_index(v, n3[d]);
    }
    atom_cell[i] = (c3[0] * ny + c3[1]) * nz + c3[2];
  }
  uint32_t *cursor = (uint32_t *)calloc(ncell, sizeof(uint32_t));
  if (!cursor) return 1;
  memset(cell_count, 0, ncell * sizeof(uint32_t));
  for (uint32_t i = 0; i < natom; i++) cell_count[atom_cell[i]]++;
  uint32_t run = 0;
  for (size_t c = 0; c < ncell; c++) {
    cell_point[c] = run;
    cursor[c] = run;
    run += cell_count[c];
  }
  /* stable: original order within a cell (CUB radix sort is stable) */
  for (uint32_t i = 0; i < natom; i++) sorted_idx[cursor[atom_cell[i]]++] = i;
  free(cursor);
  return 0;
}

/* ---- the full step -------------------------------------------------------- */

/* summed-volume index helper */
#define CID(x, y, z) (((size_t)(x) * ny + (y)) * nz + (z))

/* box_sum / cube_sum / shell_sum: src/glst_force.cu:774-878 */
static void box_sum(const double *Pr, const double *Pi, size_t ld, size_t k,
                    uint32_t x0, uint32_t y0, uint32_t z0, uint32_t x1,
                    uint32_t y1, uint32_t z1, uint32_t ny, uint32_t nz,
                    double *br, double *bi) {
  const int xb = x0 == 0, yb = y0 == 0, zb = z0 == 0;
  const uint32_t xm = xb ? 0 : x0 - 1, ym = yb ? 0 : y0 - 1,
                 zm = zb ? 0 : z0 - 1;
  const size_t c0 = CID(x1, y1, z1), c1 = CID(xm, y1, z1),
               c2 = CID(x1, ym, z1), c3 = CID(x1, y1, zm),
               c4 = CID(xm, ym, z1), c5 = CID(xm, y1, zm),
               c6 = CID(x1, ym, zm), c7 = CID(xm, ym, zm);
#define G(P, c, skip) ((skip) ? 0.0 : P[(c)*ld + k])
  *br = G(Pr, c0, 0) - G(Pr, c1, xb) - G(Pr, c2, yb) - G(Pr, c3, zb) +
        G(Pr, c4, xb || yb) + G(Pr, c5, xb || zb) + G(Pr, c6, yb || zb) -
        G(Pr, c7, xb || yb || zb);
  *bi = G(Pi, c0, 0) - G(Pi, c1, xb) - G(Pi, c2, yb) - G(Pi, c3, zb) +
        G(Pi, c4, xb || yb) + G(Pi, c5, xb || zb) + G(Pi, c6, yb || zb) -
        G(Pi, c7, xb || yb || zb);
#undef G
}

static void cube_sum(const double *Pr, const double *Pi, size_t ld, size_t k,
                     uint32_t x, uint32_t y, uint32_t z, uint32_t r,
                     uint32_t nx, uint32_t ny, uint32_t nz, double *cr,
                     double *ci) {
  const uint32_t x0 = x < r ? 0 : x - r, y0 = y < r ? 0 : y - r,
                 z0 = z < r ? 0 : z - r;
  uint32_t x1 = x + r, y1 = y + r, z1 = z + r;
  if (x1 >= nx) x1 = nx - 1;
  if (y1 >= ny) y1 = ny - 1;
  if (z1 >= nz) z1 = nz - 1;
  box_sum(Pr, Pi, ld, k, x0, y0, z0, x1, y1, z1, ny, nz, cr, ci);
}

/* Wall seconds of the stages of the last orc_glst call, labelled as the
 * reference driver times them (test/full_example.cu:215-233): assign_atoms,
 * calc_sf, sum_rmt_sf, calc_lr_ef, calc_sr_ef, get_ef. */
static double g_stage_s[6];

void orc_last_stage_times(double *out) {
  for (int i = 0; i < 6; i++) out[i] = g_stage_s[i];
}

int orc_glst(int ct, const double *rx, const double *ry, const double *rz,
             const double *qc, uint32_t natom, double tol, double lx,
             double ly, double lz, double rcut, uint32_t node_stride,
             int far, int near, double *fx, double *fy, double *fz,
             double *en, double *s_re, double *s_im, double *r_re,
             double *r_im) {
  uint32_t nx, ny, nz;
  if (node_stride == 0) node_stride = 1;
  if (orc_grid_dims(ct, lx, ly, lz, rcut, &nx, &ny, &nz)) return 1;
  const double cell = ct == 1 ? (double)(float)rcut : rcut;
  const size_t ncell = (size_t)nx * ny * nz;
  uint32_t *atom_cell = (uint32_t *)malloc(sizeof(uint32_t) * (natom + 1));
  uint32_t *sidx = (uint32_t *)malloc(sizeof(uint32_t) * (natom + 1));
  uint32_t *cpnt = (uint32_t *)malloc(sizeof(uint32_t) * ncell);
  uint32_t *ccnt = (uint32_t *)malloc(sizeof(uint32_t) * ncell);
  double *X = (double *)malloc(sizeof(double) * 4 * (natom + 1));
  double *F = (double *)calloc(4 * (size_t)(natom + 1), sizeof(double));
  if (!atom_cell || !sidx || !cpnt || !ccnt || !X || !F) return 1;
  for (int i = 0; i < 6; i++) g_stage_s[i] = 0.0;
  double t0 = omp_get_wtime(), t1;
  orc_bin(ct, rx, ry, rz, natom, nx, ny, nz, rcut, atom_cell, sidx, cpnt,
          ccnt);
  for (uint32_t s = 0; s < natom; s++) {
    const uint32_t o = sidx[s];
    X[4 * s] = rx[o];
    X[4 * s + 1] = ry[o];
    X[4 * s + 2] = rz[o];
    X[4 * s + 3] = qc[o];
  }
  t1 = omp_get_wtime();
  g_stage_s[0] = t1 - t0;
  t0 = t1;
  int status = 0;
  const int has_lr = nx > 2 && ny > 2 && nz > 2; /* src/glst_force.cu:667 */
  if (far && has_lr) {
    uint32_t K = 0, ng = 0;
    uint32_t r_in[64], r_out[64];
    status = orc_alpha_groups(nx, ny, nz, cell, cell, cell, tol, 64, &ng,
                              NULL, NULL, NULL, NULL, NULL, r_in, r_out);
    if (!status) status = orc_build_nodes(tol, nx, ny, nz, cell, 0, &K, NULL,
                                          NULL, NULL, NULL, NULL);
    double *kx = NULL, *ky = NULL, *kz = NULL, *wk = NULL;
    uint32_t *grp = NULL;
    if (!status && K > 0) {
      kx = (double *)malloc(sizeof(double) * K);
      ky = (double *)malloc(sizeof(double) * K);
      kz = (double *)malloc(sizeof(double) * K);
      wk = (double *)malloc(sizeof(double) * K);
      grp = (uint32_t *)malloc(sizeof(uint32_t) * K);
      status = orc_build_nodes(tol, nx, ny, nz, cell, K, &K, kx, ky, kz, wk,
                               grp);
    }
    if (!status && K > 0) {
      const size_t Ks = (K + node_stride - 1) / node_stride; /* sampled */
      double *Sr = (double *)calloc(ncell * Ks, sizeof(double));
      double *Si = (double *)calloc(ncell * Ks, sizeof(double));
      double *Rr = (double *)calloc(ncell * Ks, sizeof(double));
      double *Ri = (double *)calloc(ncell * Ks, sizeof(double));
      /* calc_sf_kernel: src/glst_force.cu:601-664 */
#pragma omp parallel for schedule(dynamic, 1)
      for (long c = 0; c < (long)ncell; c++) {
        for (size_t ks = 0; ks < Ks; ks++) {
          const size_t k = ks * node_stride;
          double re = 0.0, im = 0.0;
          for (uint32_t j = cpnt[c]; j < cpnt[c] + ccnt[c]; j++) {
            const double th = kx[k] * X[4 * j] + ky[k] * X[4 * j + 1] +
                              kz[k] * X[4 * j + 2];
            re += X[4 * j + 3] * cos(th);
            im -= X[4 * j + 3] * sin(th);
          }
          Sr[c * Ks + ks] = re;
          Si[c * Ks + ks] = im;
        }
      }
      t1 = omp_get_wtime();
      g_stage_s[1] = t1 - t0;
      t0 = t1;
      if (s_re && node_stride == 1) memcpy(s_re, Sr, sizeof(double) * ncell * K);
      if (s_im && node_stride == 1) memcpy(s_im, Si, sizeof(double) * ncell * K);
      t0 = omp_get_wtime();
      /* prefix sums z, y, x: src/glst_force.cu:702-772 */
#pragma omp parallel for schedule(static)
      for (long ks = 0; ks < (long)Ks; ks++) {
        for (uint32_t x = 0; x < nx; x++)
          for (uint32_t y = 0; y < ny; y++) {
            double ar = 0, ai = 0;
            for (uint32_t z = 0; z < nz; z++) {
              const size_t o = CID(x, y, z) * Ks + ks;
              ar += Sr[o];
              ai += Si[o];
              Sr[o] = ar;
              Si[o] = ai;
            }
          }
        for (uint32_t x = 0; x < nx; x++)
          for (uint32_t z = 0; z < nz; z++) {
            double ar = 0, ai = 0;
            for (uint32_t y = 0; y < ny; y++) {
              const size_t o = CID(x, y, z) * Ks + ks;
              ar += Sr[o];
              ai += Si[o];
              Sr[o] = ar;
              Si[o] = ai;
            }
          }
        for (uint32_t y = 0; y < ny; y++)
          for (uint32_t z = 0; z < nz; z++) {
            double ar = 0, ai = 0;
            for (uint32_t x = 0; x < nx; x++) {
              const size_t o = CID(x, y, z) * Ks + ks;
              ar += Sr[o];
              ai += Si[o];
              Sr[o] = ar;
              Si[o] = ai;
            }
          }
      }
      /* calc_rmt_sum_kernel: src/glst_force.cu:880-913 */
#pragma omp parallel for schedule(static)
      for (long c = 0; c < (long)ncell; c++) {
        const uint32_t x = (uint32_t)(c / ((size_t)ny * nz));
        const uint32_t y = (uint32_t)((c / nz) % ny);
        const uint32_t z = (uint32_t)(c % nz);
        for (size_t ks = 0; ks < Ks; ks++) {
          const size_t k = ks * node_stride;
          const uint32_t g = grp[k];
          double orr, oi, ir, ii;
          cube_sum(Sr, Si, Ks, ks, x, y, z, r_out[g], nx, ny, nz, &orr, &oi);
          cube_sum(Sr, Si, Ks, ks, x, y, z, r_in[g], nx, ny, nz, &ir, &ii);
          Rr[c * Ks + ks] = (orr - ir) * wk[k];
          Ri[c * Ks + ks] = (oi - ii) * wk[k];
        }
      }
      t1 = omp_get_wtime();
      g_stage_s[2] = t1 - t0;
      t0 = t1;
      if (r_re && node_stride == 1) memcpy(r_re, Rr, sizeof(double) * ncell * K);
      if (r_im && node_stride == 1) memcpy(r_im, Ri, sizeof(double) * ncell * K);
      t0 = omp_get_wtime();
      /* calc_lr_ef_kernel: src/glst_force.cu:1060-1135 */
#pragma omp parallel for schedule(dynamic, 64)
      for (long s = 0; s < (long)natom; s++) {
        const size_t c = atom_cell[sidx[s]];
        const double xa = X[4 * s], ya = X[4 * s + 1], za = X[4 * s + 2],
                     qa = X[4 * s + 3];
        double f0 = 0, f1 = 0, f2 = 0, e0 = 0;
        for (size_t ks = 0; ks < Ks; ks++) {
          const size_t k = ks * node_stride;
          const double th = kx[k] * xa + ky[k] * ya + kz[k] * za;
          const double re = cos(th), im = sin(th);
          const double rr = Rr[c * Ks + ks], ri = Ri[c * Ks + ks];
          const double dre = qa * (re * rr - im * ri);
          const double dim = qa * (re * ri + im * rr);
          f0 += dim * kx[k];
          f1 += dim * ky[k];
          f2 += dim * kz[k];
          e0 += dre;
        }
        F[4 * s] += f0;
        F[4 * s + 1] += f1;
        F[4 * s + 2] += f2;
        F[4 * s + 3] += e0;
      }
      t1 = omp_get_wtime();
      g_stage_s[3] = t1 - t0;
      free(Sr);
      free(Si);
      free(Rr);
      free(Ri);
    }
    free(kx);
    free(ky);
    free(kz);
    free(wk);
    free(grp);
  }
  t0 = omp_get_wtime();
  if (near && !status) {
    /* calc_sr_ef_intra/inter: src/glst_force.cu:1248-1324, 1436-1544 */
#pragma omp parallel for schedule(dynamic, 64)
    for (long s = 0; s < (long)natom; s++) {
      const uint32_t c = atom_cell[sidx[s]];
      const uint32_t cx = c / (ny * nz), cy = (c / nz) % ny, cz = c % nz;
      const double xi = X[4 * s], yi = X[4 * s + 1], zi = X[4 * s + 2],
                   qi = X[4 * s + 3];
      double f0 = 0, f1 = 0, f2 = 0, e0 = 0;
      for (int a = -1; a <= 1; a++) {
        if ((cx == 0 && a == -1) || (cx == nx - 1 && a == 1)) continue;
        for (int b = -1; b <= 1; b++) {
          if ((cy == 0 && b == -1) || (cy == ny - 1 && b == 1)) continue;
          for (int d = -1; d <= 1; d++) {
            if ((cz == 0 && d == -1) || (cz == nz - 1 && d == 1)) continue;
            const size_t nb = CID(cx + a, cy + b, cz + d);
            for (uint32_t j = cpnt[nb]; j < cpnt[nb] + ccnt[nb]; j++) {
              if (j == (uint32_t)s) continue;
              const double qij = qi * X[4 * j + 3];
              const double xij = xi - X[4 * j], yij = yi - X[4 * j + 1],
                           zij = zi - X[4 * j + 2];
              const double rij2 = xij * xij + yij * yij + zij * zij;
              const double rij = sqrt(rij2);
              const double irij = 1.0 / rij;
              const double dudr = qij / rij2;
              f0 += dudr * (xij * irij);
              f1 += dudr * (yij * irij);
              f2 += dudr * (zij * irij);
              e0 += qij * irij;
            }
          }
        }
      }
      F[4 * s] += f0;
      F[4 * s + 1] += f1;
      F[4 * s + 2] += f2;
      F[4 * s + 3] += e0;
    }
  }
  t1 = omp_get_wtime();
  g_stage_s[4] = t1 - t0;
  t0 = t1;
  /* get_ef: back to original order (src/glst_force.cu:142-176) */
  for (uint32_t s = 0; s < natom; s++) {
    const uint32_t o = sidx[s];
    fx[o] = F[4 * s];
    fy[o] = F[4 * s + 1];
    fz[o] = F[4 * s + 2];
    en[o] = F[4 * s + 3];
  }
  g_stage_s[5] = omp_get_wtime() - t0;
  free(atom_cell);
  free(sidx);
  free(cpnt);
  free(ccnt);
  free(X);
  free(F);
  return status;
}

/* ---- direct Coulomb: utils/coulomb.cu:14-81 ------------------------------- */
int orc_coulomb(const double *rx, const double *ry, const double *rz,
                const double *qc, uint32_t natom, double *fx, double *fy,
                double *fz, double *en) {
#pragma omp parallel for schedule(dynamic, 64)
  for (long i = 0; i < (long)natom; i++) {
    double f0 = 0, f1 = 0, f2 = 0, e0 = 0;
    for (uint32_t j = 0; j < natom; j++) {
      if ((uint32_t)i == j) continue;
      const double qij = qc[i] * qc[j];
      const double xij = rx[i] - rx[j], yij = ry[i] - ry[j],
                   zij = rz[i] - rz[j];
      const double rij2 = xij * xij + yij * yij + zij * zij;
      const double rij = sqrt(rij2);
      const double irij = 1.0 / rij;
      const double dudr = qij / rij2;
      f0 += dudr * (xij * irij);
      f1 += dudr * (yij * irij);
      f2 += dudr * (zij * irij);
      e0 += qij * irij;
    }
    fx[i] = f0;
    fy[i] = f1;
    fz[i] = f2;
    en[i] = e0;
  }
  return 0;
}
