/*
 * glst_oracle.h — CPU restatement of the reference GLST force path, used as
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  The product never links, loads or
 * calls anything here.
 *
 * Everything runs in FP64 on the host (OpenMP over cells / nodes / atoms),
 * following the reference files cited at each function.  The spherical rule is
 * the antipodally reduced product Gauss-Legendre x trapezoid rule (the
 * reference's t-design files are absent), restated independently of the
 * product's implementation.
 */
#ifndef GLST_ORACLE_H
#define GLST_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status: 0 ok, 1 invalid argument, 3 numerics failure */

/* src/glst_force.cu:1616-1640 */
int orc_erfc_inv(double y, double tol, double *x);

/* src/glst_force.cu:1642-1724; cap = capacity of the arrays */
int orc_alpha_groups(uint32_t nx, uint32_t ny, uint32_t nz, double cdx,
                     double cdy, double cdz, double tol, uint32_t cap,
                     uint32_t *ngroup, uint32_t *width, double *rmin,
                     double *rmax, double *alpha, double *zcut,
                     uint32_t *r_in, uint32_t *r_out);

/* src/cubature.cu:161-193 */
void orc_legendre(double x, uint32_t n, double *p, double *dp);

/* src/cubature.cu:195-351: non-negative roots (ascending, 0 first for odd n)
 * by the reference's scan + Newton; weights 2/((1-x^2) P_n'^2). */
int orc_gauss_legendre(uint32_t n, double *x, double *w, uint32_t *count);

/* src/cubature.cu:353-402 */
int orc_select_gl(double tol, double dist, double alpha, double zcut,
                  uint32_t *order);

/* Half-sphere product rule of degree 2m+1 (nodes ordered by ascending
 * cos(theta), then azimuth). xyz: 3*count. */
int orc_sph_rule(uint32_t m, double *xyz, double *w, uint32_t *count);

/* src/cubature.cu:461-523 with the product rule in place of design files */
int orc_select_sph(uint32_t gl_order, double tol, double dist, double alpha,
                   double zcut, uint32_t *m);

/* src/cubature.cu:99-159.  Returns the node count in *K; arrays may be NULL
 * to query the size first (cap = capacity). */
int orc_build_nodes(double tol, uint32_t nx, uint32_t ny, uint32_t nz,
                    double cell, size_t cap, uint32_t *K, double *kx,
                    double *ky, double *kz, double *w, uint32_t *group);

/* src/glst_force.cu:223-241 in precision ct (0 double, 1 float) */
int orc_grid_dims(int ct, double lx, double ly, double lz, double rcut,
                  uint32_t *nx, uint32_t *ny, uint32_t *nz);

/* assign_atoms binning + stable sort (src/glst_force.cu:356-534). */
int orc_bin(int ct, const double *rx, const double *ry, const double *rz,
            uint32_t natom, uint32_t nx, uint32_t ny, uint32_t nz,
            double rcut, uint32_t *atom_cell, uint32_t *sorted_idx,
            uint32_t *cell_point, uint32_t *cell_count);

/* Full GLST step in FP64 (src/glst_force.cu:1604-1614), results in original
 * atom order.  Node table and groups are computed internally from tol.
 * node_stride > 1 keeps only every node_stride-th node (a bounded sample of
 * the long-range work for timing); far = 0 skips the long-range part, near =
 * 0 skips the near field.  Optional outputs S / R (ncell x K, cell-major,
 * re/im) when node_stride == 1. */
int orc_glst(int ct, const double *rx, const double *ry, const double *rz,
             const double *qc, uint32_t natom, double tol, double lx,
             double ly, double lz, double rcut, uint32_t node_stride,
             int far, int near, double *fx, double *fy, double *fz,
             double *en, double *s_re, double *s_im, double *r_re,
             double *r_im);

/* Wall seconds of the six stages of the last orc_glst call, in the order
 * the reference driver labels them (test/full_example.cu:215-233):
 * assign_atoms, calc_sf, sum_rmt_sf, calc_lr_ef, calc_sr_ef, get_ef. */
void orc_last_stage_times(double *out6);

/* Direct FP64 all-pairs Coulomb (utils/coulomb.cu:14-81). */
int orc_coulomb(const double *rx, const double *ry, const double *rz,
                const double *qc, uint32_t natom, double *fx, double *fy,
                double *fz, double *en);

/* Threads OpenMP will use (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
