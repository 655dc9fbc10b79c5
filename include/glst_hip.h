/*
 * glst_hip.h — C ABI of the MI355X-native GLST (Gauss-Legendre-Spherical-t)
 * electrostatics engine.
 *
 * This is the drop-in boundary for the reference's force object
 * `glst_force<CT>` (include/glst_force.hcu:24-128, src/glst_force.cu).  Every
 * entry point takes plain pointers and sizes; device pointers are HIP device
 * addresses on the current device.  Nothing here names a torch type.
 *
 * Physics (identical to the reference): open boundaries, bare 1/r, k_e = 1.
 * Per-atom energy en_i = q_i * phi_i (every pair counted on both atoms, so the
 * total energy is 0.5 * sum(en)); fx/fy/fz are forces (-grad E).
 *
 * Error convention: every function returns a glst_status_t.  A non-zero status
 * leaves a message retrievable with glst_last_error(handle) (or
 * glst_global_last_error() for functions without a handle).  The C++ facade in
 * glst_force.hpp rethrows these as std::runtime_error, which is what the
 * reference throws for host failures (src/glst_force.cu:42-45, 1633-1637;
 * src/cubature.cu:393-396, 427-428, 519-520).
 *
 * Threading: one host thread per handle, one HIP stream per handle; a handle is
 * not re-entrant.  Outputs accumulate (+=) exactly as in the reference
 * (src/glst_force.cu:1050-1053, 1238-1241); glst_zero_ef clears them.
 */
#ifndef GLST_HIP_H
#define GLST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum glst_status {
  GLST_OK = 0,
  GLST_ERR_INVALID = 1,  /* bad argument (null pointer, zero size, ...) */
  GLST_ERR_HIP = 2,      /* a HIP runtime call failed */
  GLST_ERR_NUMERICS = 3, /* erfc_inv / Gauss-Legendre / spherical rule failed */
  GLST_ERR_ALLOC = 4,    /* device allocation failed */
  GLST_ERR_IO = 5,       /* file could not be opened or parsed */
  GLST_ERR_STATE = 6     /* call out of order (e.g. calc_sf before assign) */
} glst_status_t;

/* Calculation type CT of the reference (include/glst_force.hcu:127-128).
 * GLST_CT_DOUBLE: plain FP64 path (CT = double).
 * GLST_CT_FLOAT : mixed path (CT = float): FP32 phases, S(k) accumulation and
 *                 pair math; FP64 range reduction of k.r (cell-origin phases),
 *                 FP64 S(k) fields, energies and per-atom force reductions. */
typedef enum glst_ct { GLST_CT_DOUBLE = 0, GLST_CT_FLOAT = 1 } glst_ct_t;

typedef struct glst_engine *glst_handle_t;

/* Setup summary, the values the reference prints at src/glst_force.cu:245-293. */
typedef struct glst_info {
  uint32_t natom;
  uint32_t ncell[3];
  uint32_t ncell_total;
  double cell_dim[3];   /* (double)(CT)rcut, as the reference stores it */
  uint32_t ngroup;
  uint32_t num_nodes;   /* total cubature nodes K */
  uint32_t has_long_range; /* 1 iff all ncell_d > 2 (src/glst_force.cu:667-670) */
  uint32_t max_atoms_cell; /* from the last assign_atoms (0 before) */
  int ct;               /* glst_ct_t */
  int spherical_rule;   /* 0 = product Gauss-Legendre x trapezoid, 1 = design files */
  size_t device_bytes;  /* device memory held by the handle */
} glst_info_t;

/* Per-group setup (group g of ngroup). */
typedef struct glst_group_info {
  uint32_t width;      /* shell width in cells */
  uint32_t r_in, r_out;/* Chebyshev shell (r_in, r_out] */
  double rmin, rmax, alpha, zcut;
  uint32_t gl_order;   /* Gauss-Legendre order n (even); n/2 positive roots */
  uint32_t sph_degree; /* spherical rule degree t = 2*m+1 */
  uint32_t num_dirs;   /* half-sphere directions M */
  uint32_t first_node; /* offset of the group's nodes in the table */
  uint32_t num_nodes;  /* (n/2) * M */
} glst_group_info_t;

/* ---- object lifetime ---------------------------------------------------- */

/* glst_force(natom, tol, Lx, Ly, Lz, rcut) + init():
 * include/glst_force.hcu:28-29, src/glst_force.cu:72-78, 178-327.
 * `stream` is a hipStream_t (NULL = the handle creates its own stream). */
glst_status_t glst_create(int ct, uint32_t natom, double tol, double box_x,
                          double box_y, double box_z, double rcut,
                          void *stream, glst_handle_t *out);

/* glst_force(natom, tol, Lx, Ly, Lz, ncell_x, ncell_y, ncell_z):
 * include/glst_force.hcu:30-33, src/glst_force.cu:80-95
 * (rcut = min_d L_d / n_d, then init). */
glst_status_t glst_create_ncell(int ct, uint32_t natom, double tol,
                                double box_x, double box_y, double box_z,
                                uint32_t ncell_x, uint32_t ncell_y,
                                uint32_t ncell_z, void *stream,
                                glst_handle_t *out);

/* ~glst_force(): src/glst_force.cu:97-100, 1768-1796. */
glst_status_t glst_destroy(glst_handle_t h);

/* Change the stream subsequent work is issued on (hipStream_t). */
glst_status_t glst_set_stream(glst_handle_t h, void *stream);

/* ---- the force step (src/glst_force.cu:1604-1614) ----------------------- */

/* assign_atoms(d_rx, d_ry, d_rz, d_qc): src/glst_force.cu:431-534.
 * Device FP64 SoA in original order, coordinates in Angstrom with minimum >= 0
 * (the caller recenters, utils/utils.cpp:35-58).  Inputs are borrowed and must
 * stay valid until the stream reaches the end of the step. */
glst_status_t glst_assign_atoms(glst_handle_t h, const double *d_rx,
                                const double *d_ry, const double *d_rz,
                                const double *d_qc);
/* calc_sf(): src/glst_force.cu:666-700. */
glst_status_t glst_calc_sf(glst_handle_t h);
/* sum_rmt_sf(): src/glst_force.cu:915-981. */
glst_status_t glst_sum_rmt_sf(glst_handle_t h);
/* calc_lr_ef(): src/glst_force.cu:1137-1168. */
glst_status_t glst_calc_lr_ef(glst_handle_t h);
/* calc_sr_ef(): src/glst_force.cu:1546-1582. */
glst_status_t glst_calc_sr_ef(glst_handle_t h);
/* comm_ef(): src/glst_force.cu:1584-1602.  One GPU holds the whole system, so
 * the cross-device reduction is the identity; like the reference it returns
 * only after the step's work has completed on the stream. */
glst_status_t glst_comm_ef(glst_handle_t h);
/* calc_ener_force(d_rx, d_ry, d_rz, d_qc): src/glst_force.cu:1604-1614.
 * Runs assign -> sf -> sum_rmt -> lr -> sr -> comm. */
glst_status_t glst_calc_ener_force(glst_handle_t h, const double *d_rx,
                                   const double *d_ry, const double *d_rz,
                                   const double *d_qc);
/* calc_ener_force without the final stream synchronisation (work is queued on
 * the handle's stream; results are valid once that stream is synchronised). */
glst_status_t glst_calc_ener_force_async(glst_handle_t h, const double *d_rx,
                                         const double *d_ry,
                                         const double *d_rz,
                                         const double *d_qc);

/* ---- results ------------------------------------------------------------ */

/* get_ef(fx, fy, fz, en): src/glst_force.cu:142-176.  Writes the accumulated
 * energies and forces in ORIGINAL atom order into caller-owned device arrays
 * of natom doubles each, then synchronises the stream. */
glst_status_t glst_get_ef(glst_handle_t h, double *d_fx, double *d_fy,
                          double *d_fz, double *d_en);
/* fx()/fy()/fz()/en() accessors: include/glst_force.hcu:36-43.  Device
 * pointers to the handle's accumulators, in cell-SORTED order. */
glst_status_t glst_sorted_ef_ptrs(glst_handle_t h, double **fx, double **fy,
                                  double **fz, double **en);
/* The driver's fx()[dev].set(0.0) (test/full_example.cu:201-209). */
glst_status_t glst_zero_ef(glst_handle_t h);
/* Original atom index of each sorted slot (sorted_idx_, device pointer). */
glst_status_t glst_sorted_index_ptr(glst_handle_t h, const uint32_t **idx);

/* ---- setup inspection --------------------------------------------------- */

glst_status_t glst_get_info(glst_handle_t h, glst_info_t *info);
glst_status_t glst_get_group_info(glst_handle_t h, uint32_t g,
                                  glst_group_info_t *info);
/* Copies the node table to host arrays of num_nodes entries (any may be NULL):
 * k vectors (1/Angstrom), weights W, group ids (cubature<CT>::x/y/z/w/group,
 * include/cubature.hcu:34-40). */
glst_status_t glst_get_nodes(glst_handle_t h, double *kx, double *ky,
                             double *kz, double *w, uint32_t *group);
/* Host copies of the binning result of the last assign_atoms:
 * atom_cell[natom] (cell of each ORIGINAL atom), sorted_idx[natom],
 * cell_point[ncell] and cell_count[ncell] (cell_atom_point_/count_). */
glst_status_t glst_get_cells(glst_handle_t h, uint32_t *atom_cell,
                             uint32_t *sorted_idx, uint32_t *cell_point,
                             uint32_t *cell_count);
/* Host copy of the structure factors S_c(k) = sum_j q_j exp(-i k.r_j) after
 * calc_sf (sf_re_/sf_im_, ncell x num_nodes, cell-major). */
glst_status_t glst_get_sf(glst_handle_t h, double *re, double *im);
/* Host copy of the remote sums R_c(k) = W_k * sum_{c' in shell} S_c'(k)
 * after sum_rmt_sf (rmt_sum_re_/rmt_sum_im_, ncell x num_nodes).  CT = float
 * keeps them on the device in polar form; they are converted back here. */
glst_status_t glst_get_rmt(glst_handle_t h, double *re, double *im);

/* Tuning options of a handle:
 *   "overlap_near" (default 1): calc_ener_force runs the near field on a
 *   second stream of the handle, started with the structure factors (1) or
 *   after them, overlapping the shell sums and the back-projection (2); 0
 *   runs the stages one after another.  The results are bitwise identical.
 *   "rmt_shape" (default 4): shell sums with LDS-resident summed-volume
 *   tables of as many nodes per workgroup as fit (a power of two up to 64)
 *   (0); always one node per workgroup (1); the global-memory tables
 *   otherwise used when one node's table does not fit in LDS (2); for
 *   CT = float on grids at most 16 cells wide in x and y, separable window
 *   sums on the matrix cores (FP32 direct sums; elsewhere as 0) (3); 3
 *   where nx * ny >= 128 (a z slice fills at least half of the matrix
 *   cores' 16 x 16 tile), else 5 where it applies, else 0 (4); for
 *   CT = float on grids at most 8 cells a side, FP32 window sums along
 *   lines held in registers, one wave per four nodes (elsewhere as 0) (5).
 *   Shapes 0-2 agree to FP64 rounding, shapes 3 and 5 with them to FP32
 *   rounding of the remote sums.
 *   "pipeline" (default 1): CT = float with LDS-resident tables only;
 *   calc_ener_force overlaps the stages of successive alpha groups: the
 *   shell sums of group g run on a third stream once S(k) of group g is
 *   complete, alongside S(k) of later groups and the back-projection of
 *   earlier ones (0: one stage after another).  Bitwise identical.
 *   "lr_blocks" (default 16384): back-projection workgroups aimed for; the
 *   direction ranges of each alpha group are split accordingly (on
 *   multiples of 8 directions).
 *   "lr_atom_groups" (default 1): CT = float back-projection over atom
 *   groups (up to 4 same-cell atoms share each remote-sum load, 8 lanes per
 *   group) (1), or one atom per lane (0).  Results agree to FP32 rounding.
 *   "sf_cheb" (default 1): CT = float structure factors as the Chebyshev
 *   moments of each cell's atoms times a matrix of Bessel values (the
 *   Jacobi-Anger expansion of exp(-i k.d) over the cell, truncated below
 *   3e-8), in the alpha groups where that takes fewer operations than the
 *   direct sum (1), or direct sums everywhere (0).  Results agree to FP32
 *   rounding.
 *   "graph" (default 0): calc_ener_force captures the whole step as a HIP
 *   graph for the input arrays of its first call and replays it while the
 *   same arrays are passed (per-kernel timing is then not recorded).
 * Every option leaves the results unchanged up to the FP rounding noted. */
glst_status_t glst_set_option(glst_handle_t h, const char *name,
                              int64_t value);
glst_status_t glst_get_option(glst_handle_t h, const char *name,
                              int64_t *value);

/* Autotuner (no counterpart in the reference), for CT = double handles:
 * times calc_ener_force on these inputs under candidate schedules, `reps`
 * steps each, and keeps the fastest, one knob at a time: "graph" (0, 1;
 * tuned first and again last), "overlap_near" (0, 1, 2), "lr_blocks"
 * (powers of two from 1024 to 65536), "rmt_shape" (0, 1).  Clears the
 * accumulators afterwards; *best_ms (may be NULL) receives the chosen
 * schedule's time per step.  Refused with GLST_ERR_STATE, the schedule
 * left as it is: CT = float handles and handles of more than 300,000
 * atoms (a host segmentation fault, seen once while tuning a 1.0M-charge
 * handle and once while tuning a 3,000-charge CT = float handle, is not
 * explained); set their schedule with glst_set_option. */
glst_status_t glst_autotune(glst_handle_t h, const double *d_rx,
                            const double *d_ry, const double *d_rz,
                            const double *d_qc, uint32_t reps,
                            double *best_ms);

/* Per-kernel timing with HIP events on the handle's stream.  When enabled,
 * every launch of the named kernel is bracketed by events; glst_kernel_time
 * returns the summed milliseconds and launch count since the last reset.
 * Names: "bin", "sort", "sf", "rmt", "lr", "lr_reduce", "sr", "unpermute". */
glst_status_t glst_set_timing(glst_handle_t h, int enable);
glst_status_t glst_kernel_time(glst_handle_t h, const char *name,
                               double *total_ms, uint32_t *launches);
glst_status_t glst_reset_timing(glst_handle_t h);

const char *glst_last_error(glst_handle_t h);
const char *glst_global_last_error(void);
const char *glst_version(void);

/* ---- direct Coulomb sum (utils/coulomb.cu:14-81, 84-172) ----------------- */

/* en_i = sum_{j!=i} q_i q_j / r_ij, F_i = sum_{j!=i} q_i q_j (r_i-r_j)/r_ij^3,
 * all pairs, FP64, free space.  Device arrays; outputs are overwritten.
 * Below about 524,000 atoms the source atoms are split into up to 64 ranges
 * whose partial sums (a temporary device buffer of at most 32 MiB, freed
 * before return) are added in range order.  Synchronises `stream`. */
glst_status_t glst_coulomb_direct(const double *d_rx, const double *d_ry,
                                  const double *d_rz, const double *d_qc,
                                  uint32_t natom, double *d_fx, double *d_fy,
                                  double *d_fz, double *d_en, void *stream);

/* No counterpart in the reference.  glst_coulomb_direct for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): the same kernels and split plan,
 * with the partial sums of the source splits in a fixed 32 MiB array of the
 * library (1,048,576 double4 entries per GPU, held until the process ends)
 * instead of a temporary buffer.  The plan never needs more than that
 * (natom * splits <= 1,048,576; larger systems run unsplit), and a plan
 * beyond it would be refused with GLST_ERR_INVALID.  Calls no hipMalloc or
 * hipFree and does not synchronise: the outputs are ready when `stream` is.
 * Every call on one GPU shares the array, so calls that overlap on one GPU
 * (other streams or host threads) race; serialise them. */
glst_status_t glst_coulomb_direct_noalloc(const double *d_rx,
                                          const double *d_ry,
                                          const double *d_rz,
                                          const double *d_qc, uint32_t natom,
                                          double *d_fx, double *d_fy,
                                          double *d_fz, double *d_en,
                                          void *stream);

/* ---- host numerics (no GPU needed) -------------------------------------- */

/* erfc_inv(y, tol): src/glst_force.cu:1616-1640. */
glst_status_t glst_erfc_inv(double y, double tol, double *x);
/* init_alpha_groups: src/glst_force.cu:1642-1724.  Writes at most `cap`
 * groups; *ngroup receives the count.  Arrays may be NULL when cap == 0. */
glst_status_t glst_alpha_groups(uint32_t nx, uint32_t ny, uint32_t nz,
                                double cdx, double cdy, double cdz,
                                double tol, uint32_t cap, uint32_t *ngroup,
                                uint32_t *width, double *rmax, double *alpha,
                                double *zcut, uint32_t *r_in,
                                uint32_t *r_out);
/* Positive Legendre roots and Gauss-Legendre weights of order n
 * (src/cubature.cu:161-351): n/2 values (n even), (n+1)/2 (n odd, root 0
 * first). */
glst_status_t glst_gauss_legendre(uint32_t n, double *x, double *w,
                                  uint32_t *count);
/* get_gauss_legendre order selection (src/cubature.cu:353-402): the smallest
 * even n in [2, 324] passing the radial test at distance `dist`. */
glst_status_t glst_select_gl_order(double tol, double dist, double alpha,
                                   double zcut, uint32_t *order);
/* Half-sphere spherical rule of degree 2m+1 used in place of the absent
 * t-design files: unit vectors (x,y,z interleaved, 3*count doubles) and
 * weights summing to 1.  Pass NULL arrays to query *count. */
glst_status_t glst_spherical_rule(uint32_t m, double *xyz, double *w,
                                  uint32_t *count);
/* get_tdesign_props + read_tdesign (src/cubature.cu:422-459): the first N/2
 * points of the symmetric design of degree t in `dir` (a file named
 * ...TTT.NNNNN, one point per line in three 25-character columns), as unit
 * vectors (x,y,z interleaved).  Among several files of degree t the one
 * whose name sorts first is read; other file names are skipped (the
 * reference takes the first in directory order and throws on names it
 * cannot parse).  Pass xyz = NULL to query *count; otherwise *count holds
 * the capacity in points on entry. */
glst_status_t glst_read_tdesign(const char *dir, uint32_t t, double *xyz,
                                uint32_t *count);
/* get_spherical_tdesign (src/cubature.cu:461-523) for the Gauss-Legendre
 * rule of even order gl_order: the degree index m (rule of degree 2m+1) the
 * three-axis test accepts at distance `dist`, from the design files of the
 * current directory (glst_set_tdesign_dir) or, with none set, the built-in
 * product rule; *count (may be NULL) receives its number of half-sphere
 * directions. */
glst_status_t glst_select_spherical_degree(uint32_t gl_order, double tol,
                                           double dist, double alpha,
                                           double zcut, uint32_t *m,
                                           uint32_t *count);
/* Directory of Womersley symmetric design files (ssTTT.NNNNN); when set (and
 * non-empty) new handles use those designs, as src/cubature.cu:422-541 does.
 * NULL or "" selects the built-in product rule.  The environment variable
 * GLST_TDESIGN_DIR sets the initial value. */
glst_status_t glst_set_tdesign_dir(const char *dir);

/* Grid dimensions exactly as init() derives them in precision CT
 * (src/glst_force.cu:223-241). */
glst_status_t glst_grid_dims(int ct, double box_x, double box_y, double box_z,
                             double rcut, uint32_t *nx, uint32_t *ny,
                             uint32_t *nz);

/* Synthetic rigid-water system (O -0.8476 e, H +0.4238 e, O-H 1.0 A,
 * H-O-H 109.47 deg): oxygen centres uniform in [0,L)^3, uniformly random
 * orientations, from SplitMix64-seeded xoshiro256**; then `nion` Na+ and
 * `nion` Cl- at uniform positions; finally recentred so each coordinate's
 * minimum is 0 (utils/utils.cpp:35-58).  Arrays hold 3*nwater + 2*nion
 * entries, host memory. */
glst_status_t glst_make_water_box(uint32_t nwater, uint32_t nion, double L,
                                  uint64_t seed, double *rx, double *ry,
                                  double *rz, double *qc);

#ifdef __cplusplus
}
#endif

#endif /* GLST_HIP_H */
