/*
 * glst_batch.h — batched multi-system engine.
 *
 * The reference has no counterpart: it runs one system per glst_force object
 * (include/glst_force.hcu:24-128).  A batch holds many independent systems,
 * each with exactly the semantics of one glst_force<CT> (its own box, rcut,
 * cell grid, alpha groups and cubature; open boundaries; outputs accumulate),
 * and runs every stage of calc_ener_force for all of them in one launch per
 * stage (one per radial node count for S(k) and the back-projection) or, on
 * the mixed path, the whole step in one persistent launch (option "fused"),
 * so that small systems fill the GPU together.
 *
 * Atoms are packed: system s owns atoms [offset_s, offset_s + natom_s) of
 * every input and output array, offset_s = natom_0 + ... + natom_{s-1}
 * (glst_batch_atom_offset).  Inputs are device FP64 arrays in original order,
 * coordinates with minimum >= 0 per system (src/glst_force.cu:431-534);
 * outputs are device FP64 arrays in the same packed original order.
 *
 * Errors follow glst_hip.h: status codes plus glst_batch_last_error (or
 * glst_global_last_error for calls without a batch).
 */
#ifndef GLST_BATCH_H
#define GLST_BATCH_H

#include "glst_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glst_batch *glst_batch_t;

/* Create a batch of nsys systems sharing ct and tol.  natom[s] >= 1 atoms,
 * box[3s..3s+2] the box edges and rcut[s] the cell edge of system s.
 * `stream` is a hipStream_t (NULL = the batch creates its own).  Every
 * system's cell table of one node must fit in LDS (about 10,000 cells);
 * larger systems belong in a glst_create handle. */
glst_status_t glst_batch_create(int ct, uint32_t nsys, const uint32_t *natom,
                                const double *box, const double *rcut,
                                double tol, void *stream, glst_batch_t *out);
glst_status_t glst_batch_destroy(glst_batch_t b);

/* Device bytes a batch of these systems would hold, computed on the host
 * without touching a GPU (admission control). */
glst_status_t glst_batch_device_bytes(int ct, uint32_t nsys,
                                      const uint32_t *natom, const double *box,
                                      const double *rcut, double tol,
                                      size_t *bytes);

glst_status_t glst_batch_num_systems(glst_batch_t b, uint32_t *nsys);
glst_status_t glst_batch_atom_offset(glst_batch_t b, uint32_t sys,
                                     uint32_t *offset);
/* Setup summary of system `sys` (device_bytes: the whole batch). */
glst_status_t glst_batch_get_info(glst_batch_t b, uint32_t sys,
                                  glst_info_t *info);

/* calc_ener_force for every system (src/glst_force.cu:1604-1614 per
 * system); returns after the work is complete. */
glst_status_t glst_batch_calc_ener_force(glst_batch_t b, const double *d_rx,
                                         const double *d_ry,
                                         const double *d_rz,
                                         const double *d_qc);
/* The same without the final stream synchronisation. */
glst_status_t glst_batch_calc_ener_force_async(glst_batch_t b,
                                               const double *d_rx,
                                               const double *d_ry,
                                               const double *d_rz,
                                               const double *d_qc);
/* Accumulated energies and forces of all systems in packed original order
 * (get_ef, src/glst_force.cu:142-176, per system); synchronises. */
glst_status_t glst_batch_get_ef(glst_batch_t b, double *d_fx, double *d_fy,
                                double *d_fz, double *d_en);
glst_status_t glst_batch_zero_ef(glst_batch_t b);

/* Options: "overlap_near" (0 or 1, default 1) as for glst_set_option;
 * "lr_blocks" (default 4096): back-projection workgroups aimed for over the
 * whole batch (re-packs the batch); "graph" (default 0): 1 captures the
 * whole step as a HIP graph for the input arrays of the first step and
 * replays it while they are passed again, 2 captures only on the second
 * consecutive step with the same inputs; "fused" (default 0; mixed path):
 * 0 launches S(k) and the back-projection once per radial node count, 3
 * once per stage for all counts, and 1 or 2 runs the whole step of all
 * systems as one persistent launch whose phases (binning, S(k) with the near
 * field, shell sums, back-projection, reduction) are separated by grid-wide
 * barriers, 1 with the kernel's registers held to 2 waves per SIMD and 2 to
 * 3 waves.  Its workgroups are as many as are resident at once; should a
 * barrier nevertheless wait 10 s, the step ends and the next synchronising
 * call returns GLST_ERR_STATE.  Mixed path, each re-packs the batch:
 * "sf_cheb" (default 1) S(k) in Chebyshev form where the grid shapes
 * planned it, 0 direct sums; "rmt_wave" (default 1) the shell sums of grids
 * up to 8 cells a side as FP32 window sums, 0 the LDS summed-volume tables.
 * The persistent launch (fused 1, 2) sums directly and on the tables. */
glst_status_t glst_batch_set_option(glst_batch_t b, const char *name,
                                    int64_t value);
/* Autotuner: times the step on these inputs under each "fused" (mixed
 * path), "overlap_near" and "lr_blocks" candidate (`reps` steps each, one
 * knob at a time), keeps the
 * fastest and clears the accumulators; *best_ms (may be NULL) receives its
 * time per step. */
glst_status_t glst_batch_autotune(glst_batch_t b, const double *d_rx,
                                  const double *d_ry, const double *d_rz,
                                  const double *d_qc, uint32_t reps,
                                  double *best_ms);
glst_status_t glst_batch_set_stream(glst_batch_t b, void *stream);
const char *glst_batch_last_error(glst_batch_t b);

#ifdef __cplusplus
}
#endif

#endif /* GLST_BATCH_H */
