/*
 * glst_serve.h — serving loop over the batched engine.
 *
 * The reference has no counterpart (it computes one system per call of
 * glst_force::calc_ener_force, src/glst_force.cu:1604-1614).  A server takes
 * force requests for independent systems, each with the semantics of one
 * glst_force<CT> step (own box and rcut; the server's CT and tol), and
 * answers them by continuous batching: queued requests are admitted, oldest
 * first, into a batch as soon as one of the server's two batch slots is
 * free, so a new batch is packed and launched while the previous one still
 * runs on the device.  A batch holds at most max_batch requests and at most
 * half of the server's device-memory budget (two batches may be in flight).
 *
 * Grid setups (alpha groups, cubature and its device order) are cached by
 * (grid dimensions, cell edge), which is all they depend on, and batch
 * systems by (grid, cell edge, natom): repeat requests for a resident system
 * (an MD trajectory: same shape, new positions) skip all host setup, and a
 * batch with the same composition as its slot's previous batch reuses the
 * packed tables as well.  Requests whose own footprint exceeds a batch's
 * share of the budget are rejected at submission.
 *
 * Inputs and outputs are caller-owned device FP64 arrays of natom entries,
 * borrowed until the request completes; outputs are overwritten (fx, fy, fz,
 * en in the request's atom order).  One host thread drives a server.
 */
#ifndef GLST_SERVE_H
#define GLST_SERVE_H

#include "glst_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glst_server *glst_server_t;

typedef struct glst_server_stats {
  uint64_t submitted;     /* requests accepted */
  uint64_t rejected;      /* requests refused at submission (budget) */
  uint64_t completed;     /* requests answered */
  uint64_t batches;       /* glst_server_step calls that computed a batch */
  uint64_t reconfigures;  /* batches whose composition changed */
  uint64_t setups;        /* per-shape setups built */
  uint64_t setup_hits;    /* submissions served from the setup cache */
  uint64_t atoms;         /* atoms of all completed requests */
  uint64_t max_in_flight; /* most batches in flight at once (at most 2) */
} glst_server_stats_t;

/* budget_bytes: device memory the server's batches may hold together;
 * max_batch: most requests per batch (0 = no limit). */
glst_status_t glst_server_create(int ct, double tol, size_t budget_bytes,
                                 uint32_t max_batch, void *stream,
                                 glst_server_t *out);
glst_status_t glst_server_destroy(glst_server_t srv);

/* Queue one request; *ticket identifies it in glst_server_completed.
 * GLST_ERR_ALLOC: the request alone needs more than a batch's share of the
 * budget. */
glst_status_t glst_server_submit(glst_server_t srv, uint32_t natom,
                                 const double *box, double rcut,
                                 const double *d_rx, const double *d_ry,
                                 const double *d_rz, const double *d_qc,
                                 double *d_fx, double *d_fy, double *d_fz,
                                 double *d_en, uint64_t *ticket);

/* Launch batches of queued requests while a slot is free, retire finished
 * batches, and, if none had finished, wait for the oldest batch in flight.
 * *ncompleted = requests answered by this call (their outputs are
 * written). */
glst_status_t glst_server_step(glst_server_t srv, uint32_t *ncompleted);
/* The same without waiting: never blocks on the device. */
glst_status_t glst_server_poll(glst_server_t srv, uint32_t *ncompleted);

/* Tickets answered since the last call, oldest first (up to cap; *n set). */
glst_status_t glst_server_completed(glst_server_t srv, uint64_t *tickets,
                                    uint32_t cap, uint32_t *n);
/* Requests not yet answered (queued or in flight). */
glst_status_t glst_server_pending(glst_server_t srv, uint32_t *n);
glst_status_t glst_server_get_stats(glst_server_t srv,
                                    glst_server_stats_t *stats);
/* Options: "fused" (0 .. 3, default 0) and "graph" (0 .. 2, default 2) as
 * for glst_batch_set_option, for both batch slots (each slot's persistent
 * launch takes half of the device, so that the two may run at once; graph
 * mode 2 replays a slot's step while its composition repeats);
 * "max_batch" (requests per batch, 0 = no limit). */
glst_status_t glst_server_set_option(glst_server_t srv, const char *name,
                                     int64_t value);
const char *glst_server_last_error(glst_server_t srv);

#ifdef __cplusplus
}
#endif

#endif /* GLST_SERVE_H */
