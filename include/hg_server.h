/*
 * hg_server.h -- scene server: many independent MPM, DEM and hybrid scenes
 * kept resident on one GPU and stepped side by side (continuous batching).
 * No reference counterpart: the reference steps one scene per process.
 *
 * Requests are admitted oldest first while their device footprint
 * (hg_*_estimate_bytes) fits the memory budget and an active slot is free; a
 * request that does not fit is passed by smaller ones a bounded number of
 * times, then holds admission until it fits.  Each active scene runs on its
 * own stream in slices of slice_steps steps (hipGraph replays), so requests
 * that arrive while others run start as soon as memory and a slot free up and
 * their work overlaps on the device.  A finished scene's state is copied to
 * host memory, its device memory is released and its latency (submit to
 * completion) is recorded.  A scene whose state becomes invalid (a point off
 * the grid, a non-finite value) is retired alone with its error.
 *
 * Thread safety: all calls may be made from any thread; a worker thread of
 * the server makes every GPU call.
 */
#ifndef HG_SERVER_H
#define HG_SERVER_H

#include "hg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hg_server hg_server;

typedef struct hg_server_config {
  uint64_t budget_bytes;  /* device memory for resident scenes (0: 240 GiB) */
  uint32_t max_active;    /* resident scenes at once (0: 64) */
  uint32_t slice_steps;   /* steps queued per slice (0: 50) */
  uint32_t max_inflight;  /* slices queued per scene (0: 2) */
  uint32_t max_bypass;    /* times a waiting request may be passed (0: 8) */
} hg_server_config;

enum { HG_REQ_QUEUED = 0, HG_REQ_RUNNING = 1, HG_REQ_DONE = 2, HG_REQ_FAILED = 3,
       HG_REQ_CANCELLED = 4 };

typedef struct hg_server_stats {
  uint64_t submitted, completed, failed, cancelled, active, queued, bytes_in_use;
  double p50_latency_s, p99_latency_s, mean_latency_s, max_latency_s;
  double work_steps;  /* particle-steps + body-steps of completed requests */
  double span_s;      /* first submission to last completion */
} hg_server_stats;

hg_status hg_server_create(hg_ctx *ctx, const hg_server_config *cfg, hg_server **out);
/* Cancels what is still queued or running and stops the worker. */
hg_status hg_server_destroy(hg_server *s);

/* The descriptors' arrays are copied; the caller may free them on return.
 * nsteps steps of size dt are taken (the request's horizon). */
hg_status hg_server_submit_mpm(hg_server *s, const hg_mpm_desc *desc, double dt, uint64_t nsteps,
                               uint64_t *id);
hg_status hg_server_submit_dem(hg_server *s, const hg_dem_desc *desc, const double *weights,
                               double dt, uint64_t nsteps, uint64_t *id);
hg_status hg_server_submit_hybrid(hg_server *s, const hg_mpm_desc *mpm, const hg_dem_desc *dem,
                                  const double *weights, double dt, uint64_t nsteps,
                                  uint64_t *id);
hg_status hg_server_cancel(hg_server *s, uint64_t id);
/* state: HG_REQ_*; steps_done may be NULL. */
hg_status hg_server_poll(hg_server *s, uint64_t id, int32_t *state, uint64_t *steps_done);
/* Waits until the request ends or timeout_s passes (state is then still
 * queued or running). */
hg_status hg_server_wait(hg_server *s, uint64_t id, double timeout_s, int32_t *state);
/* Waits until nothing is queued or running; HG_ERR_STATE on timeout. */
hg_status hg_server_drain(hg_server *s, double timeout_s);
/* Results of a finished request: the continuum points (MPM, hybrid) and the
 * bodies (DEM, hybrid), in the request's creation order; NULL skips a field. */
hg_status hg_server_result_points(hg_server *s, uint64_t id, double *q, double *v,
                                  double *be_bar, double *J);
hg_status hg_server_result_bodies(hg_server *s, uint64_t id, double *q, double *v);
/* Submission-to-completion latency of a finished request, seconds. */
hg_status hg_server_latency(hg_server *s, uint64_t id, double *seconds);
/* The failure message of a failed request. */
hg_status hg_server_error(hg_server *s, uint64_t id, char *buf, size_t len);
/* Drops a finished request's stored results. */
hg_status hg_server_release(hg_server *s, uint64_t id);
hg_status hg_server_stats_get(hg_server *s, hg_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif
