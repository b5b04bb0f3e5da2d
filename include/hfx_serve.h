/*
 * hfx_serve.h - serving loop for concurrent acquisition streams on one GPU.
 *
 * No counterpart in the reference's hot path: the reference runs one acquisition per
 * process (Manager::start_pipeline, manager.cc), feeding one BatchQueue ring
 * (batch_queue.cc:92-174) per graph.  Here a server holds many streams at once: each admitted
 * stream owns a device-resident frame ring, a moment plan (hfx.h) with its propagation
 * tables, and an output slot per window position; its windows are replayed from hipGraphs
 * captured once per window position, on one of several HIP streams ("lanes") so that the
 * windows of different streams overlap on the device.
 *
 * Flow: admit (by free HBM and a free slot) -> ingest frames (synthetic camera into the
 * ring; refused while it would overwrite frames of windows not yet retired) -> schedule
 * (launch every ready window of every stream) -> poll (retire finished windows in order,
 * with their frame-to-moment latency) -> release.  Setting a stream up and tearing it down
 * (allocations, tables, graph capture, frees) run on a worker thread of the server, so the
 * thread that serves windows never waits for them.
 *
 * Latency of a window = time from its due time (the camera time of its last frame, given to
 * hfx_server_ingest_due, or the host time of the hfx_server_ingest call) to the completion
 * of its moments on the GPU.  GPU completion times are read on the server's host clock
 * through an epoch event recorded at hfx_server_create; they are late by at most
 * hfx_server_info.clock_error_us, never early, so a latency is never under-reported.  Time a
 * window waits for ring space before it can be ingested counts, since its due time is kept.
 *
 * One host thread drives a server; calls on one server must not overlap (the server's own
 * worker thread is internal).
 */
#ifndef HFX_SERVE_H_
#define HFX_SERVE_H_

#include <stdint.h>

#include "hfx.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HFX_NOT_READY
#define HFX_NOT_READY 5 /* retry later: no slot / memory (admit), ring full (ingest) */
#endif

typedef struct hfx_server hfx_server;

typedef struct hfx_server_desc {
  int64_t hbm_budget_bytes;  /* device bytes streams may hold; 0 = free HBM minus reserve */
  int64_t hbm_reserve_bytes; /* device bytes always left free */
  int32_t max_streams;       /* concurrent stream slots */
  int32_t n_lanes;           /* HIP streams for window work (default priority) */
  int32_t use_graphs;        /* 1: replay window positions from captured hipGraphs */
  int32_t policy;            /* enum hfx_sched_policy */
  int32_t lane_depth;        /* HFX_SCHED_EDF: windows queued per lane (>= 1) */
} hfx_server_desc;

/* Launch policy of hfx_server_schedule.
 * FIFO: every ready window goes at once to its stream's lane, oldest due time first across
 *       the streams (a stream's windows keep their order on its lane).
 * EDF: ready windows wait on the host; whenever a lane holds fewer than lane_depth windows,
 *      the ready window with the earliest deadline (due time + the stream's window period
 *      stride / fs_hz) of a stream with no window on the GPU is launched there.  A stream has
 *      at most one window in flight (its plan workspace), so its windows keep their order.
 * GROUP: the ready window with the oldest due time goes out together with the oldest ready
 *       windows of up to HFX_GROUP_MAX - 1 other streams of the same plan shape (frame size,
 *       window, band, weights, propagation kind; z, wavelength, pitch may differ) as one
 *       hfx_moments_group launch, on the lanes in turn; each window waits for its stream's
 *       previous one, so a stream's windows still run in order.  Single windows replay their
 *       captured graphs.
 * FIFO is the default: under the mixed 70 %-load workload EDF kept the same share of windows
 * within their period but doubled p99 (bursty streams wait on their own single slot). */
enum hfx_sched_policy { HFX_SCHED_FIFO = 0, HFX_SCHED_EDF = 1, HFX_SCHED_GROUP = 2 };

typedef struct hfx_stream_desc {
  hfx_plan_desc plan;        /* the stream's moment plan (frame size, T, S, band, optics) */
  int32_t queue_windows;     /* ready-but-unretired windows the ring holds (ring = T + q*S frames) */
  hfx_synth_desc camera;     /* synthetic camera feeding the ring */
} hfx_stream_desc;

typedef struct hfx_window_done {
  int32_t stream_id;
  int32_t n_outputs;         /* images per window (3 or 5) */
  int64_t window;            /* window index within the stream (0, 1, ...) */
  float latency_us;          /* due time (last frame delivered) -> moments complete */
  float gpu_us;              /* launch -> complete on its lane */
  float wait_us;             /* due time -> start of its work on its lane */
} hfx_window_done;

typedef struct hfx_server_info {
  int32_t active_streams;
  int32_t max_streams;
  int64_t bytes_held;        /* device bytes held by admitted streams */
  int64_t bytes_free;        /* hipMemGetInfo free bytes */
  int64_t windows_launched;
  int64_t windows_retired;
  double clock_error_us;     /* bound on how late GPU times read on the server clock are */
  int64_t group_launches;    /* HFX_SCHED_GROUP: launches that carried windows of several streams */
  int64_t grouped_windows;   /* windows those launches carried */
} hfx_server_info;

void hfx_server_desc_init(hfx_server_desc *d);
void hfx_stream_desc_init(hfx_stream_desc *d, int32_t height, int32_t width);
/* Device bytes a stream of this description holds once admitted (no device work). */
hfx_status hfx_server_stream_bytes(const hfx_stream_desc *d, int64_t *bytes);

hfx_status hfx_server_create(const hfx_server_desc *d, hfx_server **out);
/* Waits for all work, frees every stream. */
hfx_status hfx_server_destroy(hfx_server *s);
/* HFX_NOT_READY when no slot is free or the stream does not fit the memory budget now.
 * On HFX_OK the slot and memory are the stream's; its device state (allocations, plan tables,
 * the camera's source pool, captured graphs) is built by the server's worker thread, so the
 * call returns at once: the stream takes frames once hfx_server_stream_state reports 1. */
hfx_status hfx_server_admit(hfx_server *s, const hfx_stream_desc *d, int32_t *stream_id);
/* *state: 0 being set up, 1 ready, 2 set-up failed (the call then returns HFX_E_HIP with the
 * reason; release the stream). */
hfx_status hfx_server_stream_state(hfx_server *s, int32_t stream_id, int32_t *state);
/* Seconds on the server's clock (host steady clock, zero at hfx_server_create). */
hfx_status hfx_server_clock(hfx_server *s, double *t_s);
/* The camera writes the next n_frames frames of the stream into its ring; HFX_NOT_READY
 * (nothing written) while that would overwrite frames of a window not yet retired.
 * hfx_server_ingest_due stamps the windows these frames complete with due_s (server clock):
 * the time the camera delivered the last of them; hfx_server_ingest stamps them with the time
 * of the call. */
hfx_status hfx_server_ingest(hfx_server *s, int32_t stream_id, int64_t n_frames);
hfx_status hfx_server_ingest_due(hfx_server *s, int32_t stream_id, int64_t n_frames, double due_s);
/* Launches every window whose frames are in (all streams); *launched = count. */
hfx_status hfx_server_schedule(hfx_server *s, int32_t *launched);
/* Retires finished windows of all streams, oldest first per stream: up to cap records. */
hfx_status hfx_server_poll(hfx_server *s, hfx_window_done *out, int32_t cap, int32_t *n);
/* Device pointer to the moments of a retired window ([n_out][OH][OW] float); valid until
 * that stream ingests the frames of window + queue_windows + T/S. */
hfx_status hfx_server_window_output(hfx_server *s, int32_t stream_id, int64_t window, const float **out);
/* Returns the slot at once; the worker thread waits for the stream's last copies and kernels
 * and frees its memory (admission sees that memory once it is freed). */
hfx_status hfx_server_release(hfx_server *s, int32_t stream_id);
hfx_status hfx_server_info_get(hfx_server *s, hfx_server_info *info);

/* ------------------------------------------------------------------------------------
 * Native workload driver: streams arrive at given times; their cameras deliver windows.
 * Stream i arrives at arrival_s[i] (seconds from start) and waits until admitted; then its
 * camera delivers windows at window_rate_hz[i]: PERIODIC, one stride of frames every
 * 1/rate (window 0 due T/S strides after admission), as a camera running at
 * rate * stride frames/s does; POISSON, exponential gaps of mean 1/rate.  Each window is
 * ingested at or after its due time (later while the ring is full) and its latency counts
 * from the due time.  After windows_per_stream windows are retired the stream is released.
 * The loop admits, ingests, schedules and polls continuously on the calling thread.
 * ---------------------------------------------------------------------------------- */
enum hfx_arrivals { HFX_ARRIVALS_PERIODIC = 0, HFX_ARRIVALS_POISSON = 1 };

typedef struct hfx_workload_desc {
  hfx_server_desc server;
  int32_t n_streams;
  const hfx_stream_desc *streams;  /* n_streams */
  const double *arrival_s;         /* n_streams, non-decreasing */
  const double *window_rate_hz;    /* n_streams */
  int32_t windows_per_stream;
  uint64_t seed;                   /* window inter-arrival draws */
  int32_t capture_stream;          /* index whose window outputs are copied to capture_out, or -1 */
  float *capture_out;              /* host, windows_per_stream * n_out * OH * OW floats */
  double time_limit_s;             /* the loop stops (HFX_E_HIP) if exceeded */
  int32_t arrivals;                /* enum hfx_arrivals: window arrival process */
} hfx_workload_desc;

typedef struct hfx_workload_result {
  double wall_s;                   /* first arrival -> last retirement */
  int64_t windows_retired;
  int32_t max_concurrent;          /* most streams admitted at once */
  int32_t admissions_deferred;     /* admit attempts answered HFX_NOT_READY */
  int64_t ingest_deferred;         /* ingest attempts answered HFX_NOT_READY */
  float *latency_us;               /* caller array n_streams * windows_per_stream (or NULL) */
  float *gpu_us;                   /* same shape (or NULL) */
  double *admit_s;                 /* caller array n_streams (or NULL): admission times */
  double *finish_s;                /* caller array n_streams (or NULL): release times */
  float *wait_us;                  /* caller array n_streams * windows_per_stream (or NULL) */
  int64_t windows_ingested_late;   /* windows ingested more than one period after their due time */
  double clock_error_us;           /* the server's GPU-time bound (hfx_server_info.clock_error_us) */
  int64_t group_launches;          /* as in hfx_server_info */
  int64_t grouped_windows;
} hfx_workload_result;

hfx_status hfx_serve_workload(const hfx_workload_desc *d, hfx_workload_result *r);

#ifdef __cplusplus
}
#endif

#endif /* HFX_SERVE_H_ */
