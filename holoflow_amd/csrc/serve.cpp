// Serving loop for concurrent acquisition streams on one GPU (include/hfx_serve.h).
//
// Per admitted stream the device holds: a frame ring of R = T + q*S frames (q = queued
// windows), a source pool of P = T frames the synthetic camera plays cyclically (generated
// once at set-up by hfx_synth_frames, so ingest is a device-to-device copy, as a frame
// grabber's DMA would be; the pool is kept to one window because generating it is GPU work
// that a real camera does not add, and a 4T pool measurably delayed the windows of the
// streams already being served), the moment plan (tables + workspace) and one output slot per window
// position (n_pos = R / S positions).  Window w reads ring frames [w*S, w*S + T) mod R,
// which for R = n_pos*S is position p = w mod n_pos starting at p*S: every position is
// captured once into a hipGraph (front + back kernels of hfx_moments) and replayed.
//
// Set-up and tear-down of a stream (device allocations, plan tables, the camera's source
// pool, graph capture; waiting for its last work, frees) run on the server's worker thread,
// so the serving loop never blocks on them: a stream is usable once its state is ready.
//
// Ordering: the camera copies on one ingest stream and records `ready[p]` after the last
// frame of window w; the window's lane waits on it, records `launch[p]`, replays the graph
// and records `done[p]`.  Ingest refuses (HFX_NOT_READY) to overwrite frames that a window
// not yet retired reads, and a window is retired only after its done event completed, so
// no frame is overwritten while read and no event is re-recorded while still needed.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "hfx.h"
#include "hfx_serve.h"

namespace hfx {
void set_last_error(const std::string &msg);  // libhfx's thread-local last error (api.hip)
}

namespace {

hfx_status fail(hfx_status code, const std::string &msg) {
  hfx::set_last_error(msg);
  return code;
}

#define SRV_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(HFX_E_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

#define SRV_TRY(expr)            \
  do {                           \
    hfx_status s_ = (expr);      \
    if (s_ != HFX_OK) return s_; \
  } while (0)

struct Geometry {
  int64_t T, S, R, n_pos, P, npix, out_elems, n_planes, oh, ow, n_out;
  int64_t ring_bytes, pool_bytes, out_bytes, plan_bytes;
  int64_t total() const { return ring_bytes + pool_bytes + out_bytes + plan_bytes + npix * 8; }
};

hfx_status geometry(const hfx_stream_desc &d, Geometry &g) {
  const hfx_plan_desc &p = d.plan;
  if (p.window <= 0 || p.stride <= 0 || p.height <= 0 || p.width <= 0)
    return fail(HFX_E_INVALID, "stream: frame size, window and stride must be positive");
  if (p.window % p.stride != 0) return fail(HFX_E_INVALID, "stream: window must be a multiple of stride");
  if (d.queue_windows < 1) return fail(HFX_E_INVALID, "stream: queue_windows must be >= 1");
  if (d.camera.height != p.height || d.camera.width != p.width)
    return fail(HFX_E_INVALID, "stream: camera frame size differs from the plan's");
  g.T = p.window;
  g.S = p.stride;
  g.n_pos = g.T / g.S + d.queue_windows;
  g.R = g.n_pos * g.S;
  g.P = g.T;  // the synthetic camera plays a pool of T frames cyclically
  g.npix = (int64_t)p.height * p.width;
  g.oh = p.pad_height > 0 ? p.pad_height : p.height;
  g.ow = p.pad_width > 0 ? p.pad_width : p.width;
  g.n_out = p.split_bin > 0 ? 5 : 3;
  g.out_elems = g.n_out * g.oh * g.ow;
  const int64_t nb = (int64_t)p.bin_end - p.bin_begin;
  g.n_planes = (p.flags & HFX_FLAG_BOTH_SIDEBANDS) ? 2 * nb : nb;
  g.ring_bytes = g.R * g.npix * 2;
  g.pool_bytes = g.P * g.npix * 2;
  g.out_bytes = g.n_pos * g.out_elems * 4;
  // plan workspace upper bound: blocked + natural band planes, transfer table, small tables
  g.plan_bytes = 2 * g.n_planes * g.oh * g.ow * 8 + g.oh * g.ow * 8 + (1 << 20);
  return HFX_OK;
}

struct Stream {
  int32_t id = -1;
  hfx_stream_desc d{};
  Geometry g{};
  int lane = 0;
  bool groupable = false;  // its windows can join grouped launches (hfx_moments_group)
  hfx_plan *plan = nullptr;
  uint16_t *ring = nullptr, *pool = nullptr;
  float *outs = nullptr;
  double *phase = nullptr;
  std::vector<hipGraphExec_t> graphs;
  std::vector<hipEvent_t> ev_ready, ev_launch, ev_done;
  std::vector<double> due_t;  // server-clock time (s) the camera delivered each position's last frame
  int64_t frames_in = 0, launched = 0, retired = 0;
  // set up by the worker thread: 0 building, 1 ready, 2 failed (error holds the reason); the
  // worker writes every other member before it stores 1 or 2
  std::atomic<int> state{0};
  std::string error;

  int64_t ready_windows() const { return frames_in >= g.T ? (frames_in - g.T) / g.S + 1 : 0; }

  void free_all() {
    for (auto &x : graphs)
      if (x) (void)hipGraphExecDestroy(x);
    for (auto *v : {&ev_ready, &ev_launch, &ev_done})
      for (auto &e : *v)
        if (e) (void)hipEventDestroy(e);
    graphs.clear();
    ev_ready.clear();
    ev_launch.clear();
    ev_done.clear();
    if (plan) hfx_plan_destroy(plan);
    plan = nullptr;
    for (void *p : {(void *)ring, (void *)pool, (void *)outs, (void *)phase})
      if (p) (void)hipFree(p);
    ring = pool = nullptr;
    outs = nullptr;
    phase = nullptr;
  }
};

}  // namespace

struct hfx_server {
  hfx_server_desc d{};
  int device = 0;
  hipStream_t ingest = nullptr, capture = nullptr;
  std::vector<hipStream_t> lanes;
  std::vector<std::unique_ptr<Stream>> slots;
  std::vector<int> lane_load;
  std::vector<std::deque<std::pair<int32_t, int64_t>>> lane_windows;  // EDF: (stream, window) per lane
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  int64_t bytes_held = 0, windows_launched = 0, windows_retired = 0, group_launches = 0, grouped_windows = 0;
  int next_lane = 0;  // HFX_SCHED_GROUP: lanes taken in turn
  // worker thread: builds admitted streams (their objects stay in `slots`) and tears released
  // ones down (their objects move into the job)
  struct Job {
    Stream *build = nullptr;
    std::unique_ptr<Stream> teardown;
  };
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> jobs;
  bool stopping = false;
  // GPU timestamps on the server clock: `epoch` was recorded on an idle stream and had
  // completed by host time epoch_s, so event e completed at or before
  // epoch_s + elapsed(epoch, e), late by at most epoch_err_s (the recording's round trip)
  hipEvent_t epoch = nullptr;
  double epoch_s = 0.0, epoch_err_s = 0.0;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  hipError_t gpu_time(hipEvent_t e, double &t) const {
    float ms = 0.f;
    const hipError_t r = hipEventElapsedTime(&ms, epoch, e);
    t = epoch_s + 1e-3 * (double)ms;
    return r;
  }
};

namespace {

Stream *stream_of(hfx_server *s, int32_t id) {
  if (!s || id < 0 || id >= (int32_t)s->slots.size()) return nullptr;
  return s->slots[id].get();
}

hfx_status copy_frames(hfx_server *s, Stream &st, int64_t first, int64_t n) {
  // frame f of the stream = pool frame f mod P, written to ring slot f mod R
  const Geometry &g = st.g;
  const size_t fb = (size_t)g.npix * 2;
  int64_t f = first;
  while (f < first + n) {
    const int64_t rs = f % g.R, ps = f % g.P;
    const int64_t len = std::min({first + n - f, g.R - rs, g.P - ps});
    SRV_HIP(hipMemcpyAsync(st.ring + rs * g.npix, st.pool + ps * g.npix, (size_t)len * fb,
                           hipMemcpyDeviceToDevice, s->ingest));
    f += len;
  }
  return HFX_OK;
}

hfx_status build_stream(hfx_server *s, Stream &st) {
  const Geometry &g = st.g;
  hipStream_t cs = s->capture;
  hfx_plan_desc pd = st.d.plan;
  pd.batch_lanes = 1;  // windows run one at a time through hfx_moments: no second lane
  SRV_TRY(hfx_plan_create(&pd, cs, &st.plan));
  hfx_plan_info info{};
  SRV_TRY(hfx_plan_get_info(st.plan, &info));
  st.groupable = info.groupable != 0;
  SRV_HIP(hipMalloc((void **)&st.ring, (size_t)g.ring_bytes));
  SRV_HIP(hipMalloc((void **)&st.pool, (size_t)g.pool_bytes));
  SRV_HIP(hipMalloc((void **)&st.outs, (size_t)g.out_bytes));
  SRV_HIP(hipMalloc((void **)&st.phase, (size_t)g.npix * 8));
  SRV_HIP(hipMemsetAsync(st.phase, 0, (size_t)g.npix * 8, cs));
  SRV_TRY(hfx_synth_frames(&st.d.camera, st.pool, 0, g.P, st.phase, cs));
  SRV_HIP(hipStreamSynchronize(cs));
  st.due_t.assign(g.n_pos, 0.0);
  st.ev_ready.assign(g.n_pos, nullptr);
  st.ev_launch.assign(g.n_pos, nullptr);
  st.ev_done.assign(g.n_pos, nullptr);
  for (int64_t p = 0; p < g.n_pos; ++p) {
    SRV_HIP(hipEventCreate(&st.ev_ready[p]));
    SRV_HIP(hipEventCreate(&st.ev_launch[p]));
    SRV_HIP(hipEventCreate(&st.ev_done[p]));
  }
  if (s->d.use_graphs) {
    st.graphs.assign(g.n_pos, nullptr);
    for (int64_t p = 0; p < g.n_pos; ++p) {
      hipGraph_t graph = nullptr;
      SRV_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      const hfx_status rc = hfx_moments(st.plan, st.ring, g.R, p * g.S, st.outs + p * g.out_elems, cs);
      const hipError_t e = hipStreamEndCapture(cs, &graph);
      if (rc != HFX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      SRV_HIP(e);
      const hipError_t ie = hipGraphInstantiate(&st.graphs[p], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      SRV_HIP(ie);
    }
  }
  return HFX_OK;
}

hfx_status launch_window(hfx_server *s, Stream &st, int lane_idx) {
  const int64_t w = st.launched;
  const int64_t p = w % st.g.n_pos;
  hipStream_t lane = s->lanes[lane_idx];
  SRV_HIP(hipStreamWaitEvent(lane, st.ev_ready[p], 0));
  SRV_HIP(hipEventRecord(st.ev_launch[p], lane));
  if (s->d.use_graphs) {
    SRV_HIP(hipGraphLaunch(st.graphs[p], lane));
  } else {
    SRV_TRY(hfx_moments(st.plan, st.ring, st.g.R, p * st.g.S, st.outs + p * st.g.out_elems, lane));
  }
  SRV_HIP(hipEventRecord(st.ev_done[p], lane));
  ++st.launched;
  ++s->windows_launched;
  return HFX_OK;
}

// Tear-down of a released stream: its windows are retired, so only its last ingest copies
// and kernels may still run; wait for them, then free (hipFree may wait for the device).
void teardown_stream(hfx_server *s, Stream &st) {
  (void)hipStreamSynchronize(s->ingest);
  if (st.launched > 0 && !st.ev_done.empty()) (void)hipEventSynchronize(st.ev_done[(st.launched - 1) % st.g.n_pos]);
  st.free_all();
}

void worker_main(hfx_server *s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    hfx_server::Job job;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stopping || !s->jobs.empty(); });
      if (s->jobs.empty()) return;  // stopping, nothing left
      job = std::move(s->jobs.front());
      s->jobs.pop_front();
    }
    if (job.build) {
      Stream &st = *job.build;
      const hfx_status rc = build_stream(s, st);
      if (rc != HFX_OK) {
        st.error = hfx_last_error();
        (void)hipStreamSynchronize(s->capture);
        st.free_all();
      }
      st.state.store(rc == HFX_OK ? 1 : 2, std::memory_order_release);
    } else if (job.teardown) {
      teardown_stream(s, *job.teardown);
    }
  }
}

void post_job(hfx_server *s, hfx_server::Job job) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->jobs.push_back(std::move(job));
  }
  s->cv.notify_one();
}

bool ready(const Stream &st) { return st.state.load(std::memory_order_acquire) == 1; }

// Windows of these two streams can share a grouped launch: the same plan shape (their z,
// wavelength, pitch, cameras and rings may differ); hfx_moments_group checks the same.
bool same_group(const Stream &a, const Stream &b) {
  const hfx_plan_desc &x = a.d.plan, &y = b.d.plan;
  return a.groupable && b.groupable && x.height == y.height && x.width == y.width && x.window == y.window &&
         x.bin_begin == y.bin_begin && x.bin_end == y.bin_end && x.prop == y.prop && x.precision == y.precision &&
         x.flags == y.flags && x.pad_height == y.pad_height && x.pad_width == y.pad_width &&
         x.split_bin == y.split_bin && x.weight_fs_hz == y.weight_fs_hz && x.filter_enabled == y.filter_enabled &&
         x.flatfield_enabled == y.flatfield_enabled;
}

// The next window of each member stream, in one launch on `lane`: each waits for its frames
// and for its stream's previous window (which may have run on another lane: one plan
// workspace per stream), a single window replays its captured graph, several run as one
// hfx_moments_group launch.
hfx_status launch_group(hfx_server *s, const std::vector<Stream *> &members, int lane_idx) {
  hipStream_t lane = s->lanes[lane_idx];
  const int n = (int)members.size();
  std::vector<hfx_plan *> plans(n);
  std::vector<const uint16_t *> rings(n);
  std::vector<int64_t> rframes(n), firsts(n);
  std::vector<float *> outs(n);
  for (int i = 0; i < n; ++i) {
    Stream &st = *members[i];
    const int64_t w = st.launched, p = w % st.g.n_pos;
    SRV_HIP(hipStreamWaitEvent(lane, st.ev_ready[p], 0));
    if (w > 0) SRV_HIP(hipStreamWaitEvent(lane, st.ev_done[(w - 1) % st.g.n_pos], 0));
    SRV_HIP(hipEventRecord(st.ev_launch[p], lane));
    plans[i] = st.plan;
    rings[i] = st.ring;
    rframes[i] = st.g.R;
    firsts[i] = p * st.g.S;
    outs[i] = st.outs + p * st.g.out_elems;
  }
  if (n == 1 && s->d.use_graphs) {
    SRV_HIP(hipGraphLaunch(members[0]->graphs[members[0]->launched % members[0]->g.n_pos], lane));
  } else {
    SRV_TRY(hfx_moments_group(plans.data(), rings.data(), rframes.data(), firsts.data(), outs.data(), n, lane));
  }
  for (int i = 0; i < n; ++i) {
    Stream &st = *members[i];
    SRV_HIP(hipEventRecord(st.ev_done[st.launched % st.g.n_pos], lane));
    ++st.launched;
    ++s->windows_launched;
  }
  if (n > 1) {
    ++s->group_launches;
    s->grouped_windows += n;
  }
  return HFX_OK;
}

// HFX_SCHED_GROUP: repeatedly take the ready window with the oldest due time and launch it
// together with the oldest ready windows of up to HFX_GROUP_MAX - 1 other streams of the same
// plan shape, on the next lane.
hfx_status schedule_group(hfx_server *s, int32_t &n) {
  for (;;) {
    std::vector<std::pair<double, Stream *>> rdy;
    for (auto &up : s->slots) {
      if (!up || !ready(*up)) continue;
      Stream &st = *up;
      if (st.launched < st.ready_windows()) rdy.emplace_back(st.due_t[st.launched % st.g.n_pos], &st);
    }
    if (rdy.empty()) return HFX_OK;
    std::stable_sort(rdy.begin(), rdy.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    std::vector<Stream *> members{rdy[0].second};
    // group size: four windows for frames up to 768^2 (their per-window cost drops by 30-40 %
    // in groups of four), two beyond (1024^2 gains nothing from four: the planes of four
    // windows no longer stay in the caches)
    const int cap = members[0]->g.npix <= 768 * 768 ? HFX_GROUP_MAX : std::min(2, HFX_GROUP_MAX);
    for (size_t i = 1; i < rdy.size() && (int)members.size() < cap; ++i)
      if (same_group(*members[0], *rdy[i].second)) members.push_back(rdy[i].second);
    const int lane = s->next_lane;
    s->next_lane = (s->next_lane + 1) % (int)s->lanes.size();
    SRV_TRY(launch_group(s, members, lane));
    n += (int32_t)members.size();
  }
}

}  // namespace

extern "C" {

void hfx_server_desc_init(hfx_server_desc *d) {
  std::memset(d, 0, sizeof(*d));
  d->hbm_budget_bytes = 0;
  d->hbm_reserve_bytes = (int64_t)8 << 30;
  d->max_streams = 16;
  d->n_lanes = 4;
  d->use_graphs = 1;
  d->policy = HFX_SCHED_FIFO;
  d->lane_depth = 1;
}

void hfx_stream_desc_init(hfx_stream_desc *d, int32_t height, int32_t width) {
  std::memset(d, 0, sizeof(*d));
  hfx_plan_desc_init(&d->plan);
  d->plan.height = height;
  d->plan.width = width;
  d->queue_windows = 4;
  hfx_synth_desc_init(&d->camera, height, width);
}

hfx_status hfx_server_stream_bytes(const hfx_stream_desc *d, int64_t *bytes) {
  if (!d || !bytes) return fail(HFX_E_INVALID, "null argument");
  Geometry g;
  SRV_TRY(geometry(*d, g));
  *bytes = g.total();
  return HFX_OK;
}

hfx_status hfx_server_create(const hfx_server_desc *d, hfx_server **out) {
  if (!d || !out) return fail(HFX_E_INVALID, "null argument");
  *out = nullptr;
  if (d->max_streams < 1 || d->max_streams > 1024) return fail(HFX_E_INVALID, "max_streams must be in [1, 1024]");
  if (d->n_lanes < 1 || d->n_lanes > 16) return fail(HFX_E_INVALID, "n_lanes must be in [1, 16]");
  if (d->hbm_budget_bytes < 0 || d->hbm_reserve_bytes < 0) return fail(HFX_E_INVALID, "negative memory size");
  if (d->policy != HFX_SCHED_FIFO && d->policy != HFX_SCHED_EDF && d->policy != HFX_SCHED_GROUP)
    return fail(HFX_E_INVALID, "unknown policy");
  auto s = std::make_unique<hfx_server>();
  s->d = *d;
  SRV_HIP(hipGetDevice(&s->device));
  SRV_HIP(hipStreamCreateWithFlags(&s->ingest, hipStreamNonBlocking));
  SRV_HIP(hipStreamCreateWithFlags(&s->capture, hipStreamNonBlocking));
  s->lanes.assign(d->n_lanes, nullptr);
  for (auto &l : s->lanes) SRV_HIP(hipStreamCreateWithFlags(&l, hipStreamNonBlocking));
  // epoch: of a few recordings on the idle ingest stream keep the one with the shortest
  // host round trip; its completion time lies inside [before, after] of that round trip, and
  // taking `after` makes every GPU time read through it late (never early) by at most the gap
  hipEvent_t cand[8] = {};
  double best_gap = 1e30;
  int best = -1;
  for (int i = 0; i < 8; ++i) {
    SRV_HIP(hipEventCreate(&cand[i]));
    const double before = s->now();
    SRV_HIP(hipEventRecord(cand[i], s->ingest));
    SRV_HIP(hipEventSynchronize(cand[i]));
    const double after = s->now();
    if (after - before < best_gap) {
      best_gap = after - before;
      best = i;
      s->epoch_s = after;
    }
  }
  for (int i = 0; i < 8; ++i)
    if (i != best) (void)hipEventDestroy(cand[i]);
  s->epoch = cand[best];
  s->epoch_err_s = best_gap;
  s->lane_load.assign(d->n_lanes, 0);
  s->lane_windows.resize(d->n_lanes);
  s->slots.resize(d->max_streams);
  hfx_server *raw = s.get();
  s->worker = std::thread(worker_main, raw);
  *out = s.release();
  return HFX_OK;
}

hfx_status hfx_server_destroy(hfx_server *s) {
  if (!s) return HFX_OK;
  if (s->worker.joinable()) {  // the worker finishes the queued builds and tear-downs first
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stopping = true;
    }
    s->cv.notify_all();
    s->worker.join();
  }
  (void)hipDeviceSynchronize();
  for (auto &st : s->slots)
    if (st) st->free_all();
  for (auto l : s->lanes)
    if (l) (void)hipStreamDestroy(l);
  if (s->ingest) (void)hipStreamDestroy(s->ingest);
  if (s->capture) (void)hipStreamDestroy(s->capture);
  if (s->epoch) (void)hipEventDestroy(s->epoch);
  delete s;
  return HFX_OK;
}

hfx_status hfx_server_admit(hfx_server *s, const hfx_stream_desc *d, int32_t *stream_id) {
  if (!s || !d || !stream_id) return fail(HFX_E_INVALID, "null argument");
  *stream_id = -1;
  Geometry g;
  SRV_TRY(geometry(*d, g));
  int slot = -1;
  for (int i = 0; i < (int)s->slots.size(); ++i)
    if (!s->slots[i]) {
      slot = i;
      break;
    }
  if (slot < 0) return fail(HFX_NOT_READY, "no free stream slot");
  size_t free_b = 0, total_b = 0;
  SRV_HIP(hipMemGetInfo(&free_b, &total_b));
  // streams still being built have not allocated yet: their bytes are spoken for
  int64_t building = 0;
  for (auto &up : s->slots)
    if (up && up->state.load(std::memory_order_acquire) == 0) building += up->g.total();
  const int64_t need = g.total();
  if (s->d.hbm_budget_bytes > 0) {
    if (s->bytes_held + need > s->d.hbm_budget_bytes) return fail(HFX_NOT_READY, "stream does not fit the HBM budget now");
  }
  if ((int64_t)free_b - building - need < s->d.hbm_reserve_bytes) return fail(HFX_NOT_READY, "not enough free HBM now");
  auto st = std::make_unique<Stream>();
  st->id = slot;
  st->d = *d;
  st->g = g;
  st->lane = (int)(std::min_element(s->lane_load.begin(), s->lane_load.end()) - s->lane_load.begin());
  s->lane_load[st->lane]++;
  s->bytes_held += need;
  Stream *raw = st.get();
  s->slots[slot] = std::move(st);
  hfx_server::Job job;
  job.build = raw;
  post_job(s, std::move(job));
  *stream_id = slot;
  return HFX_OK;
}

hfx_status hfx_server_stream_state(hfx_server *s, int32_t stream_id, int32_t *state) {
  Stream *st = stream_of(s, stream_id);
  if (!st || !state) return fail(HFX_E_INVALID, "unknown stream id or null argument");
  *state = st->state.load(std::memory_order_acquire);
  if (*state == 2) return fail(HFX_E_HIP, "stream set-up failed: " + st->error);
  return HFX_OK;
}

hfx_status hfx_server_clock(hfx_server *s, double *t_s) {
  if (!s || !t_s) return fail(HFX_E_INVALID, "null argument");
  *t_s = s->now();
  return HFX_OK;
}

hfx_status hfx_server_ingest(hfx_server *s, int32_t stream_id, int64_t n_frames) {
  if (!s) return fail(HFX_E_INVALID, "null server");
  return hfx_server_ingest_due(s, stream_id, n_frames, s->now());
}

hfx_status hfx_server_ingest_due(hfx_server *s, int32_t stream_id, int64_t n_frames, double due_s) {
  Stream *st = stream_of(s, stream_id);
  if (!st) return fail(HFX_E_INVALID, "unknown stream id");
  const int state = st->state.load(std::memory_order_acquire);
  if (state == 0) return fail(HFX_NOT_READY, "stream still being set up");
  if (state == 2) return fail(HFX_E_HIP, "stream set-up failed: " + st->error);
  const Geometry &g = st->g;
  if (n_frames < 0 || n_frames > g.R) return fail(HFX_E_INVALID, "n_frames must be in [0, ring frames]");
  if (n_frames == 0) return HFX_OK;
  const int64_t F = st->frames_in + n_frames;
  // windows reading a frame < F - R must be retired before those frames are overwritten
  const int64_t over = F - g.R;
  const int64_t w_need = over > 0 ? (over + g.S - 1) / g.S : 0;
  if (st->retired < w_need) return fail(HFX_NOT_READY, "ring full: windows not retired yet");
  const int64_t old_ready = st->ready_windows();
  SRV_TRY(copy_frames(s, *st, st->frames_in, n_frames));
  st->frames_in = F;
  for (int64_t w = old_ready; w < st->ready_windows(); ++w) {
    SRV_HIP(hipEventRecord(st->ev_ready[w % g.n_pos], s->ingest));
    st->due_t[w % g.n_pos] = due_s;
  }
  return HFX_OK;
}

// true while the stream's last launched window has not completed on the GPU
static bool window_on_gpu(const Stream &st, int64_t w) {
  return hipEventQuery(st.ev_done[w % st.g.n_pos]) == hipErrorNotReady;
}

static hfx_status schedule_edf(hfx_server *s, int32_t &n) {
  // drop finished windows from the lanes' queues
  for (auto &q : s->lane_windows) {
    while (!q.empty()) {
      const Stream *st = stream_of(s, q.front().first);
      if (st && q.front().second < st->launched && window_on_gpu(*st, q.front().second)) break;
      q.pop_front();
    }
  }
  const int depth = std::max(1, s->d.lane_depth);
  for (;;) {
    int lane = -1;
    for (int l = 0; l < (int)s->lanes.size(); ++l)
      if ((int)s->lane_windows[l].size() < depth && (lane < 0 || s->lane_windows[l].size() < s->lane_windows[lane].size()))
        lane = l;
    if (lane < 0) break;
    Stream *best = nullptr;
    double best_deadline = 0.0;
    for (auto &up : s->slots) {
      if (!up || !ready(*up)) continue;
      Stream &st = *up;
      if (st.launched >= st.ready_windows()) continue;
      if (st.launched > 0 && window_on_gpu(st, st.launched - 1)) continue;  // one window in flight
      const double period = (double)st.g.S / (st.d.plan.fs_hz > 0 ? st.d.plan.fs_hz : 1.0);
      const double deadline = st.due_t[st.launched % st.g.n_pos] + period;
      if (!best || deadline < best_deadline) {
        best = &st;
        best_deadline = deadline;
      }
    }
    if (!best) break;
    s->lane_windows[lane].emplace_back(best->id, best->launched);
    SRV_TRY(launch_window(s, *best, lane));
    ++n;
  }
  return HFX_OK;
}

hfx_status hfx_server_schedule(hfx_server *s, int32_t *launched) {
  if (!s) return fail(HFX_E_INVALID, "null server");
  int32_t n = 0;
  if (s->d.policy == HFX_SCHED_EDF) {
    SRV_TRY(schedule_edf(s, n));
  } else if (s->d.policy == HFX_SCHED_GROUP) {
    SRV_TRY(schedule_group(s, n));
  } else {
    // every ready window, oldest camera due time first across the streams (a stream's own
    // windows keep their order: their due times do not decrease)
    std::vector<std::pair<double, Stream *>> rdy;
    for (auto &up : s->slots) {
      if (!up || !ready(*up)) continue;
      Stream &st = *up;
      for (int64_t w = st.launched; w < st.ready_windows(); ++w) rdy.emplace_back(st.due_t[w % st.g.n_pos], &st);
    }
    std::stable_sort(rdy.begin(), rdy.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    for (auto &r : rdy) {
      SRV_TRY(launch_window(s, *r.second, r.second->lane));
      ++n;
    }
  }
  if (launched) *launched = n;
  return HFX_OK;
}

hfx_status hfx_server_poll(hfx_server *s, hfx_window_done *out, int32_t cap, int32_t *n) {
  if (!s || !n || (cap > 0 && !out)) return fail(HFX_E_INVALID, "null argument");
  int32_t k = 0;
  for (auto &up : s->slots) {
    if (!up || !ready(*up)) continue;
    Stream &st = *up;
    while (k < cap && st.retired < st.launched) {
      const int64_t p = st.retired % st.g.n_pos;
      const hipError_t q = hipEventQuery(st.ev_done[p]);
      if (q == hipErrorNotReady) break;
      SRV_HIP(q);
      float gpu = 0.f;
      double t_start = 0.0, t_done = 0.0;
      SRV_HIP(hipEventElapsedTime(&gpu, st.ev_launch[p], st.ev_done[p]));
      SRV_HIP(s->gpu_time(st.ev_launch[p], t_start));
      SRV_HIP(s->gpu_time(st.ev_done[p], t_done));
      out[k].stream_id = st.id;
      out[k].n_outputs = (int32_t)st.g.n_out;
      out[k].window = st.retired;
      out[k].latency_us = (float)((t_done - st.due_t[p]) * 1e6);
      out[k].gpu_us = gpu * 1e3f;
      out[k].wait_us = (float)((t_start - st.due_t[p]) * 1e6);
      ++k;
      ++st.retired;
      ++s->windows_retired;
    }
  }
  *n = k;
  return HFX_OK;
}

hfx_status hfx_server_window_output(hfx_server *s, int32_t stream_id, int64_t window, const float **out) {
  Stream *st = stream_of(s, stream_id);
  if (!st || !out || !ready(*st)) return fail(HFX_E_INVALID, "unknown or not set-up stream id, or null argument");
  if (window < 0 || window >= st->retired || window + st->g.n_pos <= st->launched)
    return fail(HFX_E_INVALID, "window not retired or its output slot was reused");
  *out = st->outs + (window % st->g.n_pos) * st->g.out_elems;
  return HFX_OK;
}

hfx_status hfx_server_release(hfx_server *s, int32_t stream_id) {
  Stream *st = stream_of(s, stream_id);
  if (!st) return fail(HFX_E_INVALID, "unknown stream id");
  for (auto &q : s->lane_windows)
    q.erase(std::remove_if(q.begin(), q.end(), [&](const auto &e) { return e.first == stream_id; }), q.end());
  s->lane_load[st->lane]--;
  s->bytes_held -= st->g.total();
  // the worker waits for the stream's last copies and kernels and frees it; its jobs run in
  // order, so a stream released while still being built is built first
  hfx_server::Job job;
  job.teardown = std::move(s->slots[stream_id]);
  post_job(s, std::move(job));
  return HFX_OK;
}

hfx_status hfx_server_info_get(hfx_server *s, hfx_server_info *info) {
  if (!s || !info) return fail(HFX_E_INVALID, "null argument");
  size_t free_b = 0, total_b = 0;
  SRV_HIP(hipMemGetInfo(&free_b, &total_b));
  info->active_streams = 0;
  for (auto &up : s->slots) info->active_streams += up ? 1 : 0;
  info->max_streams = s->d.max_streams;
  info->bytes_held = s->bytes_held;
  info->bytes_free = (int64_t)free_b;
  info->windows_launched = s->windows_launched;
  info->windows_retired = s->windows_retired;
  info->clock_error_us = s->epoch_err_s * 1e6;
  info->group_launches = s->group_launches;
  info->grouped_windows = s->grouped_windows;
  return HFX_OK;
}

// ---------------------------------------------------------------------------- workload
hfx_status hfx_serve_workload(const hfx_workload_desc *d, hfx_workload_result *r) {
  if (!d || !r || d->n_streams < 1 || !d->streams || !d->arrival_s || !d->window_rate_hz ||
      d->windows_per_stream < 1)
    return fail(HFX_E_INVALID, "invalid workload description");
  if (d->capture_stream >= d->n_streams || (d->capture_stream >= 0 && !d->capture_out))
    return fail(HFX_E_INVALID, "invalid capture stream");
  for (int i = 0; i < d->n_streams; ++i)
    if (!(d->window_rate_hz[i] > 0)) return fail(HFX_E_INVALID, "window rates must be positive");
  if (d->arrivals != HFX_ARRIVALS_PERIODIC && d->arrivals != HFX_ARRIVALS_POISSON)
    return fail(HFX_E_INVALID, "unknown window arrival process");
  hfx_server *srv = nullptr;
  SRV_TRY(hfx_server_create(&d->server, &srv));
  std::unique_ptr<hfx_server, hfx_status (*)(hfx_server *)> guard(srv, hfx_server_destroy);

  // Camera model: after admission a stream's frames accumulate; window w is complete at its
  // due time, the camera time of its last frame.  Periodic: the camera delivers one stride of
  // frames every 1/rate (window 0 needs T/S strides); Poisson: exponential gaps of mean 1/rate
  // (window 0 included).  A window is ingested at or after its due time (later when the ring
  // is still full), and its latency runs from the due time, so waiting behind a full ring
  // counts.
  struct Live {
    int32_t id = -1;
    bool admitted = false, started = false, released = false;
    int64_t arrived = 0, retired = 0;
    double due_s = 0;  // due time of window `arrived` (server clock)
  };
  const int n = d->n_streams, wps = d->windows_per_stream;
  std::vector<Live> live(n);
  std::mt19937_64 rng(d->seed);
  std::vector<hfx_window_done> done(4096);
  std::vector<int> by_id(d->server.max_streams, -1);
  r->windows_retired = 0;
  r->max_concurrent = 0;
  r->admissions_deferred = 0;
  r->ingest_deferred = 0;
  r->windows_ingested_late = 0;
  int active = 0, finished = 0, next_admit = 0;
  const double t0 = srv->now();  // arrival_s is relative to the loop's start
  auto now_s = [&] { return srv->now(); };
  double last_retire = t0;
  auto next_gap = [&](int i) {
    const double rate = d->window_rate_hz[i];
    if (d->arrivals == HFX_ARRIVALS_PERIODIC) return 1.0 / rate;
    std::exponential_distribution<double> ex(rate);
    return ex(rng);
  };

  while (finished < n) {
    const double now = now_s();
    if (now - t0 > d->time_limit_s) return fail(HFX_E_HIP, "workload exceeded its time limit");
    // admissions, in arrival order
    while (next_admit < n && t0 + d->arrival_s[next_admit] <= now) {
      Live &L = live[next_admit];
      const hfx_status rc = hfx_server_admit(srv, &d->streams[next_admit], &L.id);
      if (rc == HFX_NOT_READY) {
        ++r->admissions_deferred;
        break;
      }
      SRV_TRY(rc);
      L.admitted = true;
      const double t_adm = now_s();
      by_id[L.id] = next_admit;
      if (r->admit_s) r->admit_s[next_admit] = t_adm - t0;
      ++active;
      r->max_concurrent = std::max(r->max_concurrent, active);
      ++next_admit;
    }
    // cameras start once their stream is set up (the server's worker builds it)
    for (int i = 0; i < next_admit; ++i) {
      Live &L = live[i];
      if (!L.admitted || L.started || L.released) continue;
      int32_t state = 0;
      SRV_TRY(hfx_server_stream_state(srv, L.id, &state));
      if (state != 1) continue;
      L.started = true;
      const hfx_plan_desc &p = d->streams[i].plan;
      L.due_s = now_s() + (d->arrivals == HFX_ARRIVALS_PERIODIC
                               ? (double)(p.window / p.stride) / d->window_rate_hz[i]
                               : next_gap(i));
    }
    // window arrivals: the camera has delivered the window's frames by its due time
    for (int i = 0; i < next_admit; ++i) {
      Live &L = live[i];
      if (!L.admitted || !L.started || L.released) continue;
      while (L.arrived < wps && L.due_s <= now_s()) {
        const hfx_plan_desc &p = d->streams[i].plan;
        const hfx_status rc = hfx_server_ingest_due(srv, L.id, L.arrived == 0 ? p.window : p.stride, L.due_s);
        if (rc == HFX_NOT_READY) {
          ++r->ingest_deferred;
          break;
        }
        SRV_TRY(rc);
        if (now_s() - L.due_s > 1.0 / d->window_rate_hz[i]) ++r->windows_ingested_late;
        ++L.arrived;
        L.due_s += next_gap(i);
      }
    }
    int32_t launched = 0;
    SRV_TRY(hfx_server_schedule(srv, &launched));
    int32_t k = 0;
    SRV_TRY(hfx_server_poll(srv, done.data(), (int32_t)done.size(), &k));
    for (int j = 0; j < k; ++j) {
      const hfx_window_done &w = done[j];
      const int i = by_id[w.stream_id];
      Live &L = live[i];
      const int64_t at = (int64_t)i * wps + w.window;
      if (r->latency_us) r->latency_us[at] = w.latency_us;
      if (r->gpu_us) r->gpu_us[at] = w.gpu_us;
      if (r->wait_us) r->wait_us[at] = w.wait_us;
      if (i == d->capture_stream) {
        const float *src = nullptr;
        SRV_TRY(hfx_server_window_output(srv, w.stream_id, w.window, &src));
        Stream *st = stream_of(srv, w.stream_id);
        const size_t elems = (size_t)st->g.out_elems;
        // on a non-blocking stream: a legacy-stream copy would have to wait on the worker's
        // graph capture, which HIP refuses
        SRV_HIP(hipMemcpyAsync(d->capture_out + (size_t)w.window * elems, src, elems * 4, hipMemcpyDeviceToHost,
                               srv->ingest));
        SRV_HIP(hipStreamSynchronize(srv->ingest));
      }
      ++L.retired;
      ++r->windows_retired;
      last_retire = now_s();
      if (L.retired == wps) {
        SRV_TRY(hfx_server_release(srv, L.id));
        by_id[L.id] = -1;
        L.released = true;
        if (r->finish_s) r->finish_s[i] = now_s() - t0;
        --active;
        ++finished;
      }
    }
  }
  r->wall_s = last_retire - (t0 + d->arrival_s[0]);
  r->clock_error_us = srv->epoch_err_s * 1e6;
  r->group_launches = srv->group_launches;
  r->grouped_windows = srv->grouped_windows;
  return HFX_OK;
}

}  // extern "C"
