// af_server.hip -- real-time chunk server with HBM-resident DM-overlap carry.
//
// The reference processes a stream chunk by chunk: each call reads the
// chunk plus the DM-overlap tail from the file and uploads the whole window
// (python/astroflow/search.py:365,464 -> src/core/dedispered.cu:1040-1048).
// Here every beam keeps its window in device memory; a chunk uploads only its
// new rows, and the beam's plan runs on the window in place.
//
// Threading: callers submit and poll from any thread; one server thread owns
// every HIP call after set-up.  A beam's chunks run in order on the beam's
// stream (they share the window buffer and the plan's scratch); chunks of
// different beams run on different streams and overlap on the device.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "af_internal.h"

namespace af {

namespace {

double now_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Beam {
  af_plan *plan = nullptr;
  hipStream_t st = nullptr;
  size_t row_bytes = 0;     // nchan * element size
  size_t W = 0, C = 0;      // window rows, chunk rows
  size_t rcap = 0;          // buffer capacity in rows (2W + C: compaction never overlaps)
  uint8_t *d_buf = nullptr; // [rcap][nchan]
  uint64_t head = 0;        // beam row held in buffer row 0
  uint64_t total = 0;       // rows uploaded so far (dispatch order)
  uint64_t next_j = 0;      // next output chunk (window start next_j * C)
  size_t plane_bytes = 0, mask_bytes = 0;
  size_t resident = 0;      // buffer + plan scratch
};

struct Req {
  int beam;
  const void *h;
  uint64_t tag;
  double t_submit;
};

struct Flight {
  Req r;
  hipEvent_t ev = nullptr;
  double t_admit = 0;
  void *plane = nullptr;
  uint8_t *masks = nullptr;
  uint64_t j = 0;
  int status = AF_OK;
};

}  // namespace
}  // namespace af

struct af_server {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  int device = 0;
  size_t budget = 0, used = 0, peak = 0;
  int max_inflight = 16, max_queue = 4096;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread thr;
  bool stop = false;

  std::deque<af::Beam> beams;       // a deque: adding a beam moves no other beam
  std::deque<af::Req> queue;
  std::deque<af::Flight> flights;   // dispatched, not finished (server thread only)
  std::deque<af_server_result> done;
  size_t nflight = 0;               // flights.size(), readable under mu
  int inflight_peak = 0;

  // result planes: every allocation counts against the budget until destroy
  std::map<void *, size_t> plane_size;                // all planes
  std::map<size_t, std::vector<void *>> plane_free;   // free planes by size
  std::vector<hipEvent_t> ev_free;

  uint64_t submitted = 0, admitted = 0, completed = 0, planes = 0;
  std::vector<double> lat;          // submit -> done, chunks with a plane
  double first_submit = -1, last_done = 0;
  std::string err;                  // last asynchronous error
};

namespace af {
namespace {

// Server-thread helpers (called with the lock released unless noted).

void *plane_get(af_server *s, size_t bytes) {  // under lock
  auto it = s->plane_free.find(bytes);
  if (it != s->plane_free.end() && !it->second.empty()) {
    void *p = it->second.back();
    it->second.pop_back();
    return p;
  }
  // idle planes of other sizes go back to the budget before giving up
  for (auto &kv : s->plane_free) {
    while (s->used + bytes > s->budget && !kv.second.empty()) {
      void *q = kv.second.back();
      kv.second.pop_back();
      (void)hipFree(q);
      s->plane_size.erase(q);
      s->used -= kv.first;
    }
  }
  if (s->used + bytes > s->budget) return nullptr;
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  s->plane_size[p] = bytes;
  s->used += bytes;
  s->peak = std::max(s->peak, s->used);
  return p;
}

af_status dispatch(af_server *s, Flight &f) {
  Beam &b = s->beams[f.r.beam];
  const size_t in_buf = (size_t)(b.total - b.head);
  if (in_buf + b.C > b.rcap) {
    // keep rows from the next window start; source and destination are
    // disjoint because rcap >= 2W + C and fewer than W rows are kept
    const uint64_t keep = b.next_j * b.C;
    const size_t off = (size_t)(keep - b.head), len = (size_t)(b.total - keep);
    if (len)
      AF_HIP(hipMemcpyAsync(b.d_buf, b.d_buf + off * b.row_bytes, len * b.row_bytes,
                            hipMemcpyDeviceToDevice, b.st));
    b.head = keep;
  }
  AF_HIP(hipMemcpyAsync(b.d_buf + (size_t)(b.total - b.head) * b.row_bytes, f.r.h,
                        b.C * b.row_bytes, hipMemcpyHostToDevice, b.st));
  b.total += b.C;
  if (f.plane) {
    f.j = b.next_j++;
    const size_t start = (size_t)(f.j * b.C - b.head);
    AF_TRY(execute(b.plan, b.d_buf + start * b.row_bytes, f.plane, f.masks, b.st));
  }
  AF_HIP(hipEventRecord(f.ev, b.st));
  return AF_OK;
}

void finish(af_server *s, Flight &f, double t) {  // under lock
  af_server_result r{};
  r.beam = f.r.beam;
  r.has_plane = f.plane != nullptr;
  r.status = f.status;
  r.tag = f.r.tag;
  r.window = f.j;
  r.d_plane = f.plane;
  r.d_masks = f.masks;
  r.t_submit_ms = f.r.t_submit;
  r.t_admit_ms = f.t_admit;
  r.t_done_ms = t;
  s->completed++;
  if (f.plane && f.status == AF_OK) {
    s->planes++;
    s->lat.push_back(t - f.r.t_submit);
  }
  s->last_done = std::max(s->last_done, t);
  s->done.push_back(r);
  if (f.ev) s->ev_free.push_back(f.ev);
}

void run(af_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progress = false;
    // reap finished flights (in dispatch order per beam; any order overall)
    for (auto it = s->flights.begin(); it != s->flights.end();) {
      hipError_t q = hipEventQuery(it->ev);
      if (q == hipErrorNotReady) { ++it; continue; }
      if (q != hipSuccess) {
        it->status = AF_ERR_RUNTIME;
        s->err = std::string("chunk failed on the device: ") + hipGetErrorString(q);
      }
      finish(s, *it, now_ms(s->t0));
      it = s->flights.erase(it);
      progress = true;
    }
    s->nflight = s->flights.size();
    if (progress) s->cv_done.notify_all();
    // admit queued chunks, first come first served
    while (!s->queue.empty() && (int)s->flights.size() < s->max_inflight) {
      Req r = s->queue.front();
      Beam &b = s->beams[r.beam];
      // a plane is needed when this chunk completes the next window; the
      // rows of queued chunks of this beam count in dispatch order
      const bool need = b.total + b.C >= b.next_j * b.C + b.W;
      Flight f;
      f.r = r;
      if (need) {
        f.plane = plane_get(s, b.plane_bytes + b.mask_bytes);
        if (!f.plane) break;  // wait for released planes
        f.masks = b.mask_bytes ? (uint8_t *)f.plane + b.plane_bytes : nullptr;
      }
      if (s->ev_free.empty()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) break;
        s->ev_free.push_back(e);
      }
      f.ev = s->ev_free.back();
      s->ev_free.pop_back();
      s->queue.pop_front();
      f.t_admit = now_ms(s->t0);
      s->admitted++;
      lk.unlock();
      af_status st = dispatch(s, f);
      lk.lock();
      if (st != AF_OK) {
        f.status = st;
        s->err = error_text();
        (void)hipStreamSynchronize(b.st);
        finish(s, f, now_ms(s->t0));
        s->cv_done.notify_all();
      } else {
        s->flights.push_back(f);
      }
      s->inflight_peak = std::max(s->inflight_peak, (int)s->flights.size());
      progress = true;
    }
    s->nflight = s->flights.size();
    if (s->stop && s->flights.empty()) {
      // chunks still waiting for planes that nobody will release are dropped
      while (!s->queue.empty()) {
        Flight f;
        f.r = s->queue.front();
        s->queue.pop_front();
        f.status = AF_ERR_BUSY;
        f.ev = nullptr;
        finish(s, f, now_ms(s->t0));
      }
      break;
    }
    if (!progress) {
      if (s->flights.empty()) {
        s->cv_work.wait(lk);
      } else {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      }
    }
  }
  s->cv_done.notify_all();
}

}  // namespace
}  // namespace af

using namespace af;

extern "C" {

af_status af_server_create(const af_server_desc *desc, af_server **srv) {
  if (!srv) return fail(AF_ERR_INVALID, "null argument");
  *srv = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(AF_ERR_NODEV, "No HIP devices found");
  af_server *s = new af_server();
  (void)hipGetDevice(&s->device);
  if (desc && desc->max_inflight > 0) s->max_inflight = desc->max_inflight;
  if (desc && desc->max_queue > 0) s->max_queue = desc->max_queue;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    delete s;
    return fail(AF_ERR_RUNTIME, "hipMemGetInfo failed");
  }
  s->budget = (desc && desc->hbm_budget) ? desc->hbm_budget : (size_t)(0.9 * (double)fr);
  s->thr = std::thread(run, s);
  *srv = s;
  return AF_OK;
}

af_status af_server_add_beam(af_server *s, const af_plan_desc *w, size_t chunk_rows, int *beam) {
  if (!s || !w || !beam) return fail(AF_ERR_INVALID, "null argument");
  if (w->ds != 1) return fail(AF_ERR_INVALID, "served beams need ds == 1");
  if (chunk_rows == 0 || chunk_rows > w->nsamp)
    return fail(AF_ERR_INVALID, "chunk_rows must be in [1, window rows]");
  if (w->nout != chunk_rows)
    return fail(AF_ERR_INVALID, "window nout must equal chunk_rows");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->completed != s->submitted) return fail(AF_ERR_BUSY, "chunks in flight");
  }
  AF_HIP(hipSetDevice(s->device));
  Beam b;
  AF_TRY(af_plan_create(w, &b.plan));
  af_plan_info info;
  af_plan_query(b.plan, &info);
  b.row_bytes = (size_t)w->nchan * dtype_size(w->dtype);
  b.W = w->nsamp;
  b.C = chunk_rows;
  b.rcap = 2 * b.W + b.C;
  b.plane_bytes = ((size_t)info.ndm * info.nout * info.out_elem_bytes + 255) & ~(size_t)255;
  b.mask_bytes = (size_t)info.nwin * w->nchan;
  b.resident = b.rcap * b.row_bytes + info.scratch_bytes;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->used + b.resident > s->budget) {
    af_plan_destroy(b.plan);
    return fail(AF_ERR_NOMEM, "beam does not fit the HBM budget");
  }
  if (hipMalloc(&b.d_buf, b.rcap * b.row_bytes) != hipSuccess) {
    af_plan_destroy(b.plan);
    return fail(AF_ERR_NOMEM, "hipMalloc beam buffer");
  }
  if (hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking) != hipSuccess) {
    (void)hipFree(b.d_buf);
    af_plan_destroy(b.plan);
    return fail(AF_ERR_RUNTIME, "hipStreamCreate");
  }
  s->used += b.resident;
  s->peak = std::max(s->peak, s->used);
  s->beams.push_back(b);
  *beam = (int)s->beams.size() - 1;
  return AF_OK;
}

af_status af_server_submit(af_server *s, int beam, const void *h_rows, uint64_t tag) {
  if (!s || !h_rows) return fail(AF_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (beam < 0 || beam >= (int)s->beams.size()) return fail(AF_ERR_INVALID, "no such beam");
  if ((int)s->queue.size() >= s->max_queue) return fail(AF_ERR_BUSY, "server queue full");
  Req r{beam, h_rows, tag, now_ms(s->t0)};
  if (s->first_submit < 0) s->first_submit = r.t_submit;
  s->queue.push_back(r);
  s->submitted++;
  s->cv_work.notify_one();
  return AF_OK;
}

af_status af_server_poll(af_server *s, af_server_result *out, int max, int *n, int wait_ms) {
  if (!s || !n || (max > 0 && !out)) return fail(AF_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->done.empty() && wait_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
  int k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n = k;
  return AF_OK;
}

af_status af_server_release(af_server *s, void *d_plane) {
  if (!s || !d_plane) return fail(AF_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->plane_size.find(d_plane);
  if (it == s->plane_size.end()) return fail(AF_ERR_INVALID, "not a result plane of this server");
  s->plane_free[it->second].push_back(d_plane);
  s->cv_work.notify_one();
  return AF_OK;
}

af_status af_server_drain(af_server *s) {
  if (!s) return fail(AF_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    if (s->completed == s->submitted) break;
    s->cv_work.notify_one();
    s->cv_done.wait_for(lk, std::chrono::milliseconds(5));
  }
  if (!s->err.empty()) {
    set_error(s->err);
    return AF_ERR_RUNTIME;
  }
  return AF_OK;
}

af_status af_server_get_stats(af_server *s, af_server_stats *st) {
  if (!s || !st) return fail(AF_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = af_server_stats{};
  st->submitted = s->submitted;
  st->admitted = s->admitted;
  st->completed = s->completed;
  st->planes = s->planes;
  if (!s->lat.empty()) {
    std::vector<double> v = s->lat;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {  // linear interpolation between order statistics
      const double pos = q * (double)(v.size() - 1);
      const size_t i = (size_t)pos;
      const double f = pos - (double)i;
      return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.back();
    double sum = 0;
    for (double x : v) sum += x;
    st->mean_ms = sum / (double)v.size();
  }
  st->first_submit_ms = s->first_submit;
  st->last_done_ms = s->last_done;
  st->hbm_used = s->used;
  st->hbm_peak = s->peak;
  st->hbm_budget = s->budget;
  st->inflight_peak = s->inflight_peak;
  return AF_OK;
}

af_status af_server_destroy(af_server *s) {
  if (!s) return AF_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->thr.joinable()) s->thr.join();
  (void)hipSetDevice(s->device);
  for (auto &b : s->beams) {
    if (b.st) (void)hipStreamSynchronize(b.st);
    if (b.plan) af_plan_destroy(b.plan);
    if (b.d_buf) (void)hipFree(b.d_buf);
    if (b.st) (void)hipStreamDestroy(b.st);
  }
  for (auto &kv : s->plane_size) (void)hipFree(kv.first);
  for (auto e : s->ev_free) (void)hipEventDestroy(e);
  delete s;
  return AF_OK;
}

}  // extern "C"
