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
  size_t resident = 0;      // buffers + plan scratch

  // Incremental RFI state (u8, ds 1, chunk a multiple of 256 rows, IQRM off
  // or one window over the whole buffer).  Results are identical to the
  // plan's own execute on the window:
  //  - IQRM sums are exact integers, kept per chunk (first r = W mod C rows,
  //    and the rest) in a ring; a window's sums add whole chunks plus the
  //    head of the chunk that completes it;
  //  - the cleaned channel-major rows of window j-1 that overlap window j
  //    were cleaned with window j-1's mask; when window j's mask is the same
  //    (decided on the device) only the new rows are cleaned.
  bool fast = false;
  uint8_t *d_xT = nullptr;  // [nch][Lcap], window j at column xbase
  size_t Lcap = 0, xbase = 0;
  size_t r = 0;             // W mod C
  int R = 0;                // ring slots
  Part *d_ring = nullptr;   // [R][2][Csub]
  uint2 *d_cpart = nullptr; // per-item partials of one chunk
  RowItem *d_citems = nullptr;
  int *d_cseg = nullptr;    // {0, head items, all items}
  int ncitems = 0;
  RowItem *d_items_old = nullptr, *d_items_new = nullptr;
  int n_old = 0, n_new = 0;
  uint8_t *d_prev = nullptr;
  int *d_flag = nullptr;    // {valid, changed}
  std::vector<void *> allocs;
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
  bool incremental = true;

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
  std::map<void *, int> plane_held;                   // planes returned by poll, not yet released
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

// Window j of a fast beam, completed by chunk k (already uploaded).
af_status fast_window(af_server *s, Beam &b, Flight &f, uint64_t k) {
  (void)s;
  af_plan *p = b.plan;
  const af_plan_desc &d = p->d;
  f.j = b.next_j++;
  const uint64_t j = f.j;
  const size_t start = (size_t)(j * b.C - b.head);
  const uint8_t *xw = b.d_buf + start * b.row_bytes;
  if (d.use_iqrm) {
    const int nfull = (int)(k - j);
    AF_HIP(launch_ring_sum(b.d_ring, b.R, (int)(j % b.R), nfull, (int)(k % b.R), b.r == 0 ? 1 : 0,
                           p->Csub, p->d_segsum, b.st));
    AF_TRY(run_iqrm_vote(p, b.st));
  }
  AF_HIP(launch_mask_delta(p->d_masks, b.d_prev, d.use_iqrm ? d.nchan : 0, b.d_flag, b.st));
  // window j starts C columns after window j-1; move the overlap back to
  // column 0 when the next window would not fit (disjoint: xbase > Lp + C)
  if (j > 0) b.xbase += b.C;
  if (b.xbase + p->Lp > b.Lcap) {
    // the copy rounds each row up to 16 bytes: the extra bytes land in rows
    // [W - C, W), which the new-row clean below rewrites
    AF_HIP(launch_copy_rows(b.d_xT, b.d_xT + b.xbase, b.Lcap, b.W - b.C, p->nch, b.st));
    b.xbase = 0;
  }
  CleanArgs ca{};
  ca.x = xw;
  ca.xT = b.d_xT + b.xbase;
  ca.Lp = b.Lcap;
  ca.nchan = d.nchan;
  ca.segmask = d.use_iqrm ? p->d_segmask : nullptr;
  ca.stat_mask = p->d_static;
  ca.zdm = d.use_zero_dm && d.zdm_c1 > d.zdm_c0;
  ca.zc0 = d.zdm_c0;
  ca.zc1 = d.zdm_c1;
  ca.dd_c0 = d.dd_c0;
  ca.dd_c1 = d.dd_c1;
  if (b.n_old) {
    ca.items = b.d_items_old;
    ca.nitems = b.n_old;
    ca.run_if = b.d_flag + 1;   // only when the mask changed
    AF_HIP(launch_clean_u8(ca, b.st));
  }
  ca.items = b.d_items_new;
  ca.nitems = b.n_new;
  ca.run_if = nullptr;
  AF_HIP(launch_clean_u8(ca, b.st));
  AF_HIP(launch_zero_rows(b.d_xT + b.xbase + b.W, b.Lcap, p->Lp - b.W, p->nch, b.st));
  AF_TRY(run_dedisp(p, b.d_xT + b.xbase, b.Lcap, f.plane, b.st, -1));
  if (f.masks && d.use_iqrm)
    AF_HIP(hipMemcpyAsync(f.masks, p->d_masks, (size_t)p->nwin * d.nchan, hipMemcpyDeviceToDevice,
                          b.st));
  return AF_OK;
}

// b: the flight's beam, looked up under the lock (beams is appended to by
// af_server_add_beam on other threads; references to its elements stay valid).
af_status dispatch(af_server *s, Beam &b, Flight &f) {
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
  const size_t crow = (size_t)(b.total - b.head);   // buffer row of this chunk
  // host rows (page-locked: asynchronous) or rows already in device memory
  AF_HIP(hipMemcpyAsync(b.d_buf + crow * b.row_bytes, f.r.h, b.C * b.row_bytes, hipMemcpyDefault,
                        b.st));
  const uint64_t k = b.total / b.C;                  // beam chunk index
  b.total += b.C;
  if (b.fast) {
    af_plan *p = b.plan;
    if (p->d.use_iqrm) {
      // this chunk's sums: head rows [0, r) and the rest
      AF_HIP(launch_seg_stats_u8(b.d_buf + crow * b.row_bytes, p->d.nchan, p->d.iqrm_c0,
                                 p->d.iqrm_c1, b.d_citems, b.ncitems, b.d_cpart, b.st));
      AF_HIP(launch_seg_reduce_u8(b.d_cpart, b.d_cseg, 2, p->Csub,
                                  b.d_ring + (size_t)(k % b.R) * 2 * p->Csub, b.st));
    }
    if (f.plane) AF_TRY(fast_window(s, b, f, k));
  } else if (f.plane) {
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
  if (f.plane) s->plane_held[f.plane] += 1;
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
      af_status st = dispatch(s, b, f);
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
  if (desc && desc->full_rfi) s->incremental = false;
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

extern "C++" {  // helpers of af_server_add_beam (C++ linkage inside the C-ABI block)

static void release_beam(af::Beam &b) {
  if (b.st) (void)hipStreamSynchronize(b.st);
  if (b.plan) af_plan_destroy(b.plan);
  if (b.d_buf) (void)hipFree(b.d_buf);
  for (void *q : b.allocs) (void)hipFree(q);
  if (b.st) (void)hipStreamDestroy(b.st);
  b = af::Beam{};
}

// Rows [r0, r1) cut into items of at most `step` rows (segment 0).
static std::vector<af::RowItem> row_items(size_t r0, size_t r1, size_t step) {
  std::vector<af::RowItem> v;
  for (size_t r = r0; r < r1; r += step)
    v.push_back(af::RowItem{0u, (uint32_t)r, (uint32_t)std::min(r1, r + step), 0u});
  return v;
}

template <typename T>
static af_status beam_alloc(af::Beam &b, T **ptr, size_t bytes) {
  void *q = nullptr;
  if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) return fail(AF_ERR_NOMEM, "hipMalloc beam state");
  b.allocs.push_back(q);
  *ptr = (T *)q;
  return AF_OK;
}

template <typename T>
static af_status beam_upload(af::Beam &b, T **ptr, const std::vector<T> &v) {
  AF_TRY(beam_alloc(b, ptr, v.size() * sizeof(T)));
  if (!v.empty()) AF_HIP(hipMemcpy(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return AF_OK;
}

// Incremental RFI state for beams that qualify; returns its device bytes
// through `bytes` (0 when the beam runs the plan's full execute per window).
static af_status setup_fast(af::Beam &b, size_t *bytes) {
  *bytes = 0;
  af_plan *p = b.plan;
  const af_plan_desc &d = p->d;
  const bool ok = d.dtype == AF_U8 && d.ds == 1 && b.C % 256 == 0 &&
                  d.nchan <= af::kCleanU8MaxChan && p->nch == d.dd_c1 - d.dd_c0 &&
                  (!d.use_iqrm || (p->nwin == 1 && p->nseg == 1 && p->Csub > 0));
  if (!ok) return AF_OK;
  b.Lcap = ((2 * p->Lp + b.C) + 255) & ~(size_t)255;
  b.r = b.W % b.C;
  b.R = (int)((b.W + b.C - 1) / b.C) + 2;
  AF_TRY(beam_alloc(b, &b.d_xT, (size_t)p->nch * b.Lcap));
  *bytes += (size_t)p->nch * b.Lcap;
  if (d.use_iqrm) {
    // 64-row items: a chunk is a small grid, so more, shorter items
    std::vector<af::RowItem> it = row_items(0, b.r, af::kStatRowsU8 / 4);
    const int nhead = (int)it.size();
    std::vector<af::RowItem> rest = row_items(b.r, b.C, af::kStatRowsU8 / 4);
    it.insert(it.end(), rest.begin(), rest.end());
    b.ncitems = (int)it.size();
    AF_TRY(beam_upload(b, &b.d_citems, it));
    AF_TRY(beam_upload(b, &b.d_cseg, std::vector<int>{0, nhead, b.ncitems}));
    AF_TRY(beam_alloc(b, &b.d_cpart, (size_t)b.ncitems * p->Csub * sizeof(uint2)));
    AF_TRY(beam_alloc(b, &b.d_ring, (size_t)b.R * 2 * p->Csub * sizeof(af::Part)));
    *bytes += (size_t)b.ncitems * p->Csub * sizeof(uint2) + (size_t)b.R * 2 * p->Csub * sizeof(af::Part);
  }
  // the new rows' items start on a 32-row boundary at or before W - C; the
  // few overlap rows they repeat come out identical
  const size_t A = ((b.W - b.C) / af::kCleanRowsU8) * af::kCleanRowsU8;
  std::vector<af::RowItem> old_items = row_items(0, A, af::kCleanRowsU8);
  std::vector<af::RowItem> new_items = row_items(A, b.W, af::kCleanRowsU8);
  b.n_old = (int)old_items.size();
  b.n_new = (int)new_items.size();
  AF_TRY(beam_upload(b, &b.d_items_old, old_items));
  AF_TRY(beam_upload(b, &b.d_items_new, new_items));
  AF_TRY(beam_alloc(b, &b.d_prev, (size_t)d.nchan));
  AF_TRY(beam_alloc(b, &b.d_flag, 2 * sizeof(int)));
  AF_HIP(hipMemset(b.d_prev, 0, (size_t)d.nchan));
  AF_HIP(hipMemset(b.d_flag, 0, 2 * sizeof(int)));
  *bytes += (old_items.size() + new_items.size()) * sizeof(af::RowItem) + d.nchan + 8;
  b.fast = true;
  return AF_OK;
}

}  // extern "C++"

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
  // the beam keeps its own channel-major window: its plan never cleans per
  // super-tile
  af_plan_desc wd = *w;
  wd.clean_budget = ~(size_t)0;
  AF_TRY(af_plan_create(&wd, &b.plan));
  af_plan_info info;
  af_plan_query(b.plan, &info);
  b.row_bytes = (size_t)w->nchan * dtype_size(w->dtype);
  b.W = w->nsamp;
  b.C = chunk_rows;
  b.rcap = 2 * b.W + b.C;
  b.plane_bytes = ((size_t)info.ndm * info.nout * info.out_elem_bytes + 255) & ~(size_t)255;
  b.mask_bytes = (size_t)info.nwin * w->nchan;
  b.resident = b.rcap * b.row_bytes + info.scratch_bytes;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->used + b.resident > s->budget) {
      release_beam(b);
      return fail(AF_ERR_NOMEM, "beam does not fit the HBM budget");
    }
  }
  af_status st = AF_OK;
  size_t extra = 0;
  if (hipMalloc(&b.d_buf, b.rcap * b.row_bytes) != hipSuccess)
    st = fail(AF_ERR_NOMEM, "hipMalloc beam buffer");
  if (st == AF_OK && hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking) != hipSuccess)
    st = fail(AF_ERR_RUNTIME, "hipStreamCreate");
  if (st == AF_OK && s->incremental) st = setup_fast(b, &extra);
  std::lock_guard<std::mutex> lk(s->mu);
  // chunks submitted since the first check: the beam is not added
  if (st == AF_OK && s->completed != s->submitted) st = fail(AF_ERR_BUSY, "chunks in flight");
  if (st == AF_OK && s->used + b.resident + extra > s->budget)
    st = fail(AF_ERR_NOMEM, "beam does not fit the HBM budget");
  if (st != AF_OK) {
    release_beam(b);
    return st;
  }
  b.resident += extra;
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
  auto h = s->plane_held.find(d_plane);
  if (h == s->plane_held.end())
    return fail(AF_ERR_INVALID, "plane is not held by the caller (already released, or its "
                                "chunk has not been returned by af_server_poll)");
  if (--h->second == 0) s->plane_held.erase(h);
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
  for (auto &b : s->beams) release_beam(b);
  for (auto &kv : s->plane_size) (void)hipFree(kv.first);
  for (auto e : s->ev_free) (void)hipEventDestroy(e);
  delete s;
  return AF_OK;
}

}  // extern "C"
