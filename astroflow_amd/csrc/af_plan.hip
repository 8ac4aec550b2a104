// af_plan.hip -- host side of the C-ABI: plans (tables, segmentation, device
// memory), the execute sequence and the device helpers.
//
// Built with -ffp-contract=off: the FP64 delay and residual tables round
// exactly as the expressions are written (the CPU oracle uses the same
// expressions).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include <sched.h>

#include "af_internal.h"

namespace {

// The blocks af_host_alloc handed out (base -> bytes).  Never destroyed, so
// that a block freed while the process exits still finds them.
std::mutex &pinned_mu() {
  static std::mutex *m = new std::mutex;
  return *m;
}
std::map<uintptr_t, size_t> &pinned_blocks() {
  static auto *b = new std::map<uintptr_t, size_t>;
  return *b;
}
void host_pinned_add(const void *p, size_t n) {
  std::lock_guard<std::mutex> lk(pinned_mu());
  pinned_blocks()[(uintptr_t)p] = n;
}
void host_pinned_remove(const void *p) {
  std::lock_guard<std::mutex> lk(pinned_mu());
  pinned_blocks().erase((uintptr_t)p);
}
// [p, p + n) lies inside one block from af_host_alloc
bool host_pinned(const void *p, size_t n) {
  if (!p) return false;
  std::lock_guard<std::mutex> lk(pinned_mu());
  const auto &b = pinned_blocks();
  auto it = b.upper_bound((uintptr_t)p);
  if (it == b.begin()) return false;
  --it;
  return (uintptr_t)p >= it->first && (uintptr_t)p + n <= it->first + it->second;
}

// Host threads of a staged copy: the CPUs this process may run on (its
// affinity), at most 8.  Eight already move pageable memory as fast as the
// link takes it; sixteen measured no faster, and on a machine whose CPU
// quota is 16 they, with the runtime's own threads, sometimes got the
// process throttled (a 15 ms stall in a 20 ms call).
int copy_threads() {
  cpu_set_t cs;
  int n = 0;
  if (sched_getaffinity(0, sizeof(cs), &cs) == 0) n = CPU_COUNT(&cs);
  if (n <= 0) n = (int)std::max(1u, std::thread::hardware_concurrency());
  return std::min(n, 8);
}

}  // namespace

// Workers that split a memcpy between pageable user memory and page-locked
// staging (one thread does not reach the link's rate).  They live as long
// as the plan, so a host call starts no threads; copy() returns when every
// slice is done.
struct CopyPool {
  std::vector<std::thread> th;
  std::mutex mu;
  std::condition_variable go, done;
  char *dst = nullptr;
  const char *src = nullptr;
  size_t n = 0, piece = 0;
  uint64_t gen = 0;
  int pending = 0;
  bool stop = false;

  // throws (std::system_error) when a thread cannot be started, after
  // stopping the ones that were
  explicit CopyPool(int workers) {
    try {
      for (int i = 0; i < workers; ++i) th.emplace_back([this, i] { loop(i + 1); });
    } catch (...) {
      shutdown();
      throw;
    }
  }
  ~CopyPool() { shutdown(); }
  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    go.notify_all();
    for (auto &t : th)
      if (t.joinable()) t.join();
    th.clear();
  }
  void loop(int slice) {
    uint64_t seen = 0;
    for (;;) {
      char *d;
      const char *s;
      size_t nn, pc;
      {
        std::unique_lock<std::mutex> lk(mu);
        go.wait(lk, [&] { return stop || gen != seen; });
        if (stop) return;
        seen = gen;
        d = dst;
        s = src;
        nn = n;
        pc = piece;
      }
      const size_t a = (size_t)slice * pc;
      if (a < nn) std::memcpy(d + a, s + a, std::min(pc, nn - a));
      {
        std::lock_guard<std::mutex> lk(mu);
        if (--pending == 0) done.notify_one();
      }
    }
  }
  // slice 0 on the calling thread, slice i on worker i
  void copy(void *d, const void *s, size_t nn) {
    if (th.empty() || nn < ((size_t)1 << 22)) {
      std::memcpy(d, s, nn);
      return;
    }
    const size_t nt = th.size() + 1;
    const size_t pc = ((nn + nt - 1) / nt + 4095) & ~(size_t)4095;
    {
      std::lock_guard<std::mutex> lk(mu);
      dst = (char *)d;
      src = (const char *)s;
      n = nn;
      piece = pc;
      pending = (int)th.size();
      ++gen;
    }
    go.notify_all();
    std::memcpy(d, s, std::min(pc, nn));
    std::unique_lock<std::mutex> lk(mu);
    done.wait(lk, [&] { return pending == 0; });
  }
};


namespace af {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
std::string error_text() { return g_err; }

af_status fail(af_status code, const std::string &msg) {
  g_err = msg;
  return code;
}

size_t dtype_size(int dtype) {
  switch (dtype) {
  case AF_U8: return 1;
  case AF_U16: return 2;
  case AF_U32: return 4;
  case AF_F32: return 4;
  default: return 0;
  }
}

int acc_size(int dtype) { return dtype == AF_U32 ? 8 : 4; }

static constexpr double kDM = 4148.808;  // src/core/dedispered.cu:333

// Delay in native samples: round half away from zero of
// K*dm*(f^-2 - fref^-2)/tsamp (src/core/dedispered.cu:329-334, in FP64).
static long delay_native(double f, double fref, double dm, double tsamp) {
  double inv = 1.0 / (f * f) - 1.0 / (fref * fref);
  double delay = kDM * dm * inv;
  return (long)std::round(delay / tsamp);
}

// Subband residual (src/core/dedispered.cu:384-413, in FP64).
static long residual(double f, double fref, double ddm, double tsamp, int ds) {
  double inv = 1.0 / (f * f) - 1.0 / (fref * fref);
  double delay = kDM * ddm * inv;
  double bins = delay / (tsamp * (double)ds);
  return (long)std::floor(bins + 0.5);
}

// IQRM windows (src/core/iqrm.cu:368-395).
static void make_windows(size_t n, unsigned W, unsigned H, bool tail,
                         std::vector<uint32_t> &t0s, std::vector<uint32_t> &t1s) {
  t0s.clear();
  t1s.clear();
  if (n == 0) return;
  if (W == 0 || W >= n) {
    t0s.push_back(0);
    t1s.push_back((uint32_t)n);
    return;
  }
  unsigned t0 = 0;
  while ((size_t)t0 + W <= n) {
    t0s.push_back(t0);
    t1s.push_back(t0 + W);
    t0 += (H ? H : W);
  }
  if (tail && t0 < n && n > (size_t)t0 + 4) {
    t0s.push_back(t0);
    t1s.push_back((uint32_t)n);
  }
}

// IQRM lags (src/core/iqrm.cu:189-200).
static std::vector<int> make_lags(int radius, double geofactor) {
  std::vector<int> lags;
  if (radius <= 0) return lags;
  int lag = 1;
  while (lag <= radius) {
    lags.push_back(+lag);
    lags.push_back(-lag);
    int nxt = (int)(geofactor * lag);
    lag = std::max(nxt, lag + 1);
  }
  return lags;
}

// Arena mode: the next 256-byte aligned piece of the arena.
static af_status carve(af_plan *p, void **dst, size_t bytes) {
  const size_t off = (p->arena_used + 255) / 256 * 256;
  if (off + bytes > p->arena_cap && !p->arena_measure)
    return fail(AF_ERR_NOMEM, "the plan needs more device scratch than the no-allocation "
                              "entry point holds");
  p->arena_used = off + bytes;
  *dst = p->arena_measure ? nullptr : (void *)(p->arena + off);
  return AF_OK;
}

template <typename V>
static af_status upload(af_plan *p, V **dst, const std::vector<V> &src) {
  if (src.empty()) {
    *dst = nullptr;
    return AF_OK;
  }
  if (p->arena || p->arena_measure) {
    void *q = nullptr;
    AF_TRY(carve(p, &q, src.size() * sizeof(V)));
    // pageable source: the runtime copies it out before returning; the
    // device copy is ordered on the caller's stream
    if (q)
      AF_HIP(hipMemcpyAsync(q, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice,
                            p->arena_st));
    *dst = (V *)q;
    return AF_OK;
  }
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, src.size() * sizeof(V));
  if (e != hipSuccess) return fail(AF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  p->allocs.push_back(q);
  p->scratch_bytes += src.size() * sizeof(V);
  AF_HIP(hipMemcpy(q, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice));
  *dst = (V *)q;
  return AF_OK;
}

static af_status dalloc(af_plan *p, void **dst, size_t bytes) {
  if (bytes == 0) {
    *dst = nullptr;
    return AF_OK;
  }
  if (p->arena || p->arena_measure) return carve(p, dst, bytes);
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess)
    return fail(AF_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes: " +
                                  hipGetErrorString(e));
  p->allocs.push_back(q);
  p->scratch_bytes += bytes;
  *dst = q;
  return AF_OK;
}

static void release(af_plan *p) {
  for (void *q : p->allocs) (void)hipFree(q);
  p->allocs.clear();
  for (void **q : {&p->hx_in, &p->hx_out, &p->hx_masks, &p->hx_img, &p->hx_imgscr})
    if (*q) (void)hipFree(*q), *q = nullptr;
  for (int b = 0; b < 2; ++b) {
    if (p->stage[b]) (void)hipHostFree(p->stage[b]);
    if (p->stage_ev[b]) (void)hipEventDestroy(p->stage_ev[b]);
    p->stage[b] = nullptr;
    p->stage_ev[b] = nullptr;
  }
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  p->ev.clear();
  delete p->copy_pool;
  p->copy_pool = nullptr;
  if (p->stream) (void)hipStreamDestroy(p->stream);
  p->stream = nullptr;
}

static af_status validate(const af_plan_desc &d) {
  if (d.dtype < AF_U8 || d.dtype > AF_F32) return fail(AF_ERR_INVALID, "unknown dtype");
  if (d.nchan <= 0) return fail(AF_ERR_INVALID, "nchan must be > 0");
  if (d.nsamp == 0) return fail(AF_ERR_INVALID, "nsamp must be > 0");
  if (d.ds < 1) return fail(AF_ERR_INVALID, "time_downsample must be >= 1");
  if (!(d.tsamp > 0.0)) return fail(AF_ERR_INVALID, "tsamp must be > 0");
  if (!d.freq_mhz) return fail(AF_ERR_INVALID, "frequency table missing");
  if (d.dd_c0 < 0 || d.dd_c1 > d.nchan || d.dd_c0 >= d.dd_c1)
    return fail(AF_ERR_INVALID, "invalid dedispersion channel range");
  if (d.ndm < 1) return fail(AF_ERR_INVALID, "ndm must be >= 1");
  if (d.dm_low < 0.0) return fail(AF_ERR_INVALID, "dm_low must be >= 0");
  if (d.ndm > 1 && !(d.dm_step > 0.0)) return fail(AF_ERR_INVALID, "dm_step must be > 0");
  if (d.method != AF_METHOD_DIRECT && d.method != AF_METHOD_SUBBAND)
    return fail(AF_ERR_INVALID, "unknown method");
  if (d.method == AF_METHOD_SUBBAND && (d.sb_nchan < 1 || d.ndm0 < 1))
    return fail(AF_ERR_INVALID, "subband method needs sb_nchan >= 1 and ndm0 >= 1");
  if (d.use_iqrm) {
    if (d.iqrm_c0 < 0 || d.iqrm_c1 > d.nchan || d.iqrm_c0 >= d.iqrm_c1)
      return fail(AF_ERR_INVALID, "invalid IQRM channel range");
    if (d.iqrm_mode != 0 && d.iqrm_mode != 1) return fail(AF_ERR_INVALID, "IQRM mode must be 0 or 1");
  }
  if (d.use_zero_dm && (d.zdm_c0 < 0 || d.zdm_c1 > d.nchan || d.zdm_c0 > d.zdm_c1))
    return fail(AF_ERR_INVALID, "invalid zero-DM channel range");
  if (d.n_bad < 0 || (d.n_bad > 0 && !d.bad_chans)) return fail(AF_ERR_INVALID, "bad channel list");
  for (int c = 1; c < d.nchan; ++c)
    if (!(d.freq_mhz[c] >= d.freq_mhz[c - 1]))
      return fail(AF_ERR_INVALID, "frequency table must be ascending");
  return AF_OK;
}

static af_status build(af_plan *p) {
  const af_plan_desc &d = p->d;
  const int ds = d.ds;
  p->nbin = (d.nsamp + (size_t)ds - 1) / (size_t)ds;
  p->nout = d.nout ? d.nout : p->nbin;
  if (p->nout > p->nbin) return fail(AF_ERR_INVALID, "nout exceeds the binned input length");
  p->nch = d.dd_c1 - d.dd_c0;
  p->ndm = d.ndm;
  const double fref = d.f_ref_mhz;
  const double *F = p->freq.data();

  // ---- delay / residual tables (FP64)
  long maxd = 0;
  if (d.method == AF_METHOD_DIRECT) {
    p->h_delays.resize((size_t)p->ndm * p->nch);
    for (int i = 0; i < p->ndm; ++i) {
      const double dm = d.dm_low + (double)i * d.dm_step;
      for (int c = d.dd_c0; c < d.dd_c1; ++c) {
        long v = delay_native(F[c], fref, dm, d.tsamp) / ds;
        if (v < 0) return fail(AF_ERR_INVALID, "negative delay: the reference frequency is below a dedispersed channel");
        if (v > 0x3FFFFFFF) return fail(AF_ERR_INVALID, "delay too large");
        maxd = std::max(maxd, v);
        p->h_delays[(size_t)i * p->nch + (c - d.dd_c0)] = (int32_t)v;
      }
    }
  } else {
    const int sbw = d.sb_nchan, ndm0 = d.ndm0;
    p->nsb = (p->nch + sbw - 1) / sbw;
    p->ndm_nom = (p->ndm + ndm0 - 1) / ndm0;
    p->h_delays.resize((size_t)p->ndm_nom * p->nch);
    for (int j = 0; j < p->ndm_nom; ++j) {
      const double dm = d.dm_low + (double)(j * ndm0) * d.dm_step;
      for (int c = d.dd_c0; c < d.dd_c1; ++c) {
        long v = delay_native(F[c], fref, dm, d.tsamp) / ds;
        if (v < 0) return fail(AF_ERR_INVALID, "negative delay: the reference frequency is below a dedispersed channel");
        if (v > 0x3FFFFFFF) return fail(AF_ERR_INVALID, "delay too large");
        maxd = std::max(maxd, v);
        p->h_delays[(size_t)j * p->nch + (c - d.dd_c0)] = (int32_t)v;
      }
    }
    p->h_res.resize((size_t)p->ndm * p->nsb);
    p->h_maxres_j.assign(p->ndm_nom, 0);
    p->h_maxres_jsb.assign((size_t)p->ndm_nom * p->nsb, 0);
    long maxr = 0;
    for (int m = 0; m < p->ndm; ++m) {
      const int j = m / ndm0;
      const double ddm = (double)(m - j * ndm0) * d.dm_step;
      for (int sb = 0; sb < p->nsb; ++sb) {
        const int ch0 = d.dd_c0 + sb * sbw;
        const int ch1 = std::min(ch0 + sbw, d.dd_c1);
        long r = residual(F[(ch0 + ch1 - 1) / 2], fref, ddm, d.tsamp, ds);
        if (r < 0) return fail(AF_ERR_INVALID, "negative subband residual");
        if (r > 0x3FFFFFFF) return fail(AF_ERR_INVALID, "residual too large");
        p->h_res[(size_t)m * p->nsb + sb] = (int32_t)r;
        p->h_maxres_j[j] = std::max<int32_t>(p->h_maxres_j[j], (int32_t)r);
        int32_t &mj = p->h_maxres_jsb[(size_t)j * p->nsb + sb];
        mj = std::max<int32_t>(mj, (int32_t)r);
        maxr = std::max(maxr, r);
      }
    }
    p->maxres = (int)maxr;
  }
  p->maxdelay = (int)maxd;

  // ---- sizes of the dedispersion stages
  // LDS per workgroup: the largest of the device's per-block limits
  int maxlds = 65536;
  {
    const hipDeviceAttribute_t attrs[] = {hipDeviceAttributeMaxSharedMemoryPerBlock,
                                          hipDeviceAttributeSharedMemPerBlockOptin};
    for (auto at : attrs) {
      int v = 0;
      if (hipDeviceGetAttribute(&v, at, p->device) == hipSuccess && v > maxlds) maxlds = v;
    }
  }
  p->maxlds = maxlds;
  size_t n_ext;
  if (d.method == AF_METHOD_DIRECT) {
    n_ext = p->nout;
    p->adds = (double)p->ndm * p->nch * (double)p->nout;
  } else {
    const bool inter16 = d.dtype == AF_U8 && d.sb_nchan <= 256;
    int tile = 0;
    if (inter16) {
      auto lds_of = [&](int t) { return (size_t)p->nsb * (size_t)((t + p->maxres + 7) / 8 * 8) * 2; };
      // all-subband kernel: each subband's row holds the tile plus its own
      // largest residual, so rows are packed (bytes of the largest group)
      auto lds_pk = [&](int t) {
        size_t best = 0;
        for (int j = 0; j < p->ndm_nom; ++j) {
          size_t e = 0;
          for (int sb = 0; sb < p->nsb; ++sb)
            e += (size_t)((t + p->h_maxres_jsb[(size_t)j * p->nsb + sb] + 7) / 8 * 8);
          best = std::max(best, e);
        }
        return best * 2 + 16;   // + one spare 16-byte chunk
      };
      bool packed = false;
      // a requested stage-2 tile (autotuner), when it fits
      if ((d.tile_t == 1024 || d.tile_t == 512 || d.tile_t == 256) &&
          lds_pk(d.tile_t) <= (size_t)maxlds) {
        tile = d.tile_t;
        packed = true;
      }
      if ((d.tile_t == 128 || d.tile_t == 64) && lds_of(d.tile_t) <= (size_t)std::min(maxlds, 65536))
        tile = d.tile_t;
      // all-subband kernel: prefer two workgroups per CU, then one
      const int big[] = {1024, 512};
      for (int half = 2; half >= 1 && !tile; --half)
        for (int t : big)
          if (lds_pk(t) <= (size_t)maxlds / half) { tile = t; packed = true; break; }
      if (packed) {
        p->h_sboff.assign((size_t)p->ndm_nom * (p->nsb + 1), 0);
        for (int j = 0; j < p->ndm_nom; ++j) {
          int32_t *o = &p->h_sboff[(size_t)j * (p->nsb + 1)];
          for (int sb = 0; sb < p->nsb; ++sb)
            o[sb + 1] = o[sb] + (tile + p->h_maxres_jsb[(size_t)j * p->nsb + sb] + 7) / 8 * 8;
        }
        p->s2_pk = (int)((lds_pk(tile) - 16) / 2);
      }
      // many subbands: stage them in chunks (<= 64 KB per workgroup)
      if (!tile && d.ndm0 <= 32) {
        for (int t : big) {
          const size_t row = (size_t)((t + p->maxres + 7) / 8 * 8) * 2;
          const int sbc = (int)std::min<size_t>((size_t)p->nsb, 65536 / row);
          if (sbc >= 4 && (d.tile_t == 0 || d.tile_t == t)) {
            tile = t;
            p->s2_sbc = sbc;
            break;
          }
        }
      }
      // element-wise kernel within 64 KB
      if (!tile) {
        const int small[] = {256, 128, 64};
        for (int t : small)
          if (lds_of(t) <= (size_t)std::min(maxlds, 65536)) { tile = t; break; }
      }
    }
    // float32 partial sums: LDS stage 2 with the subbands in chunks
    if (d.dtype == AF_F32 && d.ndm0 <= 32) {
      const int big[] = {512, 1024};
      for (int t : big) {
        if (d.tile_t != 0 && d.tile_t != t) continue;
        const size_t row = (size_t)((t + p->maxres + 3) / 4 * 4) * 4;
        const int sbc = (int)std::min<size_t>((size_t)p->nsb, 65536 / row);
        if (sbc >= 2) {
          tile = t;
          p->s2_sbc = sbc;
          break;
        }
      }
    }
    // 0: stage 2 without LDS (u16 partial sums whose halo does not fit, or
    // wider partial sums)
    p->s2_tile = tile;
    p->s2_lds = (tile + p->maxres + 7) / 8 * 8;
    if (tile == 0) tile = 512;   // output granularity of the super-tiles
    // Time super-tiles: the subband intermediate holds one super-tile of T2
    // outputs (plus the residual halo), at most kInterBudget bytes, as the
    // reference tiles time (src/core/dedispered.cu:37-39,1199-1235).
    p->inter_elem = inter16 ? 2 : (d.dtype == AF_U32 ? 8 : 4);   // float32: float partials
    const size_t per_t = (size_t)p->ndm_nom * p->nsb * (size_t)p->inter_elem;
    const size_t budget = d.inter_budget ? d.inter_budget : kInterBudget;
    size_t tmax = budget / std::max<size_t>(per_t, 1);
    tmax = tmax > (size_t)p->maxres + 2048 ? tmax - (size_t)p->maxres - 1024 : 1024;
    tmax = std::max<size_t>(1024, tmax / 1024 * 1024);
    const size_t nout_r = (p->nout + 1023) / 1024 * 1024;
    p->T2 = std::min(nout_r, tmax);
    p->nsuper = (int)((p->nout + p->T2 - 1) / p->T2);
    size_t ntile2 = (p->T2 + tile - 1) / tile;
    n_ext = ntile2 * (size_t)tile + (size_t)p->maxres + 8;
    p->adds = (double)p->ndm_nom * p->nch * (double)(p->nout + p->maxres) +
              (double)p->ndm * p->nsb * (double)p->nout;
  }
  // LDS-staged u8 dedispersion: 16 DM rows per workgroup share each channel's
  // delay window (span of the 16 rows' delays + one tile) if it fits 16 KB.
  if (d.dtype == AF_U8 && (d.method == AF_METHOD_DIRECT || d.sb_nchan <= 256)) {
    const int rows = d.method == AF_METHOD_DIRECT ? p->ndm : p->ndm_nom;
    int span = 0;
    for (int r0 = 0; r0 < rows; r0 += 16) {
      const int r1 = std::min(r0 + 16, rows) - 1;
      for (int c = 0; c < p->nch; ++c)
        span = std::max(span, p->h_delays[(size_t)r1 * p->nch + c] - p->h_delays[(size_t)r0 * p->nch + c]);
    }
    // + up to 15 bytes of misalignment: windows start 16-byte aligned
    const int win = (span + 15 + kDdTile + 15) / 16 * 16;
    if (d.dm_per_block == 16 && win <= 16 * 1024 && rows >= 8) {
      p->dd_lds_rows = 16;
      p->dd_lds_win = win;
    }
  }
  // the aligned-load u8 kernel (channel lists sorted by delay quarter class)
  p->dd_q = d.dtype == AF_U8 && p->dd_lds_rows == 0 &&
            (d.method == AF_METHOD_DIRECT || d.sb_nchan <= 256);
  p->dd_fq = d.dtype == AF_F32;
  const int dd_tile = p->dd_q ? kDdTileQ : kDdTile;
  p->ntiles1 = (int)((n_ext + dd_tile - 1) / dd_tile);
  p->L1 = (size_t)p->ntiles1 * dd_tile;
  // Row length of the channel-major buffer: every load of the dedispersion
  // kernels (tile start + delay + 16 bytes, or + the LDS window) stays inside
  // its own channel row, which is zero beyond nbin.
  size_t need_len = (size_t)(p->nsuper - 1) * p->T2 + p->L1 + (size_t)p->maxdelay + 64 +
                    (size_t)p->dd_lds_win;
  p->Lp = std::max(p->nbin, need_len);
  p->Lp = (p->Lp + 255) / 256 * 256;
  // Cleaning per super-tile: the cleaned buffer then holds one super-tile's
  // reach (L1 + the largest delay + the load overhang) instead of the whole
  // input, at the cost of cleaning the overlap rows once per super-tile.
  if (d.method == AF_METHOD_SUBBAND && p->nsuper > 1) {
    const size_t es = dtype_size(d.dtype);
    const size_t lt = ((size_t)p->L1 + (size_t)p->maxdelay + 64 + (size_t)p->dd_lds_win + 255) /
                      256 * 256;
    const size_t whole = (size_t)p->nch * p->Lp * es, one = (size_t)p->nch * lt * es;
    bool tiled = false;
    if (d.clean_budget) {
      if (whole > d.clean_budget) {
        if (one > d.clean_budget)
          return fail(AF_ERR_NOMEM, "clean_budget is below one super-tile of cleaned rows (" +
                                        std::to_string(one) + " bytes)");
        tiled = true;
      }
    } else {
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
        const size_t other = (ds > 1 ? p->nbin * (size_t)d.nchan * es : 0) +
                             (size_t)p->ndm_nom * p->nsb * p->L1 * p->inter_elem + (size_t(1) << 30);
        tiled = one < whole && whole + other > fr;
      }
    }
    if (tiled) {
      p->clean_tiled = true;
      p->Lp = lt;
    }
  }

  // ---- IQRM windows, lags and row segmentation
  std::vector<uint32_t> bounds = {0u, (uint32_t)p->nbin};
  if (p->nbin > 0xFFFFFFF0ull) return fail(AF_ERR_INVALID, "too many rows");
  if (d.use_iqrm) {
    make_windows(p->nbin, d.iqrm_win, d.iqrm_hop, d.iqrm_include_tail != 0, p->wt0, p->wt1);
    p->nwin = (int)p->wt0.size();
    p->Csub = d.iqrm_c1 - d.iqrm_c0;
    // iqrm_omp::set_iqrm_params clamps (include/iqrm.hpp:33-37)
    const float rf = std::max(0.0f, d.iqrm_radius_frac);
    const double gf = d.iqrm_geofactor > 1.0 ? d.iqrm_geofactor : 1.5;
    const int radius = std::max(1, (int)std::floor(rf * (float)p->Csub));
    p->lags = make_lags(radius, gf);
    p->nlags = (int)p->lags.size();
    for (int l : p->lags)
      if (std::abs(l) < p->Csub) p->lag_reach = std::max(p->lag_reach, std::abs(l));
    for (int w = 0; w < p->nwin; ++w) {
      bounds.push_back(p->wt0[w]);
      bounds.push_back(p->wt1[w]);
    }
    int P2 = 1;
    while (P2 < p->Csub) P2 <<= 1;
    p->iqrm_lds = (size_t)(2 * p->Csub + std::max(P2, 1024)) * sizeof(float);
    if (p->iqrm_lds > (size_t)maxlds)
      return fail(AF_ERR_INVALID, "IQRM channel range too large for one workgroup");
    // vote merge: casting counts + one bit per lag per channel
    if ((size_t)p->Csub * (1 + (p->nlags + 31) / 32) * sizeof(int) > (size_t)maxlds)
      return fail(AF_ERR_INVALID, "IQRM channel range x lags too large for one workgroup");
  }
  std::sort(bounds.begin(), bounds.end());
  bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
  for (size_t i = 0; i + 1 < bounds.size(); ++i) {
    Seg s{bounds[i], bounds[i + 1], 0, 0};
    int wlo = -1, whi = -1;
    for (int w = 0; w < p->nwin; ++w)
      if (p->wt0[w] <= s.t0 && s.t1 <= p->wt1[w]) {
        if (wlo < 0) wlo = w;
        whi = w;
      }
    if (wlo >= 0) {
      for (int w = wlo; w <= whi; ++w)
        if (!(p->wt0[w] <= s.t0 && s.t1 <= p->wt1[w]))
          return fail(AF_ERR_RUNTIME, "internal: non-contiguous window cover");
      s.w0 = (uint32_t)wlo;
      s.w1 = (uint32_t)whi + 1;
    }
    p->segs.push_back(s);
  }
  p->nseg = (int)p->segs.size();
  // segments that are the windows themselves (no overlapping hops): the
  // per-segment mask is the window mask, with no kernel to form it
  p->seg_is_win = p->nseg == p->nwin;
  for (int s = 0; s < p->nseg && p->seg_is_win; ++s)
    p->seg_is_win = p->segs[s].w0 == (uint32_t)s && p->segs[s].w1 == (uint32_t)s + 1;
  p->wseg.assign(2 * (size_t)p->nwin, 0);
  for (int w = 0; w < p->nwin; ++w) {
    int a = -1, b = -1;
    for (int s = 0; s < p->nseg; ++s)
      if (p->segs[s].t0 >= p->wt0[w] && p->segs[s].t1 <= p->wt1[w]) {
        if (a < 0) a = s;
        b = s;
      }
    p->wseg[2 * w] = (uint32_t)a;
    p->wseg[2 * w + 1] = (uint32_t)b + 1;
  }
  // u8 statistics: two 1024-thread workgroups fit a CU; rows per item are
  // chosen so the items fill the machine in one round (up to 1024 rows per
  // item: each thread then sums at most 256 rows in its packed 16-bit lanes).
  // Measured on the 4096-ch chunk: one round of 296-row items beats two of
  // 148 and four of 74 (fewer per-item reductions).
  uint32_t stat_rows = d.dtype == AF_U8 ? kStatRowsU8 : kStatRows;
  if (d.dtype == AF_U8) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess ||
        ncu <= 0)
      ncu = 256;
    const size_t slots = 2 * (size_t)ncu;
    size_t r = (p->nbin + slots - 1) / slots;
    r = (r + 3) / 4 * 4;
    stat_rows = (uint32_t)std::min<size_t>(kStatRowsU8Max, std::max<size_t>(16, r));
  }
  const uint32_t clean_rows =
      (d.dtype == AF_U8 && d.nchan <= kCleanU8MaxChan) ? kCleanRowsU8 : kCleanRows;
  p->seg_items.assign(1, 0);
  for (int s = 0; s < p->nseg; ++s) {
    for (uint32_t r = p->segs[s].t0; r < p->segs[s].t1; r += stat_rows)
      p->stat_items.push_back(RowItem{(uint32_t)s, r, std::min<uint32_t>(r + stat_rows, p->segs[s].t1), 0});
    p->seg_items.push_back((int)p->stat_items.size());
    for (uint32_t r = p->segs[s].t0; r < p->segs[s].t1; r += clean_rows)
      p->clean_items.push_back(RowItem{(uint32_t)s, r, std::min<uint32_t>(r + clean_rows, p->segs[s].t1), 0});
  }

  // Columns of the cleaned buffer the dedispersion reads: channel c (from
  // dd_c0) up to (nsuper-1)*T2 + L1 + its largest delay (+ the load
  // overhang and the LDS window of the staged form).  Each clean item keeps
  // in .pad the number of leading channels that still reach its first row;
  // the clean pass stores only those (delays fall with frequency, so for the
  // ascending band this is most of the low-frequency half at late rows).
  if (!p->clean_tiled) {
    const int rows = d.method == AF_METHOD_DIRECT ? p->ndm : p->ndm_nom;
    std::vector<size_t> need((size_t)p->nch, 0);
    const size_t base = (size_t)(p->nsuper - 1) * p->T2 + p->L1 + 64 + (size_t)p->dd_lds_win;
    for (int c = 0; c < p->nch; ++c) {
      int32_t mx = 0;
      for (int r = 0; r < rows; ++r) mx = std::max(mx, p->h_delays[(size_t)r * p->nch + c]);
      need[c] = base + (size_t)mx;
    }
    for (RowItem &it : p->clean_items) {
      int cut = 0;
      for (int c = p->nch - 1; c >= 0; --c)
        if (need[c] > it.r0) { cut = c + 1; break; }
      it.pad = (uint32_t)cut;
    }
    p->clean_cut = true;
  }
  if (p->clean_tiled) {
    p->tile_off.assign(1, 0);
    for (int k = 0; k < p->nsuper; ++k) {
      const size_t a = (size_t)k * p->T2, b = std::min(p->nbin, a + p->Lp);
      for (const RowItem &it : p->clean_items) {
        const size_t lo = std::max<size_t>(it.r0, a), hi = std::min<size_t>(it.r1, b);
        if (lo < hi) p->tile_items.push_back(RowItem{it.seg, (uint32_t)lo, (uint32_t)hi, 0});
      }
      p->tile_off.push_back((int)p->tile_items.size());
    }
  }

  // ---- stage-1 work order.  Workgroup (group g, tile t, row block b) reads
  // channel c from t*T + d(first row of b, c) onwards.  Sorting each group's
  // workgroups by that start (at the group's middle channel) makes the
  // workgroups that run together on one XCD (consecutive launch order after
  // the XCD remap) read overlapping input, so it is fetched from HBM about
  // once instead of once per tile.
  if (d.dtype == AF_U8 || p->dd_fq) {
    const int rows = d.method == AF_METHOD_DIRECT ? p->ndm : p->ndm_nom;
    const int JB = p->dd_lds_rows ? p->dd_lds_rows
                                  : (p->dd_q || p->dd_fq) ? kDdRowsQ : kDdRowsPerBlock;
    const int nrb = (rows + JB - 1) / JB;
    const int ngrp = d.method == AF_METHOD_DIRECT ? 1 : p->nsb;
    const int gw = d.method == AF_METHOD_DIRECT ? p->nch : d.sb_nchan;
    std::vector<std::pair<int64_t, int4>> keyed;
    keyed.reserve((size_t)ngrp * p->ntiles1 * nrb);
    for (int g = 0; g < ngrp; ++g) {
      const int cmid = std::min(p->nch - 1, g * gw + (std::min(gw, p->nch - g * gw) - 1) / 2);
      for (int t = 0; t < p->ntiles1; ++t)
        for (int b = 0; b < nrb; ++b) {
          const int64_t start = (int64_t)t * (p->dd_q ? kDdTileQ : kDdTile) +
                                p->h_delays[(size_t)(b * JB) * p->nch + cmid];
          keyed.push_back({(int64_t)g << 40 | start, int4{g, t, b, 0}});
        }
    }
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const auto &x, const auto &y) { return x.first < y.first; });
    p->worklist.reserve(keyed.size());
    for (auto &kv : keyed) p->worklist.push_back(kv.second);
  }

  // ---- device memory
  AF_TRY(upload(p, &p->d_worklist, p->worklist));
  AF_TRY(upload(p, &p->d_delays, p->h_delays));
  AF_TRY(upload(p, &p->d_res, p->h_res));
  AF_TRY(upload(p, &p->d_maxres_j, p->h_maxres_j));
  if (!p->h_sboff.empty()) AF_TRY(upload(p, &p->d_sboff, p->h_sboff));
  AF_TRY(upload(p, &p->d_lags, p->lags));
  AF_TRY(upload(p, &p->d_seg_items, p->seg_items));
  AF_TRY(upload(p, &p->d_wseg, p->wseg));
  AF_TRY(upload(p, &p->d_wt0, p->wt0));
  AF_TRY(upload(p, &p->d_wt1, p->wt1));
  AF_TRY(upload(p, &p->d_segs, p->segs));
  AF_TRY(upload(p, &p->d_stat_items, p->stat_items));
  AF_TRY(upload(p, &p->d_clean_items, p->clean_items));
  if (p->clean_tiled) AF_TRY(upload(p, &p->d_tile_items, p->tile_items));
  if (!p->bad.empty()) {
    std::vector<uint8_t> st(d.nchan, 0);
    bool any = false;
    for (int c : p->bad)
      if (c >= 0 && c < d.nchan) { st[c] = 1; any = true; }
    if (any) AF_TRY(upload(p, &p->d_static, st));
  }
  const size_t es = dtype_size(d.dtype);
  if (ds > 1) AF_TRY(dalloc(p, &p->d_xb, p->nbin * (size_t)d.nchan * es));
  AF_TRY(dalloc(p, &p->d_xT, (size_t)p->nch * p->Lp * es));
  if (p->arena) {
    // the arena keeps the previous call's contents: zero the columns past the
    // input (the clean pass writes columns [0, nbin) of every row) with a
    // kernel on the caller's stream
    if (p->Lp > p->nbin)
      AF_HIP(launch_zero_rows((uint8_t *)p->d_xT + p->nbin * es, p->Lp * es,
                              (p->Lp - p->nbin) * es, p->nch, p->arena_st));
  } else if (!p->arena_measure) {
    AF_HIP(hipMemset(p->d_xT, 0, (size_t)p->nch * p->Lp * es));
  }
  if (d.method == AF_METHOD_SUBBAND) {
    AF_TRY(dalloc(p, &p->d_inter, (size_t)p->ndm_nom * p->nsb * p->L1 * p->inter_elem));
  }
  if (d.use_iqrm && p->nwin > 0) {
    void *q;
    AF_TRY(dalloc(p, &q, p->stat_items.size() * (size_t)p->Csub * sizeof(Part)));
    p->d_part = (Part *)q;
    AF_TRY(dalloc(p, &q, (size_t)p->nseg * p->Csub * sizeof(Part)));
    p->d_segsum = (Part *)q;
    AF_TRY(dalloc(p, &q, (size_t)p->nwin * std::max(1, p->nlags) * p->Csub));
    p->d_flags = (uint8_t *)q;
    AF_TRY(dalloc(p, &q, (size_t)p->nwin * d.nchan));
    p->d_masks = (uint8_t *)q;
    if (p->seg_is_win) {
      p->d_segmask = p->d_masks;
    } else {
      AF_TRY(dalloc(p, &q, (size_t)p->nseg * d.nchan));
      p->d_segmask = (uint8_t *)q;
    }
  }
  if (!p->arena && !p->arena_measure)
    AF_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  return AF_OK;
}

// Stage k of a profiled execute lies between events k and k+1; the
// dedispersion stage (first-stage kernels) between events 4 and 5.  Mode 2
// records only those two.
static af_status record(af_plan *p, int set, int k, hipStream_t st) {
  if (!p->profiling || set < 0) return AF_OK;
  if (p->profiling == 2 && k != 4 && k != 5) return AF_OK;
  AF_HIP(hipEventRecord(p->ev[(size_t)set * af_plan::kNumEv + k], st));
  return AF_OK;
}

static af_status collect(af_plan *p) {
  if (p->profiling == 2) {
    for (int s = 0; s < p->ev_used; ++s) {
      hipEvent_t *e = &p->ev[(size_t)s * af_plan::kNumEv];
      AF_HIP(hipEventSynchronize(e[5]));
      float ms = 0;
      AF_HIP(hipEventElapsedTime(&ms, e[4], e[5]));
      p->acc_ms[4] += ms;
      p->acc_n += 1;
    }
    p->ev_used = 0;
    return AF_OK;
  }
  for (int s = 0; s < p->ev_used; ++s) {
    hipEvent_t *e = &p->ev[(size_t)s * af_plan::kNumEv];
    AF_HIP(hipEventSynchronize(e[af_plan::kNumEv - 1]));
    for (int k = 0; k + 1 < af_plan::kNumEv; ++k) {
      float ms = 0;
      AF_HIP(hipEventElapsedTime(&ms, e[k], e[k + 1]));
      p->acc_ms[k] += ms;
    }
    float tot = 0;
    AF_HIP(hipEventElapsedTime(&tot, e[0], e[af_plan::kNumEv - 1]));
    p->acc_total += tot;
    p->acc_n += 1;
  }
  p->ev_used = 0;
  return AF_OK;
}

// Binning, IQRM, window masks, zero-DM and static mask into the channel-
// major buffer p->d_xT (src/core/dedispered.cu:1050-1118).
// With do_clean false, only binning, the statistics and the vote run (the
// clean pass is then issued per super-tile by execute()).
static af_status run_clean(af_plan *p, const void *d_in, hipStream_t st, int set,
                           bool do_clean = true) {
  const af_plan_desc &d = p->d;
  AF_TRY(record(p, set, 0, st));
  const void *src = d_in;
  if (d.ds > 1) {
    AF_HIP(launch_bin(d.dtype, d_in, p->d_xb, d.nsamp, d.nchan, d.ds, p->nbin, st));
    src = p->d_xb;
  }
  AF_TRY(record(p, set, 1, st));
  const bool iq = d.use_iqrm && p->nwin > 0;
  if (iq) {
    AF_HIP(launch_seg_stats(d.dtype, src, d.nchan, d.iqrm_c0, d.iqrm_c1, p->d_stat_items,
                            (int)p->stat_items.size(), p->d_part, st));
    AF_HIP(launch_seg_reduce(d.dtype, p->d_part, p->d_seg_items, p->nseg, p->Csub, p->d_segsum, st));
  }
  AF_TRY(record(p, set, 2, st));
  if (iq) AF_TRY(run_iqrm_vote(p, st));
  AF_TRY(record(p, set, 3, st));
  if (!do_clean) return AF_OK;
  CleanArgs ca{};
  ca.x = src;
  ca.xT = p->d_xT;
  ca.Lp = p->Lp;
  ca.nchan = d.nchan;
  ca.items = p->d_clean_items;
  ca.nitems = (int)p->clean_items.size();
  ca.segmask = iq ? p->d_segmask : nullptr;
  ca.stat_mask = p->d_static;
  ca.zdm = d.use_zero_dm && d.zdm_c1 > d.zdm_c0;
  ca.zc0 = d.zdm_c0;
  ca.zc1 = d.zdm_c1;
  ca.dd_c0 = d.dd_c0;
  ca.dd_c1 = d.dd_c1;
  ca.cut_in_pad = p->clean_cut ? 1 : 0;
  AF_HIP(launch_clean(d.dtype, ca, st));
  AF_TRY(record(p, set, 4, st));
  return AF_OK;
}

af_status execute(af_plan *p, const void *d_in, void *d_out, uint8_t *d_masks, hipStream_t st) {
  const af_plan_desc &d = p->d;
  int set = -1;
  if (p->profiling) {
    if ((size_t)(p->ev_used + 1) * af_plan::kNumEv > p->ev.size()) AF_TRY(collect(p));
    set = p->ev_used++;
  }
  if (!p->clean_tiled) {
    AF_TRY(run_clean(p, d_in, st, set));
    AF_TRY(run_dedisp(p, p->d_xT, p->Lp, d_out, st, set));
  } else {
    // per super-tile k: clean rows [k*T2, k*T2 + Lp) into d_xT (row r at
    // column r - k*T2), zero the columns past the input, dedisperse.  Stage
    // times: "dedisp" then holds the clean passes and both stages.
    AF_TRY(run_clean(p, d_in, st, set, false));
    AF_TRY(record(p, set, 4, st));
    const size_t es = dtype_size(d.dtype);
    const bool iq = d.use_iqrm && p->nwin > 0;
    for (int k = 0; k < p->nsuper; ++k) {
      const size_t a = (size_t)k * p->T2;
      CleanArgs ca{};
      ca.x = d.ds > 1 ? p->d_xb : d_in;
      ca.xT = (void *)((uintptr_t)p->d_xT - (uintptr_t)(a * es));
      ca.Lp = p->Lp;
      ca.nchan = d.nchan;
      ca.items = p->d_tile_items + p->tile_off[k];
      ca.nitems = p->tile_off[k + 1] - p->tile_off[k];
      ca.segmask = iq ? p->d_segmask : nullptr;
      ca.stat_mask = p->d_static;
      ca.zdm = d.use_zero_dm && d.zdm_c1 > d.zdm_c0;
      ca.zc0 = d.zdm_c0;
      ca.zc1 = d.zdm_c1;
      ca.dd_c0 = d.dd_c0;
      ca.dd_c1 = d.dd_c1;
      if (ca.nitems > 0) AF_HIP(launch_clean(d.dtype, ca, st));
      if (a + p->Lp > p->nbin) {
        // columns past the input read as zero (earlier super-tiles wrote there)
        const size_t c0 = p->nbin > a ? p->nbin - a : 0;
        AF_HIP(launch_zero((uint8_t *)p->d_xT + c0 * es, p->Lp * es, (p->Lp - c0) * es, p->nch, st));
      }
      AF_TRY(run_dedisp(p, p->d_xT, p->Lp, d_out, st, k + 1 == p->nsuper ? set : -1, k));
    }
  }
  if (d_masks && d.use_iqrm && p->nwin > 0)
    AF_HIP(launch_copy(d_masks, p->d_masks, (size_t)p->nwin * d.nchan, st));
  return AF_OK;
}

af_status run_iqrm_vote(af_plan *p, hipStream_t st) {
  const af_plan_desc &d = p->d;
  if (p->nlags > 0) {
    AF_HIP(launch_iqrm_lags(d.dtype, p->d_segsum, p->d_segs, p->d_wseg, p->d_wt0, p->d_wt1,
                            p->nwin, p->Csub, d.iqrm_mode, p->d_lags, p->nlags,
                            std::max(0.0f, d.iqrm_nsigma), p->d_flags, p->iqrm_lds, st));
  }
  AF_HIP(launch_iqrm_merge(p->d_flags, p->d_lags, p->nlags, p->nwin, p->Csub, d.iqrm_c0,
                           d.nchan, p->d_masks, p->lag_reach, st));
  if (!p->seg_is_win)
    AF_HIP(launch_segmask(p->d_masks, p->d_segs, p->nseg, d.nchan, p->d_segmask, st));
  return AF_OK;
}

// k_only >= 0: super-tile k_only alone, with xT holding its rows from column 0.
af_status run_dedisp(af_plan *p, const void *xT, size_t Lp, void *d_out, hipStream_t st, int set,
                     int k_only) {
  const af_plan_desc &d = p->d;
  DdArgs da{};
  da.xT = xT;
  da.Lp = Lp;
  da.delays = p->d_delays;
  da.nch = p->nch;
  da.ntiles = p->ntiles1;
  da.lds_rows = p->dd_lds_rows;
  da.lds_win = p->dd_lds_win;
  da.worklist = p->d_worklist;
  da.u8q = p->dd_q ? 1 : 0;
  da.f32q = p->dd_fq ? 1 : 0;
  da.nwork = (int)p->worklist.size();
  if (d.method == AF_METHOD_DIRECT) {
    da.nrows = p->ndm;
    da.ngrp = 1;
    da.grp_w = p->nch;
    da.out = d_out;
    da.ostride = p->nout;
    da.nvalid = p->nout;
    da.out16 = 0;
    AF_HIP(launch_dedisperse(d.dtype, da, st));
    AF_TRY(record(p, set, 5, st));
  } else {
    const bool inter16 = d.dtype == AF_U8 && d.sb_nchan <= 256;
    const size_t es = dtype_size(d.dtype), os = (size_t)acc_size(d.dtype);
    da.nrows = p->ndm_nom;
    da.ngrp = p->nsb;
    da.grp_w = d.sb_nchan;
    da.out = p->d_inter;
    da.ostride = p->L1;
    da.nvalid = p->L1;
    da.out16 = inter16 ? 1 : 0;
    Stage2Args sa{};
    sa.inter = p->d_inter;
    sa.L1 = p->L1;
    sa.nsb = p->nsb;
    sa.res = p->d_res;
    sa.maxres_j = p->d_maxres_j;
    sa.ndm = p->ndm;
    sa.ndm0 = d.ndm0;
    sa.ndm_nom = p->ndm_nom;
    sa.ostride = p->nout;
    sa.tile = p->s2_tile;
    sa.lds_len = p->s2_pk ? p->s2_pk : p->s2_lds;   // packed rows: elements per workgroup
    sa.inter16 = inter16 ? 1 : 0;
    sa.sbc = p->s2_sbc;
    sa.sboff = p->d_sboff;
    // super-tile k: stage 1 reads the cleaned buffer from column k*T2,
    // stage 2 writes output columns [k*T2, min(nout, (k+1)*T2))
    const int k0 = k_only >= 0 ? k_only : 0, k1 = k_only >= 0 ? k_only + 1 : p->nsuper;
    for (int k = k0; k < k1; ++k) {
      const size_t t0 = (size_t)k * p->T2;
      da.xT = (const uint8_t *)xT + (k_only >= 0 ? 0 : t0 * es);
      AF_HIP(launch_dedisperse(d.dtype, da, st));
      if (k + 1 == p->nsuper) AF_TRY(record(p, set, 5, st));
      sa.out = (uint8_t *)d_out + t0 * os;
      sa.nout = std::min(p->T2, p->nout - t0);
      AF_HIP(launch_stage2(d.dtype, sa, st));
    }
  }
  AF_TRY(record(p, set, 6, st));
  return AF_OK;
}

}  // namespace af

using namespace af;

extern "C" {

af_status af_plan_create(const af_plan_desc *desc, af_plan **plan) {
  if (!desc || !plan) return fail(AF_ERR_INVALID, "null argument");
  *plan = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(AF_ERR_NODEV, "No HIP devices found");
  AF_TRY(validate(*desc));
  af_plan *p = new af_plan();
  p->d = *desc;
  p->freq.assign(desc->freq_mhz, desc->freq_mhz + desc->nchan);
  p->d.freq_mhz = nullptr;
  if (desc->n_bad > 0) p->bad.assign(desc->bad_chans, desc->bad_chans + desc->n_bad);
  p->d.bad_chans = nullptr;
  (void)hipGetDevice(&p->device);
  af_status s = build(p);
  if (s != AF_OK) {
    std::string msg = g_err;
    release(p);
    delete p;
    g_err = msg;
    return s;
  }
  *plan = p;
  return AF_OK;
}

af_status af_plan_destroy(af_plan *plan) {
  if (!plan) return AF_OK;
  (void)hipSetDevice(plan->device);
  (void)hipStreamSynchronize(plan->stream);
  release(plan);
  delete plan;
  return AF_OK;
}

af_status af_plan_query(const af_plan *p, af_plan_info *info) {
  if (!p || !info) return fail(AF_ERR_INVALID, "null argument");
  info->nbin = p->nbin;
  info->nout = p->nout;
  info->ndm = p->ndm;
  info->out_elem_bytes = acc_size(p->d.dtype);
  info->nwin = p->d.use_iqrm ? p->nwin : 0;
  info->nsb = p->nsb;
  info->ndm_nom = p->ndm_nom;
  info->max_delay = p->maxdelay;
  info->scratch_bytes = p->scratch_bytes;
  info->adds = p->adds;
  info->inter_bytes = p->nsb > 0 ? (size_t)p->ndm_nom * p->nsb * p->L1 * p->inter_elem : 0;
  info->clean_passes = p->clean_tiled ? p->nsuper : 1;
  info->stage1_form = p->dd_lds_rows ? 16 : p->dd_q ? 8 : p->dd_fq ? 32 : 0;
  return AF_OK;
}

af_status af_plan_tables(const af_plan *p, int32_t *delays, int32_t *residuals) {
  if (!p) return fail(AF_ERR_INVALID, "null plan");
  if (delays) std::memcpy(delays, p->h_delays.data(), p->h_delays.size() * sizeof(int32_t));
  if (residuals && !p->h_res.empty())
    std::memcpy(residuals, p->h_res.data(), p->h_res.size() * sizeof(int32_t));
  return AF_OK;
}

af_status af_plan_windows(const af_plan *p, uint32_t *t0s, uint32_t *t1s) {
  if (!p) return fail(AF_ERR_INVALID, "null plan");
  for (int w = 0; w < p->nwin; ++w) {
    if (t0s) t0s[w] = p->wt0[w];
    if (t1s) t1s[w] = p->wt1[w];
  }
  return AF_OK;
}

af_status af_plan_set_profiling(af_plan *p, int enable) {
  if (!p) return fail(AF_ERR_INVALID, "null plan");
  (void)hipSetDevice(p->device);
  if (p->ev_used) AF_TRY(collect(p));
  if (enable < 0 || enable > 2) return fail(AF_ERR_INVALID, "profiling mode must be 0, 1 or 2");
  p->profiling = enable;
  std::fill(std::begin(p->acc_ms), std::end(p->acc_ms), 0.0);
  p->acc_total = 0;
  p->acc_n = 0;
  if (p->profiling && p->ev.empty()) {
    p->ev.resize((size_t)64 * af_plan::kNumEv);
    // timing-only events: no system-scope fence (cache write-back) when they
    // complete, which otherwise idles the GPU for several microseconds at
    // every stage boundary of a profiled execute
    for (auto &e : p->ev) AF_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  }
  return AF_OK;
}

af_status af_plan_stage_times(af_plan *p, af_stage_times *t) {
  if (!p || !t) return fail(AF_ERR_INVALID, "null argument");
  (void)hipSetDevice(p->device);
  AF_TRY(collect(p));
  const double n = p->acc_n ? (double)p->acc_n : 1.0;
  t->bin = (float)(p->acc_ms[0] / n);
  t->stats = (float)(p->acc_ms[1] / n);
  t->iqrm = (float)(p->acc_ms[2] / n);
  t->clean = (float)(p->acc_ms[3] / n);
  t->dedisp = (float)(p->acc_ms[4] / n);
  t->stage2 = (float)(p->acc_ms[5] / n);
  t->total = (float)(p->acc_total / n);
  t->n = p->acc_n;
  return AF_OK;
}

af_status af_execute(af_plan *p, const void *d_in, void *d_out, uint8_t *d_masks, void *stream) {
  if (!p || !d_in || !d_out) return fail(AF_ERR_INVALID, "null argument");
  AF_HIP(hipSetDevice(p->device));
  return execute(p, d_in, d_out, d_masks, (hipStream_t)stream);
}

af_status af_execute_spectrum(af_plan *p, const void *d_in, void *d_out, size_t t0, size_t t_len,
                              uint8_t *d_masks, void *stream) {
  if (!p || !d_in || !d_out) return fail(AF_ERR_INVALID, "null argument");
  const af_plan_desc &d = p->d;
  if (d.method != AF_METHOD_DIRECT || p->ndm != 1 || d.ds != 1)
    return fail(AF_ERR_INVALID, "a spectrum plan is direct, with one DM and ds == 1");
  if (t_len == 0 || t0 + t_len + (size_t)p->maxdelay > p->nbin)
    return fail(AF_ERR_INVALID, "spectrum window reaches beyond the input");
  AF_HIP(hipSetDevice(p->device));
  hipStream_t st = (hipStream_t)stream;
  AF_TRY(run_clean(p, d_in, st, -1));
  AF_HIP(launch_gather(d.dtype, p->d_xT, p->Lp, p->d_delays, p->nch, t0, t_len, d_out, st));
  if (d_masks && d.use_iqrm && p->nwin > 0)
    AF_HIP(launch_copy(d_masks, p->d_masks, (size_t)p->nwin * d.nchan, st));
  return AF_OK;
}

namespace {

// ------------------------------------------------ no-allocation scratch --
// af_execute_noalloc carves every table and buffer of its plan from this
// array.  Each device holds its own copy (allocated when the code object is
// loaded on that device, kept until the process ends).  In host code the
// name is not a device address: the address comes from hipGetSymbolAddress,
// once per device, with that device current.
constexpr size_t kNoAllocBytes = AF_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

af_status scratch_of(int dev, uint8_t **ptr) {
  if (dev < 0 || dev >= kMaxDevices) return fail(AF_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q)
      return fail(AF_ERR_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
    g_scratch_ptr[dev] = (uint8_t *)q;
  }
  *ptr = g_scratch_ptr[dev];
  return AF_OK;
}

constexpr size_t kStageBytes = (size_t)32 << 20;   // per page-locked staging buffer

af_status grow(void **buf, size_t *have, size_t need) {
  if (*have >= need && *buf) return AF_OK;
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *have = 0;
  if (hipMalloc(buf, std::max<size_t>(need, 1)) != hipSuccess) {
    *buf = nullptr;
    return fail(AF_ERR_NOMEM, "hipMalloc of " + std::to_string(need) + " bytes for a host call");
  }
  *have = need;
  return AF_OK;
}

// page-locked staging for af_execute_host (and the plan's copy workers);
// false when it cannot be had
bool staging(af_plan *p) {
  if (!p->copy_pool) {
    try {
      p->copy_pool = new CopyPool(copy_threads() - 1);
    } catch (...) {
      p->copy_pool = nullptr;   // no threads: the runtime's own pageable copies
      return false;
    }
  }
  for (int b = 0; b < 2; ++b) {
    if (!p->stage[b] && hipHostMalloc(&p->stage[b], kStageBytes, hipHostMallocDefault) != hipSuccess) {
      p->stage[b] = nullptr;
      return false;
    }
    if (!p->stage_ev[b] &&
        hipEventCreateWithFlags(&p->stage_ev[b], hipEventDisableTiming) != hipSuccess) {
      p->stage_ev[b] = nullptr;
      return false;
    }
  }
  return true;
}

// host -> device through the two staging buffers: the copy of chunk k into
// its buffer overlaps the DMA of chunk k-1
hipError_t upload_staged(af_plan *p, void *d, const void *h, size_t n, hipStream_t st) {
  for (size_t off = 0, k = 0; off < n; off += kStageBytes, ++k) {
    const size_t m = std::min(kStageBytes, n - off);
    const int b = (int)(k & 1);
    hipError_t e = hipEventSynchronize(p->stage_ev[b]);   // its previous DMA is done
    if (e != hipSuccess) return e;
    p->copy_pool->copy(p->stage[b], (const char *)h + off, m);
    e = hipMemcpyAsync((char *)d + off, p->stage[b], m, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(p->stage_ev[b], st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// device -> host: the DMA of chunk k overlaps the copy-out of chunk k-1
hipError_t download_staged(af_plan *p, void *h, const void *d, size_t n, hipStream_t st) {
  size_t prev_off = 0, prev_n = 0;
  int prev_b = -1;
  for (size_t off = 0, k = 0; off < n || prev_b >= 0; off += kStageBytes, ++k) {
    const int b = (int)(k & 1);
    const size_t m = off < n ? std::min(kStageBytes, n - off) : 0;
    if (m) {
      hipError_t e = hipMemcpyAsync(p->stage[b], (const char *)d + off, m, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipEventRecord(p->stage_ev[b], st);
      if (e != hipSuccess) return e;
    }
    if (prev_b >= 0) {
      const hipError_t e = hipEventSynchronize(p->stage_ev[prev_b]);
      if (e != hipSuccess) return e;
      p->copy_pool->copy((char *)h + prev_off, p->stage[prev_b], prev_n);
    }
    prev_b = m ? b : -1;
    prev_off = off;
    prev_n = m;
  }
  return hipSuccess;
}

}  // namespace

// Upload, execute, then download the plane (h_out) or, with h_images, the
// DM-time images of slices of sps columns made on the device from the plane.
static af_status host_call(af_plan *p, const void *h_in, void *h_out, uint8_t *h_masks,
                           size_t sps, int isize, uint8_t *h_images) {
  AF_HIP(hipSetDevice(p->device));
  const size_t in_bytes = p->d.nsamp * (size_t)p->d.nchan * dtype_size(p->d.dtype);
  const size_t out_bytes = (size_t)p->ndm * p->nout * acc_size(p->d.dtype);
  const size_t m_bytes = (size_t)p->nwin * p->d.nchan;
  // device buffers stay with the plan for the next call
  AF_TRY(grow(&p->hx_in, &p->hx_in_n, in_bytes));
  AF_TRY(grow(&p->hx_out, &p->hx_out_n, out_bytes));
  uint8_t *dm = nullptr;
  if (h_masks && m_bytes) {
    AF_TRY(grow(&p->hx_masks, &p->hx_masks_n, m_bytes));
    dm = (uint8_t *)p->hx_masks;
  }
  size_t img_bytes = 0;
  if (h_images) {
    const size_t nsl = (p->nout + sps - 1) / sps;
    img_bytes = nsl * (size_t)isize * (size_t)isize * 3;
    AF_TRY(grow(&p->hx_img, &p->hx_img_n, img_bytes));
  }
  hipStream_t st = p->stream;
  // rows in page-locked memory from af_host_alloc go straight to the device;
  // pageable ones through the staging buffers
  const bool in_pinned = host_pinned(h_in, in_bytes);
  const bool out_pinned = !h_images && host_pinned(h_out, out_bytes);
  const bool staged = !(in_pinned && (out_pinned || h_images)) && staging(p);
  hipError_t e = staged && !in_pinned
                     ? upload_staged(p, p->hx_in, h_in, in_bytes, st)
                     : hipMemcpyAsync(p->hx_in, h_in, in_bytes, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);
    return fail(AF_ERR_RUNTIME, std::string("H2D: ") + hipGetErrorString(e));
  }
  af_status s = execute(p, p->hx_in, p->hx_out, dm, st);
  if (s != AF_OK) {
    (void)hipStreamSynchronize(st);
    return s;
  }
  if (h_images) {
    const int pt = p->d.dtype == AF_F32 ? AF_F32 : p->d.dtype == AF_U32 ? AF_U64 : AF_U32;
    AF_TRY(image_slices(p->hx_out, pt, (size_t)p->ndm, p->nout, p->nout, sps, isize,
                        (uint8_t *)p->hx_img, st, &p->hx_imgscr, &p->hx_imgscr_n));
    e = hipMemcpyAsync(h_images, p->hx_img, img_bytes, hipMemcpyDeviceToHost, st);
  } else {
    e = staged && !out_pinned
            ? download_staged(p, h_out, p->hx_out, out_bytes, st)
            : hipMemcpyAsync(h_out, p->hx_out, out_bytes, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess && dm) e = hipMemcpyAsync(h_masks, dm, m_bytes, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);
    return fail(AF_ERR_RUNTIME, std::string("D2H: ") + hipGetErrorString(e));
  }
  return AF_OK;
}

af_status af_execute_host(af_plan *p, const void *h_in, void *h_out, uint8_t *h_masks) {
  if (!p || !h_in || !h_out) return fail(AF_ERR_INVALID, "null argument");
  return host_call(p, h_in, h_out, h_masks, 0, 0, nullptr);
}

af_status af_execute_host_images(af_plan *p, const void *h_in, size_t sps, int size,
                                 uint8_t *h_images, uint8_t *h_masks) {
  if (!p || !h_in || !h_images) return fail(AF_ERR_INVALID, "null argument");
  if (sps == 0 || size <= 0 || size > 8192)
    return fail(AF_ERR_INVALID, "zero slice width or bad image size");
  return host_call(p, h_in, nullptr, h_masks, sps, size, h_images);
}

af_status af_execute_noalloc(const af_plan_desc *desc, const void *d_in, void *d_out,
                            uint8_t *d_masks, void *stream) {
  if (!desc || !d_in || !d_out) return fail(AF_ERR_INVALID, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(AF_ERR_NODEV, "No HIP devices found");
  AF_TRY(validate(*desc));
  int dev = 0;
  AF_HIP(hipGetDevice(&dev));
  uint8_t *base = nullptr;
  AF_TRY(scratch_of(dev, &base));
  // the same plan af_plan_create builds, on host memory, with the whole
  // channel-major buffer (no per-super-tile cleaning) and its device buffers
  // carved from the scratch: first counted, then placed
  auto make = [&](af_plan &p) {
    p.d = *desc;
    p.freq.assign(desc->freq_mhz, desc->freq_mhz + desc->nchan);
    p.d.freq_mhz = nullptr;
    if (desc->n_bad > 0) p.bad.assign(desc->bad_chans, desc->bad_chans + desc->n_bad);
    p.d.bad_chans = nullptr;
    p.d.clean_budget = ~(size_t)0;
    p.device = dev;
    p.arena_cap = kNoAllocBytes;
    p.arena_st = (hipStream_t)stream;
  };
  size_t need = 0;
  {
    af_plan p;
    make(p);
    p.arena_measure = true;
    AF_TRY(build(&p));
    need = p.arena_used;
  }
  if (need > kNoAllocBytes)
    return fail(AF_ERR_NOMEM, "af_execute_noalloc: the call needs " + std::to_string(need) +
                                  " bytes of device scratch; the library holds " +
                                  std::to_string(kNoAllocBytes));
  af_plan p;
  make(p);
  p.arena = base;
  AF_TRY(build(&p));
  return execute(&p, d_in, d_out, d_masks, (hipStream_t)stream);
}

af_status af_time_execute(af_plan *p, const void *d_in, void *d_out, int reps, float *ms) {
  if (!p || !d_in || !d_out || !ms || reps < 1) return fail(AF_ERR_INVALID, "bad argument");
  AF_HIP(hipSetDevice(p->device));
  hipEvent_t a, b;
  AF_HIP(hipEventCreate(&a));
  AF_HIP(hipEventCreate(&b));
  af_status s = AF_OK;
  if (hipEventRecord(a, p->stream) != hipSuccess) s = fail(AF_ERR_RUNTIME, "event record");
  for (int i = 0; i < reps && s == AF_OK; ++i) s = execute(p, d_in, d_out, nullptr, p->stream);
  if (s == AF_OK && hipEventRecord(b, p->stream) != hipSuccess) s = fail(AF_ERR_RUNTIME, "event record");
  if (s == AF_OK && hipEventSynchronize(b) != hipSuccess) s = fail(AF_ERR_RUNTIME, "event sync");
  if (s == AF_OK && hipEventElapsedTime(ms, a, b) != hipSuccess) s = fail(AF_ERR_RUNTIME, "elapsed");
  (void)hipStreamSynchronize(p->stream);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return s;
}

af_status af_device_query(int dev, af_device_props *props) {
  if (!props) return fail(AF_ERR_INVALID, "null argument");
  std::memset(props, 0, sizeof(*props));
  hipDeviceProp_t pr;
  AF_HIP(hipGetDeviceProperties(&pr, dev));
  std::strncpy(props->name, pr.name, sizeof(props->name) - 1);
  std::strncpy(props->arch, pr.gcnArchName, sizeof(props->arch) - 1);
  props->cu_count = pr.multiProcessorCount;
  props->clock_khz = pr.clockRate;
  int a = 0, b = 0;
  (void)hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, dev);
  props->lds_per_block = std::max(a, b);
  props->lds_per_cu = (int)pr.maxSharedMemoryPerMultiProcessor;
  props->l2_bytes = pr.l2CacheSize;
  props->hbm_bytes = pr.totalGlobalMem;
  return AF_OK;
}

af_status af_device_count(int *n) {
  if (!n) return fail(AF_ERR_INVALID, "null argument");
  *n = 0;
  hipError_t e = hipGetDeviceCount(n);
  if (e != hipSuccess) {
    *n = 0;
    return AF_OK;
  }
  return AF_OK;
}

af_status af_set_device(int dev) {
  AF_HIP(hipSetDevice(dev));
  return AF_OK;
}

af_status af_malloc(void **p, size_t bytes) {
  if (!p) return fail(AF_ERR_INVALID, "null argument");
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) return fail(AF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return AF_OK;
}

af_status af_free(void *p) {
  if (p) AF_HIP(hipFree(p));
  return AF_OK;
}

af_status af_host_alloc(void **p, size_t bytes) {
  if (!p) return fail(AF_ERR_INVALID, "null argument");
  hipError_t e = hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) return fail(AF_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  host_pinned_add(*p, bytes ? bytes : 1);
  return AF_OK;
}

af_status af_host_free(void *p) {
  if (p) {
    host_pinned_remove(p);
    AF_HIP(hipHostFree(p));
  }
  return AF_OK;
}

af_status af_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
  AF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  AF_HIP(hipStreamSynchronize((hipStream_t)stream));
  return AF_OK;
}

af_status af_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
  AF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  AF_HIP(hipStreamSynchronize((hipStream_t)stream));
  return AF_OK;
}

af_status af_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
  AF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return AF_OK;
}

af_status af_memset(void *dst, int value, size_t bytes, void *stream) {
  AF_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
  return AF_OK;
}

af_status af_stream_create(void **stream) {
  if (!stream) return fail(AF_ERR_INVALID, "null argument");
  hipStream_t s;
  AF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return AF_OK;
}

af_status af_stream_destroy(void *stream) {
  if (stream) AF_HIP(hipStreamDestroy((hipStream_t)stream));
  return AF_OK;
}

af_status af_stream_sync(void *stream) {
  AF_HIP(hipStreamSynchronize((hipStream_t)stream));
  return AF_OK;
}

af_status af_sync(void) {
  AF_HIP(hipDeviceSynchronize());
  return AF_OK;
}

const char *af_last_error(void) { return g_err.c_str(); }

const char *af_version(void) { return "astroflow_amd 0.1 (gfx950)"; }

}  // extern "C"
