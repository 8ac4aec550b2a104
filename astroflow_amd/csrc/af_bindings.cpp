// af_bindings.cpp -- the `_astroflow_core` Python module on top of the C-ABI.
//
// Mirrors the reference's pybind11 surface for the dedispersion path
// (src/bindpy/bindings.cpp:13-74, src/bindpy/pyapi.cpp:53-282): the same
// module name, class names, constructors and positional argument order, so
// python/astroflow's callers (dedispered.py:59-72,360-373) run unchanged.
// Every call goes through libastroflow_amd (HIP kernels for gfx950); there is
// no CPU compute path.  Additions: `_dedisperse_spec_raw` returns the plane
// itself (and the IQRM masks), which the reference never exposes.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "af_cpu.h"
#include "af_image.h"
#include "af_snr.h"
#include "astroflow_amd.h"

namespace py = pybind11;

namespace {

// ------------------------------------------------------------ config types

// include/data.h:24-50
struct Header {
  float mjd = 0, tsamp = 0, fch1 = 0, foff = 0;
  int nbits = 0, nchans = 0;
  long ndata = 0;
  std::string filename;
  Header() = default;
  Header(float mjd_, float tsamp_, float fch1_, float foff_, int nbits_, int nchans_, long ndata_,
         std::string fn)
      : mjd(mjd_), tsamp(tsamp_), fch1(fch1_), foff(foff_), nbits(nbits_), nchans(nchans_),
        ndata(ndata_), filename(std::move(fn)) {}
  explicit Header(py::object h)
      : mjd(h.attr("mjd").cast<float>()), tsamp(h.attr("tsamp").cast<float>()),
        fch1(h.attr("fch1").cast<float>()), foff(h.attr("foff").cast<float>()),
        nbits(h.attr("nbits").cast<int>()), nchans(h.attr("nchans").cast<int>()),
        ndata(h.attr("ndata").cast<long>()), filename(h.attr("filename").cast<std::string>()) {}
};

// include/rfi.h:12-57 (same defaults and positional order)
struct IQRMConfig {
  int mode = 0;
  float radius_frac = 0.10f, nsigma = 3.0f;
  double geofactor = 1.5, win_sec = 5.0, hop_sec = 2.5;
  bool include_tail = true;
  IQRMConfig() = default;
  IQRMConfig(int m, float r, float n, double g, double w, double h, bool t)
      : mode(m), radius_frac(r), nsigma(n), geofactor(g), win_sec(w), hop_sec(h), include_tail(t) {}
  explicit IQRMConfig(py::object c)
      : mode(c.attr("mode").cast<int>()), radius_frac(c.attr("radius_frac").cast<float>()),
        nsigma(c.attr("nsigma").cast<float>()), geofactor(c.attr("geofactor").cast<double>()),
        win_sec(c.attr("win_sec").cast<double>()), hop_sec(c.attr("hop_sec").cast<double>()),
        include_tail(c.attr("include_tail").cast<bool>()) {}
};

struct RFIConfig {
  bool use_mask = false, use_zero_dm = false, use_iqrm = false;
  IQRMConfig iqrm_cfg;
  RFIConfig() = default;
  RFIConfig(bool m, bool z, bool i, IQRMConfig c)
      : use_mask(m), use_zero_dm(z), use_iqrm(i), iqrm_cfg(c) {}
  explicit RFIConfig(py::object c)
      : use_mask(c.attr("use_mask").cast<bool>()), use_zero_dm(c.attr("use_zero_dm").cast<bool>()),
        use_iqrm(c.attr("use_iqrm").cast<bool>()), iqrm_cfg(IQRMConfig(c.attr("iqrm_cfg"))) {}
};

// include/data.h:52-83
struct DedispersedDataUint8 {
  py::list dm_times;               // (512, 512, 3) uint8 BGR images
  std::vector<size_t> shape;
  int dm_ndata = 0, downtsample_ndata = 0;
  float dm_low = 0, dm_high = 0, dm_step = 0, tsample = 0;
  std::string filname;
};

struct DedisperedData {
  py::list dm_times;               // one flat plane [dm * time]
  std::vector<size_t> shape;
  int dm_ndata = 0, downtsample_ndata = 0;
  float dm_low = 0, dm_high = 0, dm_step = 0, tsample = 0;
  std::string filname;
};

// ------------------------------------------------------------------ errors

[[noreturn]] void raise_status(af_status s, const char *what) {
  std::string msg = std::string(what) + ": " + af_last_error();
  if (s == AF_ERR_INVALID) throw py::value_error(msg);
  if (s == AF_ERR_NOMEM) throw std::bad_alloc();
  throw std::runtime_error(msg);
}

void check(af_status s, const char *what) {
  if (s != AF_OK) raise_status(s, what);
}

template <typename T> constexpr int dtype_code();
template <> constexpr int dtype_code<uint8_t>() { return AF_U8; }
template <> constexpr int dtype_code<uint16_t>() { return AF_U16; }
template <> constexpr int dtype_code<uint32_t>() { return AF_U32; }
template <> constexpr int dtype_code<float>() { return AF_F32; }

template <typename T> struct OutOf { using type = uint32_t; };
template <> struct OutOf<uint32_t> { using type = uint64_t; };
template <> struct OutOf<float> { using type = float; };

// Whitespace-separated channel numbers (src/rfi/rfimarker.cu:441-466); a
// missing or empty file means no bad channels.
std::vector<int> read_mask_file(const std::string &path) {
  std::vector<int> v;
  std::ifstream f(path);
  if (!f.is_open()) return v;
  std::string line;
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    int c;
    while (ss >> c) v.push_back(c);
  }
  return v;
}

// ------------------------------------------------------- the plain call --

struct Prepared {
  af_plan_desc d{};
  std::vector<double> freq;
  std::vector<int> bad;
  size_t chan_start = 0, chan_end = 0, dm_steps = 0, down_ndata = 0;
  // d with its table pointers on this object's vectors (a copy of a
  // Prepared keeps its own vectors)
  af_plan_desc desc() const {
    af_plan_desc x = d;
    x.freq_mhz = freq.data();
    x.bad_chans = bad.empty() ? nullptr : bad.data();
    x.n_bad = (int)bad.size();
    return x;
  }
};

// Channel range, DM steps, reference frequency and RFI set-up exactly as
// src/core/dedispered.cu:1003-1036,1085-1118 derive them from the call.
Prepared prepare(int dtype, const Header &h, size_t nsamp, float dm_low, float dm_high,
                 float freq_start, float freq_end, float dm_step, int ds,
                 const std::string &mask_file, const RFIConfig &rfi) {
  Prepared p;
  if (h.nchans <= 0) throw py::value_error("header.nchans must be positive");
  if (ds < 1) throw py::value_error("time_downsample must be >= 1");
  if (!(dm_step > 0.0f)) throw py::value_error("dm_step must be positive");
  if (dm_high < dm_low) throw py::value_error("dm_high must be >= dm_low");
  const size_t nchans = (size_t)h.nchans;
  p.freq.resize(nchans);
  for (size_t i = 0; i < nchans; ++i) {
    const float f = h.fch1 + (float)i * h.foff;   // float Header fields
    p.freq[i] = f;
  }
  const float fmin = (float)p.freq.front(), fmax = (float)p.freq.back();
  if (freq_start < fmin || freq_end > fmax) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "Frequency range [%.3f-%.3f MHz] out of spectrum range [%.3f-%.3f MHz]",
                  freq_start, freq_end, fmin, fmax);
    throw py::value_error(msg);
  }
  const float span = fmax - fmin;
  size_t cs = (size_t)((freq_start - fmin) / span * (float)(nchans - 1));
  size_t ce = (size_t)((freq_end - fmin) / span * (float)(nchans - 1));
  ce = std::min(nchans - 1, ce);
  if (cs >= nchans || ce >= nchans || cs > ce) {
    char msg[256];
    std::snprintf(msg, sizeof(msg), "Invalid channel range [%zu-%zu] for %zu channels", cs, ce,
                  nchans);
    throw py::value_error(msg);
  }
  p.chan_start = cs;
  p.chan_end = ce;
  p.dm_steps = (size_t)((dm_high - dm_low) / dm_step) + 1;
  p.down_ndata = (nsamp + (size_t)ds - 1) / (size_t)ds;
  af_plan_desc &d = p.d;
  d.dtype = dtype;
  d.nchan = h.nchans;
  d.nsamp = nsamp;
  d.ds = ds;
  d.tsamp = h.tsamp;
  d.f_ref_mhz = p.freq[ce];
  d.dd_c0 = (int)cs;
  d.dd_c1 = (int)ce + 1;
  d.dm_low = dm_low;
  d.dm_step = dm_step;
  d.ndm = (int)p.dm_steps;
  d.nout = 0;
  d.method = AF_METHOD_SUBBAND;
  d.sb_nchan = 32;   // src/core/dedispered.cu:29-35
  d.ndm0 = 32;
  // IQRM takes the inclusive end as exclusive (:1089); an empty range gives
  // no window masks in the reference (src/core/iqrm.cu:363-364)
  d.use_iqrm = rfi.use_iqrm && ce > cs;
  d.iqrm_c0 = (int)cs;
  d.iqrm_c1 = (int)ce;
  d.iqrm_mode = rfi.iqrm_cfg.mode;
  d.iqrm_radius_frac = rfi.iqrm_cfg.radius_frac;
  d.iqrm_nsigma = rfi.iqrm_cfg.nsigma;
  d.iqrm_geofactor = rfi.iqrm_cfg.geofactor;
  // window and hop in binned samples (src/core/iqrm.cu:368-369, with the
  // binned sample time header.tsamp * time_downsample in float)
  const double tb = (double)(h.tsamp * (float)ds);
  const double win = std::max(0.0, rfi.iqrm_cfg.win_sec), hop = std::max(0.0, rfi.iqrm_cfg.hop_sec);
  const unsigned W = win > 0.0 ? (unsigned)std::floor(win / tb) : 0u;
  const unsigned H = hop > 0.0 ? (unsigned)std::floor(hop / tb) : W;
  d.iqrm_win = (W >= p.down_ndata) ? 0u : W;
  d.iqrm_hop = H;
  d.iqrm_include_tail = rfi.iqrm_cfg.include_tail;
  d.use_zero_dm = rfi.use_zero_dm;
  d.zdm_c0 = (int)cs;
  d.zdm_c1 = (int)ce;
  if (rfi.use_mask) p.bad = read_mask_file(mask_file);
  d.bad_chans = p.bad.empty() ? nullptr : p.bad.data();
  d.n_bad = (int)p.bad.size();
  d.freq_mhz = p.freq.data();
  return p;
}

struct Raw {
  py::array plane;
  py::object masks;
  Prepared prep;
};

// Plans of recent calls, keyed by their full description (scalar fields,
// frequency table, static mask, device), most recent last.  A call takes its
// plan out while it runs (plans are confined to one thread) and puts it back
// afterwards; the oldest beyond kPlanCacheMax are destroyed.  Astroflow's
// search loop calls with the same parameters chunk after chunk: the plan,
// its tables and its host-call buffers are then built once.
constexpr size_t kPlanCacheMax = 2;
std::mutex g_plan_mu;
std::vector<std::pair<std::string, af_plan *>> g_plans;

std::string plan_key(const Prepared &pr, int dev) {
  af_plan_desc d = pr.d;
  d.freq_mhz = nullptr;
  d.bad_chans = nullptr;
  std::string k((const char *)&d, sizeof d);
  k.append((const char *)pr.freq.data(), pr.freq.size() * sizeof(double));
  k.append((const char *)pr.bad.data(), pr.bad.size() * sizeof(int));
  k.append((const char *)&dev, sizeof dev);
  return k;
}

af_plan *take_plan(const std::string &key) {
  std::lock_guard<std::mutex> g(g_plan_mu);
  for (size_t i = g_plans.size(); i-- > 0;)
    if (g_plans[i].first == key) {
      af_plan *p = g_plans[i].second;
      g_plans.erase(g_plans.begin() + (std::ptrdiff_t)i);
      return p;
    }
  return nullptr;
}

void give_back(std::string key, af_plan *p) {
  std::vector<af_plan *> drop;
  {
    std::lock_guard<std::mutex> g(g_plan_mu);
    g_plans.emplace_back(std::move(key), p);
    while (g_plans.size() > kPlanCacheMax) {
      drop.push_back(g_plans.front().second);
      g_plans.erase(g_plans.begin());
    }
  }
  for (af_plan *q : drop) af_plan_destroy(q);
}

void clear_plan_cache() {
  std::vector<std::pair<std::string, af_plan *>> all;
  {
    std::lock_guard<std::mutex> g(g_plan_mu);
    all.swap(g_plans);
  }
  for (auto &e : all) af_plan_destroy(e.second);
}

// A cached or new plan for one call: back to the cache if the call
// succeeded, destroyed otherwise.
struct PlanLease {
  std::string key;
  af_plan *plan = nullptr;
  bool ok = false;
  PlanLease(const Prepared &pr, int dev) : key(plan_key(pr, dev)) {
    plan = take_plan(key);
    if (!plan) {
      const af_plan_desc d = pr.desc();
      check(af_plan_create(&d, &plan), "af_plan_create");
    }
  }
  ~PlanLease() {
    if (!plan) return;
    if (ok) give_back(std::move(key), plan);
    else af_plan_destroy(plan);
  }
  PlanLease(const PlanLease &) = delete;
  PlanLease &operator=(const PlanLease &) = delete;
};

template <typename T>
Raw run(py::array_t<T, py::array::c_style> data, const Header &h, float dm_low, float dm_high,
        float freq_start, float freq_end, float dm_step, int ds, int target_id,
        const std::string &mask_file, const RFIConfig &rfi, bool want_masks) {
  const py::buffer_info bi = data.request();
  if (bi.ndim < 1 || bi.ndim > 3) throw std::runtime_error("data must be 1D, 2D, or 3D");
  if (h.ndata <= 0) throw py::value_error("header.ndata must be positive");
  if ((size_t)bi.size < (size_t)h.ndata * (size_t)std::max(h.nchans, 0))
    throw py::value_error("data holds fewer than header.ndata * header.nchans samples");
  int ndev = 0;
  if (af_device_count(&ndev) != AF_OK || ndev == 0) throw std::runtime_error("No HIP devices found");
  if (target_id < 0 || target_id >= ndev) throw py::value_error("target_id names no device");
  check(af_set_device(target_id), "af_set_device");
  Raw r;
  r.prep = prepare(dtype_code<T>(), h, (size_t)h.ndata, dm_low, dm_high, freq_start, freq_end,
                   dm_step, ds, mask_file, rfi);
  PlanLease lease(r.prep, target_id);
  af_plan *plan = lease.plan;
  af_plan_info info;
  check(af_plan_query(plan, &info), "af_plan_query");
  using O = typename OutOf<T>::type;
  py::array_t<O> plane({(py::ssize_t)info.ndm, (py::ssize_t)info.nout});
  py::object masks = py::none();
  uint8_t *mp = nullptr;
  if (want_masks && info.nwin > 0) {
    py::array_t<uint8_t> m({(py::ssize_t)info.nwin, (py::ssize_t)h.nchans});
    mp = m.mutable_data();
    masks = m;
  }
  const T *src = static_cast<const T *>(bi.ptr);
  O *dst = plane.mutable_data();
  af_status st;
  {
    py::gil_scoped_release nogil;
    st = af_execute_host(plan, src, dst, mp);
  }
  check(st, "af_execute_host");
  lease.ok = true;
  r.plane = plane;
  r.masks = masks;
  return r;
}

// Columns per image slice: t_sample seconds of (downsampled) samples
// (include/data.h:619-727).
size_t slice_cols(float tsamp, int ds, float t_sample) {
  const float tb = tsamp * (float)ds;
  const size_t sps = (size_t)(t_sample / tb);
  if (sps == 0) throw py::value_error("t_sample is shorter than one (downsampled) sample");
  return sps;
}

template <typename T>
DedispersedDataUint8 dedisperse_spec(py::array_t<T, py::array::c_style> data, Header header,
                                     float dm_low, float dm_high, float freq_start, float freq_end,
                                     float dm_step, int time_downsample, float t_sample,
                                     int target_id, std::string mask_file, RFIConfig rficfg) {
  // the run() checks and plan, but the plane stays on the device: the
  // images are made there and only they come back
  const py::buffer_info bi = data.request();
  if (bi.ndim < 1 || bi.ndim > 3) throw std::runtime_error("data must be 1D, 2D, or 3D");
  if (header.ndata <= 0) throw py::value_error("header.ndata must be positive");
  if ((size_t)bi.size < (size_t)header.ndata * (size_t)std::max(header.nchans, 0))
    throw py::value_error("data holds fewer than header.ndata * header.nchans samples");
  int ndev = 0;
  if (af_device_count(&ndev) != AF_OK || ndev == 0) throw std::runtime_error("No HIP devices found");
  if (target_id < 0 || target_id >= ndev) throw py::value_error("target_id names no device");
  check(af_set_device(target_id), "af_set_device");
  const size_t sps = slice_cols(header.tsamp, time_downsample, t_sample);
  const Prepared prep = prepare(dtype_code<T>(), header, (size_t)header.ndata, dm_low, dm_high,
                                freq_start, freq_end, dm_step, time_downsample, mask_file, rficfg);
  PlanLease lease(prep, target_id);
  af_plan_info info;
  check(af_plan_query(lease.plan, &info), "af_plan_query");
  constexpr int kSize = 512;
  const size_t nslices = ((size_t)info.nout + sps - 1) / sps;
  const size_t img_bytes = (size_t)kSize * kSize * 3;
  std::vector<uint8_t> buf(nslices * img_bytes);
  af_status st;
  {
    py::gil_scoped_release nogil;
    st = af_execute_host_images(lease.plan, bi.ptr, sps, kSize, buf.data(), nullptr);
  }
  check(st, "af_execute_host_images");
  lease.ok = true;
  py::list images;
  for (size_t s = 0; s < nslices; ++s) {
    py::array_t<uint8_t> a({kSize, kSize, 3});
    std::memcpy(a.mutable_data(), buf.data() + s * img_bytes, img_bytes);
    images.append(a);
  }
  DedispersedDataUint8 out;
  out.dm_times = images;
  out.shape = {512, 512, 3};
  out.dm_ndata = 512;
  out.downtsample_ndata = 512;
  out.dm_low = dm_low;
  out.dm_high = dm_high;
  out.dm_step = dm_step;
  out.tsample = t_sample;
  out.filname = header.filename;
  return out;
}

template <typename T>
py::object dedisperse_spec_raw(py::array_t<T, py::array::c_style> data, Header header,
                               float dm_low, float dm_high, float freq_start, float freq_end,
                               float dm_step, int time_downsample, int target_id,
                               std::string mask_file, RFIConfig rficfg, bool return_masks) {
  Raw r = run<T>(data, header, dm_low, dm_high, freq_start, freq_end, dm_step, time_downsample,
                 target_id, mask_file, rficfg, return_masks);
  if (return_masks) return py::make_tuple(r.plane, r.masks);
  return r.plane;
}

// ---------------------------------------------------- single-DM spectrum --
// include/data.h:85-95
struct SpectrumOut {
  py::object data;              // [ntimes][nchans], the input's element type
  size_t nchans = 0, ntimes = 0;
  float tstart = 0, tend = 0, dm = 0;
  int nbits = 0;
  float freq_start = 0, freq_end = 0;
};

struct DevBuf {
  void *p = nullptr;
  explicit DevBuf(size_t n) { check(af_malloc(&p, n), "af_malloc"); }
  ~DevBuf() { if (p) af_free(p); }
};

// cpucal::dedisperse_spec_with_dm<T> (include/cpucal.hpp:341-483): the time
// window and channel selection as the reference derives them, the RFI
// stages (IQRM over [chan_start, chan_end_excl), window masks, static mask;
// no zero-DM) on the local slice the reference cleans, then every channel
// read at its delay for this DM.  The input array is not modified (the
// reference cleans it in place).
template <typename T>
SpectrumOut spectrum_at_dm(py::array_t<T, py::array::c_style> data, Header h, float tstart,
                           float tend, float dm, float freq_start, float freq_end,
                           std::string maskfile, RFIConfig rfi) {
  const py::buffer_info bi = data.request();
  if (tend <= tstart) throw py::value_error("tend must be > tstart");
  if (h.nchans <= 0 || h.ndata <= 0) throw py::value_error("empty header");
  if ((size_t)bi.size < (size_t)h.ndata * (size_t)h.nchans)
    throw py::value_error("data holds fewer than header.ndata * header.nchans samples");
  const size_t ndata = (size_t)h.ndata, nchans = (size_t)h.nchans;
  const size_t t_start = (size_t)(tstart / h.tsamp), t_end = (size_t)(tend / h.tsamp);
  const size_t t_len_req = t_end > t_start ? t_end - t_start : 0;
  if (t_len_req == 0 || t_start >= ndata) throw py::value_error("Invalid time window for this file.");
  if (freq_start >= freq_end) throw py::value_error("freq_end must be > freq_start");
  if (dm < 0.0f) throw py::value_error("dm must be >= 0");
  if (h.foff < 0)
    throw py::value_error("frequency channels are in descending order, which is not supported yet.");
  std::vector<double> freq(nchans);
  std::vector<float> ff(nchans);
  for (size_t i = 0; i < nchans; ++i) {
    ff[i] = h.fch1 + (float)i * h.foff;
    freq[i] = ff[i];
  }
  const float f0 = ff.front(), fN = ff.back();
  const float fmin = std::min(f0, fN), fmax = std::max(f0, fN);
  if (freq_start < fmin || freq_end > fmax) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "Frequency range [%.3f, %.3f] MHz out of file range [%.3f, %.3f] MHz",
                  freq_start, freq_end, fmin, fmax);
    throw py::value_error(buf);
  }
  auto to_index = [&](float f) -> size_t {
    const double pos = (double)(f - f0) / (double)(fN - f0);
    double raw = pos * (double)(nchans - 1);
    raw = std::min(std::max(raw, 0.0), (double)(nchans - 1));
    return (size_t)std::floor(raw + 1e-9);
  };
  size_t cs = to_index(freq_start), ce = to_index(freq_end);
  if (ce < cs) std::swap(cs, ce);
  const size_t cx = std::min(ce + 1, nchans);
  const size_t sel = cx > cs ? cx - cs : 0;
  if (sel == 0) throw py::value_error("Empty channel selection.");
  const double f_high = std::max(freq[cs], freq[cx - 1]);
  int ndev = 0;
  if (af_device_count(&ndev) != AF_OK || ndev == 0) throw std::runtime_error("No HIP devices found");
  // delays first (a one-sample probe plan gives max_delay), then the slice
  af_plan_desc d{};
  d.dtype = dtype_code<T>();
  d.nchan = (int)nchans;
  d.ds = 1;
  d.tsamp = h.tsamp;
  d.freq_mhz = freq.data();
  d.f_ref_mhz = f_high;
  d.dd_c0 = (int)cs;
  d.dd_c1 = (int)cx;
  d.dm_low = dm;
  d.dm_step = 1.0;
  d.ndm = 1;
  d.method = AF_METHOD_DIRECT;
  d.nsamp = 1;   // the delays do not depend on the length
  af_plan *probe = nullptr;
  check(af_plan_create(&d, &probe), "af_plan_create");
  af_plan_info pi;
  af_plan_query(probe, &pi);
  af_plan_destroy(probe);
  const size_t dmax = (size_t)pi.max_delay;
  const size_t cap = ndata - t_start;
  const size_t t_len = cap > dmax ? std::min(t_len_req, cap - dmax) : 0;
  if (t_len == 0) throw py::value_error("Time window too short for this DM and band.");
  const size_t r0 = t_start > dmax ? t_start - dmax : 0;
  const size_t off = t_start - r0;
  const size_t slen = std::min(t_len + off + 2 * dmax, ndata - r0);
  std::vector<int> bad;
  if (rfi.use_mask) bad = read_mask_file(maskfile);
  d.nsamp = slen;
  // this path passes the exclusive end (include/cpucal.hpp:425-430); an
  // empty range gives no window masks (src/core/iqrm.cu:363-364)
  d.use_iqrm = rfi.use_iqrm && cx > cs;
  d.iqrm_c0 = (int)cs;
  d.iqrm_c1 = (int)cx;
  d.iqrm_mode = rfi.iqrm_cfg.mode;
  d.iqrm_radius_frac = rfi.iqrm_cfg.radius_frac;
  d.iqrm_nsigma = rfi.iqrm_cfg.nsigma;
  d.iqrm_geofactor = rfi.iqrm_cfg.geofactor;
  const double tb = (double)h.tsamp;
  const double win = std::max(0.0, rfi.iqrm_cfg.win_sec), hop = std::max(0.0, rfi.iqrm_cfg.hop_sec);
  const unsigned W = win > 0.0 ? (unsigned)std::floor(win / tb) : 0u;
  const unsigned H = hop > 0.0 ? (unsigned)std::floor(hop / tb) : W;
  d.iqrm_win = (W >= slen) ? 0u : W;
  d.iqrm_hop = H;
  d.iqrm_include_tail = rfi.iqrm_cfg.include_tail;
  d.use_zero_dm = 0;
  d.bad_chans = bad.empty() ? nullptr : bad.data();
  d.n_bad = (int)bad.size();
  af_plan *plan = nullptr;
  check(af_plan_create(&d, &plan), "af_plan_create");
  std::unique_ptr<af_plan, decltype(&af_plan_destroy)> guard(plan, af_plan_destroy);
  py::array_t<T> out({(py::ssize_t)t_len, (py::ssize_t)sel});
  const T *src = static_cast<const T *>(bi.ptr) + r0 * nchans;
  T *dst = out.mutable_data();
  af_status st;
  {
    py::gil_scoped_release nogil;
    DevBuf din(slen * nchans * sizeof(T)), dout(t_len * sel * sizeof(T));
    st = af_memcpy_h2d(din.p, src, slen * nchans * sizeof(T), nullptr);
    if (st == AF_OK) st = af_execute_spectrum(plan, din.p, dout.p, off, t_len, nullptr, nullptr);
    if (st == AF_OK) st = af_memcpy_d2h(dst, dout.p, t_len * sel * sizeof(T), nullptr);
  }
  check(st, "af_execute_spectrum");
  SpectrumOut r;
  r.data = out;
  r.ntimes = t_len;
  r.nchans = sel;
  r.tstart = tstart;
  r.tend = tstart + (float)t_len * h.tsamp;
  r.dm = dm;
  r.nbits = h.nbits;
  r.freq_start = std::min(ff[cs], ff[cx - 1]);
  r.freq_end = std::max(ff[cs], ff[cx - 1]);
  return r;
}

// Rows read from a file, in page-locked memory from af_host_alloc when a
// device is present, the rows take at most a quarter of the machine's
// memory and the allocation succeeds (a host call then sends them, or any
// row range of them, to the device without staging), else in ordinary
// memory.
struct HostRows {
  uint8_t *ptr = nullptr;
  size_t n = 0;
  bool pinned = false;
  std::vector<uint8_t> heap;

  static size_t pin_limit() {
    const long pages = sysconf(_SC_PHYS_PAGES), psz = sysconf(_SC_PAGE_SIZE);
    return pages > 0 && psz > 0 ? (size_t)pages * (size_t)psz / 4 : 0;
  }
  explicit HostRows(size_t bytes) : n(bytes) {
    int ndev = 0;
    void *p = nullptr;
    if (bytes >= ((size_t)1 << 20) && bytes <= pin_limit() && af_device_count(&ndev) == AF_OK &&
        ndev > 0 && af_host_alloc(&p, bytes) == AF_OK) {
      ptr = (uint8_t *)p;
      pinned = true;
    } else {
      heap.resize(bytes);
      ptr = heap.data();
    }
  }
  ~HostRows() {
    if (pinned) af_host_free(ptr);
  }
  HostRows(const HostRows &) = delete;
  HostRows &operator=(const HostRows &) = delete;
  uint8_t *data() { return ptr; }
  const uint8_t *data() const { return ptr; }
  size_t size() const { return n; }
};

// ---------------------------------------------------- filterbank ingest --
// SIGPROC filterbank: a keyword header between HEADER_START and HEADER_END
// (each keyword an int32 length + text, followed by its value), then
// [ndata][nifs][nchans] samples.  Data with a negative channel offset is
// flipped to ascending frequency, as src/data/filterbank.cpp:793-842 does.
struct Filterbank {
  std::string filename;
  int nchans = 0, nifs = 1, nbits = 8, telescope_id = 0, machine_id = 0, data_type = 1;
  double fch1 = 0, foff = 0, tstart = 0, tsamp = 0;
  long ndata = 0;
  std::string source_name;
  std::shared_ptr<HostRows> raw;   // samples as stored (after the flip)

  Filterbank() = default;
  explicit Filterbank(const std::string &path) : filename(path) { read(); }

  // the library's SIGPROC reader (af_fil_open / af_fil_read): header, rows,
  // and the flip of a descending band
  void read() {
    af_fil_reader *r = nullptr;
    af_fil_info info;
    check(af_fil_open(filename.c_str(), &r, &info), "af_fil_open");
    nchans = info.nchans;
    nifs = 1;
    nbits = info.nbits;
    fch1 = info.fch1;
    foff = info.foff;
    tstart = info.tstart;
    tsamp = info.tsamp;
    ndata = (long)info.ndata;
    raw = std::make_shared<HostRows>((size_t)ndata * nchans * (size_t)(nbits / 8));
    af_status st;
    {
      py::gil_scoped_release nogil;
      st = af_fil_read(r, 0, (size_t)ndata, raw->data());
    }
    af_fil_close(r);
    check(st, "af_fil_read");
  }

  Header header() const {
    return Header((float)tstart, (float)tsamp, (float)fch1, (float)foff, nbits, nchans, ndata,
                  filename);
  }

  py::array data() const {
    const std::vector<py::ssize_t> shape{ndata, nifs, nchans};
    auto keep = new std::shared_ptr<HostRows>(raw);
    py::capsule cap(keep, [](void *p) { delete static_cast<std::shared_ptr<HostRows> *>(p); });
    switch (nbits) {
    case 8: return py::array_t<uint8_t>(shape, (const uint8_t *)raw->data(), cap);
    case 16: return py::array_t<uint16_t>(shape, (const uint16_t *)raw->data(), cap);
    default: return py::array_t<uint32_t>(shape, (const uint32_t *)raw->data(), cap);
    }
  }
};

// PSRFITS search-mode file (the reference's Python PsrFits,
// python/astroflow/io/psrfits.py:19-90, over af_psrfits_open / read):
// pol-averaged, ascending band, [ndata][nchans].
struct PsrFits {
  std::string filename;
  af_fil_info info{};
  std::shared_ptr<HostRows> raw;

  explicit PsrFits(const std::string &path) : filename(path) {
    af_psrfits_reader *r = nullptr;
    check(af_psrfits_open(path.c_str(), &r, &info), "af_psrfits_open");
    const size_t es = info.dtype == AF_U8 ? 1 : info.dtype == AF_U16 ? 2 : 4;
    raw = std::make_shared<HostRows>((size_t)info.ndata * info.nchans * es);
    af_status st;
    {
      py::gil_scoped_release nogil;
      st = af_psrfits_read(r, 0, (size_t)info.ndata, raw->data());
    }
    af_psrfits_close(r);
    check(st, "af_psrfits_read");
  }

  Header header() const {
    return Header((float)info.tstart, (float)info.tsamp, (float)info.fch1, (float)info.foff,
                  info.nbits, info.nchans, (long)info.ndata, filename);
  }

  py::array data() const {
    const std::vector<py::ssize_t> shape{(py::ssize_t)info.ndata, (py::ssize_t)info.nchans};
    auto keep = new std::shared_ptr<HostRows>(raw);
    py::capsule cap(keep, [](void *p) { delete static_cast<std::shared_ptr<HostRows> *>(p); });
    switch (info.dtype) {
    case AF_U8: return py::array_t<uint8_t>(shape, (const uint8_t *)raw->data(), cap);
    case AF_U16: return py::array_t<uint16_t>(shape, (const uint16_t *)raw->data(), cap);
    default: return py::array_t<float>(shape, (const float *)raw->data(), cap);
    }
  }
};

// target 0 of _dedispered_fil: the host cores (include/cpucal.hpp:75-201,
// checks and slicing as there; delays, layout and imaging in af_cpu.h /
// af_image.h).  Every slice is imaged at its own width (the reference
// images the last, shorter slice at the first slice's width, reading past
// its buffer).
template <typename T>
DedispersedDataUint8 dedispered_fil_cpu(const Filterbank &fil, float dm_low, float dm_high,
                                        float freq_start, float freq_end, float dm_step, int ds,
                                        float t_sample) {
  const Header h = fil.header();
  const size_t nchans = (size_t)h.nchans, ndata = (size_t)h.ndata;
  if (nchans == 0 || ndata == 0) throw py::value_error("empty filterbank");
  std::vector<double> freq(nchans);
  for (size_t i = 0; i < nchans; ++i) freq[i] = h.fch1 + (float)i * h.foff;   // float fields
  const float fmin = (float)freq.front(), fmax = (float)freq.back();
  if (freq_start < fmin || freq_end > fmax) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "Frequency range [%.3f-%.3f MHz] out of filterbank range [%.3f-%.3f MHz]",
                  freq_start, freq_end, fmin, fmax);
    throw py::value_error(msg);
  }
  const float span = fmax - fmin;
  size_t cs = (size_t)((freq_start - fmin) / span * (float)(nchans - 1));
  size_t ce = (size_t)((freq_end - fmin) / span * (float)(nchans - 1));
  ce = std::min(nchans - 1, ce);
  if (cs >= nchans || ce >= nchans || cs >= ce) throw py::value_error("Invalid chan parameters");
  if (ds < 1) throw py::value_error("time_downsample must be >= 1");
  if (dm_low > dm_high || dm_low < 0 || !(dm_step > 0.0f))
    throw py::value_error("Invalid DM parameters");
  if ((double)t_sample > (double)ndata * fil.tsamp)
    throw py::value_error("t_sample exceeds total observation time");
  const size_t sps = (size_t)std::round((double)t_sample / fil.tsamp);
  if (sps == 0) throw py::value_error("t_sample is shorter than one sample");
  const size_t ndm = (size_t)((dm_high - dm_low) / dm_step) + 1;
  const size_t nslices = (ndata + sps - 1) / sps;
  constexpr int kSize = 512;
  const size_t img_bytes = (size_t)kSize * kSize * 3;
  std::vector<uint8_t> buf(nslices * img_bytes);
  {
    py::gil_scoped_release nogil;
    const std::vector<int32_t> delays =
        afcpu::delay_table(freq, cs, ce, dm_low, dm_step, ndm, (double)h.tsamp);
    const std::vector<T> xt = afcpu::channel_major((const T *)fil.raw->data(), ndata, nchans, cs, ce);
    std::vector<uint64_t> plane;
    for (size_t s = 0; s < nslices; ++s) {
      const size_t start = s * sps, len = std::min(sps, ndata - start);
      const size_t down = (len + (size_t)ds - 1) / (size_t)ds;
      plane.resize(ndm * down);
      afcpu::dedisperse_slice(xt, ndata, ce - cs, delays, ndm, start, len, ds, plane.data());
      afimg::slice_image(plane.data(), down, 0, (int)ndm, (int)down, kSize,
                         buf.data() + s * img_bytes);
    }
  }
  py::list images;
  for (size_t s = 0; s < nslices; ++s) {
    py::array_t<uint8_t> a({kSize, kSize, 3});
    std::memcpy(a.mutable_data(), buf.data() + s * img_bytes, img_bytes);
    images.append(a);
  }
  DedispersedDataUint8 out;
  out.dm_times = images;
  out.shape = {512, 512, 3};
  out.dm_ndata = 512;
  out.downtsample_ndata = 512;
  out.dm_low = dm_low;
  out.dm_high = dm_high;
  out.dm_step = dm_step;
  out.tsample = t_sample;
  out.filname = h.filename;
  return out;
}

DedispersedDataUint8 dedispered_fil(std::string filename, float dm_low, float dm_high,
                                    float freq_start, float freq_end, float dm_step,
                                    int time_downsample, float t_sample, int target, int target_id,
                                    std::string mask_file, RFIConfig rficfg) {
  if (target != 0 && target != 1)
    throw py::value_error("target must be 0 (host cores) or 1 (GPU)");
  Filterbank fil(filename);
  if (target == 0) {
    switch (fil.nbits) {
    case 8:
      return dedispered_fil_cpu<uint8_t>(fil, dm_low, dm_high, freq_start, freq_end, dm_step,
                                         time_downsample, t_sample);
    case 16:
      return dedispered_fil_cpu<uint16_t>(fil, dm_low, dm_high, freq_start, freq_end, dm_step,
                                          time_downsample, t_sample);
    default:
      return dedispered_fil_cpu<uint32_t>(fil, dm_low, dm_high, freq_start, freq_end, dm_step,
                                          time_downsample, t_sample);
    }
  }
  const Header h = fil.header();
  py::array a = fil.data();
  switch (fil.nbits) {
  case 8:
    return dedisperse_spec<uint8_t>(a.cast<py::array_t<uint8_t, py::array::c_style>>(), h, dm_low,
                                    dm_high, freq_start, freq_end, dm_step, time_downsample,
                                    t_sample, target_id, mask_file, rficfg);
  case 16:
    return dedisperse_spec<uint16_t>(a.cast<py::array_t<uint16_t, py::array::c_style>>(), h,
                                     dm_low, dm_high, freq_start, freq_end, dm_step,
                                     time_downsample, t_sample, target_id, mask_file, rficfg);
  default:
    return dedisperse_spec<uint32_t>(a.cast<py::array_t<uint32_t, py::array::c_style>>(), h,
                                     dm_low, dm_high, freq_start, freq_end, dm_step,
                                     time_downsample, t_sample, target_id, mask_file, rficfg);
  }
}

template <typename T>
SpectrumOut fil_spectrum_at_dm(const Filterbank &fil, float tstart, float tend, float dm,
                               float freq_start, float freq_end, std::string maskfile,
                               RFIConfig rfi) {
  if (fil.nbits != (int)sizeof(T) * 8) throw py::value_error("filterbank nbits does not match");
  py::array a = fil.data();
  return spectrum_at_dm<T>(a.cast<py::array_t<T, py::array::c_style>>(), fil.header(), tstart,
                           tend, dm, freq_start, freq_end, maskfile, rfi);
}

template <typename T>
void def_spec(py::module_ &m) {
  m.def("_dedisperse_spec", &dedisperse_spec<T>, py::arg("data"), py::arg("header"),
        py::arg("dm_low"), py::arg("dm_high"), py::arg("freq_start"), py::arg("freq_end"),
        py::arg("dm_step") = 1.0f, py::arg("time_downsample") = 2, py::arg("t_sample") = 0.5f,
        py::arg("target_id") = 0, py::arg("mask_file") = "mask.txt",
        py::arg("rficfg") = RFIConfig());
  m.def("_dedisperse_spec_with_dm", &spectrum_at_dm<T>, py::arg("data"), py::arg("header"),
        py::arg("tstart"), py::arg("tend"), py::arg("dm"), py::arg("freq_start"),
        py::arg("freq_end"), py::arg("maskfile") = "mask.txt", py::arg("rficfg") = RFIConfig());
  m.def("_dedisperse_spec_raw", &dedisperse_spec_raw<T>, py::arg("data"), py::arg("header"),
        py::arg("dm_low"), py::arg("dm_high"), py::arg("freq_start"), py::arg("freq_end"),
        py::arg("dm_step") = 1.0f, py::arg("time_downsample") = 1, py::arg("target_id") = 0,
        py::arg("mask_file") = "mask.txt", py::arg("rficfg") = RFIConfig(),
        py::arg("return_masks") = false);
}

template <typename D>
void def_result(py::module_ &m, const char *name) {
  py::class_<D>(m, name)
      .def(py::init<>())
      .def_readonly("dm_times", &D::dm_times)
      .def_readonly("shape", &D::shape)
      .def_readonly("dm_ndata", &D::dm_ndata)
      .def_readonly("downtsample_ndata", &D::downtsample_ndata)
      .def_readonly("dm_low", &D::dm_low)
      .def_readonly("dm_high", &D::dm_high)
      .def_readonly("dm_step", &D::dm_step)
      .def_readonly("tsample", &D::tsample)
      .def_readonly("filname", &D::filname);
}

// _calculate_frb_snr (src/bindpy/frb_snr.cpp:144-398): the same argument
// order, defaults and return tuple (snr, width, peak, (mean, std, info)).
py::tuple calculate_frb_snr(py::array spec, py::object noise_range, double threshold_sigma,
                            py::object toa_sample_idx, py::object fitting_window_samples,
                            py::object tsamp, py::object target_time_us) {
  auto a = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(spec);
  if (!a) throw std::runtime_error("spec must be a 2D numpy array");
  if (a.ndim() != 2) throw std::runtime_error("spec must be 2D (time x frequency)");
  afsnr::SnrInput in;
  in.spec = a.data();
  in.ntime = (int)a.shape(0);
  in.nfreq = (int)a.shape(1);
  in.sigma = threshold_sigma;
  if (!noise_range.is_none() && py::len(noise_range) > 0) {
    in.has_noise_ranges = true;
    for (py::handle item : noise_range) {
      py::sequence pr = py::reinterpret_borrow<py::sequence>(item);
      if (py::len(pr) != 2) continue;
      in.noise_ranges.emplace_back(pr[0].cast<int>(), pr[1].cast<int>());
    }
  }
  if (!toa_sample_idx.is_none()) {
    in.has_toa = true;
    in.toa = toa_sample_idx.cast<int>();
  }
  if (!fitting_window_samples.is_none()) {
    in.has_window = true;
    in.window = fitting_window_samples.cast<int>();
  }
  if (!tsamp.is_none()) {
    in.has_tsamp = true;
    in.tsamp = tsamp.cast<double>();
  }
  if (!target_time_us.is_none()) {
    in.has_target = true;
    in.target_us = target_time_us.cast<double>();
  }
  const afsnr::SnrResult r = afsnr::boxcar_snr(in);
  py::dict q;
  q["fit_converged"] = r.converged;
  if (r.empty) return py::make_tuple(-1.0, 0, 0, py::make_tuple(0.0, 1.0, q));
  if (r.fallback) {
    q["method"] = "fallback";
    return py::make_tuple(r.snr, r.width, r.peak, py::make_tuple(r.noise_mean, r.noise_std, q));
  }
  q["method"] = "boxcar";
  q["width_samples"] = r.width;
  q["search_start"] = r.search_start;
  q["search_end"] = r.search_end;
  q["downsample"] = r.downsample;
  return py::make_tuple(r.snr, r.width, r.peak, py::make_tuple(r.noise_mean, r.noise_std, q));
}

}  // namespace

PYBIND11_MODULE(_astroflow_core, m) {
  m.doc() = "astroflow dedispersion core on MI355X (gfx950), over libastroflow_amd";

  m.def("_plan_cache_size", [] {
    std::lock_guard<std::mutex> g(g_plan_mu);
    return g_plans.size();
  }, "plans kept from recent _dedisperse_spec / _dedispered_fil calls");
  m.def("_plan_cache_clear", &clear_plan_cache, "destroy the kept plans (frees their device memory)");
  // destroy kept plans before interpreter shutdown, while the HIP runtime is up
  py::module_::import("atexit").attr("register")(py::cpp_function(&clear_plan_cache));

  def_result<DedispersedDataUint8>(m, "DedispersedDataUint8");
  def_result<DedisperedData>(m, "DedisperedData");

  py::class_<Header>(m, "Header")
      .def(py::init<>())
      .def(py::init<float, float, float, float, int, int, long, std::string>())
      .def(py::init<py::object>())
      .def_readonly("mjd", &Header::mjd)
      .def_readonly("tsamp", &Header::tsamp)
      .def_readonly("fch1", &Header::fch1)
      .def_readonly("foff", &Header::foff)
      .def_readonly("nbits", &Header::nbits)
      .def_readonly("nchans", &Header::nchans)
      .def_readonly("ndata", &Header::ndata)
      .def_readonly("filename", &Header::filename);

  py::class_<IQRMConfig>(m, "IQRMconfig")
      .def(py::init<>())
      .def(py::init<int, float, float, double, double, double, bool>())
      .def(py::init<py::object>())
      .def_readonly("mode", &IQRMConfig::mode)
      .def_readonly("radius_frac", &IQRMConfig::radius_frac)
      .def_readonly("nsigma", &IQRMConfig::nsigma)
      .def_readonly("geofactor", &IQRMConfig::geofactor)
      .def_readonly("win_sec", &IQRMConfig::win_sec)
      .def_readonly("hop_sec", &IQRMConfig::hop_sec)
      .def_readonly("include_tail", &IQRMConfig::include_tail);

  py::class_<RFIConfig>(m, "RFIconfig")
      .def(py::init<>())
      .def(py::init<bool, bool, bool, IQRMConfig>())
      .def(py::init<py::object>())
      .def_readonly("use_zero_dm", &RFIConfig::use_zero_dm)
      .def_readonly("use_mask", &RFIConfig::use_mask)
      .def_readonly("use_iqrm", &RFIConfig::use_iqrm)
      .def_readonly("iqrm_cfg", &RFIConfig::iqrm_cfg);

  py::class_<Filterbank>(m, "Filterbank")
      .def(py::init<>())
      .def(py::init<const std::string &>())
      .def_readonly("filename", &Filterbank::filename)
      .def_readonly("nchans", &Filterbank::nchans)
      .def_readonly("nifs", &Filterbank::nifs)
      .def_readonly("nbits", &Filterbank::nbits)
      .def_readonly("fch1", &Filterbank::fch1)
      .def_readonly("foff", &Filterbank::foff)
      .def_readonly("tstart", &Filterbank::tstart)
      .def_readonly("tsamp", &Filterbank::tsamp)
      .def_readonly("ndata", &Filterbank::ndata)
      .def_property_readonly("data", &Filterbank::data)
      .def_property_readonly("header", &Filterbank::header);

  py::class_<PsrFits>(m, "PsrFits")
      .def(py::init<const std::string &>())
      .def_readonly("filename", &PsrFits::filename)
      .def_property_readonly("nchans", [](const PsrFits &f) { return f.info.nchans; })
      .def_property_readonly("nbits", [](const PsrFits &f) { return f.info.nbits; })
      .def_property_readonly("npol", [](const PsrFits &f) { return f.info.npol; })
      .def_property_readonly("ndata", [](const PsrFits &f) { return (long)f.info.ndata; })
      .def_property_readonly("fch1", [](const PsrFits &f) { return f.info.fch1; })
      .def_property_readonly("foff", [](const PsrFits &f) { return f.info.foff; })
      .def_property_readonly("tsamp", [](const PsrFits &f) { return f.info.tsamp; })
      .def_property_readonly("mjd", [](const PsrFits &f) { return f.info.tstart; })
      .def_property_readonly("data", &PsrFits::data)
      .def("get_spectrum", &PsrFits::data)
      .def_property_readonly("header", &PsrFits::header);

  py::class_<SpectrumOut>(m, "Spectrum")
      .def(py::init<>())
      .def_readonly("data", &SpectrumOut::data)
      .def_readonly("ntimes", &SpectrumOut::ntimes)
      .def_readonly("nchans", &SpectrumOut::nchans)
      .def_readonly("tstart", &SpectrumOut::tstart)
      .def_readonly("tend", &SpectrumOut::tend)
      .def_readonly("dm", &SpectrumOut::dm)
      .def_readonly("nbits", &SpectrumOut::nbits)
      .def_readonly("freq_start", &SpectrumOut::freq_start)
      .def_readonly("freq_end", &SpectrumOut::freq_end);
  // the reference's per-type names (src/bindpy/bindings.cpp:22-24) share one class
  m.attr("Spectrum8") = m.attr("Spectrum");
  m.attr("Spectrum16") = m.attr("Spectrum");
  m.attr("Spectrum32") = m.attr("Spectrum");

  m.def("_dedispered_fil_with_dm_uint8", &fil_spectrum_at_dm<uint8_t>, py::arg("fil"),
        py::arg("tstart"), py::arg("tend"), py::arg("dm"), py::arg("freq_start"),
        py::arg("freq_end"), py::arg("maskfile"), py::arg("rficfg"));
  m.def("_dedispered_fil_with_dm_uint16", &fil_spectrum_at_dm<uint16_t>, py::arg("fil"),
        py::arg("tstart"), py::arg("tend"), py::arg("dm"), py::arg("freq_start"),
        py::arg("freq_end"), py::arg("maskfile"), py::arg("rficfg"));
  m.def("_dedispered_fil_with_dm_uint32", &fil_spectrum_at_dm<uint32_t>, py::arg("fil"),
        py::arg("tstart"), py::arg("tend"), py::arg("dm"), py::arg("freq_start"),
        py::arg("freq_end"), py::arg("maskfile"), py::arg("rficfg"));

  // float32 first: pybind11 tries exact dtype matches before converting, and
  // a float array must not be silently converted to an integer overload
  def_spec<float>(m);
  def_spec<uint8_t>(m);
  def_spec<uint16_t>(m);
  def_spec<uint32_t>(m);

  m.def("_dedispered_fil", &dedispered_fil, py::arg("filename"), py::arg("dm_low"),
        py::arg("dm_high"), py::arg("freq_start"), py::arg("freq_end"), py::arg("dm_step"),
        py::arg("time_downsample"), py::arg("t_sample"), py::arg("target") = 1,
        py::arg("target_id") = 0, py::arg("maskfile") = "mask.txt",
        py::arg("rficfg") = RFIConfig());

  m.def("_calculate_frb_snr", &calculate_frb_snr, py::arg("spec"),
        py::arg("noise_range") = py::none(), py::arg("threshold_sigma") = 5.0,
        py::arg("toa_sample_idx") = py::none(), py::arg("fitting_window_samples") = py::none(),
        py::arg("tsamp") = py::none(), py::arg("target_time_us") = 256.0);

  m.def("_describe_spec", [](size_t nsamp, const Header &h, float dm_low, float dm_high,
                             float freq_start, float freq_end, float dm_step, int ds,
                             const std::string &mask_file, const RFIConfig &rfi) {
    // the plan description _dedisperse_spec derives from its arguments (no
    // device needed; for tests of the parameter derivation)
    const Prepared pr = prepare(AF_U8, h, nsamp, dm_low, dm_high, freq_start, freq_end, dm_step,
                                ds, mask_file, rfi);
    const af_plan_desc &d = pr.d;
    py::dict o;
    o["chan_start"] = pr.chan_start;
    o["chan_end"] = pr.chan_end;
    o["dm_steps"] = pr.dm_steps;
    o["down_ndata"] = pr.down_ndata;
    o["f_ref_mhz"] = d.f_ref_mhz;
    o["dd_c0"] = d.dd_c0;
    o["dd_c1"] = d.dd_c1;
    o["use_iqrm"] = d.use_iqrm;
    o["iqrm_c0"] = d.iqrm_c0;
    o["iqrm_c1"] = d.iqrm_c1;
    o["iqrm_win"] = d.iqrm_win;
    o["iqrm_hop"] = d.iqrm_hop;
    o["use_zero_dm"] = d.use_zero_dm;
    o["zdm_c0"] = d.zdm_c0;
    o["zdm_c1"] = d.zdm_c1;
    o["bad_chans"] = pr.bad;
    return o;
  }, py::arg("nsamp"), py::arg("header"), py::arg("dm_low"), py::arg("dm_high"),
     py::arg("freq_start"), py::arg("freq_end"), py::arg("dm_step") = 1.0f,
     py::arg("time_downsample") = 1, py::arg("mask_file") = "mask.txt",
     py::arg("rficfg") = RFIConfig());

  m.def("_image_slice", [](py::array_t<float, py::array::c_style> plane, int size) {
    // one DM-time image from a float plane (exposed for tests of the imaging)
    if (plane.ndim() != 2) throw py::value_error("plane must be 2-D");
    py::array_t<uint8_t> out({size, size, 3});
    afimg::slice_image(plane.data(), (size_t)plane.shape(1), 0, (int)plane.shape(0),
                       (int)plane.shape(1), size, out.mutable_data());
    return out;
  }, py::arg("plane"), py::arg("size") = 512);
}
