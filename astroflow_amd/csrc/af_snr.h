// af_snr.h -- boxcar S/N of a dedispersed dynamic spectrum (host C++).
//
// Behaviour of the reference's `_calculate_frb_snr`
// (src/bindpy/frb_snr.cpp:144-398, helpers :12-140): the spectrum's rows are
// summed over frequency in double, optionally decimated to a target time
// resolution, the noise level is a sigma-clipped median/MAD estimate, and a
// boxcar of widths 1..16 plus 32 log-spaced widths up to the fitting window
// is slid over the series (within a search radius around a given TOA).  The
// best window's noise is then re-estimated without the samples near it.
// The reference runs this on the host after a candidate is found; so does
// the drop-in `_calculate_frb_snr`.  af_snr.hip runs the same steps for a
// spectrum already on the device (the per-sample sums and the boxcar search
// there, the order statistics here).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <utility>
#include <vector>

namespace afsnr {

// median with the two middle values averaged for an even count
inline double median_of(std::vector<double> v) {
  const size_t n = v.size();
  if (n == 0) return 0.0;
  const size_t h = n / 2;
  std::nth_element(v.begin(), v.begin() + h, v.end());
  const double upper = v[h];
  if (n & 1) return upper;
  const double lower = *std::max_element(v.begin(), v.begin() + h);
  return 0.5 * (lower + upper);
}

inline double mean_of(const std::vector<double> &v) {
  if (v.empty()) return 0.0;
  double s = 0.0;
  for (double x : v) s += x;
  return s / (double)v.size();
}

// population standard deviation about `mu`
inline double pop_std(const std::vector<double> &v, double mu) {
  if (v.empty()) return 0.0;
  double s = 0.0;
  for (double x : v) s += (x - mu) * (x - mu);
  return std::sqrt(s / (double)v.size());
}

// Sigma-clipped (mean, std): up to `rounds` passes that keep the samples
// within sigma * (1.4826 MAD, or the std when the MAD is zero) of the
// median; a pass that removes nothing, or would leave fewer than
// max(10, 20 %) samples, ends the clipping.  Non-finite samples are dropped
// first; an empty set gives (0, 1), and a zero spread falls back to the
// unclipped std and then to 1.
inline std::pair<double, double> clipped_stats(const std::vector<double> &data, double sigma,
                                               int rounds = 3) {
  std::vector<double> keep;
  keep.reserve(data.size());
  for (double x : data)
    if (std::isfinite(x)) keep.push_back(x);
  if (keep.empty()) return {0.0, 1.0};
  const std::vector<double> finite = keep;
  for (int r = 0; r < rounds; ++r) {
    const double med = median_of(keep);
    std::vector<double> dev(keep.size());
    for (size_t i = 0; i < keep.size(); ++i) dev[i] = std::fabs(keep[i] - med);
    const double mad = median_of(std::move(dev));
    const double spread = mad > 0.0 ? 1.4826 * mad : pop_std(keep, mean_of(keep));
    if (!(spread > 0.0)) break;
    const double lim = sigma * spread;
    std::vector<double> next;
    next.reserve(keep.size());
    for (double x : keep)
      if (std::fabs(x - med) <= lim) next.push_back(x);
    if (next.size() == keep.size()) break;
    if (next.size() < std::max<size_t>(10, (size_t)(0.2 * (double)keep.size()))) break;
    keep.swap(next);
  }
  const double mu = mean_of(keep);
  double sd = pop_std(keep, mu);
  if (!(sd > 0.0)) sd = pop_std(finite, mean_of(finite));
  if (!(sd > 0.0)) sd = 1.0;
  return {mu, sd};
}

// Boxcar widths: 1..min(16, wmax), then the distinct roundings of 32
// log-spaced values from 17 to wmax that lie in [17, wmax].
inline std::vector<int> boxcar_widths(int wmax) {
  std::vector<int> w;
  if (wmax <= 1) return {1};
  for (int i = 1; i <= std::min(16, wmax); ++i) w.push_back(i);
  if (wmax <= 16) return w;
  const double a = std::log10(17.0), b = std::log10((double)wmax);
  std::vector<int> g;
  for (int i = 0; i < 32; ++i) {
    const long long r = std::llround(std::pow(10.0, a + (double)i / 31.0 * (b - a)));
    if (r >= 17 && r <= wmax) g.push_back((int)r);
  }
  std::sort(g.begin(), g.end());
  g.erase(std::unique(g.begin(), g.end()), g.end());
  w.insert(w.end(), g.begin(), g.end());
  return w;
}

struct SnrInput {
  const float *spec = nullptr;   // [ntime][nfreq], row-major
  int ntime = 0, nfreq = 0;
  // a non-empty range list was given (entries that are not pairs are
  // dropped from noise_ranges, but the list still counts as given)
  bool has_noise_ranges = false;
  std::vector<std::pair<int, int>> noise_ranges;   // native-rate [start, end)
  double sigma = 5.0;
  bool has_toa = false;
  int toa = 0;                   // native-rate sample index
  bool has_window = false;
  int window = 0;                // fitting window, native-rate samples
  bool has_tsamp = false, has_target = false;
  double tsamp = 0.0, target_us = 0.0;
};

struct SnrResult {
  double snr = -1.0;
  int width = 0, peak = 0;       // native-rate samples
  double noise_mean = 0.0, noise_std = 1.0;
  bool converged = false;
  bool fallback = false;         // method "fallback" (no boxcar found)
  bool empty = false;            // no data: no method reported
  int search_start = 0, search_end = 0, downsample = 1;
};

// The series after the frequency sums, decimated, with everything the
// boxcar search needs: noise statistics, prefix sums, widths, search range.
struct SnrPrep {
  std::vector<double> ts, cs;
  std::vector<int> widths;   // ascending; those >= n are not searched
  int ds = 1, n = 0, s0 = 0, s1 = 0, argmax = 0;
  double mu = 0.0, sd = 1.0;
};

// Best boxcar of one width: the first start (lowest index) with the largest
// S/N among centres inside the search range.
struct WidthBest {
  bool any = false;
  double snr = 0.0, sum = 0.0;
  int centre = 0;
};

// ts: the frequency-summed series at the native rate (ntime values)
inline SnrPrep snr_prepare(const SnrInput &in, std::vector<double> ts) {
  SnrPrep p;
  // decimation to the target resolution (whole groups only)
  int ds = 1;
  if (in.has_tsamp && in.has_target) {
    if (in.tsamp > 0.0) ds = (int)std::llround(in.target_us * 1e-6 / in.tsamp);
    ds = std::max(ds, 1);
  }
  if (ds > 1) {
    const int ng = in.ntime / ds;
    if (ng >= 1) {
      std::vector<double> d((size_t)ng);
      for (int g = 0; g < ng; ++g) {
        double s = 0.0;
        for (int i = 0; i < ds; ++i) s += ts[(size_t)g * ds + i];
        d[(size_t)g] = s;
      }
      ts.swap(d);
    } else {
      ds = 1;
    }
  }
  const int n = (int)ts.size();
  const int span = n * ds;   // native-rate samples covered
  int wmax = in.has_window ? std::max(4, std::min(in.window, span / 2))
                           : std::max(8, std::min(span / 4, 512));
  wmax = std::max(1, std::min(wmax, span));
  const int wmax_ds = std::min(n, (int)std::ceil((double)wmax / ds));

  int s0 = 0, s1 = n;
  if (in.has_toa) {
    const int rad = std::max(wmax_ds, (int)std::ceil(100.0 / ds));
    const int c = in.toa / ds;
    s0 = std::max(0, c - rad);
    s1 = std::min(n, c + rad);
  }

  std::vector<double> noise;
  auto append = [&](int a, int b) { noise.insert(noise.end(), ts.begin() + a, ts.begin() + b); };
  if (in.has_noise_ranges) {
    for (const auto &r : in.noise_ranges) {
      const int a = std::max(0, r.first / ds);
      const int b = std::min(n, (int)std::ceil((double)r.second / ds));
      if (b > a) append(a, b);
    }
    if (noise.empty()) noise = ts;
  } else {
    if (in.has_toa && (s0 > 0 || s1 < n)) {
      if (s0 > 0) append(0, s0);
      if (s1 < n) append(s1, n);
    }
    if (noise.size() < 10) noise = ts;
  }
  const auto st = clipped_stats(noise, in.sigma);
  p.mu = st.first;
  p.sd = st.second;
  p.cs.assign((size_t)n + 1, 0.0);
  for (int i = 0; i < n; ++i) p.cs[(size_t)i + 1] = p.cs[(size_t)i] + ts[(size_t)i];
  p.argmax = (int)(std::max_element(ts.begin(), ts.end()) - ts.begin());
  p.widths = boxcar_widths(wmax_ds);
  p.ds = ds;
  p.n = n;
  p.s0 = s0;
  p.s1 = s1;
  p.ts.swap(ts);
  return p;
}

// the boxcar of width w at every start, on the host
inline WidthBest search_width(const SnrInput &in, const SnrPrep &p, int w) {
  const double ninf = -std::numeric_limits<double>::infinity();
  WidthBest b;
  b.snr = ninf;
  for (int i = 0; i + w <= p.n; ++i) {
    const int c = i + w / 2;
    if (in.has_toa && (c < p.s0 || c >= p.s1)) continue;
    const double sum = p.cs[(size_t)i + w] - p.cs[(size_t)i];
    const double snr = p.sd > 0.0 ? (sum - p.mu * w) / (p.sd * std::sqrt((double)w)) : ninf;
    if (!b.any || snr > b.snr) {
      b.any = true;
      b.snr = snr;
      b.centre = c;
      b.sum = sum;
    }
  }
  return b;
}

// best[k]: the search at widths[k] (for the widths below n, in order)
inline SnrResult snr_finish(const SnrInput &in, SnrPrep &p, const std::vector<WidthBest> &best_w) {
  SnrResult res;
  const int n = p.n, ds = p.ds;
  const double ninf = -std::numeric_limits<double>::infinity();
  double best = ninf, best_sum = std::numeric_limits<double>::quiet_NaN();
  int bw = 1, best_c = p.argmax;
  for (size_t k = 0; k < p.widths.size() && k < best_w.size(); ++k) {
    const int w = p.widths[k];
    if (w >= n) break;
    const WidthBest &b = best_w[k];
    if (b.any && b.snr > best) {
      best = b.snr;
      bw = w;
      best_c = b.centre;
      best_sum = b.sum;
    }
  }
  res.downsample = ds;
  if (!std::isfinite(best_sum)) {
    // no usable boxcar: peak sample against the whole series' clipped stats
    auto [m2, s2] = clipped_stats(p.ts, in.sigma);
    res.noise_mean = m2;
    res.noise_std = s2;
    res.snr = s2 > 0.0 ? (p.ts[(size_t)p.argmax] - m2) / s2 : -1.0;
    res.width = std::max(1, ds);
    res.peak = std::min(in.ntime - 1, p.argmax * ds + ds / 2);
    res.fallback = true;
    return res;
  }
  // noise again, without the samples within two widths of the best window
  double mu = p.mu, sd = p.sd;
  const int rl = std::max(0, best_c - 2 * bw);
  const int rr = std::min(n, best_c + 2 * bw + 1);
  std::vector<double> rest;
  rest.reserve((size_t)n);
  for (int i = 0; i < n; ++i)
    if (i < rl || i >= rr) rest.push_back(p.ts[(size_t)i]);
  if (rest.size() >= std::max<size_t>(10, (size_t)(0.1 * (double)n))) {
    auto st = clipped_stats(rest, in.sigma);
    mu = st.first;
    sd = st.second;
  }
  res.snr = sd > 0.0 ? (best_sum - mu * bw) / (sd * std::sqrt((double)bw)) : -1.0;
  res.noise_mean = mu;
  res.noise_std = sd;
  res.width = std::max(1, bw * ds);
  res.peak = std::min(in.ntime - 1, best_c * ds + ds / 2);
  res.converged = true;
  res.search_start = p.s0;
  res.search_end = p.s1;
  return res;
}

// frequency-summed series, in row order, in double
inline std::vector<double> freq_sums(const SnrInput &in) {
  std::vector<double> ts((size_t)in.ntime);
  for (int t = 0; t < in.ntime; ++t) {
    const float *row = in.spec + (size_t)t * (size_t)in.nfreq;
    double s = 0.0;
    for (int f = 0; f < in.nfreq; ++f) s += row[f];
    ts[(size_t)t] = s;
  }
  return ts;
}

inline SnrResult boxcar_snr(const SnrInput &in) {
  if (in.ntime <= 0 || in.nfreq <= 0) {
    SnrResult res;
    res.empty = true;
    return res;
  }
  SnrPrep p = snr_prepare(in, freq_sums(in));
  std::vector<WidthBest> bw;
  for (int w : p.widths) {
    if (w >= p.n) break;
    bw.push_back(search_width(in, p, w));
  }
  return snr_finish(in, p, bw);
}

}  // namespace afsnr
