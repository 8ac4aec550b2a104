// af_cpu.h -- the host-core target of _dedispered_fil (target = 0).
//
// Behaviour of cpucal::dedispered_fil_omp (include/cpucal.hpp:75-201):
// direct dedispersion of channels [cs, ce) (the top selected channel is left
// out, as there), DM trials dm_low + i*dm_step, in slices of
// round(t_sample / tsamp) native samples; time_downsample > 1 reads every
// ds-th sample of the slice (decimation, not binning); u64 sums; a sample
// past the end of the data adds zero.  Each slice's [ndm][ceil(len/ds)] plane
// is then imaged (af_image.h).  No RFI stage, as in the reference.
//
// The delays are the FP64 table of the GPU paths (round half away from zero
// of 4148.808 dm (f^-2 - fref^-2) / tsamp, fref the frequency of channel ce)
// rather than the reference's float32 expression, so at ds = 1 the plane
// equals the GPU direct method's on the same channels.
//
// Layout: the selected channels are first copied channel-major, so each
// (DM, channel) pair adds one contiguous (ds = 1) or strided run of a channel
// row into the DM row -- the inner loop has no gather and no bounds test
// (the run is cut at the data end beforehand).  DM rows are spread over the
// host threads (OpenMP).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace afcpu {

// delays[i][c - cs] for c in [cs, ce), native samples
inline std::vector<int32_t> delay_table(const std::vector<double> &freq, size_t cs, size_t ce,
                                        double dm_low, double dm_step, size_t ndm,
                                        double tsamp) {
  const size_t nc = ce - cs;
  std::vector<int32_t> d(ndm * nc);
  const double fref = freq[ce];
  for (size_t i = 0; i < ndm; ++i) {
    const double dm = dm_low + (double)i * dm_step;
    for (size_t c = cs; c < ce; ++c) {
      const double inv = 1.0 / (freq[c] * freq[c]) - 1.0 / (fref * fref);
      d[i * nc + (c - cs)] = (int32_t)std::round(4148.808 * dm * inv / tsamp);
    }
  }
  return d;
}

// x: [ndata][nchan] time-major; returns [ce - cs][ndata] channel-major
template <typename T>
std::vector<T> channel_major(const T *x, size_t ndata, size_t nchan, size_t cs, size_t ce) {
  const size_t nc = ce - cs;
  std::vector<T> xt(nc * ndata);
  constexpr size_t kB = 256;   // time block: rows stay in cache while read across
#pragma omp parallel for schedule(static)
  for (long long tb = 0; tb < (long long)((ndata + kB - 1) / kB); ++tb) {
    const size_t t0 = (size_t)tb * kB, t1 = std::min(ndata, t0 + kB);
    for (size_t c = 0; c < nc; ++c) {
      T *o = xt.data() + c * ndata;
      for (size_t t = t0; t < t1; ++t) o[t] = x[t * nchan + cs + c];
    }
  }
  return xt;
}

// One slice [start, start + len): out[i][ti] = sum_c xt[c][start + ti*ds + d[i][c]]
template <typename T>
void dedisperse_slice(const std::vector<T> &xt, size_t ndata, size_t nc,
                      const std::vector<int32_t> &delays, size_t ndm, size_t start, size_t len,
                      int ds, uint64_t *out) {
  const size_t down = (len + (size_t)ds - 1) / (size_t)ds;
#pragma omp parallel for schedule(dynamic)
  for (long long i = 0; i < (long long)ndm; ++i) {
    uint64_t *row = out + (size_t)i * down;
    std::fill(row, row + down, (uint64_t)0);
    const int32_t *d = delays.data() + (size_t)i * nc;
    for (size_t c = 0; c < nc; ++c) {
      const long long base = (long long)start + d[c];
      if (base < 0 || base >= (long long)ndata) continue;
      // ti < n: start + ti*ds + d < ndata
      const size_t n = std::min(down, ((size_t)((long long)ndata - base) + ds - 1) / (size_t)ds);
      const T *src = xt.data() + c * ndata + (size_t)base;
      if (ds == 1) {
        for (size_t ti = 0; ti < n; ++ti) row[ti] += src[ti];
      } else {
        for (size_t ti = 0; ti < n; ++ti) row[ti] += src[ti * (size_t)ds];
      }
    }
  }
}

}  // namespace afcpu
