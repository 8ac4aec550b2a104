// ref_harness.cpp -- C entry points around the reference's own CPU RFI code,
// compiled from the reference sources in place (include/iqrm.hpp,
// src/rfi/rfimarker_cpu.cpp) into oracle/_ref/libafref.so by oracle/Makefile.
// Test infrastructure only: it exists to pin the oracle's IQRM and mask
// application against the reference's own implementation.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "iqrm.hpp"
#include "rfimarker_cpu.h"

namespace {
template <typename T>
int run_iqrm(const T *data, size_t nsample, int nchan, int c0, int c1,
             double tsamp, int mode, float radius_frac, float nsigma,
             double geofactor, double win_sec, double hop_sec, int include_tail,
             uint8_t *masks, uint32_t *t0s, uint32_t *t1s, int maxw) {
  rficonfig cfg;
  cfg.use_iqrm = true;
  cfg.iqrm_cfg = iqrmconfig(mode, radius_frac, nsigma, geofactor, win_sec,
                            hop_sec, include_tail != 0);
  auto out = rfi_iqrm<T>(const_cast<T *>(data), (size_t)c0, (size_t)c1,
                         nsample, (size_t)nchan, tsamp, cfg);
  int n = (int)out.size();
  for (int w = 0; w < n && w < maxw; ++w) {
    t0s[w] = out[w].t0;
    t1s[w] = out[w].t1;
    std::memcpy(masks + (size_t)w * nchan, out[w].mask.data(), (size_t)nchan);
  }
  return n;
}

template <typename T>
void run_mask(T *data, int nchan, size_t nsample, const uint8_t *masks,
              const uint32_t *t0s, const uint32_t *t1s, int nwin) {
  std::vector<iqrm_omp::WindowMask> wm;
  for (int w = 0; w < nwin; ++w) {
    iqrm_omp::WindowMask m;
    m.t0 = t0s[w];
    m.t1 = t1s[w];
    m.mask.assign(masks + (size_t)w * nchan, masks + (size_t)(w + 1) * nchan);
    wm.push_back(std::move(m));
  }
  RfiMarkerCPU<T> rm("");
  rm.mask(data, (unsigned)nchan, (unsigned)nsample, wm);
}
} // namespace

extern "C" {

int afref_iqrm_u8(const uint8_t *data, size_t nsample, int nchan, int c0, int c1,
                  double tsamp, int mode, float radius_frac, float nsigma,
                  double geofactor, double win_sec, double hop_sec,
                  int include_tail, uint8_t *masks, uint32_t *t0s, uint32_t *t1s,
                  int maxw) {
  return run_iqrm<uint8_t>(data, nsample, nchan, c0, c1, tsamp, mode, radius_frac,
                           nsigma, geofactor, win_sec, hop_sec, include_tail,
                           masks, t0s, t1s, maxw);
}

int afref_iqrm_u16(const uint16_t *data, size_t nsample, int nchan, int c0,
                   int c1, double tsamp, int mode, float radius_frac,
                   float nsigma, double geofactor, double win_sec, double hop_sec,
                   int include_tail, uint8_t *masks, uint32_t *t0s,
                   uint32_t *t1s, int maxw) {
  return run_iqrm<uint16_t>(data, nsample, nchan, c0, c1, tsamp, mode,
                            radius_frac, nsigma, geofactor, win_sec, hop_sec,
                            include_tail, masks, t0s, t1s, maxw);
}

void afref_iqrm_on_stat(const float *x, int C, float radius_frac, float nsigma,
                        double geofactor, uint8_t *mask) {
  std::vector<float> xv(x, x + C);
  std::vector<uint8_t> m;
  iqrm_omp::iqrm_on_stat(xv, radius_frac, nsigma, geofactor, m);
  std::memcpy(mask, m.data(), (size_t)C);
}

void afref_mask_u8(uint8_t *data, int nchan, size_t nsample, const uint8_t *masks,
                   const uint32_t *t0s, const uint32_t *t1s, int nwin) {
  run_mask<uint8_t>(data, nchan, nsample, masks, t0s, t1s, nwin);
}

// Static bad-channel mask: load the file with the reference's parser and
// zero the listed channels with the reference's mark_rfi.
int afref_mark_rfi_u8(uint8_t *data, int nchan, size_t nsample,
                      const char *mask_file, int *bad_out, int max_bad) {
  RfiMarkerCPU<uint8_t> rm(mask_file);
  rm.mark_rfi(data, (unsigned)nchan, (unsigned)nsample);
  const auto &b = rm.get_bad_channels();
  int n = (int)b.size();
  for (int i = 0; i < n && i < max_bad; ++i) bad_out[i] = b[i];
  return n;
}
}
