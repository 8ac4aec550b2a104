// af_snr.hip -- boxcar S/N of a dynamic spectrum held on the device.
// Compiled with -ffp-contract=off.
//
// The same result as af_snr.h's boxcar_snr (the behaviour of the
// reference's _calculate_frb_snr, src/bindpy/frb_snr.cpp:144-398) for a
// spectrum already in HBM, e.g. the output of af_execute_spectrum: the two
// parts that touch every sample run on the device, the order statistics on
// the host.
//   k_snr_rows   ts[t] = sum_f (double)(float)spec[t][f], f in order (one
//                thread per row: the same additions in the same order as the
//                host, so the series is identical)
//   host         decimation, search range, sigma-clipped noise (medians),
//                prefix sums cs, width list (af_snr.h)
//   k_snr_widths one workgroup per width w: the S/N (cs[i+w] - cs[i] -
//                mu w) / (sd sqrt(w)) of every start i whose centre lies in
//                the search range, reduced to the largest value, ties to the
//                lowest start (the host loop's first maximum)
//   host         best width, noise re-estimated without the pulse
// sqrt(w) comes from the host, so every S/N is the host's double expression.
// A series with a non-finite value (NaN or inf data) is searched on the host:
// the reduction above assumes ordered values.
#include "af_internal.h"
#include "af_snr.h"

#include <cmath>
#include <vector>

namespace af {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_snr_rows(const T *__restrict__ spec, int ntime,
                                                  int nfreq, double *__restrict__ ts) {
  const int t = (int)(blockIdx.x * 256 + threadIdx.x);
  if (t >= ntime) return;
  const T *row = spec + (size_t)t * (size_t)nfreq;
  double s = 0.0;
  for (int f = 0; f < nfreq; ++f) s += (double)(float)row[f];
  ts[t] = s;
}

struct WidthOut {
  double snr, sum;
  int centre, any;
};

constexpr int kSnrThreads = 256;

__global__ __launch_bounds__(kSnrThreads) void k_snr_widths(
    const double *__restrict__ cs, int n, const int *__restrict__ widths,
    const double *__restrict__ sqw, double mu, double sd, int has_toa, int s0, int s1,
    WidthOut *__restrict__ out) {
  __shared__ double s_snr[kSnrThreads], s_sum[kSnrThreads];
  __shared__ int s_i[kSnrThreads];
  const int w = widths[blockIdx.x];
  const double rw = sqw[blockIdx.x];
  const double ninf = -__builtin_inf();
  double best = ninf, bsum = 0.0;
  int bi = INT32_MAX;   // start of the best window; INT32_MAX: none seen
  for (int i = (int)threadIdx.x; i + w <= n; i += kSnrThreads) {
    const int c = i + w / 2;
    if (has_toa && (c < s0 || c >= s1)) continue;
    const double sum = cs[i + w] - cs[i];
    const double snr = sd > 0.0 ? (sum - mu * w) / (sd * rw) : ninf;
    // starts rise within a thread: only a larger value replaces
    if (bi == INT32_MAX || snr > best) {
      best = snr;
      bsum = sum;
      bi = i;
    }
  }
  s_snr[threadIdx.x] = best;
  s_sum[threadIdx.x] = bsum;
  s_i[threadIdx.x] = bi;
  __syncthreads();
  for (int h = kSnrThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      const int j = threadIdx.x + h;
      const int ia = s_i[threadIdx.x], ib = s_i[j];
      // b wins when it exists and (a does not, or b is larger, or equal
      // with a lower start)
      const bool take = ib != INT32_MAX &&
                        (ia == INT32_MAX || s_snr[j] > s_snr[threadIdx.x] ||
                         (s_snr[j] == s_snr[threadIdx.x] && ib < ia));
      if (take) {
        s_snr[threadIdx.x] = s_snr[j];
        s_sum[threadIdx.x] = s_sum[j];
        s_i[threadIdx.x] = ib;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    WidthOut o;
    o.any = s_i[0] != INT32_MAX;
    o.snr = s_snr[0];
    o.sum = s_sum[0];
    o.centre = o.any ? s_i[0] + w / 2 : 0;
    out[blockIdx.x] = o;
  }
}

}  // namespace

af_status frb_snr(const void *d_spec, int dtype, int ntime, int nfreq, const af_snr_params *prm,
                  af_snr_result *res, hipStream_t st) {
  if (!res || !prm) return fail(AF_ERR_INVALID, "null argument");
  if (ntime < 0 || nfreq < 0) return fail(AF_ERR_INVALID, "negative spectrum size");
  if (dtype < AF_U8 || dtype > AF_F32) return fail(AF_ERR_INVALID, "unknown element type");
  if (prm->n_noise_ranges < 0 || (prm->n_noise_ranges > 0 && !prm->noise_ranges))
    return fail(AF_ERR_INVALID, "noise ranges");
  afsnr::SnrInput in;
  in.ntime = ntime;
  in.nfreq = nfreq;
  in.sigma = prm->threshold_sigma;
  in.has_noise_ranges = prm->has_noise_ranges != 0;
  for (int k = 0; k < prm->n_noise_ranges; ++k)
    in.noise_ranges.emplace_back(prm->noise_ranges[2 * k], prm->noise_ranges[2 * k + 1]);
  in.has_toa = prm->has_toa != 0;
  in.toa = prm->toa;
  in.has_window = prm->has_window != 0;
  in.window = prm->window;
  in.has_tsamp = prm->has_tsamp != 0;
  in.tsamp = prm->tsamp;
  in.has_target = prm->has_target != 0;
  in.target_us = prm->target_time_us;
  *res = af_snr_result{};
  res->snr = -1.0;
  res->noise_std = 1.0;
  res->downsample = 1;
  if (ntime == 0 || nfreq == 0) return AF_OK;   // method 0: no data
  if (!d_spec) return fail(AF_ERR_INVALID, "null spectrum");

  // frequency sums on the device
  double *d_ts = nullptr;
  AF_HIP(hipMalloc(&d_ts, (size_t)ntime * sizeof(double)));
  const unsigned g = (unsigned)((ntime + 255) / 256);
  switch (dtype) {
  case AF_U8: hipLaunchKernelGGL(k_snr_rows<uint8_t>, dim3(g), dim3(256), 0, st, (const uint8_t *)d_spec, ntime, nfreq, d_ts); break;
  case AF_U16: hipLaunchKernelGGL(k_snr_rows<uint16_t>, dim3(g), dim3(256), 0, st, (const uint16_t *)d_spec, ntime, nfreq, d_ts); break;
  case AF_U32: hipLaunchKernelGGL(k_snr_rows<uint32_t>, dim3(g), dim3(256), 0, st, (const uint32_t *)d_spec, ntime, nfreq, d_ts); break;
  default: hipLaunchKernelGGL(k_snr_rows<float>, dim3(g), dim3(256), 0, st, (const float *)d_spec, ntime, nfreq, d_ts); break;
  }
  std::vector<double> ts((size_t)ntime);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(ts.data(), d_ts, ts.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d_ts);
  if (e != hipSuccess) return fail(AF_ERR_RUNTIME, std::string("S/N sums: ") + hipGetErrorString(e));

  afsnr::SnrPrep p = afsnr::snr_prepare(in, std::move(ts));
  std::vector<int> ws;
  for (int w : p.widths)
    if (w < p.n) ws.push_back(w);
  bool finite = true;
  for (double v : p.ts) finite = finite && std::isfinite(v);
  std::vector<afsnr::WidthBest> best(ws.size());
  if (!ws.empty() && finite) {
    std::vector<double> sqw(ws.size());
    for (size_t k = 0; k < ws.size(); ++k) sqw[k] = std::sqrt((double)ws[k]);
    const size_t b_cs = p.cs.size() * sizeof(double), b_sq = sqw.size() * sizeof(double);
    const size_t b_ws = ws.size() * sizeof(int), b_out = ws.size() * sizeof(WidthOut);
    uint8_t *buf = nullptr;
    AF_HIP(hipMalloc(&buf, b_cs + b_sq + b_out + b_ws));
    double *d_cs = (double *)buf;
    double *d_sq = (double *)(buf + b_cs);
    WidthOut *d_out = (WidthOut *)(buf + b_cs + b_sq);
    int *d_ws = (int *)(buf + b_cs + b_sq + b_out);
    std::vector<WidthOut> h(ws.size());
    e = hipMemcpyAsync(d_cs, p.cs.data(), b_cs, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sq, sqw.data(), b_sq, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ws, ws.data(), b_ws, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_snr_widths, dim3((unsigned)ws.size()), dim3(kSnrThreads), 0, st, d_cs,
                         p.n, d_ws, d_sq, p.mu, p.sd, in.has_toa ? 1 : 0, p.s0, p.s1, d_out);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_out, b_out, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(buf);
    if (e != hipSuccess)
      return fail(AF_ERR_RUNTIME, std::string("S/N search: ") + hipGetErrorString(e));
    for (size_t k = 0; k < ws.size(); ++k) {
      best[k].any = h[k].any != 0;
      best[k].snr = h[k].snr;
      best[k].sum = h[k].sum;
      best[k].centre = h[k].centre;
    }
  } else {
    for (size_t k = 0; k < ws.size(); ++k) best[k] = afsnr::search_width(in, p, ws[k]);
  }
  const afsnr::SnrResult r = afsnr::snr_finish(in, p, best);
  res->snr = r.snr;
  res->width = r.width;
  res->peak = r.peak;
  res->noise_mean = r.noise_mean;
  res->noise_std = r.noise_std;
  res->method = r.fallback ? 2 : 1;
  res->converged = r.converged ? 1 : 0;
  res->search_start = r.search_start;
  res->search_end = r.search_end;
  res->downsample = r.downsample;
  return AF_OK;
}

}  // namespace af

extern "C" af_status af_frb_snr(const void *d_spec, int dtype, int ntime, int nfreq,
                                const af_snr_params *params, af_snr_result *result,
                                void *stream) {
  return af::frb_snr(d_spec, dtype, ntime, nfreq, params, result, (hipStream_t)stream);
}
