/*
 * af_oracle.c -- CPU oracle (test infrastructure only; see af_oracle.h).
 *
 * Compiled with -ffp-contract=off so every floating-point expression rounds
 * exactly as written, which is what the reference's source text specifies.
 */
#include "af_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static size_t dsize(int dtype) {
  switch (dtype) {
  case AFO_U8: return 1;
  case AFO_U16: return 2;
  case AFO_U32: return 4;
  case AFO_F32: return 4;
  default: return 0;
  }
}

static double vmax_of(int dtype) {
  switch (dtype) {
  case AFO_U8: return 255.0;
  case AFO_U16: return 65535.0;
  case AFO_U32: return 4294967295.0;
  default: return 0.0;
  }
}

/* Load element i of a typed buffer as an exact integer (unsigned types). */
static uint64_t ld_u(const void *p, int dtype, size_t i) {
  switch (dtype) {
  case AFO_U8: return ((const uint8_t *)p)[i];
  case AFO_U16: return ((const uint16_t *)p)[i];
  case AFO_U32: return ((const uint32_t *)p)[i];
  default: return 0;
  }
}

static void st_u(void *p, int dtype, size_t i, uint64_t v) {
  switch (dtype) {
  case AFO_U8: ((uint8_t *)p)[i] = (uint8_t)v; break;
  case AFO_U16: ((uint16_t *)p)[i] = (uint16_t)v; break;
  case AFO_U32: ((uint32_t *)p)[i] = (uint32_t)v; break;
  default: break;
  }
}

/* ---------------------------------------------------------------- setup -- */

/* src/core/dedispered.cu:1003-1005: h_freq[i] = header.fch1 + i*header.foff,
 * Header fields are float (include/data.h:24-31), so the product and the sum
 * are float32 operations. */
void afo_freq_table_f32(float fch1, float foff, int nchan, double *out) {
  for (int i = 0; i < nchan; ++i) {
    float prod = (float)i * foff;
    float f = fch1 + prod;
    out[i] = (double)f;
  }
}

/* src/core/dedispered.cu:1007-1027. */
int afo_chan_range(const double *freq, int nchan, float fs, float fe,
                   long *chan_start, long *chan_end) {
  float fmin = (float)freq[0];
  float fmax = (float)freq[nchan - 1];
  if (fs < fmin || fe > fmax) return -1;
  float span = fmax - fmin;
  float nm1 = (float)(nchan - 1);
  float a = (fs - fmin) / span;
  float b = (fe - fmin) / span;
  long cs = (long)(size_t)(a * nm1);
  long ce = (long)(size_t)(b * nm1);
  if (ce > nchan - 1) ce = nchan - 1;
  if (cs < 0 || cs >= nchan || ce >= nchan) return -1;
  *chan_start = cs;
  *chan_end = ce;
  return 0;
}

/* src/core/dedispered.cu:1029 (float arithmetic, truncation, +1). */
long afo_dm_steps(float dm_low, float dm_high, float dm_step) {
  float q = (dm_high - dm_low) / dm_step;
  return (long)(size_t)q + 1;
}

/* ------------------------------------------------------------- delays -- */

/* Delay in samples at native rate: round half away from zero of
 * K*dm*(f^-2 - fref^-2)/tsamp, evaluated left to right in FP64.
 * Reference: src/core/dedispered.cu:329-334 (float32 there). */
static long delay_native(double f, double fref, double dm, double tsamp) {
  double inv = 1.0 / (f * f) - 1.0 / (fref * fref);
  double delay = AFO_KDM * dm * inv;
  return (long)round(delay / tsamp);
}

void afo_delay_table(const double *freq, int c0, int c1, double f_ref,
                     double tsamp, int ds, const double *dms, int ndm,
                     int32_t *out) {
  int nc = c1 - c0;
  for (int i = 0; i < ndm; ++i)
    for (int c = c0; c < c1; ++c) {
      long d = delay_native(freq[c], f_ref, dms[i], tsamp);
      /* src/core/dedispered.cu:379-382: integer division truncates */
      out[(size_t)i * nc + (c - c0)] = (int32_t)(d / ds);
    }
}

void afo_residual_table(const double *freq, int c0, int c1, int sbw,
                        double f_ref, double tsamp, int ds, double dm_low,
                        double dm_step, int ndm, int ndm0, int32_t *out) {
  (void)dm_low; /* residual depends on dm_m - dm_j only */
  int nsb = (c1 - c0 + sbw - 1) / sbw;
  for (int m = 0; m < ndm; ++m) {
    int j = m / ndm0;
    double ddm = (double)(m - j * ndm0) * dm_step;
    for (int sb = 0; sb < nsb; ++sb) {
      int ch0 = c0 + sb * sbw;
      int ch1 = ch0 + sbw < c1 ? ch0 + sbw : c1;
      double f = freq[(ch0 + ch1 - 1) / 2];
      double inv = 1.0 / (f * f) - 1.0 / (f_ref * f_ref);
      double delay = AFO_KDM * ddm * inv;
      double bins = delay / (tsamp * (double)ds);
      out[(size_t)m * nsb + sb] = (int32_t)floor(bins + 0.5);
    }
  }
}

/* -------------------------------------------------------------- binning -- */

/* src/core/dedispered.cu:337-376 with USE_SQRT_COMPRESSION. */
void afo_bin(const void *in, int dtype, size_t nsamp, int nchan, int ds, void *out) {
  size_t nout = (nsamp + ds - 1) / ds;
  /* rows are independent: split over threads (each element's arithmetic is
   * unchanged) */
#pragma omp parallel for schedule(static)
  for (size_t s = 0; s < nout; ++s) {
    size_t t0 = s * (size_t)ds;
    size_t t1 = t0 + (size_t)ds < nsamp ? t0 + (size_t)ds : nsamp;
    size_t cnt = t1 - t0;
    for (int c = 0; c < nchan; ++c) {
      if (dtype == AFO_F32) {
        double acc = 0.0;
        for (size_t t = t0; t < t1; ++t) acc += (double)((const float *)in)[t * nchan + c];
        ((float *)out)[s * nchan + c] = (float)acc;
        continue;
      }
      uint64_t sum = 0;
      for (size_t t = t0; t < t1; ++t) sum += ld_u(in, dtype, t * nchan + c);
      double vmax = vmax_of(dtype);
      uint64_t vmax_i = (uint64_t)vmax;
      double sq = sqrt((double)sum);
      double smax = sqrt((double)(vmax_i * (uint64_t)cnt));
      double v = (sq / (smax > 0.0 ? smax : 1.0)) * vmax;
      double r = v < vmax ? v : vmax;
      st_u(out, dtype, s * nchan + c, (uint64_t)r);
    }
  }
}

/* ----------------------------------------------------------------- IQRM -- */

/* src/core/iqrm.cu:368-395 */
int afo_iqrm_windows(size_t nsample, unsigned W, unsigned H, int include_tail,
                     uint32_t *t0s, uint32_t *t1s, int maxw) {
  int n = 0;
#define PUSHW(a, b)                                                          \
  do {                                                                       \
    if (n < maxw) { t0s[n] = (uint32_t)(a); t1s[n] = (uint32_t)(b); }        \
    ++n;                                                                     \
  } while (0)
  if (nsample == 0) return 0;
  if (W == 0 || W >= nsample) {
    PUSHW(0, nsample);
  } else {
    unsigned t0 = 0;
    while ((size_t)t0 + W <= nsample) {
      PUSHW(t0, t0 + W);
      t0 += (H ? H : W);
    }
    if (include_tail && t0 < nsample) {
      size_t t1 = nsample;
      if (t1 > (size_t)t0 + 4) PUSHW(t0, t1);
    }
  }
#undef PUSHW
  return n;
}

/* src/core/iqrm.cu:189-200 */
int afo_iqrm_lags(int radius, double geofactor, int *lags, int maxl) {
  int n = 0;
  if (radius <= 0) return 0;
  int lag = 1;
  while (lag <= radius) {
    if (n < maxl) lags[n] = +lag;
    ++n;
    if (n < maxl) lags[n] = -lag;
    ++n;
    int nxt = (int)(geofactor * lag);
    lag = nxt > lag + 1 ? nxt : lag + 1;
  }
  return n;
}

/* src/core/iqrm.cu:255 / include/iqrm.hpp:106 */
int afo_iqrm_radius(float radius_frac, int C) {
  float p = radius_frac * (float)C;
  int r = (int)floorf(p);
  return r > 1 ? r : 1;
}

void afo_iqrm_stat(const void *data, int dtype, int nchan, uint32_t t0,
                   uint32_t t1, int c0, int c1, int mode, int flavour,
                   float *x) {
  /* Row-major traversal with per-channel accumulators: each channel still
   * sums its samples in ascending time order.  Large windows split their
   * channels into blocks over threads (the same order per channel). */
  const uint32_t W = t1 - t0;
  const int C = c1 - c0;
  const int nblk = (C + 255) / 256;
  if (nblk > 1 && (uint64_t)W * (uint64_t)C > (1u << 22)) {
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < nblk; ++b) {
      const int a0 = c0 + 256 * b, a1 = a0 + 256 < c1 ? a0 + 256 : c1;
      afo_iqrm_stat(data, dtype, nchan, t0, t1, a0, a1, mode, flavour, x + (a0 - c0));
    }
    return;
  }
  double *ds = (double *)calloc((size_t)C, sizeof(double));
  double *dq = (double *)calloc((size_t)C, sizeof(double));
  uint64_t *us = (uint64_t *)calloc((size_t)C, sizeof(uint64_t));
  u128 *uq = (u128 *)calloc((size_t)C, sizeof(u128));
  if (dtype == AFO_F32) {
    const float *f = (const float *)data;
    for (uint32_t t = t0; t < t1; ++t)
      for (int c = c0; c < c1; ++c) {
        double v = (double)f[(size_t)t * nchan + c];
        ds[c - c0] += v;
        dq[c - c0] += v * v;
      }
    for (int k = 0; k < C; ++k) {
      double mean = ds[k] / (double)W;
      if (mode == 0) { x[k] = (float)mean; continue; }
      double var = dq[k] / (double)W - mean * mean;
      if (var < 0.0) var = 0.0;
      x[k] = (float)sqrt(var);
    }
    if (mode == 1 && flavour == AFO_IQRM_CPU) {
      memset(dq, 0, sizeof(double) * C);
      for (uint32_t t = t0; t < t1; ++t)
        for (int c = c0; c < c1; ++c) {
          double v = (double)f[(size_t)t * nchan + c] - ds[c - c0] / (double)W;
          dq[c - c0] += v * v;
        }
      for (int k = 0; k < C; ++k) {
        double var = dq[k] / (double)W;
        x[k] = (float)sqrt(var > 0.0 ? var : 0.0);
      }
    }
  } else if (flavour == AFO_IQRM_CPU) {
    /* include/iqrm.hpp:189-217: double accumulation, two-pass std */
    for (uint32_t t = t0; t < t1; ++t)
      for (int c = c0; c < c1; ++c) ds[c - c0] += (double)ld_u(data, dtype, (size_t)t * nchan + c);
    for (int k = 0; k < C; ++k) ds[k] = ds[k] / (double)W;   /* mean */
    if (mode == 0) {
      for (int k = 0; k < C; ++k) x[k] = (float)ds[k];
    } else {
      for (uint32_t t = t0; t < t1; ++t)
        for (int c = c0; c < c1; ++c) {
          double v = (double)ld_u(data, dtype, (size_t)t * nchan + c) - ds[c - c0];
          dq[c - c0] += v * v;
        }
      for (int k = 0; k < C; ++k) {
        double s2 = dq[k] / (double)W;
        x[k] = (float)sqrt(s2 > 0.0 ? s2 : 0.0);
      }
    }
  } else {
    /* src/core/iqrm.cu:21-91: sum and sum of squares (exact integers here,
     * which is what the reference's double sums equal for 8/16-bit data),
     * then mean = S/W and var = S2/W - mean^2 in double. */
    for (uint32_t t = t0; t < t1; ++t)
      for (int c = c0; c < c1; ++c) {
        uint64_t v = ld_u(data, dtype, (size_t)t * nchan + c);
        us[c - c0] += v;
        uq[c - c0] += (u128)v * v;
      }
    for (int k = 0; k < C; ++k) {
      double mean = (double)us[k] / (double)W;
      if (mode == 0) { x[k] = (float)mean; continue; }
      double var = (double)uq[k] / (double)W - mean * mean;
      if (var < 0.0) var = 0.0;
      x[k] = (float)sqrt(var);
    }
  }
  free(ds); free(dq); free(us); free(uq);
}

static int cmp_float(const void *a, const void *b) {
  float x = *(const float *)a, y = *(const float *)b;
  return (x > y) - (x < y);
}

/* include/iqrm.hpp:56-63 (float) */
static float pct_cpu(const float *v, size_t m, float q) {
  if (m == 0) return NAN;
  if (m == 1) return v[0];
  float pos = q * (float)(m - 1);
  size_t lo = (size_t)floorf(pos), hi = (size_t)ceilf(pos);
  float w = pos - (float)lo;
  float diff = v[hi] - v[lo];
  float prod = diff * w;
  return v[lo] + prod;
}

/* src/core/iqrm.cu:295-303 (double position and weight) */
static float pct_gpu(const float *v, size_t m, float q) {
  if (m == 1) return v[0];
  double pos = (double)q * (double)(m - 1);
  size_t lo = (size_t)floor(pos), hi = (size_t)ceil(pos);
  double w = pos - (double)lo;
  float diff = v[hi] - v[lo];
  double prod = (double)diff * w;
  double r = (double)v[lo] + prod;
  return (float)r;
}

/* Quartiles and median of m lagged differences: the CPU flavour drops
 * non-finite values and interpolates in float (include/iqrm.hpp:56-63,
 * 118-123), the GPU flavour sorts all m and interpolates with a double
 * position and weight (src/core/iqrm.cu:288-306).  Returns the number of
 * values used (0: no quartiles). */
size_t afo_iqrm_quartiles(const float *d, size_t m_in, int flavour, float *srt, float q[3]) {
  size_t m = 0;
  if (flavour == AFO_IQRM_CPU) {
    for (size_t k = 0; k < m_in; ++k)
      if (isfinite(d[k])) srt[m++] = d[k];
    if (m == 0) return 0;
    qsort(srt, m, sizeof(float), cmp_float);
    q[0] = pct_cpu(srt, m, 0.25f);
    q[1] = pct_cpu(srt, m, 0.50f);
    q[2] = pct_cpu(srt, m, 0.75f);
  } else {
    memcpy(srt, d, sizeof(float) * m_in);
    m = m_in;
    if (m == 0) return 0;
    qsort(srt, m, sizeof(float), cmp_float);
    q[0] = pct_gpu(srt, m, 0.25f);
    q[1] = pct_gpu(srt, m, 0.50f);
    q[2] = pct_gpu(srt, m, 0.75f);
  }
  return m;
}

/* Partner channel of a flagged channel k for a lag
 * (src/core/iqrm.cu:138-145, include/iqrm.hpp:134-141). */
static int partner(int k, int lag, int C) {
  if (lag < 0) return (k < C + lag) ? k - lag : C - 1;
  return (k < lag) ? 0 : k - lag;
}

void afo_iqrm_on_stat(const float *x, int C, float radius_frac, float nsigma,
                      double geofactor, int flavour, uint8_t *mask) {
  memset(mask, 0, (size_t)C);
  if (C <= 0) return;
  int radius = afo_iqrm_radius(radius_frac, C);
  int lags[256];
  int L = afo_iqrm_lags(radius, geofactor, lags, 256);
  if (L > 256) L = 256;
  float *d = (float *)malloc(sizeof(float) * C);
  float *srt = (float *)malloc(sizeof(float) * C);
  uint8_t *flags = (uint8_t *)calloc((size_t)L * C, 1);
  int *recv = (int *)calloc((size_t)C, sizeof(int));
  int *cast = (int *)calloc((size_t)C, sizeof(int));
  for (int li = 0; li < L; ++li) {
    int lag = lags[li];
    /* A lag of magnitude >= C makes the reference index outside the
     * statistic vector; such lags are skipped here. */
    if (lag >= C || -lag >= C) continue;
    /* src/core/iqrm.cu:94-117, include/iqrm.hpp:77-99 */
    if (lag < 0) {
      int cut = C + lag;
      float edge = x[cut] - x[C - 1];
      for (int k = 0; k < C; ++k) d[k] = (k < cut) ? x[k] - x[k - lag] : edge;
    } else {
      float edge = x[lag] - x[0];
      for (int k = 0; k < C; ++k) d[k] = (k < lag) ? edge : x[k] - x[k - lag];
    }
    float q[3];
    if (afo_iqrm_quartiles(d, (size_t)C, flavour, srt, q) == 0) continue;
    const float q1 = q[0], med = q[1], q3 = q[2];
    float s = fabsf((q3 - q1) / 1.349f);
    if (!(s > 0.f) || !isfinite(s)) continue;
    float thr = nsigma * s;
    for (int k = 0; k < C; ++k) {
      float val = d[k] - med;
      if (val > thr) {
        flags[(size_t)li * C + k] = 1;
        recv[k] += 1;
        cast[partner(k, lag, C)] += 1;
      }
    }
  }
  /* src/core/iqrm.cu:154-185 / include/iqrm.hpp:143-158 */
  for (int k = 0; k < C; ++k) {
    if (recv[k] <= 0) continue;
    for (int li = 0; li < L; ++li) {
      if (!flags[(size_t)li * C + k]) continue;
      if (cast[partner(k, lags[li], C)] < recv[k]) { mask[k] = 1; break; }
    }
  }
  free(d); free(srt); free(flags); free(recv); free(cast);
}

int afo_iqrm(const void *data, int dtype, size_t nsample, int nchan, int c0,
             int c1, int mode, float radius_frac, float nsigma,
             double geofactor, unsigned W, unsigned H, int include_tail,
             int flavour, uint8_t *masks, uint32_t *t0s, uint32_t *t1s,
             int maxw) {
  if (nsample == 0 || c0 >= c1 || c1 > nchan) return 0;
  int nw = afo_iqrm_windows(nsample, W, H, include_tail, t0s, t1s, maxw);
  int C = c1 - c0;
  float *x = (float *)malloc(sizeof(float) * C);
  uint8_t *sub = (uint8_t *)malloc((size_t)C);
  for (int w = 0; w < nw && w < maxw; ++w) {
    afo_iqrm_stat(data, dtype, nchan, t0s[w], t1s[w], c0, c1, mode, flavour, x);
    afo_iqrm_on_stat(x, C, radius_frac, nsigma, geofactor, flavour, sub);
    uint8_t *row = masks + (size_t)w * nchan;
    memset(row, 0, (size_t)nchan);
    memcpy(row + c0, sub, (size_t)C);
  }
  free(x); free(sub);
  return nw;
}

void afo_apply_window_masks(void *data, int dtype, size_t nsample, int nchan,
                            const uint8_t *masks, const uint32_t *t0s,
                            const uint32_t *t1s, int nwin) {
  size_t es = dsize(dtype);
  for (int w = 0; w < nwin; ++w) {
    size_t t0 = t0s[w] < nsample ? t0s[w] : nsample;
    size_t t1 = t1s[w] < nsample ? t1s[w] : nsample;
    for (int c = 0; c < nchan; ++c) {
      if (!masks[(size_t)w * nchan + c]) continue;
      for (size_t t = t0; t < t1; ++t)
        memset((char *)data + (t * nchan + c) * es, 0, es);
    }
  }
}

/* ---------------------------------------------------------------- zero-DM -- */

void afo_zero_dm(void *data, int dtype, size_t nsample, int nchan, int c0, int c1) {
  if (c0 < 0) c0 = 0;
  if (c1 > nchan) c1 = nchan;
  if (c0 >= c1) return;
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < nsample; ++t) {
    size_t row = t * (size_t)nchan;
    if (dtype == AFO_F32) {
      float *f = (float *)data;
      double S = 0.0;
      unsigned n = 0;
      for (int c = c0; c < c1; ++c)
        if (f[row + c] != 0.0f) { S += (double)f[row + c]; ++n; }
      float m = n ? (float)(S / (double)n) : 0.0f;
      for (int c = c0; c < c1; ++c) f[row + c] = f[row + c] - m;
      continue;
    }
    uint64_t S = 0;
    unsigned n = 0;
    for (int c = c0; c < c1; ++c) {
      uint64_t v = ld_u(data, dtype, row + c);
      if (v) { S += v; ++n; }
    }
    float m = n ? (float)S / (float)n : 0.0f;
    float vmax = (float)vmax_of(dtype);
    for (int c = c0; c < c1; ++c) {
      float r = (float)ld_u(data, dtype, row + c) - m;
      if (r < 0.0f) r = 0.0f;
      if (r > vmax) r = vmax;
      uint64_t o = (r >= 4294967296.0f) ? 4294967295ull : (uint64_t)r;
      st_u(data, dtype, row + c, o);
    }
  }
}

void afo_static_mask(void *data, int dtype, size_t nsample, int nchan,
                     const int *bad, int nbad) {
  size_t es = dsize(dtype);
  for (int b = 0; b < nbad; ++b) {
    int c = bad[b];
    if (c < 0 || c >= nchan) continue;
    for (size_t t = 0; t < nsample; ++t)
      memset((char *)data + ((size_t)t * nchan + c) * es, 0, es);
  }
}

/* ---------------------------------------------------------- dedispersion -- */

void afo_dedisperse_direct(const void *data, int dtype, size_t nsamp, int nchan,
                           int c0, int c1, const int32_t *delays, int ndm,
                           size_t nout, void *out) {
  int nc = c1 - c0;
  /* (DM row, block of 4096 outputs) pairs over threads; every output is
   * summed in channel order as written */
  const size_t nb = (nout + 4095) / 4096;
  const long long npair = (long long)ndm * (long long)nb;
#pragma omp parallel for schedule(dynamic)
  for (long long q = 0; q < npair; ++q) {
    const int i = (int)(q / (long long)nb);
    const size_t tb = (size_t)(q % (long long)nb) * 4096;
    const size_t te = tb + 4096 < nout ? tb + 4096 : nout;
    const int32_t *d = delays + (size_t)i * nc;
    for (size_t t = tb; t < te; ++t) {
      if (dtype == AFO_F32) {
        double acc = 0.0;
        for (int c = c0; c < c1; ++c) {
          long long idx = (long long)t + d[c - c0];
          if (idx >= 0 && (size_t)idx < nsamp)
            acc += (double)((const float *)data)[(size_t)idx * nchan + c];
        }
        ((double *)out)[(size_t)i * nout + t] = acc;
      } else {
        uint64_t acc = 0;
        for (int c = c0; c < c1; ++c) {
          long long idx = (long long)t + d[c - c0];
          if (idx >= 0 && (size_t)idx < nsamp) acc += ld_u(data, dtype, (size_t)idx * nchan + c);
        }
        if (dtype == AFO_U32) ((uint64_t *)out)[(size_t)i * nout + t] = acc;
        else ((uint32_t *)out)[(size_t)i * nout + t] = (uint32_t)acc;
      }
    }
  }
}

void afo_dedisperse_subband(const void *data, int dtype, size_t nsamp, int nchan,
                            int c0, int c1, int sbw, const int32_t *delays1,
                            const int32_t *res, int ndm, int ndm0, size_t nout,
                            void *out) {
  afo_dedisperse_subband_rows(data, dtype, nsamp, nchan, c0, c1, sbw, delays1, res,
                              ndm, ndm0, nout, NULL, ndm, out);
}

void afo_dedisperse_subband_rows(const void *data, int dtype, size_t nsamp, int nchan,
                                 int c0, int c1, int sbw, const int32_t *delays1,
                                 const int32_t *res, int ndm, int ndm0, size_t nout,
                                 const int *rows, int nrows, void *out) {
  int nc = c1 - c0;
  int nsb = (nc + sbw - 1) / sbw;
  (void)ndm;
  /* (row, block of 4096 outputs) pairs over threads; every output is summed
   * subband by subband, each subband in channel order, as written */
  const size_t nb = (nout + 4095) / 4096;
  const long long npair = (long long)nrows * (long long)nb;
#pragma omp parallel for schedule(dynamic)
  for (long long q = 0; q < npair; ++q) {
    const int ri = (int)(q / (long long)nb);
    const size_t tb = (size_t)(q % (long long)nb) * 4096;
    const size_t te = tb + 4096 < nout ? tb + 4096 : nout;
    int m = rows ? rows[ri] : ri;
    size_t orow = (size_t)ri;
    int j = m / ndm0;
    const int32_t *d1 = delays1 + (size_t)j * nc;
    const int32_t *r = res + (size_t)m * nsb;
    for (size_t t = tb; t < te; ++t) {
      uint64_t acc = 0;
      double accf = 0.0;
      for (int sb = 0; sb < nsb; ++sb) {
        long long ts = (long long)t + r[sb];
        if (ts < 0 || (size_t)ts >= nsamp) continue;
        int ch0 = c0 + sb * sbw;
        int ch1 = ch0 + sbw < c1 ? ch0 + sbw : c1;
        uint64_t part = 0;
        double partf = 0.0;
        for (int c = ch0; c < ch1; ++c) {
          long long idx = ts + d1[c - c0];
          if (idx < 0 || (size_t)idx >= nsamp) continue;
          if (dtype == AFO_F32) partf += (double)((const float *)data)[(size_t)idx * nchan + c];
          else part += ld_u(data, dtype, (size_t)idx * nchan + c);
        }
        acc += part;
        accf += partf;
      }
      if (dtype == AFO_F32) ((double *)out)[orow * nout + t] = accf;
      else if (dtype == AFO_U32) ((uint64_t *)out)[orow * nout + t] = acc;
      else ((uint32_t *)out)[orow * nout + t] = (uint32_t)acc;
    }
  }
}
