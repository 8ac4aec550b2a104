/*
 * af_oracle.h -- CPU oracle for the astroflow dedispersion + RFI hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every function restates the reference algorithm from its source text
 * (lintian233/astroflow), with the citations given per function, and with one
 * documented change: the delay and residual tables are computed in FP64.
 *
 * Pinning (what checks each function, see tests/):
 *   afo_freq_table, afo_chan_range, afo_dm_steps  : hand-computed values + closed forms
 *   afo_delay_table, afo_residual_table           : closed form, hand-computed spot values,
 *                                                   monotonicity, zero at f_ref
 *   afo_bin                                       : exhaustive small cases against exact
 *                                                   integer floor(sqrt(vmax*S/cnt))
 *   afo_iqrm_windows, afo_iqrm_lags               : closed forms
 *   afo_iqrm (flavour CPU)                        : bit-exact vs the reference's own
 *                                                   include/iqrm.hpp compiled into oracle/_ref
 *   afo_iqrm (flavour GPU)                        : equal to the CPU flavour wherever the
 *                                                   stats/percentile precisions cannot differ
 *                                                   (mode 0, exact quartile positions), plus
 *                                                   special cases (constant stat, lone outlier)
 *   afo_apply_window_masks                        : bit-exact vs the reference's
 *                                                   src/rfi/rfimarker_cpu.cpp in oracle/_ref
 *   afo_zero_dm, afo_static_mask                  : closed forms per sample
 *   afo_dedisperse_direct                         : constant input, DM=0 channel sums,
 *                                                   impulse placement, linearity, brute force
 *   afo_dedisperse_subband                        : equal to direct when subband width=1 and
 *                                                   NDM0=1; brute force on tiny inputs
 *   afo_cpucal_direct                             : equal to afo_dedisperse_direct for ds=1
 *   afo_iqrm_quartiles                            : closed forms at every C mod 4 (exact
 *                                                   interpolation weights 0, 1/4, 1/2, 3/4),
 *                                                   and a case where float and double
 *                                                   interpolation round differently
 *
 * Floating-point contraction.  Every C file here is compiled with
 * -ffp-contract=off: a*b+c rounds twice, as written.  "Bit-exact to the
 * reference" for the floating-point steps (IQRM statistic Sum x^2/W - mean^2
 * and percentile interpolation, zero-DM mean and subtraction, delay tables)
 * rests on that reading of the source text; nvcc's default (--fmad=true)
 * may contract such expressions into fused multiply-adds, which would change
 * the last bit of a statistic in rare cases.
 */
#ifndef AF_ORACLE_H
#define AF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { AFO_U8 = 1, AFO_U16 = 2, AFO_U32 = 3, AFO_F32 = 4 };

/* Dispersion constant used by every delay formula in the reference
 * (src/core/dedispered.cu:333, include/cpucal.hpp:135). */
#define AFO_KDM 4148.808

/* h_freq[i] = fch1 + i*foff evaluated in float32, as
 * src/core/dedispered.cu:1003-1005 does with the float Header fields. */
void afo_freq_table_f32(float fch1, float foff, int nchan, double *out);

/* Channel range from a frequency range, float32 arithmetic and truncation as
 * src/core/dedispered.cu:1007-1027.  Returns 0 on success, -1 when the
 * frequency range lies outside the table (reference: std::invalid_argument). */
int afo_chan_range(const double *freq, int nchan, float fs, float fe,
                   long *chan_start, long *chan_end);

/* dm_steps = (size_t)((dm_high - dm_low)/dm_step) + 1 in float32
 * (src/core/dedispered.cu:1029). */
long afo_dm_steps(float dm_low, float dm_high, float dm_step);

/* FP64 delay table d[i][c-c0] = trunc(round_half_away(K*dm_i*(f_c^-2 - f_ref^-2)/tsamp)/ds)
 * for c in [c0,c1).  Follows src/core/dedispered.cu:318-335 (delay) and
 * :379-382 (truncating division by the downsample factor). */
void afo_delay_table(const double *freq, int c0, int c1, double f_ref,
                     double tsamp, int ds, const double *dms, int ndm,
                     int32_t *out);

/* Subband residual table res[m][sb] = floor(K*(dm_m - dm_j)*(f_sb^-2 - f_ref^-2)/(tsamp*ds) + 0.5),
 * dm_j the coarse DM of group j=m/ndm0, f_sb = freq[(ch0+ch1-1)/2].
 * Follows src/core/dedispered.cu:384-413 and the subband centre choice at
 * :1152-1163.  Subband sb covers channels [c0+sb*sbw, min(c0+(sb+1)*sbw, c1)). */
void afo_residual_table(const double *freq, int c0, int c1, int sbw,
                        double f_ref, double tsamp, int ds, double dm_low,
                        double dm_step, int ndm, int ndm0, int32_t *out);

/* Time binning by ds, src/core/dedispered.cu:337-376 (sqrt compression) for
 * unsigned types; float32 data is summed (no reference semantics exist). */
void afo_bin(const void *in, int dtype, size_t nsamp, int nchan, int ds, void *out);

/* IQRM window list, src/core/iqrm.cu:368-395.  Returns the window count;
 * writes at most maxw windows. */
int afo_iqrm_windows(size_t nsample, unsigned W, unsigned H, int include_tail,
                     uint32_t *t0s, uint32_t *t1s, int maxw);

/* IQRM lag list, src/core/iqrm.cu:189-200 / include/iqrm.hpp:44-55. */
int afo_iqrm_lags(int radius, double geofactor, int *lags, int maxl);

/* IQRM radius from the fraction: max(1, floor(radius_frac*(float)C)). */
int afo_iqrm_radius(float radius_frac, int C);

/* IQRM flavours.  GPU: the shipped path (src/core/iqrm.cu): one-pass std from
 * exact sums, percentile position/weight in double.  CPU: include/iqrm.hpp
 * (two-pass std, float percentile, non-finite values dropped before sorting). */
enum { AFO_IQRM_GPU = 0, AFO_IQRM_CPU = 1 };

/* Per-channel statistic of one window (mode 0 mean, 1 std) for channels
 * [c0,c1); flavour selects the formula (src/core/iqrm.cu:72-91 or
 * include/iqrm.hpp:189-217). */
void afo_iqrm_stat(const void *data, int dtype, int nchan, uint32_t t0,
                   uint32_t t1, int c0, int c1, int mode, int flavour,
                   float *x);

/* Robust lag voting on a statistic vector, src/core/iqrm.cu:244-327 and
 * include/iqrm.hpp:102-159.  mask[C] gets 0/1. */
/* Quartiles q[0..2] = (q1, median, q3) of m lagged differences, sorted into
 * srt[m]: flavour CPU (finite values, float interpolation) or GPU (all
 * values, double position and weight).  Returns the count used (0: none). */
size_t afo_iqrm_quartiles(const float *d, size_t m, int flavour, float *srt, float q[3]);

void afo_iqrm_on_stat(const float *x, int C, float radius_frac, float nsigma,
                      double geofactor, int flavour, uint8_t *mask);

/* Full IQRM over all windows.  masks is [nwin][nchan] with zeros outside
 * [c0,c1).  Returns the window count (<= maxw). */
int afo_iqrm(const void *data, int dtype, size_t nsample, int nchan, int c0,
             int c1, int mode, float radius_frac, float nsigma,
             double geofactor, unsigned W, unsigned H, int include_tail,
             int flavour, uint8_t *masks, uint32_t *t0s, uint32_t *t1s,
             int maxw);

/* Zero x[t][c] for every window w, t in [t0_w,t1_w), mask_w[c] != 0
 * (src/rfi/rfimarker.cu:76-100, :309-390). */
void afo_apply_window_masks(void *data, int dtype, size_t nsample, int nchan,
                            const uint8_t *masks, const uint32_t *t0s,
                            const uint32_t *t1s, int nwin);

/* Zero-DM over channels [c0,c1): per sample m = S/n over non-zero values
 * (float), x = clamp((float)x - m, 0, vmax) truncated.  The range form of
 * src/rfi/rfimarker.cu:172-305 with alpha = 1, clip_to_zero = true.  For
 * float32 data the mean is over all values and there is no clipping.
 * The per-sample sums are exact integers converted to float once. */
void afo_zero_dm(void *data, int dtype, size_t nsample, int nchan, int c0, int c1);

/* Zero listed channels for all samples (src/rfi/rfimarker.cu:9-30,441-488). */
void afo_static_mask(void *data, int dtype, size_t nsample, int nchan,
                     const int *bad, int nbad);

/* Direct dedispersion over channels [c0,c1) with delay rows d[i][c-c0]:
 * out[i][t] = sum_c [0 <= t+d < nsamp] x[t+d][c], t in [0,nout).
 * Accumulator: u32 for u8/u16 input, u64 for u32, double (FP64) for f32.
 * Loop nest of include/cpucal.hpp:167-186. */
void afo_dedisperse_direct(const void *data, int dtype, size_t nsamp, int nchan,
                           int c0, int c1, const int32_t *delays, int ndm,
                           size_t nout, void *out);

/* Subband dedispersion, src/core/dedispered.cu:416-476 with the tile loop of
 * :1199-1235 (the tiling does not change the result):
 *   out[m][t] = sum_sb [0 <= t+r < nsamp] sum_{c in sb} [0 <= t+r+d1 < nsamp] x[t+r+d1][c]
 * with r = res[m][sb], d1 = d1[m/ndm0][c-c0].  delays1 is [ndm_nom][c1-c0];
 * res is [ndm][nsb].  Accumulators as afo_dedisperse_direct. */
void afo_dedisperse_subband(const void *data, int dtype, size_t nsamp, int nchan,
                            int c0, int c1, int sbw, const int32_t *delays1,
                            const int32_t *res, int ndm, int ndm0, size_t nout,
                            void *out);

/* The same for a subset of fine-DM rows: out is [nrows][nout], row ri holds
 * DM index rows[ri]. */
void afo_dedisperse_subband_rows(const void *data, int dtype, size_t nsamp, int nchan,
                                 int c0, int c1, int sbw, const int32_t *delays1,
                                 const int32_t *res, int ndm, int ndm0, size_t nout,
                                 const int *rows, int nrows, void *out);

/* CPU baseline: the include/cpucal.hpp:117-186 loop nest (OpenMP dynamic over
 * DMs, SIMD reduction over channels, u64 accumulation, decimation by ds) on
 * one slice [start, start+len) of u8 data.  delays[i][c-c0] at native rate. */
void afo_cpucal_direct_u8(const uint8_t *data, size_t ndata, int nchan, int c0,
                          int c1, const int32_t *delays, int ndm, size_t start,
                          size_t len, int ds, uint64_t *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
