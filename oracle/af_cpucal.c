/*
 * af_cpucal.c -- CPU baseline (test/bench infrastructure only).
 *
 * Restates the loop nest of cpucal::dedispered_fil_omp
 * (include/cpucal.hpp:167-186): OpenMP dynamic schedule over DM trials, a
 * time loop, and a SIMD sum reduction over channels with u64 accumulation;
 * samples beyond the data end contribute zero; time_downsample > 1
 * decimates (reads every ds-th native sample, include/cpucal.hpp:174).
 * Built with the reference's host optimisation flags (-Ofast, OpenMP).
 */
#include "af_oracle.h"

#include <omp.h>

void afo_cpucal_direct_u8(const uint8_t *data, size_t ndata, int nchan, int c0,
                          int c1, const int32_t *delays, int ndm, size_t start,
                          size_t len, int ds, uint64_t *out, int nthreads) {
  const size_t nc = (size_t)(c1 - c0);
  const size_t down = (len + (size_t)ds - 1) / (size_t)ds;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
  for (int step = 0; step < ndm; ++step) {
    const int32_t *d = delays + (size_t)step * nc;
    uint64_t *row = out + (size_t)step * down;
    for (size_t ti = 0; ti < down; ++ti) {
      const size_t base = start + ti * (size_t)ds;
      uint64_t sum = 0;
#pragma omp simd reduction(+ : sum)
      for (size_t ch = 0; ch < nc; ++ch) {
        size_t target = base + (size_t)(long)d[ch];
        if (target < ndata) sum += data[target * (size_t)nchan + (ch + (size_t)c0)];
      }
      row[ti] = sum;
    }
  }
}
