// af_kernels.hip -- gfx950 kernels of the dedispersion + RFI path.
//
// Data layouts in HBM:
//   input        x[nsamp][nchan]       time-major, as read from a filterbank
//   binned       xb[nbin][nchan]       time-major (ds > 1 only)
//   cleaned      xT[nch][Lp]           channel-major, channels [dd_c0, dd_c1),
//                                      zero-padded from nbin to Lp so that the
//                                      dedispersion loads need no bounds checks
//   intermediate inter[ndm_nom][nsb][L1]  subband partial sums (u16 for u8 data)
//   output       out[ndm][nout]        DM-major
//
// Compiled with -ffp-contract=off: every floating-point expression rounds
// exactly as written, as the CPU oracle does.
#include "af_internal.h"

namespace af {

// ----------------------------------------------------------------- helpers --

template <typename T> struct Tr;
template <> struct Tr<uint8_t> { static constexpr double vmax = 255.0; };
template <> struct Tr<uint16_t> { static constexpr double vmax = 65535.0; };
template <> struct Tr<uint32_t> { static constexpr double vmax = 4294967295.0; };
template <> struct Tr<float> { static constexpr double vmax = 0.0; };

// Block index remap so that consecutive logical blocks land on the same XCD
// (the dispatcher hands block b to XCD b % 8): logical ranges stay inside one
// XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
  const uint32_t X = 8;
  uint32_t q = n / X, r = n % X;
  uint32_t xcd = b % X, k = b / X;
  uint32_t start = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return start + k;
}

// Bytes 1 and 3 of w as two 16-bit lanes: (w >> 8) & 0x00FF00FF in one
// v_perm_b32 (selector byte 0x0c yields zero).
__device__ __forceinline__ uint32_t odd_bytes(uint32_t w) {
  return __builtin_amdgcn_perm(0u, w, 0x0c030c01u);
}

__device__ __forceinline__ uint4 ldu4(const uint8_t *p) {
  uint4 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}

// Correctly rounded conversion of a 128-bit unsigned integer to double
// (round to nearest, ties to even), as the host's (double)(unsigned __int128).
__host__ __device__ inline double u128_to_double(uint64_t hi, uint64_t lo) {
  if (hi == 0) return (double)lo;
  int lz = __builtin_clzll(hi);
  int sh = 64 - lz;  // bits of hi
  // value = hi:lo ; keep the top 64 bits in m, remember the dropped bits
  uint64_t m = (hi << lz) | (lo >> sh);
  uint64_t dropped = lo << lz;  // bits shifted out below m (as a fraction)
  // m has its top bit set; round m to 53 bits
  uint64_t keep = m >> 11;
  uint64_t rem = m & 0x7FF;
  bool sticky = dropped != 0;
  bool round_up = rem > 0x400 || (rem == 0x400 && (sticky || (keep & 1)));
  if (round_up) keep += 1;
  // keep is in [2^52, 2^53]; value = keep * 2^(11 + sh)
  return ldexp((double)keep, 11 + sh);
}

// Copies nrows rows of `ext` 16-byte chunks (row r starts at src + r*pitch
// elements of E per chunk) into LDS rows of `ldsw` chunks.  Each thread keeps
// four loads in flight before its LDS stores.
template <typename E>
__device__ __forceinline__ void fill_rows16(uint4 *__restrict__ lds, int ldsw,
                                            const E *__restrict__ src, size_t pitch, int epc,
                                            int nrows, int ext) {
  const int total = nrows * ext;
  if (total <= 0) return;
  const int step = (int)blockDim.x;
  // a thread past the end repeats the last chunk (same value, same slot):
  // loads and stores stay unconditional, so the four loads stay in flight
  // (a load under a per-lane branch is waited for at the join)
  for (int base = threadIdx.x; base < total; base += 4 * step) {
    int slot[4];
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int idx = min(base + u * step, total - 1);
      const int r = idx / ext, k = idx - r * ext;
      slot[u] = r * ldsw + k;
      v[u] = *(const uint4 *)(src + (size_t)r * pitch + (size_t)epc * k);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) lds[slot[u]] = v[u];
  }
}

// ------------------------------------------------------------------ binning --
// src/core/dedispered.cu:337-376 (sqrt compression of the ds-sample sum).
template <typename T>
__device__ __forceinline__ void bin_body(const BinArgs &a, uint32_t bid, uint32_t nblk) {
  const T *__restrict__ in = (const T *)a.in;
  T *__restrict__ out = (T *)a.out;
  const size_t nsamp = a.nsamp, nbin = a.nbin;
  const int nchan = a.nchan, ds = a.ds;
  const size_t total = nbin * (size_t)nchan;
  for (size_t idx = (size_t)bid * blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)nblk * blockDim.x) {
    const size_t c = idx % (size_t)nchan, s = idx / (size_t)nchan;
    const size_t t0 = s * (size_t)ds;
    const size_t t1 = t0 + (size_t)ds < nsamp ? t0 + (size_t)ds : nsamp;
    if constexpr (std::is_same<T, float>::value) {
      double acc = 0.0;
      for (size_t t = t0; t < t1; ++t) acc += (double)in[t * nchan + c];
      out[idx] = (float)acc;
    } else {
      uint64_t sum = 0;
      for (size_t t = t0; t < t1; ++t) sum += (uint64_t)in[t * nchan + c];
      const double vmax = Tr<T>::vmax;
      const uint64_t cnt = (uint64_t)(t1 - t0);
      double sq = sqrt((double)sum);
      double smax = sqrt((double)((uint64_t)vmax * cnt));
      double v = (sq / smax) * vmax;
      double r = v < vmax ? v : vmax;
      out[idx] = (T)r;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bin(BinArgs a) { bin_body<T>(a, blockIdx.x, gridDim.x); }

// ------------------------------------------------------- IQRM statistics --
// Per (row block, channel) sums over the block's rows: the first step of
// src/core/iqrm.cu:21-69, with exact integer sums (coalesced: a wave reads 64
// consecutive channels of a row).
template <typename T>
__device__ __forceinline__ void stats_body(const StatArgs &a, uint32_t bid, uint32_t) {
  const T *__restrict__ x = (const T *)a.x;
  const int nchan = a.nchan, c0 = a.c0, c1 = a.c1;
  const RowItem *__restrict__ items = a.items;
  Part *__restrict__ part = (Part *)a.part;
  const int Csub = c1 - c0;
  const int nchunk = (Csub + 255) / 256;
  const int item = (int)bid / nchunk;
  const int c = c0 + ((int)bid % nchunk) * 256 + threadIdx.x;
  if (c >= c1) return;
  const RowItem it = items[item];
  Part p{0, 0, 0};
  if constexpr (std::is_same<T, float>::value) {
    double s = 0.0, q = 0.0;
    for (uint32_t r = it.r0; r < it.r1; ++r) {
      double v = (double)x[(size_t)r * nchan + c];
      s += v;
      q += v * v;
    }
    p.s = __double_as_longlong(s);
    p.q_lo = __double_as_longlong(q);
  } else {
    uint64_t s = 0, lo = 0, hi = 0;
    for (uint32_t r = it.r0; r < it.r1; ++r) {
      uint64_t v = (uint64_t)x[(size_t)r * nchan + c];
      uint64_t q = v * v;
      s += v;
      lo += q;
      hi += (lo < q) ? 1u : 0u;
    }
    p.s = s;
    p.q_lo = lo;
    p.q_hi = hi;
  }
  part[(size_t)item * Csub + (c - c0)] = p;
}

template <typename T>
__global__ __launch_bounds__(256) void k_seg_stats(StatArgs a) {
  stats_body<T>(a, blockIdx.x, gridDim.x);
}

// Sum the row-block partials of each segment in item order.
template <typename T>
__device__ __forceinline__ void reduce_body(const ReduceArgs &a, uint32_t bid, uint32_t) {
  const Part *__restrict__ part = (const Part *)a.part;
  const int *__restrict__ seg_items = a.seg_items;
  const int nseg = a.nseg, Csub = a.Csub;
  Part *__restrict__ segsum = a.segsum;
  const size_t idx = (size_t)bid * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nseg * Csub) return;
  const int s = (int)(idx / Csub), c = (int)(idx % Csub);
  Part acc{0, 0, 0};
  if constexpr (std::is_same<T, float>::value) {
    double a = 0.0, b = 0.0;
    for (int i = seg_items[s]; i < seg_items[s + 1]; ++i) {
      a += __longlong_as_double(part[(size_t)i * Csub + c].s);
      b += __longlong_as_double(part[(size_t)i * Csub + c].q_lo);
    }
    acc.s = __double_as_longlong(a);
    acc.q_lo = __double_as_longlong(b);
  } else {
    for (int i = seg_items[s]; i < seg_items[s + 1]; ++i) {
      const Part p = part[(size_t)i * Csub + c];
      acc.s += p.s;
      uint64_t lo = acc.q_lo + p.q_lo;
      acc.q_hi += p.q_hi + (lo < p.q_lo ? 1u : 0u);
      acc.q_lo = lo;
    }
  }
  segsum[(size_t)s * Csub + c] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_seg_reduce(ReduceArgs a) {
  reduce_body<T>(a, blockIdx.x, gridDim.x);
}

// --------------------------------------------------- IQRM robust voting --
// One bitonic pass at distance J < E between a thread's own elements.
template <int E, int J>
__device__ __forceinline__ void sort_pass_regs(float (&v)[E], int t, int kk) {
  if constexpr (J < E) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      if ((r & J) == 0) {
        const bool asc = ((E * t + r) & kk) == 0;
        const float a = v[r], b = v[r | J];
        const float mn = fminf(a, b), mx = fmaxf(a, b);
        v[r] = asc ? mn : mx;
        v[r | J] = asc ? mx : mn;
      }
    }
  }
}

// Lagged differences of the statistic (src/core/iqrm.cu:94-117) into ds_
// and, sorted ascending, into srt.  Bitonic network over P2 values held E per
// thread (1024 threads; element E*t + r in thread t's register r): distances
// below E are compare-exchanges in registers, partners inside the wave come
// by lane shuffles, and only the 10 longest distances (of 78 passes at
// 4096) exchange through LDS between workgroup barriers.  (DPP moves for the
// partners up to 8 lanes away measured slower than the shuffles.)  srt has
// room for max(P2, 1024) floats.  P2C: P2 known at compile time (the
// network is then unrolled, with no per-pass branches; on the 4096-channel
// chunk the vote took 35 us instead of 45 us), or 0.
template <int E, int P2C>
__device__ __forceinline__ void lag_sort(const float *xs, float *ds_, float *srt, int Csub, int lag,
                                         int P2r) {
  const int P2 = P2C > 0 ? P2C : P2r;
  const int t = threadIdx.x;
  float v[E];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    const int k = E * t + r;
    float dv = INFINITY;
    if (k < Csub) {
      if (lag < 0) {
        const int cut = Csub + lag;
        dv = (k < cut) ? xs[k] - xs[k - lag] : xs[cut] - xs[Csub - 1];
      } else {
        dv = (k < lag) ? xs[lag] - xs[0] : xs[k] - xs[k - lag];
      }
      ds_[k] = dv;
    }
    v[r] = dv;
  }
  // one pass of the network: sequence length kk, distance j
  auto pass = [&](const int kk, const int j) {
    if (j < E) {
      // compile-time distances keep v[] in registers
      if (j == 1) sort_pass_regs<E, 1>(v, t, kk);
      else if (j == 2) sort_pass_regs<E, 2>(v, t, kk);
      else if (j == 4) sort_pass_regs<E, 4>(v, t, kk);
      else sort_pass_regs<E, 8>(v, t, kk);
      return;
    }
    const int m = j / E;   // partner thread t ^ m holds the other element
    float pv[E];
    if (m < 64) {
#pragma unroll
      for (int r = 0; r < E; ++r) pv[r] = __shfl_xor(v[r], m);
    } else {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < E; ++r) srt[E * t + r] = v[r];
      __syncthreads();
#pragma unroll
      for (int r = 0; r < E; ++r) pv[r] = srt[E * (t ^ m) + r];
    }
    const bool lower = (t & m) == 0;
#pragma unroll
    for (int r = 0; r < E; ++r) {
      const bool asc = ((E * (lower ? t : (t ^ m)) + r) & kk) == 0;
      const float mn = fminf(v[r], pv[r]), mx = fmaxf(v[r], pv[r]);
      v[r] = (lower == asc) ? mn : mx;
    }
  };
  if constexpr (P2C > 0) {
#pragma unroll
    for (int kk = 2; kk <= P2C; kk <<= 1)
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1) pass(kk, j);
  } else {
    for (int kk = 2; kk <= P2; kk <<= 1)
      for (int j = kk >> 1; j > 0; j >>= 1) pass(kk, j);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < E; ++r) srt[E * t + r] = v[r];
}

// One workgroup per (window, lag): statistic from the window's segment sums
// (src/core/iqrm.cu:72-91), lagged difference with edge clipping (:94-117),
// quartiles by a bitonic sort (lag_sort) and linear interpolation in double
// (:288-306), and the flag test (:120-135).
template <typename T>
__device__ __forceinline__ void lag_body(const LagArgs &a, uint32_t bid, uint32_t) {
  const Part *__restrict__ segsum = a.segsum;
  const uint32_t *__restrict__ wseg = a.wseg;
  const uint32_t *__restrict__ wt0 = a.wt0;
  const uint32_t *__restrict__ wt1 = a.wt1;
  const int Csub = a.Csub, mode = a.mode, nlags = a.nlags;
  const int *__restrict__ lags = a.lags;
  const float nsigma = a.nsigma;
  uint8_t *__restrict__ flags = a.flags;
  extern __shared__ float lds[];
  const int w = (int)bid / nlags, li = (int)bid % nlags;
  const int lag = lags[li];
  int P2 = 1;
  while (P2 < Csub) P2 <<= 1;
  float *xs = lds;
  float *ds_ = lds + Csub;
  float *srt = lds + 2 * Csub;
  __shared__ float sh_med, sh_thr;
  __shared__ int sh_valid;
  uint8_t *frow = flags + ((size_t)w * nlags + li) * Csub;
  const bool lag_ok = lag < Csub && -lag < Csub;
  if (!lag_ok) {
    for (int k = threadIdx.x; k < Csub; k += blockDim.x) frow[k] = 0;
    return;
  }
  const uint32_t s0 = wseg[2 * w], s1 = wseg[2 * w + 1];
  const double W = (double)(wt1[w] - wt0[w]);
  for (int c = threadIdx.x; c < Csub; c += blockDim.x) {
    float xv;
    if constexpr (std::is_same<T, float>::value) {
      double S = 0.0, Q = 0.0;
      for (uint32_t s = s0; s < s1; ++s) {
        S += __longlong_as_double(segsum[(size_t)s * Csub + c].s);
        Q += __longlong_as_double(segsum[(size_t)s * Csub + c].q_lo);
      }
      double mean = S / W;
      if (mode == 0) {
        xv = (float)mean;
      } else {
        double var = Q / W - mean * mean;
        if (var < 0.0) var = 0.0;
        xv = (float)sqrt(var);
      }
    } else {
      uint64_t S = 0, lo = 0, hi = 0;
      for (uint32_t s = s0; s < s1; ++s) {
        const Part p = segsum[(size_t)s * Csub + c];
        S += p.s;
        uint64_t nl = lo + p.q_lo;
        hi += p.q_hi + (nl < p.q_lo ? 1u : 0u);
        lo = nl;
      }
      double mean = (double)S / W;
      if (mode == 0) {
        xv = (float)mean;
      } else {
        double var = u128_to_double(hi, lo) / W - mean * mean;
        if (var < 0.0) var = 0.0;
        xv = (float)sqrt(var);
      }
    }
    xs[c] = xv;
  }
  __syncthreads();
  // lagged differences: ds_ (for the flags) and the sort input, E per
  // thread (thread t holds elements E*t .. E*t+E-1; +inf past Csub)
  switch (P2) {
  case 1024: lag_sort<1, 1024>(xs, ds_, srt, Csub, lag, P2); break;
  case 2048: lag_sort<2, 2048>(xs, ds_, srt, Csub, lag, P2); break;
  case 4096: lag_sort<4, 4096>(xs, ds_, srt, Csub, lag, P2); break;
  case 8192: lag_sort<8, 8192>(xs, ds_, srt, Csub, lag, P2); break;
  default:
    if (P2 < 1024) lag_sort<1, 0>(xs, ds_, srt, Csub, lag, P2);
    else lag_sort<16, 0>(xs, ds_, srt, Csub, lag, P2);
    break;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float qv[3];
    const float qs[3] = {0.25f, 0.50f, 0.75f};
    for (int i = 0; i < 3; ++i) {
      if (Csub == 1) {
        qv[i] = srt[0];
      } else {
        double pos = (double)qs[i] * (double)(Csub - 1);
        int lo = (int)floor(pos), hi = (int)ceil(pos);
        double wgt = pos - (double)lo;
        float diff = srt[hi] - srt[lo];
        double prod = (double)diff * wgt;
        qv[i] = (float)((double)srt[lo] + prod);
      }
    }
    float s = fabsf((qv[2] - qv[0]) / 1.349f);
    sh_valid = (s > 0.f) && isfinite(s);
    sh_med = qv[1];
    sh_thr = nsigma * s;
  }
  __syncthreads();
  const int valid = sh_valid;
  const float med = sh_med, thr = sh_thr;
  for (int k = threadIdx.x; k < Csub; k += blockDim.x) {
    float val = ds_[k] - med;
    frow[k] = (valid && val > thr) ? 1 : 0;
  }
}

template <typename T>
__global__ __launch_bounds__(1024) void k_iqrm_lag(LagArgs a) {
  lag_body<T>(a, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ int iqrm_partner(int k, int lag, int C) {
  if (lag < 0) return (k < C + lag) ? k - lag : C - 1;
  return (k < lag) ? 0 : k - lag;
}

// Votes and final window mask (src/core/iqrm.cu:138-185), one workgroup per
// (window, block of kMergeBlock channels).  A channel's mask needs the casting
// counts of its partners (at most R = max |lag| away), and a count needs the
// flags of channels at most R from it (the edge partners 0 and C-1 collect
// channels within R of the edge), so the block's flags are gathered over a
// halo of 2R: one bit per lag per channel in LDS (eight byte loads in flight
// per thread; nw = ceil(nlags / 32) words per channel).  The counts over the
// block and a halo of R then come from LDS atomics, and the test reads LDS.
constexpr int kMergeBlock = 1024;
__device__ __forceinline__ void merge_body(const MergeArgs &a, uint32_t bid, uint32_t) {
  const uint8_t *__restrict__ flags = a.flags;
  const int *__restrict__ lags = a.lags;
  const int nlags = a.nlags, Csub = a.Csub, c0 = a.c0, nchan = a.nchan;
  const int nw = (nlags + 31) / 32;
  uint8_t *__restrict__ masks = a.masks;
  extern __shared__ int cast[];
  const int nblk = (Csub + kMergeBlock - 1) / kMergeBlock;
  const int w = (int)bid / nblk, blk = (int)bid % nblk;
  int R = 0;
  for (int l = 0; l < nlags; ++l) {
    const int al = abs(lags[l]);
    if (al < Csub) R = max(R, al);
  }
  const int b0 = blk * kMergeBlock, b1 = min(b0 + kMergeBlock, Csub);
  const int f0 = max(0, b0 - 2 * R), f1 = min(Csub, b1 + 2 * R);   // flags gathered
  const int g0 = max(0, b0 - R), g1 = min(Csub, b1 + R);           // counts formed
  uint32_t *bits = (uint32_t *)(cast + (g1 - g0));                 // [nw][f1 - f0]
  const int nf = f1 - f0;
  const uint8_t *f = flags + (size_t)w * nlags * Csub;
  for (int k = f0 + threadIdx.x; k < f1; k += blockDim.x) {
    for (int l0 = 0; l0 < nlags; l0 += 32) {
      const int l1 = min(l0 + 32, nlags);
      uint32_t b = 0;
      int l = l0;
      for (; l + 8 <= l1; l += 8) {
        uint8_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = f[(size_t)(l + u) * Csub + k];
#pragma unroll
        for (int u = 0; u < 8; ++u) b |= (v[u] ? 1u : 0u) << (l + u - l0);
      }
      for (; l < l1; ++l) b |= (f[(size_t)l * Csub + k] ? 1u : 0u) << (l - l0);
      bits[(size_t)(l0 >> 5) * nf + (k - f0)] = b;
    }
  }
  for (int j = g0 + threadIdx.x; j < g1; j += blockDim.x) cast[j - g0] = 0;
  __syncthreads();
  for (int k = f0 + threadIdx.x; k < f1; k += blockDim.x)
    for (int l = 0; l < nlags; ++l)
      if ((bits[(size_t)(l >> 5) * nf + (k - f0)] >> (l & 31)) & 1u) {
        const int j = iqrm_partner(k, lags[l], Csub);
        if (j >= g0 && j < g1) atomicAdd(&cast[j - g0], 1);
      }
  __syncthreads();
  uint8_t *mrow = masks + (size_t)w * nchan;
  // channels outside the IQRM range are 0: the first block writes those
  // below it, the last block those above it
  const int cl = blk == 0 ? 0 : c0 + b0;
  const int ch = blk == nblk - 1 ? nchan : c0 + b1;
  for (int c = cl + threadIdx.x; c < ch; c += blockDim.x) {
    const int k = c - c0;
    uint8_t m = 0;
    if (k >= 0 && k < Csub) {
      int recv = 0;
      for (int q = 0; q < nw; ++q) recv += __builtin_popcount(bits[(size_t)q * nf + (k - f0)]);
      if (recv > 0)
        for (int l = 0; l < nlags; ++l)
          if (((bits[(size_t)(l >> 5) * nf + (k - f0)] >> (l & 31)) & 1u) &&
              cast[iqrm_partner(k, lags[l], Csub) - g0] < recv) {
            m = 1;
            break;
          }
    }
    mrow[c] = m;
  }
}

__global__ __launch_bounds__(1024) void k_iqrm_merge(MergeArgs a) {
  merge_body(a, blockIdx.x, gridDim.x);
}

// OR of the masks of the windows covering each segment.
__device__ __forceinline__ void segmask_body(const SegmaskArgs &a, uint32_t bid, uint32_t) {
  const uint8_t *__restrict__ masks = a.masks;
  const Seg *__restrict__ segs = a.segs;
  const int nseg = a.nseg, nchan = a.nchan;
  uint8_t *__restrict__ segmask = a.segmask;
  const size_t idx = (size_t)bid * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nseg * nchan) return;
  const int s = (int)(idx / nchan), c = (int)(idx % nchan);
  uint8_t m = 0;
  for (uint32_t w = segs[s].w0; w < segs[s].w1; ++w) m |= masks[(size_t)w * nchan + c];
  segmask[idx] = m;
}

__global__ __launch_bounds__(256) void k_segmask(SegmaskArgs a) {
  segmask_body(a, blockIdx.x, gridDim.x);
}

// --------------------------------------------- clean + transpose pass --
// Per row block: IQRM window mask, zero-DM (src/rfi/rfimarker.cu:172-254:
// mean of the non-zero values over [zc0,zc1), x = clamp(x - mean, 0, vmax)),
// static bad channels (src/rfi/rfimarker.cu:9-30), then a transpose through
// LDS into the channel-major buffer the dedispersion kernels read.
template <typename T>
__device__ __forceinline__ T zdm_sub(T v, float m) {
  if constexpr (std::is_same<T, float>::value) {
    return v - m;
  } else {
    float r = (float)v - m;
    if (r < 0.0f) r = 0.0f;
    const float vmax = (float)Tr<T>::vmax;
    if (r > vmax) r = vmax;
    if (r >= 4294967296.0f) return (T)0xFFFFFFFFu;
    return (T)(uint32_t)r;
  }
}

template <typename T>
__device__ __forceinline__ void clean_body(const CleanArgs &a, uint32_t bid, uint32_t) {
  __shared__ float m_s[kCleanRows];
  __shared__ T tile[64][kCleanRows + 4];
  const RowItem it = a.items[bid];
  const int nrows = (int)(it.r1 - it.r0);
  const T *__restrict__ x = (const T *)a.x;
  T *__restrict__ xT = (T *)a.xT;
  const uint8_t *smask = a.segmask ? a.segmask + (size_t)it.seg * a.nchan : nullptr;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (a.zdm) {
    for (int rr = wave; rr < nrows; rr += 4) {
      const T *row = x + (size_t)(it.r0 + rr) * a.nchan;
      if constexpr (std::is_same<T, float>::value) {
        double S = 0.0;
        unsigned n = 0;
        for (int c = a.zc0 + lane; c < a.zc1; c += 64) {
          float v = row[c];
          if (smask && smask[c]) v = 0.0f;
          if (v != 0.0f) { S += (double)v; ++n; }
        }
        for (int o = 32; o > 0; o >>= 1) {
          S += __shfl_xor(S, o);
          n += __shfl_xor(n, o);
        }
        if (lane == 0) m_s[rr] = n ? (float)(S / (double)n) : 0.0f;
      } else {
        uint64_t S = 0;
        unsigned n = 0;
        for (int c = a.zc0 + lane; c < a.zc1; c += 64) {
          T v = row[c];
          if (smask && smask[c]) v = 0;
          if (v != 0) { S += (uint64_t)v; ++n; }
        }
        for (int o = 32; o > 0; o >>= 1) {
          S += __shfl_xor(S, o);
          n += __shfl_xor(n, o);
        }
        if (lane == 0) m_s[rr] = n ? (float)S / (float)n : 0.0f;
      }
    }
    __syncthreads();
  }
  // channels past the cut are never read at these rows (item.pad)
  const int cend = a.cut_in_pad ? min(a.dd_c1, a.dd_c0 + (int)it.pad) : a.dd_c1;
  for (int cb = a.dd_c0; cb < cend; cb += 64) {
    const int c = cb + lane;
    for (int rr = wave; rr < kCleanRows; rr += 4) {
      T v = 0;
      if (c < cend && rr < nrows) {
        v = x[(size_t)(it.r0 + rr) * a.nchan + c];
        if (smask && smask[c]) v = 0;
        if (a.zdm && c >= a.zc0 && c < a.zc1) v = zdm_sub<T>(v, m_s[rr]);
        if (a.stat_mask && a.stat_mask[c]) v = 0;
      }
      tile[lane][rr] = v;
    }
    __syncthreads();
    for (int k = wave; k < 64; k += 4) {
      const int cc = cb + k;
      if (cc < cend && lane < nrows)
        xT[(size_t)(cc - a.dd_c0) * a.Lp + it.r0 + lane] = tile[k][lane];
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_clean(CleanArgs a) {
  clean_body<T>(a, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------- dedispersion --
// u8 data: each wave owns one DM row and 1024 consecutive outputs (16 per
// lane); per channel every lane loads 16 consecutive bytes of the channel-
// major buffer at its delay (one unaligned 16-byte load; the delay is
// wave-uniform and comes from a scalar load).  The bytes are summed as
// packed 16-bit lanes: S += w accumulates the raw words, H += (w>>8)&0x00FF00FF
// the odd bytes; after at most 256 channels the even-byte sums are
// S - (H<<8).  That is two adds per four samples (three with the byte
// extraction), exact because no 16-bit lane can exceed 256*255.
__device__ __forceinline__ uint32_t ab(uint32_t h, uint32_t l, uint32_t r) {
  return __builtin_amdgcn_alignbyte(h, l, r);
}

// Channels [c0, c1) of one DM row, 16 outputs per lane, into packed 16-bit
// lane sums S/H.  Per channel every lane issues one 16-byte load at the
// delay rounded down to a dword (a dword-aligned lane load costs the L1 no
// more than a 16-byte-aligned one, a byte-misaligned one about one access
// per lane); the dword after it is the right neighbour's first dword (DPP
// wave shift by one lane), and the row's delay mod 4 (wave-uniform, in an
// SGPR) is the byte realignment: no branch on the delay.  Lane 63 has no
// right neighbour: it only loads, for lane 62, and its sums are discarded (a
// wave keeps 63 x 16 outputs).  Eight loads are issued before any is used.  SYNC: the workgroup's rows, which walk the same
// channels, meet at a barrier every eight channels, so that their
// overlapping windows are fetched into the L1 once.  O32: the channel-major
// buffer is below 4 GiB, so raw buffer loads take the row offset in an SGPR.
template <bool O32, bool SYNC>
__device__ __forceinline__ void dd_q_channels(const uint8_t *__restrict__ tb, size_t Lp,
                                              uint32_t loff, const int32_t *__restrict__ drow,
                                              int c0, int c1, uint32_t (&S)[4],
                                              uint32_t (&H)[4], bool live = true) {
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)tb, (short)0, -1, 0x00020000);
  auto load = [&](int c, int d, uint4 &lo) {
    if (O32) {
      const uint32_t off = (uint32_t)c * (uint32_t)Lp + ((uint32_t)d & ~3u);
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)loff, (int)off, 0);
      lo = make_uint4(v[0], v[1], v[2], v[3]);
    } else {
      const uint8_t *rb = tb + (size_t)c * Lp + (d & ~3);
      __builtin_memcpy(&lo, __builtin_assume_aligned(rb + loff, 4), 16);
    }
  };
  // (bound_ctrl: lane 63 reads zero)
  auto dpp = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xF, 0xF, true); };
  auto add = [&](const uint4 lo, const uint32_t r) {
    const uint32_t h0 = dpp(lo.x);
    const uint4 w =
        make_uint4(ab(lo.y, lo.x, r), ab(lo.z, lo.y, r), ab(lo.w, lo.z, r), ab(h0, lo.w, r));
    S[0] += w.x; S[1] += w.y; S[2] += w.z; S[3] += w.w;
    H[0] += odd_bytes(w.x);
    H[1] += odd_bytes(w.y);
    H[2] += odd_bytes(w.z);
    H[3] += odd_bytes(w.w);
  };
  // the delay table is read-only for the kernel's lifetime: constant address
  // space, so the wave-uniform reads are scalar loads
  typedef __attribute__((address_space(4))) const int32_t cint32;
  cint32 *dc = (cint32 *)drow;
  int c = c0;
  // a wave past the last DM row (live false, wave-uniform) only meets the
  // others at the barriers
  if (SYNC && !live) {
    for (; c + 8 <= c1; c += 8) __builtin_amdgcn_s_barrier();
    return;
  }
  for (; c + 8 <= c1; c += 8) {
    uint4 lo[8];
    uint32_t sh[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int d = dc[c + u];
      load(c + u, d, lo[u]);
      sh[u] = (uint32_t)d & 3u;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) add(lo[u], sh[u]);
    if (SYNC) __builtin_amdgcn_s_barrier();
  }
  for (; c < c1; ++c) {
    const int d = dc[c];
    uint4 lo;
    load(c, d, lo);
    add(lo, (uint32_t)d & 3u);
  }
}

// u8 dedispersion, aligned loads: a workgroup of kDdRowsQ waves owns
// kDdRowsQ consecutive DM rows of one channel group and one tile of
// kDdTileQ outputs; the rows walk the group's channels in step (see
// dd_q_channels).  Channel blocks of <= 256 channels are flushed from the
// packed lanes into 32-bit totals.
template <bool O32>
__device__ __forceinline__ void dd_u8_q_body(const DdArgs &a, uint32_t lid, uint32_t) {
  const int nrowblk = (a.nrows + kDdRowsQ - 1) / kDdRowsQ;
  int rowblk, tile, grp;
  if (a.worklist) {
    const int4 wi = a.worklist[lid];
    // workgroup-uniform: in SGPRs (the buffer descriptor is built from them)
    grp = __builtin_amdgcn_readfirstlane(wi.x);
    tile = __builtin_amdgcn_readfirstlane(wi.y);
    rowblk = __builtin_amdgcn_readfirstlane(wi.z);
  } else {
    rowblk = (int)(lid % (uint32_t)nrowblk);
    const uint32_t rest = lid / (uint32_t)nrowblk;
    tile = (int)(rest % (uint32_t)a.ntiles);
    grp = (int)(rest / (uint32_t)a.ntiles);
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // every wave runs to the end (the rows meet at barriers): rows past the
  // last compute row nrows - 1 again and store nothing
  const int row_w = rowblk * kDdRowsQ + wave;
  const bool row_ok = row_w < a.nrows;
  const int row = row_ok ? row_w : a.nrows - 1;
  const int ca = grp * a.grp_w;
  const int cb = min(ca + a.grp_w, a.nch);
  const int blk_w = a.ngrp == 1 ? 256 : a.grp_w;
  const uint8_t *__restrict__ tb = (const uint8_t *)a.xT + (size_t)tile * kDdTileQ;
  const uint32_t loff = (uint32_t)lane * kDdLanesOut;
  const int32_t *__restrict__ drow = a.delays + (size_t)row * a.nch;
  uint32_t tot[16];
  if (cb - ca <= blk_w) {
    // one block (a subband): the totals are not live during the channel loop
    uint32_t S[4] = {0, 0, 0, 0}, H[4] = {0, 0, 0, 0};
    dd_q_channels<O32, true>(tb, a.Lp, loff, drow, ca, cb, S, H, row_ok);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t L = S[k] - (H[k] << 8);
      tot[4 * k + 0] = L & 0xFFFFu;
      tot[4 * k + 1] = H[k] & 0xFFFFu;
      tot[4 * k + 2] = L >> 16;
      tot[4 * k + 3] = H[k] >> 16;
    }
  } else {
    // several blocks (the direct method over more than 256 channels): the
    // running totals wait in LDS ([16][threads], launched with kDdTotLds
    // bytes), which keeps them out of the registers the channel loop needs
    extern __shared__ uint32_t dd_tot[];
    uint32_t *tl = dd_tot + threadIdx.x;
    constexpr int NT = 64 * kDdRowsQ;
#pragma unroll
    for (int i = 0; i < 16; ++i) tl[i * NT] = 0;
    for (int b0 = ca; b0 < cb; b0 += blk_w) {
      uint32_t S[4] = {0, 0, 0, 0}, H[4] = {0, 0, 0, 0};
      dd_q_channels<O32, true>(tb, a.Lp, loff, drow, b0, min(b0 + blk_w, cb), S, H, row_ok);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t L = S[k] - (H[k] << 8);
        tl[(4 * k + 0) * NT] += L & 0xFFFFu;
        tl[(4 * k + 1) * NT] += H[k] & 0xFFFFu;
        tl[(4 * k + 2) * NT] += L >> 16;
        tl[(4 * k + 3) * NT] += H[k] >> 16;
      }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) tot[i] = tl[i * NT];
  }
  if (lane == 63 || !row_ok) return;
  const size_t t = (size_t)tile * kDdTileQ + (size_t)lane * kDdLanesOut;
  const size_t orow = ((size_t)row * a.ngrp + grp) * a.ostride;
  if (a.out16) {
    uint16_t *o = (uint16_t *)a.out + orow + t;
    uint4 v0 = make_uint4(tot[0] | (tot[1] << 16), tot[2] | (tot[3] << 16),
                          tot[4] | (tot[5] << 16), tot[6] | (tot[7] << 16));
    uint4 v1 = make_uint4(tot[8] | (tot[9] << 16), tot[10] | (tot[11] << 16),
                          tot[12] | (tot[13] << 16), tot[14] | (tot[15] << 16));
    if (t + 16 <= a.nvalid) {
      __builtin_memcpy(o, &v0, 16);
      __builtin_memcpy(o + 8, &v1, 16);
    } else {
      for (int i = 0; i < 16; ++i)
        if (t + i < a.nvalid) o[i] = (uint16_t)tot[i];
    }
  } else {
    uint32_t *o = (uint32_t *)a.out + orow + t;
    if (t + 16 <= a.nvalid) {
      if (a.ngrp == 1) {
        // the direct method's output is the plane, not read again on the
        // device: streaming stores
        typedef uint32_t v4u_any __attribute__((ext_vector_type(4), aligned(4)));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const v4u_any v = {tot[4 * k], tot[4 * k + 1], tot[4 * k + 2], tot[4 * k + 3]};
          __builtin_nontemporal_store(v, (v4u_any *)(o + 4 * k));
        }
      } else {
        __builtin_memcpy(o, &tot[0], 64);
      }
    } else {
      for (int i = 0; i < 16; ++i)
        if (t + i < a.nvalid) o[i] = tot[i];
    }
  }
}

// eight waves per SIMD (64 VGPRs): four workgroups per CU hide more of the
// L1 miss latency than three
__global__ __launch_bounds__(64 * kDdRowsQ) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_dd_u8_q(DdArgs a) {
  const uint32_t lid = xcd_remap(blockIdx.x, gridDim.x);
  if ((uint64_t)a.nch * a.Lp <= 0xFFFFFFFFull) dd_u8_q_body<true>(a, lid, gridDim.x);
  else dd_u8_q_body<false>(a, lid, gridDim.x);
}

// float32 dedispersion (direct, or subband partial sums): a workgroup of
// kDdRowsQ waves owns kDdRowsQ consecutive DM rows of one channel group and
// one tile of kDdTile outputs, the rows walking the group's channels in step
// (a barrier every 8 channels, so overlapping windows are fetched into the L1
// once, as in k_dd_u8_q).  Load instruction i (0..3) of lane l reads the 4
// floats at 256 i + 4 l of the row's window, so each instruction reads 1 KB
// contiguous; sums in FP64, stored as float32.  Work list order and XCD
// placement as for u8.
template <bool O32>
__device__ __forceinline__ void dd_f32_q_body(const DdArgs &a, uint32_t lid, uint32_t) {
  const int nrowblk = (a.nrows + kDdRowsQ - 1) / kDdRowsQ;
  int rowblk, tile, grp;
  if (a.worklist) {
    const int4 wi = a.worklist[lid];
    grp = __builtin_amdgcn_readfirstlane(wi.x);
    tile = __builtin_amdgcn_readfirstlane(wi.y);
    rowblk = __builtin_amdgcn_readfirstlane(wi.z);
  } else {
    rowblk = (int)(lid % (uint32_t)nrowblk);
    const uint32_t rest = lid / (uint32_t)nrowblk;
    tile = (int)(rest % (uint32_t)a.ntiles);
    grp = (int)(rest / (uint32_t)a.ntiles);
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // every wave runs to the end (the rows meet at barriers): rows past the
  // last compute row nrows - 1 again and store nothing
  const int row_w = rowblk * kDdRowsQ + wave;
  const bool row_ok = row_w < a.nrows;
  const int row = row_ok ? row_w : a.nrows - 1;
  const int ca = grp * a.grp_w;
  const int cb = min(ca + a.grp_w, a.nch);
  const float *__restrict__ tb = (const float *)a.xT + (size_t)tile * kDdTile;
  typedef __attribute__((address_space(4))) const int32_t cint32;
  cint32 *dc = (cint32 *)(a.delays + (size_t)row * a.nch);
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)tb, (short)0, -1, 0x00020000);
  const uint32_t loff = (uint32_t)lane * 16u;   // bytes
  auto load = [&](int c, int d, int i) {
    if (O32) {
      const uint32_t off = ((uint32_t)c * (uint32_t)a.Lp + (uint32_t)d + 256u * (uint32_t)i) * 4u;
      const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)loff, (int)off, 0);
      return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                         __uint_as_float(v[3]));
    } else {
      float4 v;
      __builtin_memcpy(&v, tb + (size_t)c * a.Lp + d + 256 * i + 4 * lane, 16);
      return v;
    }
  };
  double acc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0;
  auto add = [&](const float4 (&v)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc[4 * i + 0] += (double)v[i].x;
      acc[4 * i + 1] += (double)v[i].y;
      acc[4 * i + 2] += (double)v[i].z;
      acc[4 * i + 3] += (double)v[i].w;
    }
  };
  int c = ca;
  if (!row_ok) {
    // past the last DM row: only meet the other rows at the barriers
    for (; c + 8 <= cb; c += 8) __builtin_amdgcn_s_barrier();
    return;
  }
  for (; c + 8 <= cb; c += 8) {
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
      float4 v0[4], v1[4];
      const int d0 = dc[c + u], d1 = dc[c + u + 1];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v0[i] = load(c + u, d0, i);
        v1[i] = load(c + u + 1, d1, i);
      }
      add(v0);
      add(v1);
    }
    __builtin_amdgcn_s_barrier();
  }
  for (; c < cb; ++c) {
    float4 v[4];
    const int d = dc[c];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = load(c, d, i);
    add(v);
  }
  if (!row_ok) return;
  const size_t t0 = (size_t)tile * kDdTile;
  float *o = (float *)a.out + ((size_t)row * a.ngrp + grp) * a.ostride + t0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const size_t t = (size_t)256 * i + 4 * (size_t)lane;
    const float4 w = make_float4((float)acc[4 * i], (float)acc[4 * i + 1], (float)acc[4 * i + 2],
                                 (float)acc[4 * i + 3]);
    if (t0 + t + 4 <= a.nvalid) {
      __builtin_memcpy(o + t, &w, 16);
    } else {
      const float wv[4] = {w.x, w.y, w.z, w.w};
      for (int k = 0; k < 4; ++k)
        if (t0 + t + k < a.nvalid) o[t + k] = wv[k];
    }
  }
}

__global__ __launch_bounds__(64 * kDdRowsQ) void k_dd_f32_q(DdArgs a) {
  const uint32_t lid = xcd_remap(blockIdx.x, gridDim.x);
  if ((uint64_t)a.nch * a.Lp * 4 <= 0xFFFFFFFFull) dd_f32_q_body<true>(a, lid, gridDim.x);
  else dd_f32_q_body<false>(a, lid, gridDim.x);
}

__device__ __forceinline__ void dd_u8_body(const DdArgs &a, uint32_t lid, uint32_t) {
  const int nrowblk = (a.nrows + kDdRowsPerBlock - 1) / kDdRowsPerBlock;
  int rowblk, tile, grp;
  if (a.worklist) {
    // host-sorted order: workgroups that read overlapping input run together
    const int4 wi = a.worklist[lid];
    grp = wi.x;
    tile = wi.y;
    rowblk = wi.z;
  } else {
    rowblk = (int)(lid % (uint32_t)nrowblk);
    const uint32_t rest = lid / (uint32_t)nrowblk;
    tile = (int)(rest % (uint32_t)a.ntiles);
    grp = (int)(rest / (uint32_t)a.ntiles);
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = rowblk * kDdRowsPerBlock + wave;
  if (row >= a.nrows) return;
  const int ca = grp * a.grp_w;
  const int cb = min(ca + a.grp_w, a.nch);
  const size_t t = (size_t)tile * kDdTile + (size_t)lane * kDdLanesOut;
  // uniform (scalar) part of the address: tile start; per-lane part: 32-bit offset
  const uint8_t *__restrict__ tbase = (const uint8_t *)a.xT + (size_t)tile * kDdTile;
  const uint32_t loff = (uint32_t)lane * kDdLanesOut;
  const int32_t *__restrict__ drow = a.delays + (size_t)row * a.nch;

  uint32_t tot[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) tot[i] = 0;

  for (int c0 = ca; c0 < cb; c0 += 256) {
    const int c1 = min(c0 + 256, cb);
    uint32_t S0 = 0, S1 = 0, S2 = 0, S3 = 0, H0 = 0, H1 = 0, H2 = 0, H3 = 0;
#pragma unroll 8
    for (int c = c0; c < c1; ++c) {
      // Dword-aligned loads (16 bytes + the next dword) and a byte
      // realignment by the wave-uniform delay remainder: a byte-misaligned
      // 16-byte lane load costs the L1 about one access per lane.
      const int d = drow[c];
      const uint8_t *rp = tbase + (size_t)c * a.Lp + (d & ~3) + loff;
      uint4 lo;
      uint32_t nx;
      __builtin_memcpy(&lo, __builtin_assume_aligned(rp, 4), 16);
      __builtin_memcpy(&nx, __builtin_assume_aligned(rp + 16, 4), 4);
      const uint32_t r = (uint32_t)(d & 3);
      const uint4 w = make_uint4(ab(lo.y, lo.x, r), ab(lo.z, lo.y, r), ab(lo.w, lo.z, r), ab(nx, lo.w, r));
      S0 += w.x; S1 += w.y; S2 += w.z; S3 += w.w;
      H0 += odd_bytes(w.x);
      H1 += odd_bytes(w.y);
      H2 += odd_bytes(w.z);
      H3 += odd_bytes(w.w);
    }
    const uint32_t L0 = S0 - (H0 << 8), L1 = S1 - (H1 << 8);
    const uint32_t L2 = S2 - (H2 << 8), L3 = S3 - (H3 << 8);
    tot[0] += L0 & 0xFFFFu;  tot[1] += H0 & 0xFFFFu;  tot[2] += L0 >> 16;  tot[3] += H0 >> 16;
    tot[4] += L1 & 0xFFFFu;  tot[5] += H1 & 0xFFFFu;  tot[6] += L1 >> 16;  tot[7] += H1 >> 16;
    tot[8] += L2 & 0xFFFFu;  tot[9] += H2 & 0xFFFFu;  tot[10] += L2 >> 16; tot[11] += H2 >> 16;
    tot[12] += L3 & 0xFFFFu; tot[13] += H3 & 0xFFFFu; tot[14] += L3 >> 16; tot[15] += H3 >> 16;
  }

  const size_t orow = ((size_t)row * a.ngrp + grp) * a.ostride;
  if (a.out16) {
    uint16_t *o = (uint16_t *)a.out + orow + t;
    uint4 v0 = make_uint4(tot[0] | (tot[1] << 16), tot[2] | (tot[3] << 16),
                          tot[4] | (tot[5] << 16), tot[6] | (tot[7] << 16));
    uint4 v1 = make_uint4(tot[8] | (tot[9] << 16), tot[10] | (tot[11] << 16),
                          tot[12] | (tot[13] << 16), tot[14] | (tot[15] << 16));
    if (t + 16 <= a.nvalid) {
      __builtin_memcpy(o, &v0, 16);
      __builtin_memcpy(o + 8, &v1, 16);
    } else {
      for (int i = 0; i < 16; ++i)
        if (t + i < a.nvalid) o[i] = (uint16_t)tot[i];
    }
  } else {
    uint32_t *o = (uint32_t *)a.out + orow + t;
    if (t + 16 <= a.nvalid) {
      __builtin_memcpy(o, &tot[0], 64);
    } else {
      for (int i = 0; i < 16; ++i)
        if (t + i < a.nvalid) o[i] = tot[i];
    }
  }
}

__global__ __launch_bounds__(256) void k_dd_u8(DdArgs a) {
  dd_u8_body(a, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
}

// u8 dedispersion with the delay windows staged in LDS.  A workgroup owns JB
// consecutive DM rows (one per wave) and 1024 outputs of one channel group.
// For each channel the union of the JB rows' windows -- from the first row's
// delay (the smallest; delays grow with DM) to the last row's plus 1024 -- is
// loaded once into LDS (16 bytes per thread, double-buffered, one barrier per
// channel) and every wave reads its 16 bytes per lane from there.  The
// packed-lane accumulation is that of k_dd_u8.
template <int JB>
__global__ __launch_bounds__(JB * 64) void k_dd_u8_lds(DdArgs a) {
  extern __shared__ uint4 win_lds[];
  const int nrowblk = (a.nrows + JB - 1) / JB;
  const uint32_t lid = xcd_remap(blockIdx.x, gridDim.x);
  int rowblk, tile, grp;
  if (a.worklist) {
    const int4 wi = a.worklist[lid];
    grp = wi.x;
    tile = wi.y;
    rowblk = wi.z;
  } else {
    rowblk = (int)(lid % (uint32_t)nrowblk);
    const uint32_t rest = lid / (uint32_t)nrowblk;
    tile = (int)(rest % (uint32_t)a.ntiles);
    grp = (int)(rest / (uint32_t)a.ntiles);
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row0 = rowblk * JB;
  const int row = row0 + wave;
  const bool rvalid = row < a.nrows;
  const int ca = grp * a.grp_w;
  const int cb = min(ca + a.grp_w, a.nch);
  const size_t t = (size_t)tile * kDdTile;
  const uint8_t *__restrict__ tbase = (const uint8_t *)a.xT + t;
  const int32_t *__restrict__ d0row = a.delays + (size_t)row0 * a.nch;
  const int32_t *__restrict__ drow = a.delays + (size_t)(rvalid ? row : row0) * a.nch;
  const int nchunk = a.lds_win >> 4;           // 16-byte chunks per window
  const uint8_t *lbytes = (const uint8_t *)win_lds;
  // Every thread loads and stores one 16-byte chunk per channel (threads past
  // the window re-read its last chunk and write LDS nobody reads), and past
  // the group's last channel the loads repeat that channel: no branch around
  // the loads, so the wait before each LDS store covers only the oldest load.
  const int chunk = min((int)threadIdx.x, nchunk - 1);
  const uint8_t *gsrc = tbase + 16 * chunk;
  const int wstride_all = blockDim.x;          // uint4 per LDS buffer
  auto chan_src = [&](int x) {
    const int xc = min(x, cb - 1);
    // 16-byte aligned window start: aligned lane loads (an unaligned
    // 16-byte lane load costs about one L1 access per lane)
    return gsrc + (size_t)xc * a.Lp + (d0row[xc] & ~15);
  };

  // Register pipeline: channel x's window chunk lives in pre[(x - ca) % PD]
  // from the load issued PD channels ahead until it is stored to LDS.
  constexpr int PD = 4;
  uint4 pre[PD];
#pragma unroll
  for (int k = 0; k < PD; ++k) pre[k] = ldu4(chan_src(ca + k));
  win_lds[threadIdx.x] = pre[0];
  pre[0] = ldu4(chan_src(ca + PD));
  __syncthreads();

  uint32_t tot[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) tot[i] = 0;
  uint32_t S0 = 0, S1 = 0, S2 = 0, S3 = 0, H0 = 0, H1 = 0, H2 = 0, H3 = 0;
  int since_flush = 0;
  // The trip count is padded to a multiple of PD (padding channels only load
  // and synchronise) so the unrolled body has no exit.
  const int cpad = ca + (cb - ca + PD - 1) / PD * PD;
  for (int cc = ca; cc < cpad; cc += PD) {
#pragma unroll
    for (int k = 0; k < PD; ++k) {
      const int c = cc + k;
      const int nb = (c - ca + 1) & 1;   // buffer of channel c+1
      win_lds[nb * wstride_all + threadIdx.x] = pre[(k + 1) % PD];
      pre[(k + 1) % PD] = ldu4(chan_src(c + 1 + PD));
      if (rvalid && c < cb) {
        const int off = drow[c] - (d0row[c] & ~15);
        uint4 w;
        __builtin_memcpy(&w, lbytes + (size_t)(nb ^ 1) * wstride_all * 16 + off + 16 * lane, 16);
        S0 += w.x; S1 += w.y; S2 += w.z; S3 += w.w;
        H0 += odd_bytes(w.x);
        H1 += odd_bytes(w.y);
        H2 += odd_bytes(w.z);
        H3 += odd_bytes(w.w);
      }
      __syncthreads();
      if (++since_flush == 256 || c + 1 == cb) {
        since_flush = 0;
        const uint32_t L0 = S0 - (H0 << 8), L1 = S1 - (H1 << 8);
        const uint32_t L2 = S2 - (H2 << 8), L3 = S3 - (H3 << 8);
        tot[0] += L0 & 0xFFFFu;  tot[1] += H0 & 0xFFFFu;  tot[2] += L0 >> 16;  tot[3] += H0 >> 16;
        tot[4] += L1 & 0xFFFFu;  tot[5] += H1 & 0xFFFFu;  tot[6] += L1 >> 16;  tot[7] += H1 >> 16;
        tot[8] += L2 & 0xFFFFu;  tot[9] += H2 & 0xFFFFu;  tot[10] += L2 >> 16; tot[11] += H2 >> 16;
        tot[12] += L3 & 0xFFFFu; tot[13] += H3 & 0xFFFFu; tot[14] += L3 >> 16; tot[15] += H3 >> 16;
        S0 = S1 = S2 = S3 = H0 = H1 = H2 = H3 = 0;
      }
    }
  }
  if (!rvalid) return;
  const size_t to = t + (size_t)lane * kDdLanesOut;
  const size_t orow = ((size_t)row * a.ngrp + grp) * a.ostride;
  if (a.out16) {
    uint16_t *o = (uint16_t *)a.out + orow + to;
    uint4 v0 = make_uint4(tot[0] | (tot[1] << 16), tot[2] | (tot[3] << 16),
                          tot[4] | (tot[5] << 16), tot[6] | (tot[7] << 16));
    uint4 v1 = make_uint4(tot[8] | (tot[9] << 16), tot[10] | (tot[11] << 16),
                          tot[12] | (tot[13] << 16), tot[14] | (tot[15] << 16));
    if (to + 16 <= a.nvalid) {
      __builtin_memcpy(o, &v0, 16);
      __builtin_memcpy(o + 8, &v1, 16);
    } else {
      for (int i = 0; i < 16; ++i)
        if (to + i < a.nvalid) o[i] = (uint16_t)tot[i];
    }
  } else {
    uint32_t *o = (uint32_t *)a.out + orow + to;
    if (to + 16 <= a.nvalid) {
      __builtin_memcpy(o, &tot[0], 64);
    } else {
      for (int i = 0; i < 16; ++i)
        if (to + i < a.nvalid) o[i] = tot[i];
    }
  }
}

// Element-wise dedispersion for u16/u32/f32 data (accumulator Acc, output
// OutT), same tiling as k_dd_u8.
template <typename T, typename Acc, typename OutT>
__device__ __forceinline__ void dd_generic_body(const DdArgs &a, uint32_t lid, uint32_t) {
  const int nrowblk = (a.nrows + kDdRowsPerBlock - 1) / kDdRowsPerBlock;
  const int rowblk = (int)(lid % (uint32_t)nrowblk);
  const uint32_t rest = lid / (uint32_t)nrowblk;
  const int tile = (int)(rest % (uint32_t)a.ntiles);
  const int grp = (int)(rest / (uint32_t)a.ntiles);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int row = rowblk * kDdRowsPerBlock + wave;
  if (row >= a.nrows) return;
  const int ca = grp * a.grp_w;
  const int cb = min(ca + a.grp_w, a.nch);
  const size_t t = (size_t)tile * kDdTile + (size_t)lane * kDdLanesOut;
  const T *__restrict__ base = (const T *)a.xT + t;
  const int32_t *__restrict__ drow = a.delays + (size_t)row * a.nch;
  Acc acc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0;
  for (int c = ca; c < cb; ++c) {
    const T *p = base + (size_t)c * a.Lp + drow[c];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] += (Acc)p[i];
  }
  OutT *o = (OutT *)a.out + ((size_t)row * a.ngrp + grp) * a.ostride + t;
  for (int i = 0; i < 16; ++i)
    if (t + i < a.nvalid) o[i] = (OutT)acc[i];
}

template <typename T, typename Acc, typename OutT>
__global__ __launch_bounds__(256) void k_dd_generic(DdArgs a) {
  dd_generic_body<T, Acc, OutT>(a, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
}

// Subband stage 2 for u16 partial sums (src/core/dedispered.cu:449-476):
// out[m][t] = sum_sb inter[j][sb][t + res[m][sb]], j = m / ndm0.  A workgroup
// stages the group's partial sums for its tile (tile + max residual of the
// group) in LDS, then each wave sums the subbands for one fine DM at a time.
__device__ __forceinline__ void stage2_u16_body(const Stage2Args &a, uint32_t lid, uint32_t) {
  extern __shared__ uint16_t s2[];
  const int ntile = (int)((a.nout + a.tile - 1) / a.tile);
  const int ti = (int)(lid % (uint32_t)ntile);
  const int j = (int)(lid / (uint32_t)ntile);
  const size_t t0 = (size_t)ti * a.tile;
  const int ext = a.tile + a.maxres_j[j];
  const uint16_t *__restrict__ src = (const uint16_t *)a.inter + (size_t)j * a.nsb * a.L1 + t0;
  for (int sb = 0; sb < a.nsb; ++sb)
    for (int i = threadIdx.x; i < ext; i += blockDim.x)
      s2[sb * a.lds_len + i] = src[(size_t)sb * a.L1 + i];
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int per = a.tile / 64;  // outputs per lane (tile is a multiple of 64)
  for (int mi = wave; mi < a.ndm0; mi += 4) {
    const int m = j * a.ndm0 + mi;
    if (m >= a.ndm) break;
    const int32_t *__restrict__ rrow = a.res + (size_t)m * a.nsb;
    uint32_t acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0;
    for (int sb = 0; sb < a.nsb; ++sb) {
      const uint16_t *p = s2 + sb * a.lds_len + lane * per + ldc(rrow, sb);
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i < per) acc[i] += p[i];
    }
    uint32_t *o = (uint32_t *)a.out + (size_t)m * a.ostride + t0 + (size_t)lane * per;
    for (int i = 0; i < per; ++i)
      if (t0 + (size_t)lane * per + i < a.nout) o[i] = acc[i];
  }
}

__global__ __launch_bounds__(256) void k_stage2_u16(Stage2Args a) {
  stage2_u16_body(a, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
}

// Stage 2 from partial sums of type In (u16/u32/u64/double), accumulated
// in Acc, straight from HBM/L2.
template <typename In, typename Acc, typename OutT>
__device__ __forceinline__ void stage2_generic_body(const Stage2Args &a, uint32_t bid, uint32_t) {
  const size_t idx = (size_t)bid * blockDim.x + threadIdx.x;
  if (idx >= (size_t)a.ndm * a.nout) return;
  const int m = (int)(idx / a.nout);
  const size_t t = idx % a.nout;
  const int j = m / a.ndm0;
  const In *src = (const In *)a.inter + (size_t)j * a.nsb * a.L1 + t;
  Acc acc = 0;
  for (int sb = 0; sb < a.nsb; ++sb) acc += (Acc)src[(size_t)sb * a.L1 + a.res[(size_t)m * a.nsb + sb]];
  ((OutT *)a.out)[(size_t)m * a.ostride + t] = (OutT)acc;
}

template <typename In, typename Acc, typename OutT>
__global__ __launch_bounds__(256) void k_stage2_generic(Stage2Args a) {
  stage2_generic_body<In, Acc, OutT>(a, blockIdx.x, gridDim.x);
}

// Stage 2 for float32 partial sums: a.sbc subbands at a time are staged in
// LDS (tile + the group's halo, 16-byte loads), each wave keeps the FP64
// sums of up to 4 fine DMs x P outputs per lane in registers across the
// chunks, and stores float32.  One workgroup (512 threads) per (coarse DM,
// tile of 64*P outputs); requires ndm0 <= 32.
template <int P>
__device__ __forceinline__ void stage2_f32c_body(const Stage2Args &a, uint32_t lid, uint32_t) {
  extern __shared__ uint4 s2f[];
  float *s2 = (float *)s2f;
  constexpr int TILE = 64 * P;
  constexpr int MD = 4;
  const int ntile = (int)((a.nout + TILE - 1) / TILE);
  const int ti = (int)(lid % (uint32_t)ntile);
  const int j = (int)(lid / (uint32_t)ntile);
  const size_t t0 = (size_t)ti * TILE;
  const int ext4 = (TILE + a.maxres_j[j] + 3) / 4;     // 16-byte chunks per subband row
  const int lds4 = a.lds_len / 4;
  const float *__restrict__ src = (const float *)a.inter + (size_t)j * a.nsb * a.L1 + t0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  double acc[MD][P];
#pragma unroll
  for (int d = 0; d < MD; ++d)
#pragma unroll
    for (int i = 0; i < P; ++i) acc[d][i] = 0.0;
  for (int sb0 = 0; sb0 < a.nsb; sb0 += a.sbc) {
    const int nsc = min(a.sbc, a.nsb - sb0);
    fill_rows16(s2f, lds4, src + (size_t)sb0 * a.L1, a.L1, 4, nsc, ext4);
    __syncthreads();
#pragma unroll
    for (int d = 0; d < MD; ++d) {
      const int mi = wave + d * nw;
      const int m = j * a.ndm0 + mi;
      if (mi < a.ndm0 && m < a.ndm) {
        const int32_t *__restrict__ rrow = a.res + (size_t)m * a.nsb + sb0;
        for (int sb = 0; sb < nsc; ++sb) {
          const float *p = s2 + sb * a.lds_len + lane + ldc(rrow, sb);   // aligned 4-byte reads
#pragma unroll
          for (int i = 0; i < P; ++i) acc[d][i] += (double)p[64 * i];
        }
      }
    }
    __syncthreads();   // the next chunk overwrites LDS
  }
#pragma unroll
  for (int d = 0; d < MD; ++d) {
    const int mi = wave + d * nw;
    const int m = j * a.ndm0 + mi;
    if (mi >= a.ndm0 || m >= a.ndm) continue;
    float *o = (float *)a.out + (size_t)m * a.ostride + t0 + (size_t)lane;
#pragma unroll
    for (int i = 0; i < P; ++i)
      if (t0 + (size_t)lane + 64 * i < a.nout) o[64 * i] = (float)acc[d][i];
  }
}

template <int P>
__global__ __launch_bounds__(512) void k_stage2_f32c(Stage2Args a) {
  stage2_f32c_body<P>(a, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
}

// --------------------------------------------------- single-DM spectrum --
// include/cpucal.hpp:467-481: every channel of the cleaned data read at its
// delay.  A workgroup moves a 64-channel x 64-sample block through LDS:
// channel rows are read contiguously, time rows written contiguously.
template <typename T>
__global__ __launch_bounds__(256) void k_gather(const T *__restrict__ xT, size_t Lp,
                                                const int32_t *__restrict__ delay, int nch,
                                                size_t t0, size_t tlen, T *__restrict__ out) {
  __shared__ T tile[64][65];
  const int cb = blockIdx.y * 64;
  const size_t tb = (size_t)blockIdx.x * 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = w; k < 64; k += 4) {
    const int c = cb + k;
    T v = 0;
    if (c < nch && tb + lane < tlen) v = xT[(size_t)c * Lp + t0 + tb + lane + (size_t)delay[c]];
    tile[k][lane] = v;
  }
  __syncthreads();
  for (int k = w; k < 64; k += 4) {
    const size_t t = tb + k;
    const int c = cb + lane;
    if (t < tlen && c < nch) out[t * (size_t)nch + c] = tile[lane][k];
  }
}

// ------------------------------------------------------ batched launches --
// One launch runs the same stage of many plans: workgroup g (after the XCD
// remap) belongs to argument block i with start[i] <= g < start[i+1] and runs
// as that block's workgroup g - start[i].
__device__ __forceinline__ int batch_find(const uint32_t *__restrict__ start, int n, uint32_t g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define AF_BATCHED(NAME, ARGS, BOUNDS, CALL)                                            \
  __global__ __launch_bounds__(BOUNDS) void NAME(const ARGS *__restrict__ tab,          \
                                                 const uint32_t *__restrict__ start, int n) { \
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);                                \
    const int i = batch_find(start, n, g);                                              \
    const ARGS a = tab[i];                                                              \
    const uint32_t lid = g - start[i], nb = start[i + 1] - start[i];                    \
    (void)nb;                                                                           \
    CALL;                                                                               \
  }

AF_BATCHED(kb_bin_u8, BinArgs, 256, (bin_body<uint8_t>(a, lid, nb)))
AF_BATCHED(kb_bin_u16, BinArgs, 256, (bin_body<uint16_t>(a, lid, nb)))
AF_BATCHED(kb_bin_u32, BinArgs, 256, (bin_body<uint32_t>(a, lid, nb)))
AF_BATCHED(kb_bin_f32, BinArgs, 256, (bin_body<float>(a, lid, nb)))
AF_BATCHED(kb_stat_u16, StatArgs, 256, (stats_body<uint16_t>(a, lid, nb)))
AF_BATCHED(kb_stat_u32, StatArgs, 256, (stats_body<uint32_t>(a, lid, nb)))
AF_BATCHED(kb_stat_f32, StatArgs, 256, (stats_body<float>(a, lid, nb)))
AF_BATCHED(kb_red_u16, ReduceArgs, 256, (reduce_body<uint16_t>(a, lid, nb)))
AF_BATCHED(kb_red_u32, ReduceArgs, 256, (reduce_body<uint32_t>(a, lid, nb)))
AF_BATCHED(kb_red_f32, ReduceArgs, 256, (reduce_body<float>(a, lid, nb)))
AF_BATCHED(kb_lag_u8, LagArgs, 1024, (lag_body<uint8_t>(a, lid, nb)))
AF_BATCHED(kb_lag_u16, LagArgs, 1024, (lag_body<uint16_t>(a, lid, nb)))
AF_BATCHED(kb_lag_u32, LagArgs, 1024, (lag_body<uint32_t>(a, lid, nb)))
AF_BATCHED(kb_lag_f32, LagArgs, 1024, (lag_body<float>(a, lid, nb)))
AF_BATCHED(kb_merge, MergeArgs, 1024, (merge_body(a, lid, nb)))
AF_BATCHED(kb_segmask, SegmaskArgs, 256, (segmask_body(a, lid, nb)))
AF_BATCHED(kb_clean_u16, CleanArgs, 256, (clean_body<uint16_t>(a, lid, nb)))
AF_BATCHED(kb_clean_u32, CleanArgs, 256, (clean_body<uint32_t>(a, lid, nb)))
AF_BATCHED(kb_clean_f32, CleanArgs, 256, (clean_body<float>(a, lid, nb)))
#define AF_BATCHED_REF(NAME, ARGS, BOUNDS, ATTR, CALL)                                 \
  __global__ __launch_bounds__(BOUNDS) ATTR void NAME(                                    \
      const ARGS *__restrict__ tab, const uint32_t *__restrict__ start, int n) {         \
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);                                \
    const int i = batch_find(start, n, g);                                              \
    const ARGS &a = tab[i];                                                             \
    const uint32_t lid = g - start[i], nb = start[i + 1] - start[i];                    \
    (void)nb;                                                                           \
    CALL;                                                                               \
  }

AF_BATCHED_REF(kb_dd_u8q, DdArgs, 64 * kDdRowsQ, __attribute__((amdgpu_waves_per_eu(8, 8))),
                ((uint64_t)a.nch * a.Lp <= 0xFFFFFFFFull ? dd_u8_q_body<true>(a, lid, nb)
                                                         : dd_u8_q_body<false>(a, lid, nb)))
AF_BATCHED_REF(kb_dd_f32q, DdArgs, 64 * kDdRowsQ, ,
                ((uint64_t)a.nch * a.Lp * 4 <= 0xFFFFFFFFull ? dd_f32_q_body<true>(a, lid, nb)
                                                             : dd_f32_q_body<false>(a, lid, nb)))
AF_BATCHED(kb_dd_u8, DdArgs, 256, (dd_u8_body(a, lid, nb)))
AF_BATCHED(kb_dd_g_u8, DdArgs, 256, (dd_generic_body<uint8_t, uint32_t, uint32_t>(a, lid, nb)))
AF_BATCHED(kb_dd_g_u16, DdArgs, 256, (dd_generic_body<uint16_t, uint32_t, uint32_t>(a, lid, nb)))
AF_BATCHED(kb_dd_g_u32, DdArgs, 256, (dd_generic_body<uint32_t, uint64_t, uint64_t>(a, lid, nb)))
AF_BATCHED(kb_dd_g_f32f, DdArgs, 256, (dd_generic_body<float, double, float>(a, lid, nb)))
AF_BATCHED(kb_dd_g_f32d, DdArgs, 256, (dd_generic_body<float, double, double>(a, lid, nb)))
AF_BATCHED(kb_s2_u16, Stage2Args, 256, (stage2_u16_body(a, lid, nb)))
AF_BATCHED(kb_s2_g_u16, Stage2Args, 256, (stage2_generic_body<uint16_t, uint32_t, uint32_t>(a, lid, nb)))
AF_BATCHED(kb_s2_g_u32, Stage2Args, 256, (stage2_generic_body<uint32_t, uint32_t, uint32_t>(a, lid, nb)))
AF_BATCHED(kb_s2_g_u64, Stage2Args, 256, (stage2_generic_body<uint64_t, uint64_t, uint64_t>(a, lid, nb)))
AF_BATCHED(kb_s2_g_f32, Stage2Args, 256, (stage2_generic_body<float, double, float>(a, lid, nb)))
AF_BATCHED(kb_s2_f32c8, Stage2Args, 512, (stage2_f32c_body<8>(a, lid, nb)))
AF_BATCHED(kb_s2_f32c16, Stage2Args, 512, (stage2_f32c_body<16>(a, lid, nb)))

// ---------------------------------------------------------- launchers --

thread_local std::vector<LaunchDesc> *g_record = nullptr;

static unsigned grid_for(size_t n, unsigned bs, unsigned cap = 1u << 20) {
  size_t g = (n + bs - 1) / bs;
  if (g == 0) g = 1;
  return (unsigned)(g < cap ? g : cap);
}

hipError_t emit_desc(const LaunchDesc &d, hipStream_t st) {
  if (g_record) {
    g_record->push_back(d);
    return hipSuccess;
  }
  return launch_desc(d, st);
}

template <typename A>
static hipError_t emit(int kind, int stage, size_t nblk, unsigned threads, size_t lds, const A &args,
                       hipStream_t st) {
  return emit_args(kind, stage, nblk, threads, lds, args, st);
}

#define AF_DISPATCH_T(dtype, T, ...)                                          \
  switch (dtype) {                                                            \
  case AF_U8: { using T = uint8_t; __VA_ARGS__; break; }                      \
  case AF_U16: { using T = uint16_t; __VA_ARGS__; break; }                    \
  case AF_U32: { using T = uint32_t; __VA_ARGS__; break; }                    \
  case AF_F32: { using T = float; __VA_ARGS__; break; }                       \
  default: return hipErrorInvalidValue;                                       \
  }

#define AF_KIND4(base, dtype)                                                    \
  ((dtype) == AF_U8 ? base##U8 : (dtype) == AF_U16 ? base##U16 : (dtype) == AF_U32 ? base##U32 \
                                                                                 : base##F32)

hipError_t launch_bin(int dtype, const void *in, void *out, size_t nsamp, int nchan,
                      int ds, size_t nbin, hipStream_t st) {
  if (dtype < AF_U8 || dtype > AF_F32) return hipErrorInvalidValue;
  BinArgs a{in, out, nsamp, nbin, nchan, ds};
  return emit(AF_KIND4(KB_BIN_, dtype), 0, grid_for(nbin * (size_t)nchan, 256), 256, 0, a, st);
}

hipError_t launch_seg_stats(int dtype, const void *x, int nchan, int c0, int c1,
                            const RowItem *items, int nitems, Part *part, hipStream_t st) {
  if (dtype == AF_U8)
    return launch_seg_stats_u8((const uint8_t *)x, nchan, c0, c1, items, nitems, (uint2 *)part, st);
  if (dtype < AF_U8 || dtype > AF_F32) return hipErrorInvalidValue;
  const int nchunk = (c1 - c0 + 255) / 256;
  StatArgs a{x, items, part, nchan, c0, c1, nitems * nchunk};
  return emit(AF_KIND4(KB_STAT_, dtype), 1, (size_t)nitems * nchunk, 256, 0, a, st);
}

hipError_t launch_seg_reduce(int dtype, const Part *part, const int *seg_items, int nseg,
                             int Csub, Part *segsum, hipStream_t st) {
  if (dtype == AF_U8)
    return launch_seg_reduce_u8((const uint2 *)part, seg_items, nseg, Csub, segsum, st);
  if (dtype < AF_U8 || dtype > AF_F32) return hipErrorInvalidValue;
  ReduceArgs a{part, seg_items, segsum, nseg, Csub};
  return emit(AF_KIND4(KB_RED_, dtype), 2, ((size_t)nseg * Csub + 255) / 256, 256, 0, a, st);
}

hipError_t launch_iqrm_lags(int dtype, const Part *segsum, const Seg *segs,
                            const uint32_t *wseg, const uint32_t *wt0, const uint32_t *wt1,
                            int nwin, int Csub, int mode, const int *lags, int nlags,
                            float nsigma, uint8_t *flags, size_t lds_bytes, hipStream_t st) {
  (void)segs;
  if (dtype < AF_U8 || dtype > AF_F32) return hipErrorInvalidValue;
  LagArgs a{segsum, wseg, wt0, wt1, lags, flags, Csub, mode, nlags, nsigma};
  return emit(AF_KIND4(KB_LAG_, dtype), 3, (size_t)nwin * nlags, 1024, lds_bytes, a, st);
}

hipError_t launch_iqrm_merge(const uint8_t *flags, const int *lags, int nlags, int nwin,
                             int Csub, int c0, int nchan, uint8_t *masks, int reach,
                             hipStream_t st) {
  MergeArgs a{flags, lags, masks, nlags, Csub, c0, nchan};
  // LDS: counts over the block + R each side, bits over the block + 2R
  // (R = reach, the largest |lag| below Csub)
  const size_t R = (size_t)std::max(0, reach);
  const size_t ncnt = std::min<size_t>(Csub, kMergeBlock + 2 * R);
  const size_t nbit = std::min<size_t>(Csub, kMergeBlock + 4 * R);
  const size_t lds = (ncnt + nbit * ((nlags + 31) / 32)) * sizeof(int);
  const size_t nblk = (Csub + kMergeBlock - 1) / kMergeBlock;
  return emit(KB_MERGE, 4, (size_t)nwin * nblk, 1024, lds, a, st);
}

hipError_t launch_segmask(const uint8_t *masks, const Seg *segs, int nseg, int nchan,
                          uint8_t *segmask, hipStream_t st) {
  SegmaskArgs a{masks, segs, segmask, nseg, nchan};
  return emit(KB_SEGMASK, 5, ((size_t)nseg * nchan + 255) / 256, 256, 0, a, st);
}

hipError_t launch_clean(int dtype, const CleanArgs &a, hipStream_t st) {
  if (dtype == AF_U8 && a.nchan <= kCleanU8MaxChan) return launch_clean_u8(a, st);
  if (dtype < AF_U8 || dtype > AF_F32) return hipErrorInvalidValue;
  return emit(AF_KIND4(KB_CLEAN_, dtype), 6, (size_t)a.nitems, 256, 0, a, st);
}

// dynamic LDS of k_dd_u8_q: the running totals of a channel group wider than
// one 256-channel block (only the direct method has one)
static size_t dd_u8q_lds(const DdArgs &a) {
  return (a.ngrp == 1 && a.nch > 256) ? kDdTotLds : 0;
}

hipError_t launch_dedisperse(int dtype, const DdArgs &a, hipStream_t st) {
  const int nrowblk = (a.nrows + kDdRowsPerBlock - 1) / kDdRowsPerBlock;
  const size_t nblk = (size_t)nrowblk * a.ntiles * a.ngrp;
  if (nblk == 0) return hipSuccess;
  switch (dtype) {
  case AF_U8:
    if (a.lds_rows == 16 && (a.out16 || a.ngrp == 1)) {
      const size_t nblk16 = (size_t)((a.nrows + 15) / 16) * a.ntiles * a.ngrp;
      const size_t lds = 2 * (size_t)1024 * 16;   // two buffers of one 16-byte chunk per thread
      return emit(KB_DD_U8LDS, 7, nblk16, 1024, lds, a, st);
    }
    if (a.u8q && (a.out16 || a.ngrp == 1))
      return emit(KB_DD_U8Q, 7, (size_t)((a.nrows + kDdRowsQ - 1) / kDdRowsQ) * a.ntiles * a.ngrp,
                  64 * kDdRowsQ, dd_u8q_lds(a), a, st);
    if (a.out16 || a.ngrp == 1) return emit(KB_DD_U8, 7, nblk, 256, 0, a, st);
    return emit(KB_DD_G_U8, 7, nblk, 256, 0, a, st);
  case AF_U16: return emit(KB_DD_G_U16, 7, nblk, 256, 0, a, st);
  case AF_U32: return emit(KB_DD_G_U32, 7, nblk, 256, 0, a, st);
  // float32: FP64 accumulation, float32 stores (the output, or the subband
  // partial sums)
  case AF_F32:
    if (a.f32q)
      return emit(KB_DD_F32Q, 7, (size_t)((a.nrows + kDdRowsQ - 1) / kDdRowsQ) * a.ntiles * a.ngrp,
                  64 * kDdRowsQ, 0, a, st);
    return emit(KB_DD_G_F32F, 7, nblk, 256, 0, a, st);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage2(int dtype, const Stage2Args &a, hipStream_t st) {
  if (dtype == AF_U8 && a.inter16 && (a.sboff || a.tile == 512 || a.tile == 1024))
    return launch_stage2_u16v(a, st);   // all subbands in LDS (packed rows), or in chunks
  if (dtype == AF_U8 && a.inter16 && a.tile > 0) {
    const int ntile = (int)((a.nout + a.tile - 1) / a.tile);
    const size_t lds = (size_t)a.nsb * a.lds_len * sizeof(uint16_t);
    return emit(KB_S2_U16, 8, (size_t)ntile * a.ndm_nom, 256, lds, a, st);
  }
  if (dtype == AF_F32 && a.sbc > 0 && (a.tile == 512 || a.tile == 1024)) {
    const int ntile = (int)((a.nout + a.tile - 1) / a.tile);
    const size_t lds = (size_t)a.sbc * a.lds_len * sizeof(float);
    return emit(a.tile == 1024 ? KB_S2_F32C16 : KB_S2_F32C8, 8, (size_t)ntile * a.ndm_nom, 512,
                lds, a, st);
  }
  const size_t n = (size_t)a.ndm * a.nout;
  const size_t g = (n + 255) / 256;
  if (dtype == AF_U8 && a.inter16) return emit(KB_S2_G_U16, 8, g, 256, 0, a, st);
  switch (dtype) {
  case AF_U8:
  case AF_U16: return emit(KB_S2_G_U32, 8, g, 256, 0, a, st);
  case AF_U32: return emit(KB_S2_G_U64, 8, g, 256, 0, a, st);
  case AF_F32: return emit(KB_S2_G_F32, 8, g, 256, 0, a, st);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_gather(int dtype, const void *xT, size_t Lp, const int32_t *delay, int nch,
                         size_t t0, size_t tlen, void *out, hipStream_t st) {
  const size_t gx = (tlen + 63) / 64, gy = ((size_t)nch + 63) / 64;
  if (gx == 0 || gy == 0) return hipSuccess;
  if (gx > 0x7FFFFFFFull || gy > 65535) return hipErrorInvalidValue;
  const dim3 g((unsigned)gx, (unsigned)gy), b(256);
  AF_DISPATCH_T(dtype, T,
                hipLaunchKernelGGL(k_gather<T>, g, b, 0, st, (const T *)xT, Lp, delay, nch, t0,
                                   tlen, (T *)out));
  return hipGetLastError();
}

struct ZeroArgs {
  uint8_t *base;
  size_t pitch, width;
  int nrows;
};

hipError_t launch_zero(uint8_t *base, size_t pitch, size_t width, int nrows, hipStream_t st) {
  if (width == 0 || nrows <= 0) return hipSuccess;
  ZeroArgs a{base, pitch, width, nrows};
  return emit(KB_ZERO, 6, 1, 1, 0, a, st);
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  CopyArgs a{dst, src, bytes};
  return emit(KB_COPY, 9, 1, 1, 0, a, st);
}

// u8 kernels live in af_kernels_u8.hip
hipError_t launch_desc_u8(const LaunchDesc &d, hipStream_t st);
hipError_t launch_batched_u8(int kind, const void *tab, const uint32_t *start, int n,
                             uint32_t total, uint32_t lds, hipStream_t st);

template <typename A>
static const A &argsof(const LaunchDesc &d) {
  return *reinterpret_cast<const A *>(d.args);
}

hipError_t launch_desc(const LaunchDesc &d, hipStream_t st) {
  const dim3 g(d.nblk), b(d.threads);
  switch (d.kind) {
  case KB_BIN_U8: hipLaunchKernelGGL(k_bin<uint8_t>, g, b, d.lds, st, argsof<BinArgs>(d)); break;
  case KB_BIN_U16: hipLaunchKernelGGL(k_bin<uint16_t>, g, b, d.lds, st, argsof<BinArgs>(d)); break;
  case KB_BIN_U32: hipLaunchKernelGGL(k_bin<uint32_t>, g, b, d.lds, st, argsof<BinArgs>(d)); break;
  case KB_BIN_F32: hipLaunchKernelGGL(k_bin<float>, g, b, d.lds, st, argsof<BinArgs>(d)); break;
  case KB_STAT_U16: hipLaunchKernelGGL(k_seg_stats<uint16_t>, g, b, d.lds, st, argsof<StatArgs>(d)); break;
  case KB_STAT_U32: hipLaunchKernelGGL(k_seg_stats<uint32_t>, g, b, d.lds, st, argsof<StatArgs>(d)); break;
  case KB_STAT_F32: hipLaunchKernelGGL(k_seg_stats<float>, g, b, d.lds, st, argsof<StatArgs>(d)); break;
  case KB_RED_U16: hipLaunchKernelGGL(k_seg_reduce<uint16_t>, g, b, d.lds, st, argsof<ReduceArgs>(d)); break;
  case KB_RED_U32: hipLaunchKernelGGL(k_seg_reduce<uint32_t>, g, b, d.lds, st, argsof<ReduceArgs>(d)); break;
  case KB_RED_F32: hipLaunchKernelGGL(k_seg_reduce<float>, g, b, d.lds, st, argsof<ReduceArgs>(d)); break;
  case KB_LAG_U8:
    if (d.lds > 65536)
      (void)hipFuncSetAttribute((const void *)k_iqrm_lag<uint8_t>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)d.lds);
    hipLaunchKernelGGL(k_iqrm_lag<uint8_t>, g, b, d.lds, st, argsof<LagArgs>(d));
    break;
  case KB_LAG_U16:
    if (d.lds > 65536)
      (void)hipFuncSetAttribute((const void *)k_iqrm_lag<uint16_t>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)d.lds);
    hipLaunchKernelGGL(k_iqrm_lag<uint16_t>, g, b, d.lds, st, argsof<LagArgs>(d));
    break;
  case KB_LAG_U32:
    if (d.lds > 65536)
      (void)hipFuncSetAttribute((const void *)k_iqrm_lag<uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)d.lds);
    hipLaunchKernelGGL(k_iqrm_lag<uint32_t>, g, b, d.lds, st, argsof<LagArgs>(d));
    break;
  case KB_LAG_F32:
    if (d.lds > 65536)
      (void)hipFuncSetAttribute((const void *)k_iqrm_lag<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)d.lds);
    hipLaunchKernelGGL(k_iqrm_lag<float>, g, b, d.lds, st, argsof<LagArgs>(d));
    break;
  case KB_MERGE:
    if (d.lds > 65536)
      (void)hipFuncSetAttribute((const void *)k_iqrm_merge, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)d.lds);
    hipLaunchKernelGGL(k_iqrm_merge, g, b, d.lds, st, argsof<MergeArgs>(d));
    break;
  case KB_SEGMASK: hipLaunchKernelGGL(k_segmask, g, b, d.lds, st, argsof<SegmaskArgs>(d)); break;
  case KB_CLEAN_U16: hipLaunchKernelGGL(k_clean<uint16_t>, g, b, d.lds, st, argsof<CleanArgs>(d)); break;
  case KB_CLEAN_U32: hipLaunchKernelGGL(k_clean<uint32_t>, g, b, d.lds, st, argsof<CleanArgs>(d)); break;
  case KB_CLEAN_F32: hipLaunchKernelGGL(k_clean<float>, g, b, d.lds, st, argsof<CleanArgs>(d)); break;
  case KB_DD_U8Q: hipLaunchKernelGGL(k_dd_u8_q, g, b, d.lds, st, argsof<DdArgs>(d)); break;
  case KB_DD_F32Q: hipLaunchKernelGGL(k_dd_f32_q, g, b, d.lds, st, argsof<DdArgs>(d)); break;
  case KB_DD_U8: hipLaunchKernelGGL(k_dd_u8, g, b, d.lds, st, argsof<DdArgs>(d)); break;
  case KB_DD_U8LDS: hipLaunchKernelGGL(k_dd_u8_lds<16>, g, b, d.lds, st, argsof<DdArgs>(d)); break;
  case KB_DD_G_U8:
    hipLaunchKernelGGL((k_dd_generic<uint8_t, uint32_t, uint32_t>), g, b, d.lds, st, argsof<DdArgs>(d));
    break;
  case KB_DD_G_U16:
    hipLaunchKernelGGL((k_dd_generic<uint16_t, uint32_t, uint32_t>), g, b, d.lds, st, argsof<DdArgs>(d));
    break;
  case KB_DD_G_U32:
    hipLaunchKernelGGL((k_dd_generic<uint32_t, uint64_t, uint64_t>), g, b, d.lds, st, argsof<DdArgs>(d));
    break;
  case KB_DD_G_F32F:
    hipLaunchKernelGGL((k_dd_generic<float, double, float>), g, b, d.lds, st, argsof<DdArgs>(d));
    break;
  case KB_DD_G_F32D:
    hipLaunchKernelGGL((k_dd_generic<float, double, double>), g, b, d.lds, st, argsof<DdArgs>(d));
    break;
  case KB_S2_U16: hipLaunchKernelGGL(k_stage2_u16, g, b, d.lds, st, argsof<Stage2Args>(d)); break;
  case KB_S2_G_U16:
    hipLaunchKernelGGL((k_stage2_generic<uint16_t, uint32_t, uint32_t>), g, b, d.lds, st,
                       argsof<Stage2Args>(d));
    break;
  case KB_S2_G_U32:
    hipLaunchKernelGGL((k_stage2_generic<uint32_t, uint32_t, uint32_t>), g, b, d.lds, st,
                       argsof<Stage2Args>(d));
    break;
  case KB_S2_G_U64:
    hipLaunchKernelGGL((k_stage2_generic<uint64_t, uint64_t, uint64_t>), g, b, d.lds, st,
                       argsof<Stage2Args>(d));
    break;
  case KB_S2_G_F32:
    hipLaunchKernelGGL((k_stage2_generic<float, double, float>), g, b, d.lds, st,
                       argsof<Stage2Args>(d));
    break;
  case KB_S2_F32C8:
    (void)hipFuncSetAttribute((const void *)k_stage2_f32c<8>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds);
    hipLaunchKernelGGL(k_stage2_f32c<8>, g, b, d.lds, st, argsof<Stage2Args>(d));
    break;
  case KB_S2_F32C16:
    (void)hipFuncSetAttribute((const void *)k_stage2_f32c<16>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds);
    hipLaunchKernelGGL(k_stage2_f32c<16>, g, b, d.lds, st, argsof<Stage2Args>(d));
    break;
  case KB_COPY: {
    const CopyArgs &c = argsof<CopyArgs>(d);
    return hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToDevice, st);
  }
  case KB_ZERO: {
    const ZeroArgs &z = argsof<ZeroArgs>(d);
    return launch_zero_rows(z.base, z.pitch, z.width, z.nrows, st);
  }
  default:
    return launch_desc_u8(d, st);
  }
  return hipGetLastError();
}

bool batchable(int kind) {
  return kind != KB_DD_U8LDS && kind != KB_COPY && kind != KB_ZERO && kind >= 0 &&
         kind < KB_COUNT;
}

hipError_t launch_batched(int kind, const void *tab, const uint32_t *start, int n,
                          uint32_t total, uint32_t lds, hipStream_t st) {
  const dim3 g(total);
  switch (kind) {
#define AF_BL(K, KERN, TH, A)                                                        \
  case K:                                                                            \
    if (lds > 65536)                                                                 \
      (void)hipFuncSetAttribute((const void *)KERN,                                  \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    hipLaunchKernelGGL(KERN, g, dim3(TH), lds, st, (const A *)tab, start, n);        \
    break;
    AF_BL(KB_BIN_U8, kb_bin_u8, 256, BinArgs)
    AF_BL(KB_BIN_U16, kb_bin_u16, 256, BinArgs)
    AF_BL(KB_BIN_U32, kb_bin_u32, 256, BinArgs)
    AF_BL(KB_BIN_F32, kb_bin_f32, 256, BinArgs)
    AF_BL(KB_STAT_U16, kb_stat_u16, 256, StatArgs)
    AF_BL(KB_STAT_U32, kb_stat_u32, 256, StatArgs)
    AF_BL(KB_STAT_F32, kb_stat_f32, 256, StatArgs)
    AF_BL(KB_RED_U16, kb_red_u16, 256, ReduceArgs)
    AF_BL(KB_RED_U32, kb_red_u32, 256, ReduceArgs)
    AF_BL(KB_RED_F32, kb_red_f32, 256, ReduceArgs)
    AF_BL(KB_LAG_U8, kb_lag_u8, 1024, LagArgs)
    AF_BL(KB_LAG_U16, kb_lag_u16, 1024, LagArgs)
    AF_BL(KB_LAG_U32, kb_lag_u32, 1024, LagArgs)
    AF_BL(KB_LAG_F32, kb_lag_f32, 1024, LagArgs)
    AF_BL(KB_MERGE, kb_merge, 1024, MergeArgs)
    AF_BL(KB_SEGMASK, kb_segmask, 256, SegmaskArgs)
    AF_BL(KB_CLEAN_U16, kb_clean_u16, 256, CleanArgs)
    AF_BL(KB_CLEAN_U32, kb_clean_u32, 256, CleanArgs)
    AF_BL(KB_CLEAN_F32, kb_clean_f32, 256, CleanArgs)
    AF_BL(KB_DD_U8Q, kb_dd_u8q, 64 * kDdRowsQ, DdArgs)
    AF_BL(KB_DD_F32Q, kb_dd_f32q, 64 * kDdRowsQ, DdArgs)
    AF_BL(KB_DD_U8, kb_dd_u8, 256, DdArgs)
    AF_BL(KB_DD_G_U8, kb_dd_g_u8, 256, DdArgs)
    AF_BL(KB_DD_G_U16, kb_dd_g_u16, 256, DdArgs)
    AF_BL(KB_DD_G_U32, kb_dd_g_u32, 256, DdArgs)
    AF_BL(KB_DD_G_F32F, kb_dd_g_f32f, 256, DdArgs)
    AF_BL(KB_DD_G_F32D, kb_dd_g_f32d, 256, DdArgs)
    AF_BL(KB_S2_U16, kb_s2_u16, 256, Stage2Args)
    AF_BL(KB_S2_G_U16, kb_s2_g_u16, 256, Stage2Args)
    AF_BL(KB_S2_G_U32, kb_s2_g_u32, 256, Stage2Args)
    AF_BL(KB_S2_G_U64, kb_s2_g_u64, 256, Stage2Args)
    AF_BL(KB_S2_G_F32, kb_s2_g_f32, 256, Stage2Args)
  case KB_S2_F32C8:
    (void)hipFuncSetAttribute((const void *)kb_s2_f32c8, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kb_s2_f32c8, g, dim3(512), lds, st, (const Stage2Args *)tab, start, n);
    break;
  case KB_S2_F32C16:
    (void)hipFuncSetAttribute((const void *)kb_s2_f32c16, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kb_s2_f32c16, g, dim3(512), lds, st, (const Stage2Args *)tab, start, n);
    break;
#undef AF_BL
  default:
    return launch_batched_u8(kind, tab, start, n, total, lds, st);
  }
  return hipGetLastError();
}

}  // namespace af
