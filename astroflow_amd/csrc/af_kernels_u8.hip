// af_kernels_u8.hip -- 8-bit fast paths of the RFI passes and subband stage 2.
//
// Each thread owns 16 consecutive channels (one 16-byte load per row), so a
// wave reads 1 KB of a row per instruction.  Compiled with -ffp-contract=off.
#include "af_internal.h"

namespace af {

__device__ __forceinline__ uint32_t xcd_remap_u8(uint32_t b, uint32_t n) {
  const uint32_t X = 8;
  uint32_t q = n / X, r = n % X;
  uint32_t xcd = b % X, k = b / X;
  uint32_t start = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return start + k;
}

__device__ __forceinline__ uint4 ld16(const uint8_t *p) {
  uint4 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}

// acc + low / high 16-bit half of v in one v_dot2_u32_u16 (a dot product
// with (1,0) or (0,1)): unpack and add in a single instruction.
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t add_lo16(uint32_t acc, uint32_t v) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2_t, v), (u16x2_t){1, 0}, acc, false);
}
__device__ __forceinline__ uint32_t add_hi16(uint32_t acc, uint32_t v) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2_t, v), (u16x2_t){0, 1}, acc, false);
}

// Sums over the subbands of one fine DM from LDS: P outputs per lane, lane l
// taking outputs l, l+64, ...; row sb of the tile starts at s2 + off[sb] and
// the DM's residual shifts it by rrow[sb].  Eight subbands per step: their
// offsets are scalar loads issued together, then 8 x P aligned 2-byte LDS
// reads.
template <int P>
__device__ __forceinline__ void s2_sum_dm(const uint16_t *s2, const int32_t *off,
                                          const int32_t *rrow, int nsb, int lane,
                                          uint32_t (&acc)[P]) {
  int sb = 0;
  for (; sb + 8 <= nsb; sb += 8) {
    int o[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) o[u] = ldc(off, sb + u) + ldc(rrow, sb + u);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint16_t *p = s2 + o[u] + lane;
#pragma unroll
      for (int i = 0; i < P; ++i) acc[i] += p[64 * i];
    }
  }
  for (; sb < nsb; ++sb) {
    const uint16_t *p = s2 + ldc(off, sb) + ldc(rrow, sb) + lane;
#pragma unroll
    for (int i = 0; i < P; ++i) acc[i] += p[64 * i];
  }
}

__device__ __forceinline__ uint32_t u4get(const uint4 &v, int k) {
  return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

// Loads 16 channels [cb, cb+16) of one row; channels at or beyond `lim`
// read as zero (and are never touched in memory).
__device__ __forceinline__ uint4 load_row16(const uint8_t *row, int cb, int lim) {
  if (cb + 16 <= lim) return ld16(row + cb);
  uint32_t w[4] = {0, 0, 0, 0};
  for (int k = 0; k < 16; ++k)
    if (cb + k < lim) w[k >> 2] |= (uint32_t)row[cb + k] << (8 * (k & 3));
  return make_uint4(w[0], w[1], w[2], w[3]);
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

// ------------------------------------------------------ IQRM statistics --
// Per (row block <= 1024 rows, channel): sum and sum of squares.  A 1024-
// thread block covers 4096 channels (16 per thread) of every fourth row per
// row lane (4 lanes, interleaved, so each wave reads 1 KB of one row per
// load).  Per thread, sums are packed 16-bit lanes (exact: <= 256 rows x 255)
// and squares one multiply-add per byte; the four row lanes are combined in
// LDS.  Partials: u32 sum (<= 1024*255) and u32 square sum (<= 1024*65025).
constexpr int kStatThreads = 1024;

__device__ __forceinline__ void stats_u8_body(const StatArgs &sa, uint32_t bid, uint32_t nblk) {
  const uint8_t *__restrict__ x = (const uint8_t *)sa.x;
  const int nchan = sa.nchan, c0 = sa.c0, c1 = sa.c1, nunits = sa.nunits;
  const RowItem *__restrict__ items = sa.items;
  uint2 *__restrict__ part = (uint2 *)sa.part;
  __shared__ uint32_t ls[4096], lq[4096];
  const int Csub = c1 - c0;
  const int ngrp = (Csub + 15) / 16;
  const int nchunk = (ngrp + 255) / 256;
  // persistent: a grid sized to the machine walks all (item, chunk) units
  for (int unit = (int)bid; unit < nunits; unit += (int)nblk) {
  const int item = unit / nchunk;
  const int gl = threadIdx.x & 255;
  const int rl = threadIdx.x >> 8;
  const int g = (unit % nchunk) * 256 + gl;
  const bool active = g < ngrp;
  const int cb = c0 + 16 * g;
  const RowItem it = items[item];
  uint32_t se[4] = {0, 0, 0, 0}, so[4] = {0, 0, 0, 0};
  uint32_t q[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) q[k] = 0;
  auto accum = [&](const uint4 &w) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t v = u4get(w, k);
      se[k] += v & 0x00FF00FFu;
      so[k] += __builtin_amdgcn_perm(0u, v, 0x0c030c01u);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint32_t e = (v >> (8 * b)) & 0xFFu;
        q[4 * k + b] += e * e;
      }
    }
  };
  if (active) {
    const uint8_t *row = x + (size_t)(it.r0 + rl) * nchan;
    const size_t step = (size_t)4 * nchan;
    if (cb + 16 <= nchan) {
      // the 16-byte load stays inside the row: straight-line loads, several
      // in flight, one code path for the whole wave.  Channels at or beyond
      // c1 (a group that straddles the range end) are summed in their own
      // byte lanes and never written.
      auto ldm = [&](const uint8_t *p) { return ld16(p); };
      uint32_t r = it.r0 + rl;
      for (; r + 12 < it.r1; r += 16, row += 4 * step) {
        const uint4 w0 = ldm(row + cb), w1 = ldm(row + step + cb);
        const uint4 w2 = ldm(row + 2 * step + cb), w3 = ldm(row + 3 * step + cb);
        accum(w0);
        accum(w1);
        accum(w2);
        accum(w3);
      }
      for (; r < it.r1; r += 4, row += step) accum(ldm(row + cb));
    } else {
      for (uint32_t r = it.r0 + rl; r < it.r1; r += 4, row += step) accum(load_row16(row, cb, c1));
    }
  }
  uint32_t s16[16];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    s16[4 * k + 0] = se[k] & 0xFFFFu;
    s16[4 * k + 1] = so[k] & 0xFFFFu;
    s16[4 * k + 2] = se[k] >> 16;
    s16[4 * k + 3] = so[k] >> 16;
  }
  for (int l = 0; l < 4; ++l) {
    if (rl == l) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int i = 16 * gl + k;
        ls[i] = (l ? ls[i] : 0u) + s16[k];
        lq[i] = (l ? lq[i] : 0u) + q[k];
      }
    }
    __syncthreads();
  }
  if (rl == 0 && active) {
    uint2 *o = part + (size_t)item * Csub + (cb - c0);
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (cb + k < c1) o[k] = make_uint2(ls[16 * gl + k], lq[16 * gl + k]);
  }
  __syncthreads();   // LDS is reused by the next unit
  }
}

// at most 64 VGPRs: two 1024-thread workgroups per CU (8 waves per SIMD)
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_seg_stats_u8(StatArgs a) {
  stats_u8_body(a, blockIdx.x, gridDim.x);
}

// Sum the row-block partials of each segment: 16 channels x 64 item lanes
// per block (each wave reads four 128-byte runs per step), combined in LDS.
__device__ __forceinline__ void reduce_u8_body(const ReduceArgs &ra, uint32_t bid, uint32_t) {
  const uint2 *__restrict__ part = (const uint2 *)ra.part;
  const int *__restrict__ seg_items = ra.seg_items;
  const int Csub = ra.Csub;
  Part *__restrict__ segsum = ra.segsum;
  __shared__ uint64_t la[64][16], lb[64][16];
  const int nchunk = (Csub + 15) / 16;
  const int s = (int)bid / nchunk;
  const int cl = threadIdx.x & 15;
  const int c = ((int)bid % nchunk) * 16 + cl;
  const int il = threadIdx.x >> 4;
  uint64_t a = 0, b = 0;
  if (c < Csub) {
    const int i1 = seg_items[s + 1];
    int i = seg_items[s] + il;
    // two loads in flight per step
    for (; i + 64 < i1; i += 128) {
      const uint2 p = part[(size_t)i * Csub + c], q = part[(size_t)(i + 64) * Csub + c];
      a += (uint64_t)p.x + q.x;
      b += (uint64_t)p.y + q.y;
    }
    for (; i < i1; i += 64) {
      const uint2 p = part[(size_t)i * Csub + c];
      a += p.x;
      b += p.y;
    }
  }
  la[il][cl] = a;
  lb[il][cl] = b;
  __syncthreads();
  // tree over the 64 item lanes
  for (int h = 32; h > 0; h >>= 1) {
    if (il < h) {
      la[il][cl] += la[il + h][cl];
      lb[il][cl] += lb[il + h][cl];
    }
    __syncthreads();
  }
  if (il == 0 && c < Csub) segsum[(size_t)s * Csub + c] = Part{la[0][cl], lb[0][cl], 0};
}

__global__ __launch_bounds__(1024) void k_seg_reduce_u8(ReduceArgs a) {
  reduce_u8_body(a, blockIdx.x, gridDim.x);
}

// ------------------------------------------------ clean + transpose (u8) --
// One 1024-thread block per 32-row item.  Lane l of wave w owns channel
// group g = 16*w + (l >> 2) (16 channels) and row quarter q = l & 3 (rows
// 8q..8q+7): 8 16-byte loads kept in registers.  Pass 1: IQRM window mask,
// per-row non-zero sum/count over the zero-DM range (v_dot4 byte sums; the
// count as the sum minus a masked byte SAD against 1, which skips zero
// bytes), a reduction over the row's 16 lane groups (two DPP rotates within
// each 16-lane row, then LDS atomics from the four rows).  Pass 2: x =
// trunc(max((float)x - m, 0)) inside the zero-DM range, static mask, 4x4
// byte transposes with v_perm, and one 8-byte store per channel (the four
// lanes of a group write 32 contiguous bytes of that channel's row): a
// buffer store whose channel offset j*Lp is a scalar operand, so a store
// costs no address arithmetic; a wave whose 256 channels are all stored
// takes them without per-lane tests.
__device__ __forceinline__ uint32_t zdm4(uint32_t w, float m) {
  uint32_t o = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float f = (float)((w >> (8 * b)) & 0xFFu) - m;
    f = f < 0.0f ? 0.0f : f;
    o |= ((uint32_t)f) << (8 * b);
  }
  return o;
}

// zdm4 when (uint8)trunc(max(x - m, 0)) = max(x - k, 0) for every byte x,
// k = ceil(m): true when m is an integer or frac(m) >= 2^-16.  (x - m is
// then exact or rounded by at most 2^-17 < frac(m): it never reaches the next
// integer.)  Even and odd bytes as 16-bit lanes, saturating subtract.
__device__ __forceinline__ uint32_t zsub4(uint32_t w, uint32_t kk) {
  typedef unsigned short us2 __attribute__((ext_vector_type(2)));
  const us2 k2 = __builtin_bit_cast(us2, kk);
  const us2 e = __builtin_elementwise_sub_sat(__builtin_bit_cast(us2, w & 0x00FF00FFu), k2);
  const us2 o = __builtin_elementwise_sub_sat(__builtin_bit_cast(us2, (w >> 8) & 0x00FF00FFu), k2);
  return __builtin_bit_cast(uint32_t, e) | (__builtin_bit_cast(uint32_t, o) << 8);
}

// byte masks of channels [lo, hi) among the 16 at cb: dword k covers cb+4k..
__device__ __forceinline__ uint32_t range16(int cb, int lo, int hi) {
  const int a = min(max(lo - cb, 0), 16), b = min(max(hi - cb, 0), 16);
  return b > a ? ((1u << b) - 1u) & ~((1u << a) - 1u) : 0u;
}
__device__ __forceinline__ uint32_t bytes_of(uint32_t m16, int k) {
  const uint32_t nib = (m16 >> (4 * k)) & 0xFu;
  return ((nib * 0x00204081u) & 0x01010101u) * 0xFFu;
}

template <bool O32>
__device__ __forceinline__ void clean_u8_body_t(const CleanArgs &a, uint32_t bid) {
  __shared__ uint32_t s_sum[32], s_cnt[32], s_kk[32];
  __shared__ float s_m[32];
  if (a.run_if && *a.run_if == 0) return;
  const RowItem it = a.items[bid];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane & 3;
  const int g = wave * 16 + (lane >> 2);
  const int cb = 16 * g;
  const int nrows = (int)(it.r1 - it.r0);
  const uint8_t *__restrict__ x = (const uint8_t *)a.x;
  if (threadIdx.x < 32) {
    s_sum[threadIdx.x] = 0;
    s_cnt[threadIdx.x] = 0;
  }
  // per-thread channel masks (0xFF = keep / in range).  The IQRM segment
  // mask and the static mask are 0/1 bytes: one 16-byte load each, expanded
  // to 0x00/0xFF lanes by a multiply (no carries).
  const uint32_t m_in = range16(cb, 0, a.nchan);
  const uint32_t m_z = a.zdm ? range16(cb, a.zc0, a.zc1) : 0u;
  uint32_t km[4], zr[4], st[4];
  const uint8_t *smask = a.segmask ? a.segmask + (size_t)it.seg * a.nchan : nullptr;
  uint4 sm16 = make_uint4(0, 0, 0, 0), st16 = make_uint4(0, 0, 0, 0);
  if (smask) sm16 = load_row16(smask, cb, a.nchan);
  if (a.stat_mask) st16 = load_row16(a.stat_mask, cb, a.nchan);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t inside = bytes_of(m_in, k);
    km[k] = ~(u4get(sm16, k) * 0xFFu) & inside;
    zr[k] = bytes_of(m_z, k);
    st[k] = ~(u4get(st16, k) * 0xFFu) & inside;
  }
  uint4 w[8];
  {
    const uint8_t *rowp = x + (size_t)(it.r0 + 8 * q) * a.nchan + cb;
    if (cb + 16 <= a.nchan && 8 * q + 8 <= nrows) {
      // all eight rows present: issue the eight loads back to back.  This is
      // the last read of the input: streaming loads leave the caches to the
      // channel-major columns written here, which stage 1 reads next.
      typedef uint32_t v4u_any __attribute__((ext_vector_type(4), aligned(1)));
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const v4u_any t = __builtin_nontemporal_load((const v4u_any *)(rowp + (size_t)i * a.nchan));
        w[i] = make_uint4(t.x, t.y, t.z, t.w);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        w[i] = make_uint4(0, 0, 0, 0);
        if (8 * q + i < nrows && cb < a.nchan)
          w[i] = load_row16(x + (size_t)(it.r0 + 8 * q + i) * a.nchan, cb, a.nchan);
      }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
      w[i] = make_uint4(w[i].x & km[0], w[i].y & km[1], w[i].z & km[2], w[i].w & km[3]);
  }
  __syncthreads();
  if (a.zdm) {
    // (row_ror:4 and row_ror:8 within each 16-lane row)
    auto ror4 = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, false); };
    auto ror8 = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false); };
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      uint32_t s = 0, sad = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t z = u4get(w[i], k) & zr[k];
        s = __builtin_amdgcn_udot4(z, 0x01010101u, s, false);
        // sum over the non-zero bytes of |1 - z| = z - 1
        sad = __builtin_amdgcn_msad_u8(0x01010101u, z, sad);
      }
      // per 16-lane row: s <= 16*255*4 < 2^18, n <= 64 < 2^14
      uint32_t pk = s | ((s - sad) << 18);
      pk += ror4(pk);
      pk += ror8(pk);
      if ((lane & 15) < 4) {
        atomicAdd(&s_sum[8 * q + i], pk & 0x3FFFFu);
        atomicAdd(&s_cnt[8 * q + i], pk >> 18);
      }
    }
    __syncthreads();
    if (threadIdx.x < 32) {
      const uint32_t n = s_cnt[threadIdx.x];
      const float m = n ? (float)s_sum[threadIdx.x] / (float)n : 0.0f;
      s_m[threadIdx.x] = m;
      // the byte subtract of zsub4 applies (kk != 0xFFFFFFFF)
      const float fl = floorf(m), fr = m - fl;   // exact: m < 256
      const uint32_t k = (uint32_t)fl + (fr > 0.0f ? 1u : 0u);
      s_kk[threadIdx.x] = (fr == 0.0f || fr >= 0x1p-16f) ? (k | (k << 16)) : 0xFFFFFFFFu;
    }
    __syncthreads();
    const bool zall = m_z == 0xFFFFu;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t kk = s_kk[8 * q + i];
      uint32_t o[4];
      if (kk != 0xFFFFFFFFu) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t v = u4get(w[i], k);
          o[k] = zall ? zsub4(v, kk) : (zsub4(v, kk) & zr[k]) | (v & ~zr[k]);
        }
      } else {
        const float m = s_m[8 * q + i];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t v = u4get(w[i], k);
          o[k] = (zdm4(v, m) & zr[k]) | (v & ~zr[k]);
        }
      }
      w[i] = make_uint4(o[0] & st[0], o[1] & st[1], o[2] & st[2], o[3] & st[3]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      w[i] = make_uint4(w[i].x & st[0], w[i].y & st[1], w[i].z & st[2], w[i].w & st[3]);
  }
  if (cb >= a.nchan) return;
  // channels of this lane's group that are stored: [dd_c0, dd_c1), and when
  // the item is cut, the first it.pad of them
  const int c_hi = a.cut_in_pad ? min(a.dd_c1, a.dd_c0 + (int)it.pad) : a.dd_c1;
  const uint32_t m_st = range16(cb, a.dd_c0, c_hi);
  const int rbase = 8 * q;
  const bool full = rbase + 8 <= nrows;
  // wave-uniform: every lane stores all 16 channels of all 8 rows
  const bool wave_all = __builtin_amdgcn_ballot_w64(!(full && m_st == 0xFFFFu)) == 0;
  uint8_t *__restrict__ xT = (uint8_t *)a.xT;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(xT, (short)0, -1, 0x00020000);
  // byte offset of (channel cb, row it.r0 + rbase); channel cb + j adds j*Lp
  const long long cofs = (long long)(cb - a.dd_c0) * (long long)a.Lp + it.r0 + rbase;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint32_t lo[4], hi[4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t ra = u4get(w[4 * h + 0], k), rb = u4get(w[4 * h + 1], k);
      const uint32_t rc = u4get(w[4 * h + 2], k), rd = u4get(w[4 * h + 3], k);
      const uint32_t t0 = __builtin_amdgcn_perm(rb, ra, 0x05010400u);
      const uint32_t t1 = __builtin_amdgcn_perm(rb, ra, 0x07030602u);
      const uint32_t t2 = __builtin_amdgcn_perm(rd, rc, 0x05010400u);
      const uint32_t t3 = __builtin_amdgcn_perm(rd, rc, 0x07030602u);
      uint32_t *dst = h ? hi : lo;
      dst[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
      dst[1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
      dst[2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
      dst[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = 4 * k + b;
      typedef uint32_t v2u __attribute__((ext_vector_type(2)));
      const v2u v = {lo[b], hi[b]};
      if (O32 && wave_all) {
        __builtin_amdgcn_raw_buffer_store_b64(v, rsrc, (int)(uint32_t)cofs,
                                              (int)((uint32_t)j * (uint32_t)a.Lp), 0);
        continue;
      }
      if (!((m_st >> j) & 1u)) continue;
      uint8_t *dst = xT + (size_t)(cofs + (long long)j * (long long)a.Lp);
      if (full) {
        const uint2 u = make_uint2(lo[b], hi[b]);
        __builtin_memcpy(dst, &u, 8);
      } else {
        for (int r = 0; r < 8; ++r)
          if (rbase + r < nrows) dst[r] = (uint8_t)(((r < 4 ? lo[b] : hi[b]) >> (8 * (r & 3))) & 0xFFu);
      }
    }
  }
}

// O32: the channel-major buffer is below 4 GiB (buffer stores with 32-bit offsets)
__device__ __forceinline__ void clean_u8_body(const CleanArgs &a, uint32_t bid, uint32_t) {
  if ((uint64_t)(a.dd_c1 - a.dd_c0) * a.Lp < 0xFFFFFFFFull) clean_u8_body_t<true>(a, bid);
  else clean_u8_body_t<false>(a, bid);
}

__global__ __launch_bounds__(1024) void k_clean_u8(CleanArgs a) {
  clean_u8_body(a, xcd_remap_u8(blockIdx.x, gridDim.x), gridDim.x);
}

// --------------------------------------------- subband stage 2 (u16 in) --
// One workgroup (512 threads) per (coarse DM group j, tile of 64*P outputs).
// Row sb of the group's partial sums is staged in LDS for the tile plus that
// subband's own largest residual (rows packed at a.sboff[j][sb]); each wave
// then sums the subbands of one fine DM at a time, lane l of a wave summing
// outputs l, l+64, ..., l+64(P-1).
template <int P>
__device__ __forceinline__ void stage2_u16v_body(const Stage2Args &a, uint32_t lid, uint32_t) {
  extern __shared__ uint4 s2v[];
  uint16_t *s2 = (uint16_t *)s2v;
  constexpr int TILE = 64 * P;
  const int ntile = (int)((a.nout + TILE - 1) / TILE);
  const int ti = (int)(lid % (uint32_t)ntile);
  const int j = (int)(lid / (uint32_t)ntile);
  const size_t t0 = (size_t)ti * TILE;
  const int32_t *__restrict__ off = a.sboff + (size_t)j * (a.nsb + 1);
  const uint16_t *__restrict__ src = (const uint16_t *)a.inter + (size_t)j * a.nsb * a.L1 + t0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  // fill: wave w stages rows w, w+nw, ...; a row of n 16-byte chunks takes
  // ceil(n/64) wave-wide loads.  The (row, pass) walk is wave-uniform, and
  // four loads are issued before their LDS stores.
  {
    int sb = wave, k0 = 0;
    int o8 = sb < a.nsb ? ldc(off, sb) / 8 : 0;
    int n8 = sb < a.nsb ? ldc(off, sb + 1) / 8 - o8 : 0;
    const int spare = a.lds_len / 8;
    // one (row, pass) slot of the walk (wave-uniform): the lane's chunk
    // address (a lane past its row re-reads the row's last chunk) and LDS
    // slot (-1: none)
    auto slot = [&](const uint16_t *&ptr, int &dst) {
      ptr = src;
      dst = -1;
      if (sb < a.nsb) {
        const int k = k0 + lane;
        ptr = src + (size_t)sb * a.L1 + 8 * (size_t)min(k, n8 - 1);
        dst = k < n8 ? o8 + k : -1;
        k0 += 64;
        if (k0 >= n8) {
          sb += nw;
          k0 = 0;
          if (sb < a.nsb) {
            o8 = ldc(off, sb) / 8;
            n8 = ldc(off, sb + 1) / 8 - o8;
          }
        }
      }
    };
    // four slots, then four unconditional loads, then the LDS stores (a load
    // under a per-lane branch is waited for at the join)
    while (sb < a.nsb) {
      const uint16_t *p0, *p1, *p2, *p3;
      int d0, d1, d2, d3;
      slot(p0, d0);
      slot(p1, d1);
      slot(p2, d2);
      slot(p3, d3);
      const uint4 v0 = *(const uint4 *)p0, v1 = *(const uint4 *)p1;
      const uint4 v2 = *(const uint4 *)p2, v3 = *(const uint4 *)p3;
      // unconditional stores (a spare chunk past the tile takes the unused
      // ones), so the loads are not moved into per-lane branches
      s2v[d0 >= 0 ? d0 : spare] = v0;
      s2v[d1 >= 0 ? d1 : spare] = v1;
      s2v[d2 >= 0 ? d2 : spare] = v2;
      s2v[d3 >= 0 ? d3 : spare] = v3;
    }
  }
  __syncthreads();
  for (int mi = wave; mi < a.ndm0; mi += nw) {
    const int m = j * a.ndm0 + mi;
    if (m >= a.ndm) break;
    const int32_t *__restrict__ rrow = a.res + (size_t)m * a.nsb;
    // lane l sums outputs l, l+64, ...: every LDS read is an aligned 2-byte
    // load, consecutive across the wave (a 16-byte read at the residual's
    // arbitrary offset stalls the LDS on misalignment)
    uint32_t acc[P];
#pragma unroll
    for (int i = 0; i < P; ++i) acc[i] = 0;
    s2_sum_dm<P>(s2, off, rrow, a.nsb, lane, acc);
    uint32_t *o = (uint32_t *)a.out + (size_t)m * a.ostride + t0 + (size_t)lane;
#pragma unroll
    for (int i = 0; i < P; ++i)
      // the plane is not read again on the device: streaming stores keep it
      // from displacing the next chunk's input in the caches
      if (t0 + (size_t)lane + 64 * i < a.nout) __builtin_nontemporal_store(acc[i], o + 64 * i);
  }
}

template <int P>
__global__ __launch_bounds__(512) void k_stage2_u16v(Stage2Args a) {
  stage2_u16v_body<P>(a, xcd_remap_u8(blockIdx.x, gridDim.x), gridDim.x);
}

// Stage 2 when all subbands' tile + halo do not fit LDS (many subbands, a
// large residual halo): the workgroup stages a.sbc subbands at a time, and
// each wave keeps the sums of its (up to 4) fine DMs in registers across the
// chunks.  Requires ndm0 <= 32 (8 waves x 4 DMs).
template <int P>
__device__ __forceinline__ void stage2_u16c_body(const Stage2Args &a, uint32_t lid, uint32_t) {
  extern __shared__ uint4 s2v[];
  uint16_t *s2 = (uint16_t *)s2v;
  constexpr int TILE = 64 * P;
  constexpr int MD = 4;
  const int ntile = (int)((a.nout + TILE - 1) / TILE);
  const int ti = (int)(lid % (uint32_t)ntile);
  const int j = (int)(lid / (uint32_t)ntile);
  const size_t t0 = (size_t)ti * TILE;
  const int ext8 = (TILE + a.maxres_j[j] + 7) / 8;
  const int lds8 = a.lds_len / 8;
  const uint16_t *__restrict__ src = (const uint16_t *)a.inter + (size_t)j * a.nsb * a.L1 + t0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  uint32_t acc[MD][P];
#pragma unroll
  for (int d = 0; d < MD; ++d)
#pragma unroll
    for (int i = 0; i < P; ++i) acc[d][i] = 0;
  for (int sb0 = 0; sb0 < a.nsb; sb0 += a.sbc) {
    const int nsc = min(a.sbc, a.nsb - sb0);
    fill_rows16(s2v, lds8, src + (size_t)sb0 * a.L1, a.L1, 8, nsc, ext8);
    __syncthreads();
#pragma unroll
    for (int d = 0; d < MD; ++d) {
      const int mi = wave + d * nw;
      const int m = j * a.ndm0 + mi;
      if (mi < a.ndm0 && m < a.ndm) {
        const int32_t *__restrict__ rrow = a.res + (size_t)m * a.nsb + sb0;
        for (int sb = 0; sb < nsc; ++sb) {
          const uint16_t *p = s2 + sb * a.lds_len + lane + ldc(rrow, sb);
#pragma unroll
          for (int i = 0; i < P; ++i) acc[d][i] += p[64 * i];
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
    uint32_t *o = (uint32_t *)a.out + (size_t)m * a.ostride + t0 + (size_t)lane;
#pragma unroll
    for (int i = 0; i < P; ++i)
      if (t0 + (size_t)lane + 64 * i < a.nout) __builtin_nontemporal_store(acc[d][i], o + 64 * i);
  }
}

template <int P>
__global__ __launch_bounds__(512) void k_stage2_u16c(Stage2Args a) {
  stage2_u16c_body<P>(a, xcd_remap_u8(blockIdx.x, gridDim.x), gridDim.x);
}

// ------------------------------------------------------ batched launches --
__device__ __forceinline__ int batch_find_u8(const uint32_t *__restrict__ start, int n, uint32_t g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define AF_BATCHED_U8(NAME, ARGS, BOUNDS, CALL)                                         \
  __global__ __launch_bounds__(BOUNDS) void NAME(const ARGS *__restrict__ tab,          \
                                                 const uint32_t *__restrict__ start, int n) { \
    const uint32_t g = xcd_remap_u8(blockIdx.x, gridDim.x);                             \
    const int i = batch_find_u8(start, n, g);                                           \
    const ARGS a = tab[i];                                                              \
    const uint32_t lid = g - start[i], nb = start[i + 1] - start[i];                    \
    CALL;                                                                               \
  }

AF_BATCHED_U8(kb_stat_u8, StatArgs, 1024, (stats_u8_body(a, lid, nb)))
AF_BATCHED_U8(kb_red_u8, ReduceArgs, 1024, (reduce_u8_body(a, lid, nb)))
AF_BATCHED_U8(kb_clean_u8, CleanArgs, 1024, (clean_u8_body(a, lid, nb)))
AF_BATCHED_U8(kb_s2_v4, Stage2Args, 512, (stage2_u16v_body<4>(a, lid, nb)))
AF_BATCHED_U8(kb_s2_v8, Stage2Args, 512, (stage2_u16v_body<8>(a, lid, nb)))
AF_BATCHED_U8(kb_s2_v16, Stage2Args, 512, (stage2_u16v_body<16>(a, lid, nb)))
AF_BATCHED_U8(kb_s2_c8, Stage2Args, 512, (stage2_u16c_body<8>(a, lid, nb)))
AF_BATCHED_U8(kb_s2_c16, Stage2Args, 512, (stage2_u16c_body<16>(a, lid, nb)))

// ------------------------------------------------ served-beam helpers --
// Window statistics of a served beam from its ring of per-chunk sums: one
// thread per channel, exact 64/128-bit integer adds.
__global__ __launch_bounds__(256) void k_ring_sum(const Part *__restrict__ ring, int R, int first,
                                                  int nfull, int last, int last_full, int Csub,
                                                  Part *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Csub) return;
  uint64_t s = 0, lo = 0, hi = 0;
  auto add = [&](const Part &p) {
    s += p.s;
    const uint64_t n = lo + p.q_lo;
    hi += p.q_hi + (n < lo ? 1u : 0u);
    lo = n;
  };
  for (int i = 0; i < nfull; ++i) {
    const size_t sl = (size_t)((first + i) % R);
    add(ring[(sl * 2 + 0) * Csub + c]);
    add(ring[(sl * 2 + 1) * Csub + c]);
  }
  add(ring[((size_t)last * 2 + 0) * Csub + c]);
  if (last_full) add(ring[((size_t)last * 2 + 1) * Csub + c]);
  out[c] = Part{s, lo, hi};
}

// One workgroup: did the window mask change since the last call?
__global__ __launch_bounds__(1024) void k_mask_delta(const uint8_t *__restrict__ mask,
                                                     uint8_t *__restrict__ prev, int n,
                                                     int *__restrict__ flags) {
  int diff = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint8_t m = mask[i];
    diff |= m != prev[i];
    prev[i] = m;
  }
  diff = __syncthreads_or(diff);
  if (threadIdx.x == 0) {
    flags[1] = (diff || flags[0] == 0) ? 1 : 0;
    flags[0] = 1;
  }
}

// One workgroup row per (row, 4096-byte span); byte stores (the span start
// is not aligned).
__global__ __launch_bounds__(256) void k_zero_rows(uint8_t *__restrict__ base, size_t pitch,
                                                   uint32_t width) {
  uint8_t *row = base + (size_t)blockIdx.y * pitch;
  for (uint32_t i = blockIdx.x * 4096u + threadIdx.x; i < min(width, blockIdx.x * 4096u + 4096u);
       i += 256u)
    row[i] = 0;
}

// Copies `width16` 16-byte units of each of nrows rows (row pitch `pitch`
// bytes, both bases 16-byte aligned, pitch a multiple of 16).
__global__ __launch_bounds__(256) void k_copy_rows16(uint4 *__restrict__ dst,
                                                     const uint4 *__restrict__ src,
                                                     size_t pitch16, uint32_t width16) {
  const size_t row = (size_t)blockIdx.y * pitch16;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < width16; i += gridDim.x * 256u)
    dst[row + i] = src[row + i];
}

// ---------------------------------------------------------- launchers --

hipError_t launch_copy_rows(uint8_t *dst, const uint8_t *src, size_t pitch, size_t width,
                            int nrows, hipStream_t st) {
  if (width == 0 || nrows <= 0) return hipSuccess;
  if (((uintptr_t)dst | (uintptr_t)src | pitch) & 15) return hipErrorInvalidValue;
  const size_t w16 = (width + 15) / 16;
  if (w16 > 0xFFFFFFFFull || nrows > 65535) return hipErrorInvalidValue;
  unsigned gx = (unsigned)((w16 + 255) / 256);
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(k_copy_rows16, dim3(gx, (unsigned)nrows), dim3(256), 0, st, (uint4 *)dst,
                     (const uint4 *)src, pitch / 16, (uint32_t)w16);
  return hipGetLastError();
}

hipError_t launch_ring_sum(const Part *ring, int R, int first, int nfull, int last, int last_full,
                           int Csub, Part *out, hipStream_t st) {
  if (Csub <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ring_sum, dim3((unsigned)((Csub + 255) / 256)), dim3(256), 0, st, ring, R,
                     first, nfull, last, last_full, Csub, out);
  return hipGetLastError();
}

hipError_t launch_mask_delta(const uint8_t *mask, uint8_t *prev, int n, int *flags, hipStream_t st) {
  hipLaunchKernelGGL(k_mask_delta, dim3(1), dim3(1024), 0, st, mask, prev, n, flags);
  return hipGetLastError();
}

hipError_t launch_zero_rows(uint8_t *base, size_t pitch, size_t width, int nrows, hipStream_t st) {
  if (width == 0 || nrows <= 0) return hipSuccess;
  if (width > 0xFFFFFFFFull || nrows > 65535) return hipErrorInvalidValue;
  const unsigned gx = (unsigned)((width + 4095) / 4096);
  hipLaunchKernelGGL(k_zero_rows, dim3(gx, (unsigned)nrows), dim3(256), 0, st, base, pitch,
                     (uint32_t)width);
  return hipGetLastError();
}

hipError_t launch_seg_stats_u8(const uint8_t *x, int nchan, int c0, int c1, const RowItem *items,
                               int nitems, uint2 *part, hipStream_t st) {
  const int ngrp = (c1 - c0 + 15) / 16;
  const int nchunk = (ngrp + 255) / 256;
  // one block per unit (the persistent form measured slower: its last
  // round leaves CUs idle while a full grid overlaps block tails)
  const int nunits = nitems * nchunk;
  StatArgs a{x, items, part, nchan, c0, c1, nunits};
  return emit_args(KB_STAT_U8, 1, (size_t)nunits, kStatThreads, 0, a, st);
}

hipError_t launch_seg_reduce_u8(const uint2 *part, const int *seg_items, int nseg, int Csub,
                                Part *segsum, hipStream_t st) {
  const int nchunk = (Csub + 15) / 16;   // 16 channels per block
  ReduceArgs a{part, seg_items, segsum, nseg, Csub};
  return emit_args(KB_RED_U8, 2, (size_t)nseg * nchunk, 1024, 0, a, st);
}

hipError_t launch_clean_u8(const CleanArgs &a, hipStream_t st) {
  return emit_args(KB_CLEAN_U8, 6, (size_t)a.nitems, 1024, 0, a, st);
}

hipError_t launch_stage2_u16v(const Stage2Args &a, hipStream_t st) {
  const int P = a.tile / 64;
  const int ntile = (int)((a.nout + a.tile - 1) / a.tile);
  const size_t nblk = (size_t)ntile * a.ndm_nom;
  if (a.sbc > 0 && a.sbc < a.nsb) {
    if (a.ndm0 > 32 || (P != 8 && P != 16)) return hipErrorInvalidValue;
    const size_t lds = (size_t)a.sbc * a.lds_len * sizeof(uint16_t);
    return emit_args(P == 16 ? KB_S2_C16 : KB_S2_C8, 8, nblk, 512, lds, a, st);
  }
  // all subbands: packed rows, a.lds_len u16 elements per workgroup
  if (!a.sboff || (P != 4 && P != 8 && P != 16)) return hipErrorInvalidValue;
  const size_t lds = (size_t)a.lds_len * sizeof(uint16_t) + 16;   // + the spare chunk
  return emit_args(P == 16 ? KB_S2_V16 : P == 8 ? KB_S2_V8 : KB_S2_V4, 8, nblk, 512, lds, a, st);
}

template <typename A>
static const A &argsof_u8(const LaunchDesc &d) {
  return *reinterpret_cast<const A *>(d.args);
}

hipError_t launch_desc_u8(const LaunchDesc &d, hipStream_t st) {
  const dim3 g(d.nblk), b(d.threads);
  switch (d.kind) {
  case KB_STAT_U8: hipLaunchKernelGGL(k_seg_stats_u8, g, b, d.lds, st, argsof_u8<StatArgs>(d)); break;
  case KB_RED_U8: hipLaunchKernelGGL(k_seg_reduce_u8, g, b, d.lds, st, argsof_u8<ReduceArgs>(d)); break;
  case KB_CLEAN_U8: hipLaunchKernelGGL(k_clean_u8, g, b, d.lds, st, argsof_u8<CleanArgs>(d)); break;
  case KB_S2_V4:
    (void)hipFuncSetAttribute((const void *)k_stage2_u16v<4>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds);
    hipLaunchKernelGGL(k_stage2_u16v<4>, g, b, d.lds, st, argsof_u8<Stage2Args>(d));
    break;
  case KB_S2_V8:
    (void)hipFuncSetAttribute((const void *)k_stage2_u16v<8>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds);
    hipLaunchKernelGGL(k_stage2_u16v<8>, g, b, d.lds, st, argsof_u8<Stage2Args>(d));
    break;
  case KB_S2_V16:
    (void)hipFuncSetAttribute((const void *)k_stage2_u16v<16>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds);
    hipLaunchKernelGGL(k_stage2_u16v<16>, g, b, d.lds, st, argsof_u8<Stage2Args>(d));
    break;
  case KB_S2_C8:
    (void)hipFuncSetAttribute((const void *)k_stage2_u16c<8>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds);
    hipLaunchKernelGGL(k_stage2_u16c<8>, g, b, d.lds, st, argsof_u8<Stage2Args>(d));
    break;
  case KB_S2_C16:
    (void)hipFuncSetAttribute((const void *)k_stage2_u16c<16>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds);
    hipLaunchKernelGGL(k_stage2_u16c<16>, g, b, d.lds, st, argsof_u8<Stage2Args>(d));
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batched_u8(int kind, const void *tab, const uint32_t *start, int n,
                             uint32_t total, uint32_t lds, hipStream_t st) {
  const dim3 g(total);
  switch (kind) {
  case KB_STAT_U8:
    hipLaunchKernelGGL(kb_stat_u8, g, dim3(1024), lds, st, (const StatArgs *)tab, start, n);
    break;
  case KB_RED_U8:
    hipLaunchKernelGGL(kb_red_u8, g, dim3(1024), lds, st, (const ReduceArgs *)tab, start, n);
    break;
  case KB_CLEAN_U8:
    hipLaunchKernelGGL(kb_clean_u8, g, dim3(1024), lds, st, (const CleanArgs *)tab, start, n);
    break;
  case KB_S2_V4:
    (void)hipFuncSetAttribute((const void *)kb_s2_v4, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kb_s2_v4, g, dim3(512), lds, st, (const Stage2Args *)tab, start, n);
    break;
  case KB_S2_V8:
    (void)hipFuncSetAttribute((const void *)kb_s2_v8, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kb_s2_v8, g, dim3(512), lds, st, (const Stage2Args *)tab, start, n);
    break;
  case KB_S2_V16:
    (void)hipFuncSetAttribute((const void *)kb_s2_v16, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kb_s2_v16, g, dim3(512), lds, st, (const Stage2Args *)tab, start, n);
    break;
  case KB_S2_C8:
    (void)hipFuncSetAttribute((const void *)kb_s2_c8, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kb_s2_c8, g, dim3(512), lds, st, (const Stage2Args *)tab, start, n);
    break;
  case KB_S2_C16:
    (void)hipFuncSetAttribute((const void *)kb_s2_c16, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(kb_s2_c16, g, dim3(512), lds, st, (const Stage2Args *)tab, start, n);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace af
