// af_internal.h -- shared definitions of the gfx950 dedispersion library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "astroflow_amd.h"

namespace af {

// A segment of binned rows [t0,t1) between consecutive IQRM window
// boundaries; every row of a segment is covered by the same windows.
struct Seg {
  uint32_t t0, t1;
  uint32_t w0, w1;  // covering windows [w0,w1)
};

// A block of rows [r0,r1) inside one segment: the unit of work of the
// statistics and cleaning passes.
struct RowItem {
  uint32_t seg, r0, r1, pad;
};

// Per-(row block, channel) partial sums of the IQRM statistics pass.
// Unsigned data: s = sum, (q_lo, q_hi) = 128-bit sum of squares.
// float32 data: s and q_lo hold the bit patterns of double sums.
struct Part {
  uint64_t s, q_lo, q_hi;
};

constexpr int kDdLanesOut = 16;               // outputs per lane, u8 dedispersion
constexpr int kDdTile = 64 * kDdLanesOut;     // outputs per wave tile
// The aligned-load u8 kernel: lane l + 1 hands lane l the 16 bytes after its
// own, so lane 63 only loads and its outputs are not kept.
constexpr int kDdTileQ = 63 * kDdLanesOut;
constexpr int kDdRowsPerBlock = 4;            // DM rows (waves) per workgroup
constexpr int kDdRowsQ = 8;                   // DM rows (waves) per workgroup, aligned-load u8 kernel
// LDS of the u8 stage-1 kernel when a channel group spans several 256-channel
// blocks: 16 running totals per thread
constexpr size_t kDdTotLds = (size_t)16 * 64 * kDdRowsQ * 4;
constexpr int kStatRows = 1024;               // rows per statistics item
constexpr int kCleanRows = 64;                // rows per cleaning item
constexpr size_t kInterBudget = size_t(1) << 30;   // subband intermediate per super-tile

// Read-only int table (delays, residuals, LDS row offsets) through the
// constant address space: a wave-uniform index becomes a scalar load.
__device__ __forceinline__ int32_t ldc(const int32_t *p, size_t i) {
  typedef __attribute__((address_space(4))) const int32_t ci32;
  return ((ci32 *)p)[i];
}

size_t dtype_size(int dtype);
int acc_size(int dtype);  // bytes of one output element

// Error state (thread-local message).
void set_error(const std::string &msg);
std::string error_text();  // the calling thread's last message
af_status fail(af_status code, const std::string &msg);

#define AF_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::af::fail(AF_ERR_RUNTIME, std::string(#call) + ": " +           \
                                            hipGetErrorString(e_));           \
  } while (0)

#define AF_TRY(x)                                                             \
  do {                                                                        \
    af_status s_ = (x);                                                       \
    if (s_ != AF_OK) return s_;                                               \
  } while (0)

// ----------------------------------------------------------------------
// Kernel launchers (af_kernels.hip).  All are asynchronous on `st`.
// ----------------------------------------------------------------------
hipError_t launch_bin(int dtype, const void *in, void *out, size_t nsamp,
                      int nchan, int ds, size_t nbin, hipStream_t st);

hipError_t launch_seg_stats(int dtype, const void *x, int nchan, int c0, int c1,
                            const RowItem *items, int nitems, Part *part,
                            hipStream_t st);
hipError_t launch_seg_reduce(int dtype, const Part *part, const int *seg_items,
                             int nseg, int Csub, Part *segsum, hipStream_t st);
hipError_t launch_iqrm_lags(int dtype, const Part *segsum, const Seg *segs,
                            const uint32_t *wseg, const uint32_t *wt0,
                            const uint32_t *wt1, int nwin, int Csub, int mode,
                            const int *lags, int nlags, float nsigma,
                            uint8_t *flags, size_t lds_bytes, hipStream_t st);
hipError_t launch_iqrm_merge(const uint8_t *flags, const int *lags, int nlags, int nwin,
                             int Csub, int c0, int nchan, uint8_t *masks, int reach,
                             hipStream_t st);
hipError_t launch_segmask(const uint8_t *masks, const Seg *segs, int nseg,
                          int nchan, uint8_t *segmask, hipStream_t st);

struct CleanArgs {
  const void *x;            // [nbin][nchan] time-major
  void *xT;                 // [dd_c1-dd_c0][Lp] channel-major
  size_t Lp;
  int nchan;
  const RowItem *items;
  int nitems;
  const uint8_t *segmask;   // [nseg][nchan] or null
  const uint8_t *stat_mask; // [nchan] or null
  int zdm, zc0, zc1;
  int dd_c0, dd_c1;
  const int *run_if;        // u8 kernel: skip the launch when *run_if == 0 (null = always run)
  // 1: item.pad holds how many channels from dd_c0 on are read by the
  // dedispersion at the item's rows (their delays reach that far); later
  // channels' columns are not stored there
  int cut_in_pad;
};
hipError_t launch_clean(int dtype, const CleanArgs &a, hipStream_t st);

struct DdArgs {
  const void *xT;           // [nch][Lp]
  size_t Lp;
  const int32_t *delays;    // [nrows][nch]
  int nch;                  // channels in xT / delay row length
  int nrows;                // DM rows
  int ngrp;                 // channel groups (subbands); 1 = direct
  int grp_w;                // channels per group
  void *out;                // direct: [nrows][ostride]; grouped: [nrows][ngrp][ostride]
  size_t ostride;
  size_t nvalid;            // outputs written per row (direct: nout)
  int ntiles;               // time tiles of kDdTile (kDdTileQ: aligned-load u8 kernel)
  int out16;                // 1: u16 output (u8 input, group width <= 256)
  int lds_rows;             // u8 LDS-staged kernel: DM rows per workgroup (0 = off)
  int lds_win;              // bytes of a channel's delay window staged in LDS
  const int4 *worklist;     // optional (grp, tile, rowblk, -) per workgroup, in launch order
  int nwork;
  int u8q;                  // u8: the aligned-load kernel (kDdRowsQ rows in step, kDdTileQ tiles)
  int f32q;                 // float32: kDdRowsQ rows in step, kDdTile tiles (k_dd_f32_q)
};
hipError_t launch_dedisperse(int dtype, const DdArgs &a, hipStream_t st);

struct Stage2Args {
  const void *inter;        // [ndm_nom][nsb][L1]
  size_t L1;
  int nsb;
  const int32_t *res;       // [ndm][nsb]
  const int32_t *maxres_j;  // [ndm_nom]
  int ndm, ndm0, ndm_nom;
  void *out;                // [ndm][ostride], this call's first output column
  size_t nout;              // output columns written per row by this call
  size_t ostride;           // row stride of out
  int tile;                 // outputs per workgroup tile
  int lds_len;              // LDS row length (tile + max over groups of maxres)
  int inter16;              // inter is u16 (u8 input)
  int sbc;                  // u16 LDS kernel: subbands staged at a time (0 = all)
  // All-subband u16 kernel: LDS row offsets [ndm_nom][nsb+1] in u16 elements.
  // Row sb of group j holds the tile plus that subband's own largest residual
  // over the group's fine DMs, rounded up to 8 elements; rows are packed.
  const int32_t *sboff;
};
hipError_t launch_stage2(int dtype, const Stage2Args &a, hipStream_t st);
// Single-DM spectrum: out[t][c] = xT[c][t0 + t + delay[c]], [tlen][nch].
hipError_t launch_gather(int dtype, const void *xT, size_t Lp, const int32_t *delay, int nch,
                         size_t t0, size_t tlen, void *out, hipStream_t st);
// Device-to-device copy on `st` (recorded like the kernels, see below).
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);

// Kernel argument blocks of the launchers that take scalar arguments.
struct BinArgs {
  const void *in;
  void *out;
  size_t nsamp, nbin;
  int nchan, ds;
};
struct StatArgs {
  const void *x;
  const RowItem *items;
  void *part;               // Part* (generic) or uint2* (u8)
  int nchan, c0, c1, nunits;
};
struct ReduceArgs {
  const void *part;         // Part* (generic) or uint2* (u8)
  const int *seg_items;
  Part *segsum;
  int nseg, Csub;
};
struct LagArgs {
  const Part *segsum;
  const uint32_t *wseg, *wt0, *wt1;
  const int *lags;
  uint8_t *flags;
  int Csub, mode, nlags;
  float nsigma;
};
struct MergeArgs {
  const uint8_t *flags;
  const int *lags;
  uint8_t *masks;
  int nlags, Csub, c0, nchan;
};
struct SegmaskArgs {
  const uint8_t *masks;
  const Seg *segs;
  uint8_t *segmask;
  int nseg, nchan;
};
struct CopyArgs {
  void *dst;
  const void *src;
  size_t bytes;
};

// Every launch the library makes, as data: which kernel (kind), its grid,
// workgroup size, dynamic LDS and argument block.  Used to merge the same
// stage of many plans into one launch (af_batch).
enum LaunchKind : int {
  KB_BIN_U8, KB_BIN_U16, KB_BIN_U32, KB_BIN_F32,
  KB_STAT_U8, KB_STAT_U16, KB_STAT_U32, KB_STAT_F32,
  KB_RED_U8, KB_RED_U16, KB_RED_U32, KB_RED_F32,
  KB_LAG_U8, KB_LAG_U16, KB_LAG_U32, KB_LAG_F32,
  KB_MERGE, KB_SEGMASK,
  KB_CLEAN_U8, KB_CLEAN_U16, KB_CLEAN_U32, KB_CLEAN_F32,
  KB_DD_U8Q, KB_DD_F32Q, KB_DD_U8, KB_DD_U8LDS, KB_DD_G_U8, KB_DD_G_U16, KB_DD_G_U32, KB_DD_G_F32F,
  KB_DD_G_F32D,
  KB_S2_V4, KB_S2_V8, KB_S2_V16, KB_S2_C8, KB_S2_C16, KB_S2_U16, KB_S2_G_U16, KB_S2_G_U32, KB_S2_G_U64, KB_S2_G_F32,
  KB_S2_F32C8, KB_S2_F32C16,
  KB_COPY, KB_ZERO,
  KB_COUNT
};

struct LaunchDesc {
  int kind = 0;
  int stage = 0;            // pipeline position: bin 0 .. mask copy 9
  uint32_t nblk = 0, threads = 0, lds = 0;
  uint32_t args_size = 0;
  alignas(16) unsigned char args[192];
};

// While set (per thread), the launchers append a LaunchDesc here instead of
// launching.
extern thread_local std::vector<LaunchDesc> *g_record;

// Records (when a recorder is set) or launches one described kernel.
hipError_t emit_desc(const LaunchDesc &d, hipStream_t st);

template <typename A>
inline hipError_t emit_args(int kind, int stage, size_t nblk, unsigned threads, size_t lds,
                            const A &args, hipStream_t st) {
  static_assert(sizeof(A) <= sizeof(LaunchDesc::args), "argument block too large");
  if (nblk == 0) return hipSuccess;
  if (nblk > 0x7FFFFFFFull) return hipErrorInvalidValue;
  LaunchDesc d;
  d.kind = kind;
  d.stage = stage;
  d.nblk = (uint32_t)nblk;
  d.threads = threads;
  d.lds = (uint32_t)lds;
  d.args_size = (uint32_t)sizeof(A);
  std::memcpy(d.args, &args, sizeof(A));
  return emit_desc(d, st);
}

// One launch of `kind` covering n argument blocks: workgroup b belongs to the
// block i with start[i] <= b < start[i+1] (tables in device memory).
hipError_t launch_batched(int kind, const void *d_tab, const uint32_t *d_start, int n,
                          uint32_t total_blocks, uint32_t lds, hipStream_t st);
// A recorded launch, launched on its own.
hipError_t launch_desc(const LaunchDesc &d, hipStream_t st);
bool batchable(int kind);

}  // namespace af

// The opaque plan (C-ABI type).
struct af_plan {
  af_plan_desc d{};
  std::vector<double> freq;
  std::vector<int> bad;
  int device = 0;

  size_t nbin = 0, nout = 0, Lp = 0, L1 = 0;
  int nch = 0, ndm = 0, nsb = 0, ndm_nom = 0, maxdelay = 0, maxres = 0;
  int ntiles1 = 0;          // tiles of the first dedispersion stage (per super-tile)
  size_t T2 = 0;            // outputs per time super-tile (subband method)
  int nsuper = 1;           // super-tiles
  int inter_elem = 0;       // bytes per subband partial sum
  int s2_tile = 0, s2_lds = 0, s2_sbc = 0, maxlds = 0;
  int s2_pk = 0;            // u16 elements of LDS per workgroup, packed rows (0: uniform rows)
  int dd_lds_rows = 0, dd_lds_win = 0;   // LDS-staged u8 dedispersion
  bool dd_q = false;                     // aligned-load u8 dedispersion (kDdTileQ tiles)
  bool dd_fq = false;                    // float32 rows-in-step dedispersion (kDdTile tiles)
  std::vector<int4> worklist;            // stage-1 workgroups in locality order
  int4 *d_worklist = nullptr;
  double adds = 0;
  std::vector<int32_t> h_delays, h_res, h_maxres_j, h_maxres_jsb, h_sboff;

  // IQRM / segmentation
  int nwin = 0, nseg = 0, nlags = 0, Csub = 0;
  int lag_reach = 0;        // largest |lag| below Csub (the vote's channel reach)
  size_t iqrm_lds = 0;
  std::vector<uint32_t> wt0, wt1, wseg;
  std::vector<af::Seg> segs;
  std::vector<af::RowItem> stat_items, clean_items;
  std::vector<int> seg_items, lags;
  // cleaning per super-tile: d_xT holds one super-tile's rows [k*T2,
  // k*T2 + Lp) (channel-major, row pitch Lp); tile_items[tile_off[k] ..
  // tile_off[k+1]) are the clean items clipped to those rows
  bool seg_is_win = false;               // segment s is window s: segmask aliases the masks
  bool clean_cut = false;                // clean_items[i].pad = channels read at its rows
  bool clean_tiled = false;
  std::vector<af::RowItem> tile_items;
  std::vector<int> tile_off;

  // device buffers
  int32_t *d_delays = nullptr, *d_res = nullptr, *d_maxres_j = nullptr, *d_sboff = nullptr;
  int *d_lags = nullptr, *d_seg_items = nullptr;
  uint32_t *d_wseg = nullptr, *d_wt0 = nullptr, *d_wt1 = nullptr;
  af::Seg *d_segs = nullptr;
  af::RowItem *d_stat_items = nullptr, *d_clean_items = nullptr, *d_tile_items = nullptr;
  uint8_t *d_static = nullptr;
  void *d_xb = nullptr, *d_xT = nullptr, *d_inter = nullptr;
  af::Part *d_part = nullptr, *d_segsum = nullptr;
  uint8_t *d_flags = nullptr, *d_masks = nullptr, *d_segmask = nullptr;
  size_t scratch_bytes = 0;

  hipStream_t stream = nullptr;
  struct CopyPool *copy_pool = nullptr;   // host-call copy workers (af_plan.hip)
  int profiling = 0;            // 0 off, 1 every stage, 2 the dedispersion stage only
  // one set of kNumEv events per profiled execute, recorded between stages
  static constexpr int kNumEv = 7;
  std::vector<hipEvent_t> ev;   // pool, kNumEv per set
  int ev_used = 0;              // sets recorded since the last collection
  double acc_ms[kNumEv - 1] = {};
  double acc_total = 0;
  int acc_n = 0;
  std::vector<void *> allocs;
  // Arena mode (af_execute_noalloc): every device buffer is carved from
  // [arena, arena + arena_cap) and the tables are copied in on arena_st;
  // with arena_measure set nothing is copied and only arena_used is counted.
  uint8_t *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  bool arena_measure = false;
  hipStream_t arena_st = nullptr;
  // af_execute_host: device buffers kept between calls (grow-only) and two
  // page-locked staging buffers with their events
  void *hx_in = nullptr, *hx_out = nullptr, *hx_masks = nullptr, *hx_img = nullptr, *hx_imgscr = nullptr;
  size_t hx_in_n = 0, hx_out_n = 0, hx_masks_n = 0, hx_img_n = 0, hx_imgscr_n = 0;
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

namespace af {
// Runs every stage of one plan on `st` (af_plan.hip).
af_status execute(af_plan *p, const void *d_in, void *d_out, uint8_t *d_masks, hipStream_t st);

// 8-bit fast paths (af_kernels_u8.hip)
constexpr int kStatRowsU8 = 256;   // rows per statistics item (4 row lanes x 64)
constexpr int kStatRowsU8Max = 1024;   // rows per plain-call statistics item (<= 256 per row lane)
constexpr int kCleanRowsU8 = 32;   // rows per clean item (1024 threads x 8 rows x 16 ch)
constexpr int kCleanU8MaxChan = 4096;
hipError_t launch_seg_stats_u8(const uint8_t *x, int nchan, int c0, int c1, const RowItem *items,
                               int nitems, uint2 *part, hipStream_t st);
hipError_t launch_seg_reduce_u8(const uint2 *part, const int *seg_items, int nseg, int Csub,
                                Part *segsum, hipStream_t st);
hipError_t launch_clean_u8(const CleanArgs &a, hipStream_t st);
hipError_t launch_stage2_u16v(const Stage2Args &a, hipStream_t st);

// Served-beam helpers (af_kernels_u8.hip).
// ring: [R][2][Csub] per-chunk sums (head rows, remaining rows).  out[c] =
// sum over chunks first..first+nfull-1 (mod R) of head+rest, plus chunk
// `last`'s head (and its rest when last_full).
hipError_t launch_ring_sum(const Part *ring, int R, int first, int nfull, int last, int last_full,
                           int Csub, Part *out, hipStream_t st);
// flags[1] = (flags[0] == 0 || mask != prev); prev = mask; flags[0] = 1.
hipError_t launch_mask_delta(const uint8_t *mask, uint8_t *prev, int n, int *flags, hipStream_t st);
// Copies each of nrows rows' first width bytes, rounded up to 16 (bases and
// pitch 16-byte aligned; the rounding may copy up to 15 bytes more per row).
hipError_t launch_copy_rows(uint8_t *dst, const uint8_t *src, size_t pitch, size_t width,
                            int nrows, hipStream_t st);
// Zeroes width bytes of each of nrows rows spaced pitch bytes apart.
hipError_t launch_zero_rows(uint8_t *base, size_t pitch, size_t width, int nrows, hipStream_t st);
// launch_zero_rows through the launch recorder (a batch replays it in order)
hipError_t launch_zero(uint8_t *base, size_t pitch, size_t width, int nrows, hipStream_t st);

// Pieces of the plan's execute (af_plan.hip), for callers that keep their
// own cleaned data: IQRM vote from p->d_segsum into p->d_masks / d_segmask,
// and both dedispersion stages from a channel-major cleaned buffer.
af_status run_iqrm_vote(af_plan *p, hipStream_t st);
// DM-time images of a device plane (af_image.hip); synchronises st.  With
// keep, the scratch is *keep (grown to *keep_n when needed) and stays
// allocated for the next call; without, it is allocated and freed here.
af_status image_slices(const void *d_plane, int dtype, size_t rows, size_t cols, size_t ld,
                       size_t sps, int size, uint8_t *d_images, hipStream_t st, void **keep,
                       size_t *keep_n);
af_status run_dedisp(af_plan *p, const void *xT, size_t Lp, void *d_out, hipStream_t st, int set,
                     int k_only = -1);
}  // namespace af
