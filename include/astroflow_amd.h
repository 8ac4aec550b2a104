/*
 * astroflow_amd.h -- C-ABI of the MI355X (gfx950) dedispersion + RFI library.
 *
 * Plain C types only (no torch, no C++).  Every entry point returns an
 * af_status (AF_OK == 0); on failure af_last_error() returns a message for the
 * calling thread.  No entry point exits the process.
 *
 * Reference interfaces replaced (lintian233/astroflow):
 *   af_plan_create + af_execute      : gpucal::dedisperse_spec<T>
 *                                      (src/core/dedispered.cu:988-1463) and
 *                                      gpucal::dedispered_fil_cuda<T> (:483-986):
 *                                      binning, IQRM, zero-DM, static mask and
 *                                      direct/subband dedispersion of one buffer.
 *   af_execute (use_iqrm)            : iqrm_cuda::rfi_iqrm_gpu<T> + RfiMarker<T>::mask
 *                                      (src/core/iqrm.cu:438-467, src/rfi/rfimarker.cu:309-390)
 *   af_execute (use_zero_dm)         : RfiMarker<T>::zero_dm_filter, range form
 *                                      (src/rfi/rfimarker.cu:259-305)
 *   af_execute (bad_chans)           : RfiMarker<T>::mark_rfi (src/rfi/rfimarker.cu:473-488)
 */
#ifndef ASTROFLOW_AMD_H
#define ASTROFLOW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int af_status;
#define AF_OK 0
#define AF_ERR_INVALID 1 /* bad argument (Python: ValueError)        */
#define AF_ERR_RUNTIME 2 /* HIP/runtime failure (Python: RuntimeError) */
#define AF_ERR_NOMEM 3   /* allocation or memory budget (MemoryError) */
#define AF_ERR_NODEV 4   /* no GPU visible (RuntimeError)             */
#define AF_ERR_BUSY 5    /* server queue full / not admitted yet      */

/* Element types of the filterbank samples. */
#define AF_U8 1
#define AF_U16 2
#define AF_U32 3
#define AF_F32 4
/* Element type of the plane of u32 samples (af_image_slices only). */
#define AF_U64 5

#define AF_METHOD_DIRECT 0
#define AF_METHOD_SUBBAND 1

/* Everything that defines one dedispersion call.  Channel ranges are
 * half-open [c0, c1) and explicit per stage, because the reference mixes an
 * inclusive range for dedispersion with exclusive ranges for IQRM and zero-DM
 * (src/core/dedispered.cu:1089,1107,1124). */
typedef struct af_plan_desc {
  int dtype;               /* AF_U8 .. AF_F32                                   */
  int nchan;               /* channels per time row of the input                */
  size_t nsamp;            /* native-rate time rows in the input buffer         */
  int ds;                  /* time downsample factor (>= 1)                     */
  double tsamp;            /* native sample time in seconds                     */
  const double *freq_mhz;  /* [nchan] channel centre frequencies, ascending     */
  double f_ref_mhz;        /* reference frequency for zero delay                */
  int dd_c0, dd_c1;        /* dedispersed channels                              */
  double dm_low, dm_step;  /* trial DMs dm_low + i*dm_step                      */
  int ndm;                 /* number of trial DMs                               */
  size_t nout;             /* output samples (binned rate); 0 = ceil(nsamp/ds)  */
  int method;              /* AF_METHOD_DIRECT or AF_METHOD_SUBBAND             */
  int sb_nchan;            /* channels per subband (subband method)             */
  int ndm0;                /* fine DMs per coarse DM group (subband method)     */
  /* IQRM (src/core/iqrm.cu) */
  int use_iqrm;
  int iqrm_c0, iqrm_c1;
  int iqrm_mode;           /* 0 mean, 1 std                                     */
  float iqrm_radius_frac;
  float iqrm_nsigma;
  double iqrm_geofactor;
  unsigned iqrm_win;       /* window length in binned samples, 0 = whole buffer */
  unsigned iqrm_hop;       /* hop in binned samples, 0 = window length          */
  int iqrm_include_tail;
  /* zero-DM (src/rfi/rfimarker.cu:172-305) */
  int use_zero_dm;
  int zdm_c0, zdm_c1;
  /* static bad-channel mask (applied after zero-DM) */
  const int *bad_chans;
  int n_bad;
  /* kernel tuning; 0 = library default (the result does not depend on these) */
  int tile_t;              /* stage-2 output tile: 64..1024 (0 = chosen by LDS)  */
  int dm_per_block;        /* u8 stage-1 form: 0 = 8 DM rows per workgroup, each
                              loading its windows through the L1; 16 = 16 rows
                              sharing one LDS window per channel              */
  size_t inter_budget;     /* bytes of subband intermediate per time super-tile
                              (0 = 1 GiB); longer outputs run in super-tiles  */
  size_t clean_budget;     /* bytes of the cleaned channel-major buffer (subband
                              method with several super-tiles): when the whole
                              buffer exceeds it, each super-tile's rows (plus
                              the DM-overlap reach) are cleaned just before its
                              dedispersion into a buffer of one super-tile.
                              0 = whole buffer when it fits the free device
                              memory, else per super-tile                     */
} af_plan_desc;

typedef struct af_plan af_plan;

typedef struct af_plan_info {
  size_t nbin;             /* binned rows = ceil(nsamp/ds)                      */
  size_t nout;             /* output samples per DM row                         */
  int ndm;
  int out_elem_bytes;      /* 4 (u32 for u8/u16, f32 for f32) or 8 (u64 for u32) */
  int nwin;                /* IQRM windows (0 without IQRM)                     */
  int nsb;                 /* subbands (subband method)                         */
  int ndm_nom;             /* coarse DMs (subband method)                       */
  int max_delay;           /* largest delay in binned samples over all DMs      */
  size_t scratch_bytes;    /* device memory held by the plan                    */
  double adds;             /* integer/float adds the method performs per call   */
  size_t inter_bytes;      /* subband intermediate written by stage 1 per call  */
  int clean_passes;        /* 1: one clean pass over the whole buffer; >1: one per
                              super-tile (see clean_budget)                    */
  int stage1_form;         /* dedispersion kernel of the first stage: 8 = u8, 8
                              DM rows in step loading through the L1; 16 = u8,
                              16 rows sharing an LDS window; 32 = float32 rows
                              in step; 0 = the generic kernel                  */
} af_plan_info;

/* Per-stage device time, averaged over the n executes run since profiling
 * was switched on (milliseconds; HIP events recorded on the execute stream
 * between the stages). */
typedef struct af_stage_times {
  float bin, stats, iqrm, clean, dedisp, stage2, total;
  int n;
} af_stage_times;

af_status af_plan_create(const af_plan_desc *desc, af_plan **plan);
af_status af_plan_destroy(af_plan *plan);
af_status af_plan_query(const af_plan *plan, af_plan_info *info);
/* Copy of the delay table used by the first dedispersion stage
 * ([ndm][dd_c1-dd_c0] direct, [ndm_nom][...] subband) and the residual table
 * ([ndm][nsb], subband only).  Either pointer may be NULL. */
af_status af_plan_tables(const af_plan *plan, int32_t *delays, int32_t *residuals);
/* IQRM window bounds [t0,t1) in binned samples, nwin entries each. */
af_status af_plan_windows(const af_plan *plan, uint32_t *t0s, uint32_t *t1s);
/* Stage timing with HIP events on the execute stream, reset on each call:
 * enable 0 off, 1 every stage (seven events per execute), 2 the dedispersion
 * stage only (two events per execute; the other stage times read 0).  Each
 * recorded event idles the GPU for a few microseconds at its stage boundary. */
af_status af_plan_set_profiling(af_plan *plan, int enable);
/* Synchronises with the recorded events and returns the averages. */
af_status af_plan_stage_times(af_plan *plan, af_stage_times *t);

/* Run one call on device buffers (stream may be NULL = null stream).
 *   d_in   : [nsamp][nchan] samples, time-major (as read from a filterbank)
 *   d_out  : [ndm][nout] plane, DM-major (u32 / u64 / f32)
 *   d_masks: optional [nwin][nchan] u8 IQRM masks (NULL to skip)
 * The input buffer is not modified.  Asynchronous on the stream. */
af_status af_execute(af_plan *plan, const void *d_in, void *d_out,
                     uint8_t *d_masks, void *stream);

/* Single-DM spectrum (replaces cpucal::dedisperse_spec_with_dm<T>,
 * include/cpucal.hpp:341-483): the plan's RFI stages on d_in, then
 *   d_out[t][c - dd_c0] = cleaned[t0 + t + delay(c)][c],  t < t_len,
 * channels [dd_c0, dd_c1), delay(c) the plan's delay of its one DM
 * (a direct plan with ndm == 1 and ds == 1; f_ref the highest selected
 * channel).  d_out has the input's element type, [t_len][dd_c1-dd_c0].
 * t0 + t_len + max_delay must not exceed nsamp. */
af_status af_execute_spectrum(af_plan *plan, const void *d_in, void *d_out, size_t t0,
                              size_t t_len, uint8_t *d_masks, void *stream);

/* Boxcar S/N of a dynamic spectrum on the device (replaces the host function
 * behind _calculate_frb_snr, src/bindpy/frb_snr.cpp:144-398, for a spectrum
 * already in HBM such as af_execute_spectrum's output).  d_spec is
 * [ntime][nfreq] of dtype (AF_U8 .. AF_F32; values are taken as float32, as
 * the reference's float array).  The rows are summed over frequency in
 * double, decimated to target_time_us when tsamp and the target are given,
 * the noise is a sigma-clipped median/MAD estimate (over noise_ranges, or
 * off the TOA window, or the whole series), boxcars of widths 1..16 and up
 * to 32 log-spaced wider ones are slid over the series (centres within the
 * search radius of the TOA when given), and the best window's noise is
 * re-estimated without it.  Synchronises the stream. */
typedef struct af_snr_params {
  double threshold_sigma;  /* clipping threshold (5.0)                          */
  int has_toa, toa;        /* TOA (native-rate sample index)                    */
  int has_window, window;  /* fitting window (native-rate samples)              */
  int has_tsamp, has_target;
  double tsamp;            /* seconds                                           */
  double target_time_us;   /* decimate to this resolution                       */
  int has_noise_ranges;    /* a list was given (even with no usable pair)       */
  int n_noise_ranges;
  const int *noise_ranges; /* host [n][2]: native-rate [start, end)             */
} af_snr_params;

typedef struct af_snr_result {
  double snr;              /* -1 when there is no data                          */
  int width, peak;         /* native-rate samples                               */
  double noise_mean, noise_std;
  int method;              /* 0 no data, 1 boxcar, 2 fallback (peak sample)     */
  int converged;
  int search_start, search_end, downsample;
} af_snr_result;

af_status af_frb_snr(const void *d_spec, int dtype, int ntime, int nfreq,
                     const af_snr_params *params, af_snr_result *result, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  af_plan_create + af_execute + af_plan_destroy in one
 * call that allocates nothing: the plan is built on host memory and all of
 * its device tables and scratch buffers (cleaned channel-major buffer,
 * subband intermediate, IQRM partials and masks) are carved from a static
 * device array of AF_NOALLOC_SCRATCH_BYTES that the library holds on each
 * device it runs on until the process ends.  Tables are copied in with
 * hipMemcpyAsync on `stream` from host memory; stale contents are cleared
 * by kernels on `stream`.  A call whose plan needs more than the array
 * fails with AF_ERR_NOMEM before anything is launched.
 * Every call on one device shares that device's array and leaves its last
 * contents there: calls that overlap on one device (different streams or
 * threads) race; calls in sequence on one stream do not.  Results equal
 * af_execute of af_plan_create(desc) on the same buffers. */
#define AF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
af_status af_execute_noalloc(const af_plan_desc *desc, const void *d_in, void *d_out,
                             uint8_t *d_masks, void *stream);

/* Same on host buffers: upload, execute, download, synchronise.  The plan
 * keeps device buffers for the input and the plane from its first host call
 * on (grown when needed).  Pageable data moves through two 32 MB page-locked
 * staging buffers, copied by the plan's own host threads; a buffer that lies
 * wholly inside one block from af_host_alloc is transferred directly.  All
 * are freed by af_plan_destroy. */
af_status af_execute_host(af_plan *plan, const void *h_in, void *h_out,
                          uint8_t *h_masks);

/* DM-time images of a device plane (replaces the preprocessing that turns
 * the plane into DedispersedDataUint8 images, include/data.h:619-727, and
 * its OpenCV calls cv::resize INTER_AREA / INTER_LANCZOS4, cv::normalize
 * NORM_MINMAX, cv::applyColorMap COLORMAP_VIRIDIS):
 *   d_plane : [rows][ld] plane elements, plane_dtype AF_U32 / AF_U64 /
 *             AF_F32, columns [0, cols) used
 *   sps     : columns per slice; nslices = ceil(cols / sps), the last one
 *             cols - (nslices - 1) * sps wide
 *   size    : image side (512 in the reference)
 *   d_images: [nslices][size][size][3] bytes, BGR
 * Each slice is resized to size x size (area averaging when both of its
 * sides shrink, 8-tap Lanczos a = 4 otherwise), min-max normalised to 0..255
 * (round half to even; a flat slice maps to 0) and mapped through viridis.
 * Allocates its scratch and synchronises the stream before returning. */
af_status af_image_slices(const void *d_plane, int plane_dtype, size_t rows, size_t cols,
                          size_t ld, size_t sps, int size, uint8_t *d_images, void *stream);

/* af_execute_host whose result is the DM-time images of the plane instead of
 * the plane (the drop-in for _dedisperse_spec, src/bindpy/pyapi.cpp:286-309):
 * the plane stays on the device, only h_images ([nslices][size][size][3],
 * nslices = ceil(nout / sps)) and the optional masks are downloaded. */
af_status af_execute_host_images(af_plan *plan, const void *h_in, size_t sps, int size,
                                 uint8_t *h_images, uint8_t *h_masks);

/* Properties of a device as the kernels see them. */
typedef struct af_device_props {
  char name[64];
  char arch[32];           /* e.g. "gfx950" */
  int cu_count;
  int clock_khz;
  int lds_per_block;       /* largest per-workgroup LDS the plans may use */
  int lds_per_cu;
  int l2_bytes;
  size_t hbm_bytes;
} af_device_props;

/* Device helpers so callers need no other GPU runtime. */
af_status af_device_query(int dev, af_device_props *props);
af_status af_device_count(int *n);
af_status af_set_device(int dev);
af_status af_malloc(void **p, size_t bytes);
af_status af_free(void *p);
af_status af_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
af_status af_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
af_status af_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
af_status af_memset(void *dst, int value, size_t bytes, void *stream);
af_status af_stream_create(void **stream);
af_status af_stream_destroy(void *stream);
af_status af_stream_sync(void *stream);
af_status af_sync(void);
/* Elapsed milliseconds of `reps` back-to-back af_execute calls on the plan's
 * stream, bracketed by HIP events on that stream. */
af_status af_time_execute(af_plan *plan, const void *d_in, void *d_out, int reps,
                          float *ms);

/* --------------------------------------------------------------- batch --
 * Many plans (beams, chunks of different nchan, nsamp, DM range, dtype)
 * executed together.  Replaces a loop of per-chunk calls of
 * gpucal::dedisperse_spec<T> (src/core/dedispered.cu:988-1463), each stage a
 * separate launch, with one launch per stage and kernel for the whole batch:
 * the launches every plan's af_execute would make are merged, each
 * workgroup finding its plan in a table.  Results equal af_execute of each
 * plan on its buffers.  The buffers are bound at creation; plans and
 * buffers must outlive the batch, and a plan must not run elsewhere while
 * the batch runs (they share the plan's scratch).  Each entry needs a plan
 * of its own: the same plan given twice is rejected (AF_ERR_INVALID), since
 * both entries would write its scratch at once. */
typedef struct af_batch af_batch;

/* d_masks may be NULL (or hold NULL entries). */
af_status af_batch_create(af_plan *const *plans, int nplans, const void *const *d_ins,
                          void *const *d_outs, uint8_t *const *d_masks, af_batch **batch);
/* use_graph != 0: the launch sequence is captured into a hipGraph on the
 * first call and replayed after.  Asynchronous on the stream. */
af_status af_batch_execute(af_batch *batch, int use_graph, void *stream);
/* The batch's plans, each with its own launches in its own order (nothing
 * merged), spread over nstreams (1..8) streams of the batch (largest plans
 * first, each onto the least loaded stream), forked from `stream` and
 * joined back to it, so
 * that different chunks' stages run side by side on the device (the row
 * passes of one beside the dedispersion of another).  use_graph != 0: the
 * whole fork/join is captured into a hipGraph on the first call (again
 * when nstreams changes) and replayed after.  Keep one nstreams per batch:
 * after a recapture the runtime's streams for the new graph's branches
 * can share hardware queues (measured 4.6 ms instead of 3.8 ms on 64
 * small chunks); the form without a graph is as fast on those chunks.  Asynchronous on the stream;
 * results equal the plans' own executes.  Replaces, like af_batch_execute,
 * the reference's one chunk per call (python/astroflow/search.py:365,464). */
af_status af_batch_execute_streams(af_batch *batch, int nstreams, int use_graph, void *stream);
/* Launches per execute of the batch, and of the plans run one by one. */
af_status af_batch_info(const af_batch *batch, int *launches, int *recorded);
af_status af_batch_destroy(af_batch *batch);

/* ----------------------------------------------------------- autotuner --
 * Replaces the reference's compile-time knobs AF_SUBBAND_SIZE_CH,
 * AF_SUBBAND_NDM0 and AF_SUBBAND_TBLOCK (src/core/dedispered.cu:25-39) with
 * a measured choice.  Times candidate plans on the caller's buffers:
 *  - always: the stage-1 form for u8 (dm_per_block 0 = 8 DM rows per
 *    workgroup loading their windows through the L1, 16 = 16 rows sharing
 *    an LDS window) and the stage-2 output tile (0 = the plan's own choice,
 *    256, 512, 1024 for u8; 0, 512, 1024 for float32).  These do not change
 *    the plane (tested);
 *  - when tune_subbands != 0, also the subband layout: sb_nchan {16, 32,
 *    64, 128} x ndm0 {8, 16, 32, 64}.  This changes the method's rounding
 *    of delays and therefore the plane.
 * Writes the fastest settings into *desc (tile_t, dm_per_block, sb_nchan,
 * ndm0) and returns the best time per execute.  Results are kept for the
 * life of the process, keyed by every field of *desc except the tuned ones,
 * the frequency table, the static mask and the device. */
af_status af_autotune(af_plan_desc *desc, const void *d_in, void *d_out, int tune_subbands,
                      int reps, float *best_ms);

/* ------------------------------------------------------------- ingest --
 * File readers that deliver rows as the plans and the server take them:
 * time-major, channels in ascending frequency (a descending band is flipped,
 * and the info then describes the ascending band).  Rows can be read
 * straight into page-locked memory (af_host_alloc), which the server then
 * uploads asynchronously (af_server_submit). */
typedef struct af_fil_info {
  int nchans, nbits;
  int dtype;               /* AF_U8 / AF_U16 / AF_U32 / AF_F32 of the rows read  */
  int npol;                /* PSRFITS: polarisations stored (0 for SIGPROC)      */
  int flipped;             /* the stored band descended                          */
  uint64_t ndata;          /* rows (time samples)                                */
  uint64_t data_offset;    /* byte offset of the first row (of the SUBINT table)*/
  double fch1, foff;       /* MHz, ascending after the flip                      */
  double tsamp;            /* s                                                  */
  double tstart;           /* MJD                                                */
} af_fil_info;

/* SIGPROC filterbank (replaces Filterbank::read_header, the data read and
 * the band flip of src/data/filterbank.cpp:319-472,793-842). */
typedef struct af_fil_reader af_fil_reader;
af_status af_fil_open(const char *path, af_fil_reader **reader, af_fil_info *info);
/* Rows [row0, row0 + nrows) into dst ([nrows][nchans] of nbits/8-byte
 * samples), with positioned reads over up to 16 threads. */
af_status af_fil_read(af_fil_reader *reader, uint64_t row0, size_t nrows, void *dst);
af_status af_fil_close(af_fil_reader *reader);

/* PSRFITS search mode (replaces PsrFits._load_data,
 * python/astroflow/io/psrfits.py:32-76): the SUBINT table's DATA column
 * ([NSBLK][NPOL][NCHAN] per row; 8-bit 'B', unsigned 16-bit 'I' with TZERO
 * 32768, or float 'E'), pol 0 or the mean of pols 0 and 1 for NPOL >= 2
 * (floor((a + b) / 2) for integers: the reference's uint8/uint16 sum wraps,
 * this one does not), fch1 = OBSFREQ - OBSBW/2, foff = CHAN_BW, tsamp = TBIN,
 * tstart = STT_IMJD + (STT_SMJD + STT_OFFS) / 86400. */
typedef struct af_psrfits_reader af_psrfits_reader;
af_status af_psrfits_open(const char *path, af_psrfits_reader **reader, af_fil_info *info);
/* Rows [row0, row0 + nrows) into dst ([nrows][nchans] of info.dtype). */
af_status af_psrfits_read(af_psrfits_reader *reader, uint64_t row0, size_t nrows, void *dst);
af_status af_psrfits_close(af_psrfits_reader *reader);

/* Page-locked host memory (hipHostMalloc), for asynchronous uploads.  The
 * library remembers each block until af_host_free, so that host calls on
 * any range inside it skip the staging copy. */
af_status af_host_alloc(void **p, size_t bytes);
af_status af_host_free(void *p);

/* -------------------------------------------------------------- server --
 * Real-time chunk server.  Replaces the reference's per-chunk call loop
 * (python/astroflow/search.py:365,464 calling _dedisperse_spec on chunk +
 * DM-overlap tail read from the file each time, src/core/dedispered.cu:1040
 * uploading the whole window per call) with resident state:
 *  - each beam keeps its DM-overlap carry (the last window - chunk rows) in
 *    HBM, so only the chunk's new rows cross PCIe;
 *  - a server thread admits submitted chunks while the HBM budget (beam
 *    buffers + result planes not yet released) and the in-flight limit allow,
 *    dispatches them on per-beam streams so that work of different beams
 *    overlaps on the device, and timestamps completion;
 *  - every result equals a plain af_execute of the beam's plan on the same
 *    window of rows. */
typedef struct af_server af_server;

typedef struct af_server_desc {
  size_t hbm_budget;       /* bytes for beam buffers + result planes; 0 = 90% of free memory */
  int max_inflight;        /* chunks dispatched and not finished (0 = 16)              */
  int max_queue;           /* submitted chunks waiting for admission (0 = 4096)        */
  int full_rfi;            /* 1: every window runs the whole plan (no incremental RFI) */
} af_server_desc;

/* One finished chunk.  Chunks submitted before a beam's first window is full
 * only extend the carry: they finish with has_plane = 0. */
typedef struct af_server_result {
  int beam;
  int has_plane;
  int status;              /* AF_OK or the error of this chunk                          */
  uint64_t tag;            /* caller's tag from af_server_submit                        */
  uint64_t window;         /* output chunk index j: the window starts at beam row j*chunk */
  void *d_plane;           /* [ndm][nout] result on the device; valid until released    */
  uint8_t *d_masks;        /* [nwin][nchan] IQRM masks on the device (or NULL)          */
  double t_submit_ms;      /* server clock: submit, admission, completion               */
  double t_admit_ms;
  double t_done_ms;
} af_server_result;

typedef struct af_server_stats {
  uint64_t submitted, admitted, completed, planes;
  double p50_ms, p99_ms, max_ms, mean_ms;   /* submit -> result ready, chunks with a plane */
  double first_submit_ms, last_done_ms;     /* server clock                               */
  size_t hbm_used, hbm_peak, hbm_budget;
  int inflight_peak;
} af_server_stats;

af_status af_server_create(const af_server_desc *desc, af_server **srv);
/* Stops the server thread after the work already submitted and frees everything. */
af_status af_server_destroy(af_server *srv);
/* Adds a beam.  window->nsamp is the window length W = chunk + DM-overlap
 * tail rows; window->nout is the chunk's output length (ds must be 1).
 * chunk_rows new rows arrive per chunk (chunk_rows = window->nout).  The
 * beam's buffers count against the budget (AF_ERR_NOMEM when they do not
 * fit).  Not allowed while chunks are in flight (AF_ERR_BUSY). */
af_status af_server_add_beam(af_server *srv, const af_plan_desc *window, size_t chunk_rows,
                             int *beam);
/* Queues the beam's next chunk_rows rows ([chunk_rows][nchan], time-major).
 * The rows are borrowed, not copied: keep them unchanged until this chunk's
 * result comes back from af_server_poll.  They may be host memory
 * (page-locked memory from af_host_alloc makes the upload asynchronous) or
 * device memory (rows written there by a capture path: a device copy).
 * AF_ERR_BUSY: queue full. */
af_status af_server_submit(af_server *srv, int beam, const void *h_rows, uint64_t tag);
/* Returns up to max finished chunks, in completion order; waits up to
 * wait_ms for the first one. */
af_status af_server_poll(af_server *srv, af_server_result *out, int max, int *n, int wait_ms);
/* Gives a result plane (and its masks) back to the pool.  Only a plane that
 * af_server_poll returned and that has not been released since is accepted
 * (AF_ERR_INVALID otherwise). */
af_status af_server_release(af_server *srv, void *d_plane);
/* Waits until every submitted chunk has finished. */
af_status af_server_drain(af_server *srv);
af_status af_server_get_stats(af_server *srv, af_server_stats *st);

const char *af_last_error(void);
const char *af_version(void);

#ifdef __cplusplus
}
#endif
#endif
