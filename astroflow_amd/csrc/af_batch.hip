// af_batch.hip -- many plans (beams, chunks) run as one launch per stage.
//
// The reference runs one chunk per call, each stage a separate launch
// (src/core/dedispered.cu:988-1463).  A batch records the launches every
// plan's execute would make (af_execute with a recorder set), then merges
// the launches of the same stage and kernel across plans into one: the
// merged launch carries a table of argument blocks and the prefix sums of
// their workgroup counts, and each workgroup finds its plan by a binary
// search over that prefix (af_kernels*.hip, AF_BATCHED).  Stages stay
// ordered (bin, IQRM statistics, reduction, lags, vote, segment masks,
// clean, dedispersion stage 1 and 2, mask copies), so every plan sees the
// same dependencies as in its own execute; results are identical.  The
// launch sequence can be captured once into a hipGraph and replayed.
//
// The other way to run a batch keeps each plan's launches in its own order
// and spreads the plans over a few streams (af_batch_execute_streams):
// different chunks' stages then run side by side on the device, the row
// passes (HBM-bound) of one beside the dedispersion (L1-bound) of another.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "af_internal.h"

namespace af {
namespace {

struct Merged {
  int kind = 0;
  bool batched = false;
  LaunchDesc single;              // when not batched
  size_t tab_off = 0, start_off = 0;
  int n = 0;
  uint32_t total = 0, lds = 0;
};

}  // namespace
}  // namespace af

struct af_batch {
  int device = 0;
  std::vector<af_plan *> plans;
  std::vector<af::Merged> launches;
  std::vector<std::vector<af::LaunchDesc>> own;   // each plan's launches, in its order
  int recorded = 0;               // launches the plans would make one by one
  uint8_t *d_tables = nullptr;
  hipStream_t cap = nullptr;      // capture stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // per-plan streams: the side streams, fork / join events, and the graph
  // of the fork/join captured for `side_graph_n` streams
  std::vector<hipStream_t> side;
  std::vector<int> side_of;       // plan -> side stream, for side_of_n streams
  int side_of_n = 0;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t side_graph = nullptr;
  hipGraphExec_t side_exec = nullptr;
  int side_graph_n = 0;
};

namespace af {
namespace {

af_status run_launches(af_batch *b, hipStream_t st) {
  for (const Merged &m : b->launches) {
    if (m.batched)
      AF_HIP(launch_batched(m.kind, b->d_tables + m.tab_off,
                            (const uint32_t *)(b->d_tables + m.start_off), m.n, m.total, m.lds, st));
    else
      AF_HIP(launch_desc(m.single, st));
  }
  return AF_OK;
}

void release(af_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->side_exec) (void)hipGraphExecDestroy(b->side_exec);
  if (b->side_graph) (void)hipGraphDestroy(b->side_graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  if (b->d_tables) (void)hipFree(b->d_tables);
}

// Plans onto n side streams, largest first onto the least loaded (a plan's
// load: the threads its launches start).
void assign_side(af_batch *b, int n) {
  const size_t np = b->own.size();
  std::vector<double> load(np, 0.0);
  for (size_t i = 0; i < np; ++i)
    for (const LaunchDesc &d : b->own[i]) load[i] += (double)d.nblk * (double)d.threads;
  std::vector<size_t> order(np);
  for (size_t i = 0; i < np; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return load[x] > load[y]; });
  std::vector<double> acc((size_t)n, 0.0);
  b->side_of.assign(np, 0);
  for (size_t i : order) {
    const int k = (int)(std::min_element(acc.begin(), acc.end()) - acc.begin());
    b->side_of[i] = k;
    acc[(size_t)k] += load[i];
  }
  b->side_of_n = n;
}

// The plans' own launch sequences on n side streams (assign_side), forked
// from `st` and joined back to it.
af_status run_side(af_batch *b, int n, hipStream_t st) {
  if (b->side_of_n != n) assign_side(b, n);
  AF_HIP(hipEventRecord(b->fork, st));
  for (int k = 0; k < n; ++k) AF_HIP(hipStreamWaitEvent(b->side[k], b->fork, 0));
  for (size_t i = 0; i < b->own.size(); ++i)
    for (const LaunchDesc &d : b->own[i]) AF_HIP(launch_desc(d, b->side[(size_t)b->side_of[i]]));
  for (int k = 0; k < n; ++k) {
    AF_HIP(hipEventRecord(b->join[k], b->side[k]));
    AF_HIP(hipStreamWaitEvent(st, b->join[k], 0));
  }
  return AF_OK;
}

}  // namespace
}  // namespace af

using namespace af;

namespace af {
// workgroups from which one launch alone keeps every CU busy (256 CUs x a
// few resident workgroups): such launches are not merged
constexpr uint32_t kBatchFillGrid = 2048;
// side streams of af_batch_execute_streams (the device runs 4 hardware
// queues per process by default; more streams share them)
constexpr int kBatchMaxStreams = 8;
}  // namespace af

extern "C" {

af_status af_batch_create(af_plan *const *plans, int nplans, const void *const *d_ins,
                          void *const *d_outs, uint8_t *const *d_masks, af_batch **batch) {
  if (!plans || nplans <= 0 || !d_ins || !d_outs || !batch)
    return fail(AF_ERR_INVALID, "null argument");
  *batch = nullptr;
  for (int i = 0; i < nplans; ++i)
    if (!plans[i] || !d_ins[i] || !d_outs[i]) return fail(AF_ERR_INVALID, "null plan or buffer");
  // each entry runs on its plan's own scratch (cleaned buffer, intermediate,
  // IQRM partials and masks): a plan given twice would have both entries'
  // merged launches write the same scratch at once
  {
    std::vector<af_plan *> seen(plans, plans + nplans);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
      return fail(AF_ERR_INVALID, "a plan appears more than once in the batch (one plan per "
                                  "entry: same-shape chunks need a plan each)");
  }
  af_batch *b = new af_batch();
  b->plans.assign(plans, plans + nplans);
  (void)hipGetDevice(&b->device);
  // 1. every plan's launches, in its own order
  std::vector<std::vector<LaunchDesc>> recs(nplans);
  for (int i = 0; i < nplans; ++i) {
    af_plan *p = plans[i];
    if (p->device != b->device) {
      delete b;
      return fail(AF_ERR_INVALID, "all plans of a batch must be on the current device");
    }
    const int prof = p->profiling;
    p->profiling = 0;
    g_record = &recs[i];
    af_status s = execute(p, d_ins[i], d_outs[i], d_masks ? d_masks[i] : nullptr, nullptr);
    g_record = nullptr;
    p->profiling = prof;
    if (s != AF_OK) {
      delete b;
      return s;
    }
    b->recorded += (int)recs[i].size();
  }
  // 2. merge by (stage, kind), stages in order, kinds in order of appearance.
  // A plan whose output runs in several time super-tiles launches stage 1
  // and stage 2 once per super-tile, reusing its intermediate buffer: those
  // launches stay single and in the plan's own order (placed after the
  // merged stage-2 launches; plans are independent of each other).
  std::vector<uint8_t> host;
  auto put = [&](const void *src, size_t n) {
    size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + n);
    std::memcpy(host.data() + off, src, n);
    return off;
  };
  // A plan that cleans per super-tile reuses its cleaned buffer between
  // those passes too: all of its launches stay single, in its own order,
  // after everything else.
  std::vector<bool> tiled(nplans, false), ordered(nplans, false);
  for (int i = 0; i < nplans; ++i) {
    int n7 = 0;
    for (auto &d : recs[i]) n7 += d.stage == 7;
    tiled[i] = n7 > 1;
    ordered[i] = plans[i]->clean_tiled;
  }
  auto emit_group = [&](int stage, int kind) -> af_status {
    std::vector<const LaunchDesc *> group;
    for (int i = 0; i < nplans; ++i)
      for (auto &d : recs[i])
        if (!ordered[i] && d.stage == stage && d.kind == kind &&
            !(tiled[i] && (stage == 7 || stage == 8)))
          group.push_back(&d);
    if (group.empty()) return AF_OK;
    // a launch that fills the machine by itself runs alone: merging it only
    // adds the table lookup to each of its workgroups (measured ~3% slower on
    // large chunks); the small launches of a stage are merged
    std::vector<const LaunchDesc *> small;
    for (auto *d : group) {
      if (!batchable(kind) || d->nblk >= kBatchFillGrid) {
        Merged m;
        m.kind = kind;
        m.single = *d;
        b->launches.push_back(m);
      } else {
        small.push_back(d);
      }
    }
    group.swap(small);
    if (group.empty()) return AF_OK;
    if (group.size() < 2) {
      Merged m;
      m.kind = kind;
      m.single = *group[0];
      b->launches.push_back(m);
      return AF_OK;
    }
    Merged m;
    m.kind = kind;
    m.batched = true;
    m.n = (int)group.size();
    std::vector<uint8_t> tab;
    std::vector<uint32_t> start(1, 0);
    for (auto *d : group) {
      tab.insert(tab.end(), d->args, d->args + d->args_size);
      const uint64_t nx = (uint64_t)start.back() + d->nblk;
      if (nx > 0x7FFFFFFFull) return fail(AF_ERR_INVALID, "batched launch exceeds 2^31 workgroups");
      start.push_back((uint32_t)nx);
      m.lds = std::max(m.lds, d->lds);
    }
    m.total = start.back();
    m.tab_off = put(tab.data(), tab.size());
    m.start_off = put(start.data(), start.size() * sizeof(uint32_t));
    b->launches.push_back(m);
    return AF_OK;
  };
  for (int stage = 0; stage <= 9; ++stage) {
    std::vector<int> kinds;
    for (auto &r : recs)
      for (auto &d : r)
        if (d.stage == stage && std::find(kinds.begin(), kinds.end(), d.kind) == kinds.end())
          kinds.push_back(d.kind);
    for (int kind : kinds) {
      af_status s = emit_group(stage, kind);
      if (s != AF_OK) {
        release(b);
        delete b;
        return s;
      }
    }
    if (stage == 8)
      for (int i = 0; i < nplans; ++i)
        if (tiled[i] && !ordered[i])
          for (auto &d : recs[i])
            if (d.stage == 7 || d.stage == 8) {
              Merged m;
              m.kind = d.kind;
              m.single = d;
              b->launches.push_back(m);
            }
  }
  for (int i = 0; i < nplans; ++i)
    if (ordered[i])
      for (auto &d : recs[i]) {
        Merged m;
        m.kind = d.kind;
        m.single = d;
        b->launches.push_back(m);
      }
  b->own = std::move(recs);
  if (!host.empty()) {
    if (hipMalloc(&b->d_tables, host.size()) != hipSuccess) {
      release(b);
      delete b;
      return fail(AF_ERR_NOMEM, "hipMalloc batch tables");
    }
    if (hipMemcpy(b->d_tables, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
      release(b);
      delete b;
      return fail(AF_ERR_RUNTIME, "upload of batch tables failed");
    }
  }
  *batch = b;
  return AF_OK;
}

af_status af_batch_execute(af_batch *b, int use_graph, void *stream) {
  if (!b) return fail(AF_ERR_INVALID, "null argument");
  AF_HIP(hipSetDevice(b->device));
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return run_launches(b, st);
  if (!b->exec) {
    AF_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    AF_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    af_status s = run_launches(b, b->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (s != AF_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (e != hipSuccess) return fail(AF_ERR_RUNTIME, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph = g;
    AF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  AF_HIP(hipGraphLaunch(b->exec, st));
  return AF_OK;
}

af_status af_batch_execute_streams(af_batch *b, int nstreams, int use_graph, void *stream) {
  if (!b) return fail(AF_ERR_INVALID, "null argument");
  if (nstreams < 1 || nstreams > kBatchMaxStreams)
    return fail(AF_ERR_INVALID, "nstreams must be 1.." + std::to_string(kBatchMaxStreams));
  AF_HIP(hipSetDevice(b->device));
  hipStream_t st = (hipStream_t)stream;
  const int n = std::min<int>(nstreams, (int)b->own.size());
  while ((int)b->side.size() < n) {
    hipStream_t s = nullptr;
    AF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b->side.push_back(s);
    hipEvent_t e = nullptr;
    AF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->join.push_back(e);
  }
  if (!b->fork) AF_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  if (!use_graph) return run_side(b, n, st);
  if (b->side_exec && b->side_graph_n != n) {
    (void)hipGraphExecDestroy(b->side_exec);
    (void)hipGraphDestroy(b->side_graph);
    b->side_exec = nullptr;
    b->side_graph = nullptr;
  }
  if (!b->side_exec) {
    if (!b->cap) AF_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    AF_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    af_status s = run_side(b, n, b->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (s != AF_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (e != hipSuccess) return fail(AF_ERR_RUNTIME, std::string("graph capture: ") + hipGetErrorString(e));
    b->side_graph = g;
    AF_HIP(hipGraphInstantiate(&b->side_exec, b->side_graph, nullptr, nullptr, 0));
    b->side_graph_n = n;
  }
  AF_HIP(hipGraphLaunch(b->side_exec, st));
  return AF_OK;
}

af_status af_batch_info(const af_batch *b, int *launches, int *recorded) {
  if (!b) return fail(AF_ERR_INVALID, "null argument");
  if (launches) *launches = (int)b->launches.size();
  if (recorded) *recorded = b->recorded;
  return AF_OK;
}

af_status af_batch_destroy(af_batch *b) {
  if (!b) return AF_OK;
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  release(b);
  delete b;
  return AF_OK;
}

}  // extern "C"

// ------------------------------------------------------------ autotuner --
namespace af {
namespace {

// Everything that shapes a plan and its timing except the settings being
// tuned: the scalar fields one by one (no struct padding), the frequency
// table, the static mask and the device.  tile_t and dm_per_block are
// outputs; sb_nchan and ndm0 too when the subband layout is tuned.
std::string tune_key(const af_plan_desc &d, int tune_sb, int dev) {
  std::string k;
  auto put = [&](const auto &v) { k.append((const char *)&v, sizeof v); };
  put(d.dtype); put(d.nchan); put(d.nsamp); put(d.ds); put(d.tsamp); put(d.f_ref_mhz);
  put(d.dd_c0); put(d.dd_c1); put(d.dm_low); put(d.dm_step); put(d.ndm); put(d.nout);
  put(d.method);
  const int sbw = tune_sb ? 0 : d.sb_nchan, nd0 = tune_sb ? 0 : d.ndm0;
  put(sbw); put(nd0);
  put(d.use_iqrm); put(d.iqrm_c0); put(d.iqrm_c1); put(d.iqrm_mode); put(d.iqrm_radius_frac);
  put(d.iqrm_nsigma); put(d.iqrm_geofactor); put(d.iqrm_win); put(d.iqrm_hop);
  put(d.iqrm_include_tail);
  put(d.use_zero_dm); put(d.zdm_c0); put(d.zdm_c1);
  put(d.inter_budget); put(d.clean_budget);
  put(tune_sb); put(dev);
  if (d.freq_mhz) k.append((const char *)d.freq_mhz, (size_t)std::max(0, d.nchan) * sizeof(double));
  put(d.n_bad);
  if (d.bad_chans && d.n_bad > 0) k.append((const char *)d.bad_chans, (size_t)d.n_bad * sizeof(int));
  return k;
}

struct TuneVal {
  int tile_t, dm_per_block, sb_nchan, ndm0;
  float ms;
};

std::mutex g_tune_mu;
std::map<std::string, TuneVal> g_tune;

}  // namespace
}  // namespace af

extern "C" af_status af_autotune(af_plan_desc *desc, const void *d_in, void *d_out,
                                 int tune_subbands, int reps, float *best_ms) {
  if (!desc || !d_in || !d_out) return fail(AF_ERR_INVALID, "null argument");
  if (reps < 1) reps = 1;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const std::string key = tune_key(*desc, tune_subbands, dev);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      desc->tile_t = it->second.tile_t;
      desc->dm_per_block = it->second.dm_per_block;
      desc->sb_nchan = it->second.sb_nchan;
      desc->ndm0 = it->second.ndm0;
      if (best_ms) *best_ms = it->second.ms;
      return AF_OK;
    }
  }
  const bool u8 = desc->dtype == AF_U8;
  const bool f32 = desc->dtype == AF_F32;
  const bool sub = desc->method == AF_METHOD_SUBBAND;
  // stage-1 form (u8): 0 = DM rows walking their windows through the L1,
  // 16 = 16 rows sharing an LDS window
  std::vector<int> dpbs = {0};
  if (u8) dpbs.push_back(16);
  // stage-2 output tile (0 = chosen by the plan from its LDS size)
  std::vector<int> tiles = {0};
  if (u8 && sub) tiles = {0, 256, 512, 1024};
  if (f32 && sub) tiles = {0, 512, 1024};
  // subband layout: channels per subband and fine DMs per coarse DM (these
  // change the method's delay rounding, hence the plane)
  std::vector<int> sbs = {desc->sb_nchan}, nd0s = {desc->ndm0};
  if (sub && tune_subbands) {
    sbs = {16, 32, 64, 128};
    nd0s = {8, 16, 32, 64};
  }
  TuneVal best{desc->tile_t, desc->dm_per_block, desc->sb_nchan, desc->ndm0, 1e30f};
  for (int sb : sbs)
    for (int nd0 : nd0s)
      for (int dpb : dpbs)
        for (int tile : tiles) {
          af_plan_desc d = *desc;
          d.sb_nchan = sb;
          d.ndm0 = nd0;
          d.dm_per_block = dpb;
          d.tile_t = tile;
          af_plan *p = nullptr;
          if (af_plan_create(&d, &p) != AF_OK) continue;   // a shape the kernels cannot take
          // a requested tile the plan could not honour duplicates another candidate
          if (tile != 0 && p->s2_tile != tile) {
            af_plan_destroy(p);
            continue;
          }
          float ms = 0;
          af_status s = af_time_execute(p, d_in, d_out, 1, &ms);   // warm-up
          if (s == AF_OK) s = af_time_execute(p, d_in, d_out, reps, &ms);
          af_plan_destroy(p);
          if (s != AF_OK) continue;
          ms /= (float)reps;
          if (ms < best.ms) best = TuneVal{tile, dpb, sb, nd0, ms};
        }
  if (best.ms >= 1e29f) return fail(AF_ERR_INVALID, "no candidate plan could be built");
  desc->tile_t = best.tile_t;
  desc->dm_per_block = best.dm_per_block;
  desc->sb_nchan = best.sb_nchan;
  desc->ndm0 = best.ndm0;
  if (best_ms) *best_ms = best.ms;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = best;
  return AF_OK;
}
