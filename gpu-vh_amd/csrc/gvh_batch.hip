// gvh_batch.hip -- many events on one GPU with continuous batching.
//
// The reference runs one event per process (HydroPlugin.cpp:107-434).  Here
// a batch keeps a device table of events (BatchSlot) and advances every
// running event by one KT step per batch step with one launch per kernel:
//
//   record   history slice + count of cells above e_fo   (batch_record_kernel)
//   check    nt + 1 steps or FOFREQ + 1 frozen states    (batch_check_kernel)
//   stage 1  all running events, per launch group         (kt_batch_kernel)
//   stage 2  all running events, per launch group
//   advance  buffer roles, step count                    (batch_advance_kernel)
//
// Launch groups are events of the same dimensionality and stage-kernel
// arrangement.  The sequence is captured once per batch composition into
// two HIP graphs (record and check; stages and advance) and replayed each
// step.  The freeze-out window due at a step is searched on another stream
// once the first graph has run, beside the stages and, phase by phase, the
// following steps: an event's ring has two halves, alternating by window, so
// the steps after a window record into the half the search does not read.
// Between steps the host retires finished events and
// admits queued ones while their HBM footprint (state, history ring, surface
// buffers) fits the budget: requests that arrive while others run join at
// the next step.  An event's cells are cut into tiles exactly as a launch
// of its own would cut them, so its results do not depend on the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/gpuvh.h"
#include "gvh_internal.h"

namespace gvh {
hipError_t launch_batch_stage(const BatchSlot *slots, const int *tile_slot, int begin, int end,
                              int eos, bool bulk, bool stage2, bool has_z, bool lock,
                              bool mixed, hipStream_t st);
hipError_t launch_batch_record(BatchSlot *slots, const int *tile_slot, int ntiles, bool mixed,
                               hipStream_t st);
hipError_t launch_batch_check(BatchSlot *slots, int nslots, hipStream_t st);
hipError_t launch_batch_advance(BatchSlot *slots, int nslots, hipStream_t st);
}  // namespace gvh

using namespace gvh;

namespace {

using Clock = std::chrono::steady_clock;
constexpr int TILE = GVH_BLOCK;  // cells per tile (the stage kernel's block)

struct Pending {
  long long id;
  gvh_lattice_parameters L;
  gvh_hydro_parameters H;
  std::vector<double> e;
  int surface, keep_state, fo_flags;
  int newton_budget;  // 0: MAX_ITERS
  double t_submit;
  size_t footprint;
};

// Surface rows of an event, one uninitialised block per window (no zero
// fill, no reallocation as an event's surface grows).
struct Rows {
  std::vector<std::unique_ptr<double[]>> blocks;
  std::vector<long long> counts;  // rows per block
  long long total = 0;
  double *append(long long n) {
    blocks.emplace_back(new double[(size_t)n * GVH_FO_COLUMNS]);
    counts.push_back(n);
    total += n;
    return blocks.back().get();
  }
  void copy_to(double *out, long long n) const {
    for (size_t b = 0; b < blocks.size() && n > 0; ++b) {
      const long long m = std::min(n, counts[b]);
      std::memcpy(out, blocks[b].get(), (size_t)m * GVH_FO_COLUMNS * sizeof(double));
      out += (size_t)m * GVH_FO_COLUMNS;
      n -= m;
    }
  }
};

struct Live {
  long long id = -1;
  gvh_context *ctx = nullptr;
  int surface = 0, keep = 0;
  double t_submit = 0, t_admit = 0;
  size_t footprint = 0;
  Rows rows;
  int steps = 0;       // steps at the last poll
  int max_steps = 0;
  int has_z = 0, lock = 0, ntiles = 0;
};

struct Finished {
  gvh_event_result r;
  gvh_context *ctx = nullptr;
  Rows rows;
  size_t footprint = 0;
};

// HBM bytes of one event: state arena, history ring and freeze-out buffers
// (the window buffers are sized for every cube of a window, which bounds
// them).
size_t event_footprint(const gvh_lattice_parameters &L, int precision, int surface, int fofreq) {
  const size_t nx = L.numLatticePointsX, ny = L.numLatticePointsY, nz = L.numLatticePointsRapidity;
  const size_t G = (nx + 4) * (ny + 4) * (nz + 4), nint = nx * ny * nz;
  const size_t qstride = precision == GVH_MIXED ? 4 * G + ((11 * G + 1) / 2 + 63) / 64 * 64 : 15 * G;
  size_t bytes = (3 * qstride + 16 * G) * sizeof(double) + (1u << 20);
  if (surface) {
    const size_t nrow = (size_t)fofreq * nx * ny;
    bytes += 2 * (size_t)(fofreq + 1) * 10 * nint * sizeof(double);  // ring, two halves
    bytes += 2 * nrow * sizeof(unsigned) * 5 / 4;                // row counts and offsets
    bytes += nint / 20 * (3 * sizeof(unsigned) + 80 * sizeof(double) + 200 * sizeof(double));
  }
  return bytes;
}

double seconds(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

}  // namespace

struct gvh_batch {
  gvh_options opts{};
  std::vector<double> tab_p, tab_cs2, tab_T;
  hipStream_t stream = nullptr;
  // streams of the freeze-out window searches: the events whose window is
  // due at a step are searched together, round-robin over these (default
  // priority, like every stream here)
  static constexpr int NSEARCH = 8;
  hipStream_t search[NSEARCH] = {};
  // the window search of each slot's event: phase 1 (cube counts), 2
  // (elements), 3 (rows copied to the slot's pinned staging, then into the
  // event's row blocks by host threads).  One phase is issued per step once
  // the previous one has finished, so a search runs beside the next steps:
  // the window it reads is one ring half, and those steps record into the
  // other.  It is completed before a step would record into its half again,
  // before the event's next window is searched, and before the event retires.
  struct Search {
    int phase = 0;
    int m = 0;          // window of the search (loop index / fofreq)
    double t = 0;       // the loop's time at that window
    long long ne = 0;   // rows
    hipEvent_t ev = nullptr;  // recorded after the phase's device-to-host copy
    double *stage = nullptr;
    size_t cap = 0;     // doubles
  };
  std::vector<Search> srch;
  size_t budget = 0, used = 0;
  int max_slots = 0, fofreq = 10, autotune = 0;
  BatchSlot *d_slots = nullptr;
  BatchSlot *h_slots = nullptr;  // pinned host mirror
  std::vector<Live> live;
  std::deque<Pending> queue;
  std::map<long long, Finished> done;
  std::deque<long long> unpolled;
  int *d_tiles = nullptr;
  size_t tile_cap = 0;
  struct Group {
    int has_z, lock, first, ntiles;
  };
  std::vector<Group> groups;
  int ntiles_total = 0;
  bool dirty = true;
  // a step is two graphs: the record of the history slice (and the
  // termination check), then the stages; window searches wait only for the
  // first, so they run beside the stages
  hipGraph_t graph = nullptr, graph2 = nullptr;
  hipGraphExec_t exec = nullptr, exec2 = nullptr;
  hipEvent_t recorded = nullptr;
  long long next_id = 0;
  Clock::time_point origin;
  std::map<std::tuple<int, int, int>, std::pair<int, int>> tuned;  // shape class -> (variant, strip)
};

namespace {

// Host copies dst[i][0..len[i]) = src[i][...], cut into pieces of at most
// 4 MB spread over up to 8 threads.
void parallel_copy(const std::vector<std::pair<double *, const double *>> &jobs,
                   const std::vector<size_t> &len) {
  constexpr size_t PIECE = (4u << 20) / sizeof(double);
  std::vector<std::tuple<double *, const double *, size_t>> pieces;
  for (size_t i = 0; i < jobs.size(); ++i)
    for (size_t o = 0; o < len[i]; o += PIECE)
      pieces.emplace_back(jobs[i].first + o, jobs[i].second + o, std::min(PIECE, len[i] - o));
  if (pieces.empty()) return;
  const size_t nt = std::min<size_t>({8, pieces.size(), std::max(1u, std::thread::hardware_concurrency())});
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t i = next++; i < pieces.size(); i = next++)
      std::memcpy(std::get<0>(pieces[i]), std::get<1>(pieces[i]),
                  std::get<2>(pieces[i]) * sizeof(double));
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
}

hipStream_t search_stream(gvh_batch *b, int s) { return b->search[s % gvh_batch::NSEARCH]; }

// Issue the next phase of slot s's window search once the previous phase
// has finished: with block, waiting for it; otherwise only if it has.  A
// finished third phase appends the window's rows to the event's row blocks.
int advance_search(gvh_batch *b, int s, bool block) {
  gvh_batch::Search &W = b->srch[s];
  if (W.phase == 0) return GVH_OK;
  if (block) {
    HIP_TRY(hipEventSynchronize(W.ev));
  } else {
    const hipError_t q = hipEventQuery(W.ev);
    if (q == hipErrorNotReady) return GVH_OK;
    if (q != hipSuccess) return gvh_fail(GVH_EHIP, "window search: %s", hipGetErrorString(q));
  }
  Live &L = b->live[s];
  const hipStream_t st = search_stream(b, s);
  int rc;
  if (W.phase == 1) {
    if ((rc = fo_find_elements(L.ctx, st))) return rc;
    HIP_TRY(hipEventRecord(W.ev, st));
    W.phase = 2;
  } else if (W.phase == 2) {
    if ((rc = fo_find_emit(L.ctx, W.t, st, &W.ne))) return rc;
    const size_t need = (size_t)W.ne * GVH_FO_COLUMNS;
    if (need > W.cap) {
      if (W.stage) HIP_TRY(hipHostFree(W.stage));
      W.stage = nullptr;
      W.cap = 0;
      const size_t cap = need + need / 2;
      HIP_TRY(hipHostMalloc((void **)&W.stage, cap * sizeof(double)));
      W.cap = cap;
    }
    if (W.ne > 0 && (rc = fo_rows_async(L.ctx, W.stage, W.ne, st))) return rc;
    HIP_TRY(hipEventRecord(W.ev, st));
    W.phase = 3;
  } else {
    if (W.ne > 0)
      parallel_copy({{L.rows.append(W.ne), W.stage}}, {(size_t)W.ne * GVH_FO_COLUMNS});
    W.phase = 0;
  }
  return GVH_OK;
}

int complete_search(gvh_batch *b, int s) {
  int rc;
  while (b->srch[s].phase)
    if ((rc = advance_search(b, s, true))) return rc;
  return GVH_OK;
}

// Start the search of the window due at loop index n of slot s's event; it
// reads ring half (n / fofreq) mod 2 once this step's record graph has run.
int start_search(gvh_batch *b, int s, int n) {
  gvh_batch::Search &W = b->srch[s];
  Live &L = b->live[s];
  int rc;
  if ((rc = complete_search(b, s))) return rc;
  if (!W.ev) HIP_TRY(hipEventCreateWithFlags(&W.ev, hipEventDisableTiming));
  const hipStream_t st = search_stream(b, s);
  W.m = n / b->fofreq;
  W.t = L.ctx->Hp.initialProperTimePoint + L.steps * L.ctx->L.latticeSpacingProperTime;
  W.ne = 0;
  if ((rc = fo_set_window_half(L.ctx, W.m & 1))) return rc;
  HIP_TRY(hipStreamWaitEvent(st, b->recorded, 0));
  if ((rc = fo_find_count(L.ctx, n, st))) return rc;
  HIP_TRY(hipEventRecord(W.ev, st));
  W.phase = 1;
  return GVH_OK;
}

int grow_tiles(gvh_batch *b, size_t n) {
  if (n <= b->tile_cap) return GVH_OK;
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  b->d_tiles = nullptr;
  b->tile_cap = 0;
  const size_t cap = n + n / 2 + 64;
  HIP_TRY(hipMalloc((void **)&b->d_tiles, cap * sizeof(int)));
  b->tile_cap = cap;
  return GVH_OK;
}

// Tiles of an event: the stage kernel's blocks for its lattice and strip,
// rounded up to a multiple of 8 (XCD order).
int tiles_of(const gvh_context *c) {
  const int nstrips = (c->ny + c->strip - 1) / c->strip;
  const long long ncell = (long long)c->nx * c->strip * nstrips * c->nz;
  return (int)(((ncell + TILE - 1) / TILE + 7) / 8 * 8);
}

void destroy_graph(gvh_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->exec2) (void)hipGraphExecDestroy(b->exec2);
  if (b->graph2) (void)hipGraphDestroy(b->graph2);
  b->exec = b->exec2 = nullptr;
  b->graph = b->graph2 = nullptr;
}

// Tile map and per-slot tile ranges for the current running set, then the
// step graph.
int rebuild(gvh_batch *b) {
  destroy_graph(b);
  b->groups.clear();
  std::vector<int> map;
  for (int hz = 0; hz < 2; ++hz)
    for (int lk = 0; lk < 2; ++lk) {
      gvh_batch::Group g{hz, lk, (int)map.size(), 0};
      for (int s = 0; s < b->max_slots; ++s) {
        Live &L = b->live[s];
        if (!L.ctx || b->h_slots[s].status != SLOT_RUN || L.has_z != hz || L.lock != lk) continue;
        b->h_slots[s].tile_first = (int)map.size();
        b->h_slots[s].ntiles = L.ntiles;
        for (int t = 0; t < L.ntiles; ++t) map.push_back(s);
        // only the tile fields: the device owns steps, roles and status
        HIP_TRY(hipMemcpyAsync(&b->d_slots[s].tile_first, &b->h_slots[s].tile_first,
                               2 * sizeof(int), hipMemcpyHostToDevice, b->stream));
      }
      g.ntiles = (int)map.size() - g.first;
      if (g.ntiles) b->groups.push_back(g);
    }
  b->ntiles_total = (int)map.size();
  b->dirty = false;
  if (map.empty()) return GVH_OK;
  int rc = grow_tiles(b, map.size());
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(b->d_tiles, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  // capture two graphs: record and check; stage 1 and 2 per group and advance
  const bool mixed = b->opts.precision == GVH_MIXED;
  auto capture = [&](auto &&body, hipGraph_t &graph, hipGraphExec_t &exec) -> int {
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = body();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return gvh_fail(GVH_EHIP, "batch step capture: %s",
                      hipGetErrorString(e != hipSuccess ? e : e2));
    }
    graph = g;
    HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return GVH_OK;
  };
  rc = capture([&]() {
    hipError_t e = launch_batch_record(b->d_slots, b->d_tiles, b->ntiles_total, mixed, b->stream);
    if (e == hipSuccess) e = launch_batch_check(b->d_slots, b->max_slots, b->stream);
    return e;
  }, b->graph, b->exec);
  if (rc) return rc;
  return capture([&]() {
    hipError_t e = hipSuccess;
    for (int stage = 1; stage <= 2 && e == hipSuccess; ++stage)
      for (const auto &g : b->groups) {
        e = launch_batch_stage(b->d_slots, b->d_tiles, g.first, g.first + g.ntiles,
                               b->opts.eos, b->opts.bulk != 0, stage == 2,
                               g.has_z != 0, g.lock != 0, mixed, b->stream);
        if (e != hipSuccess) break;
      }
    if (e == hipSuccess) e = launch_batch_advance(b->d_slots, b->max_slots, b->stream);
    return e;
  }, b->graph2, b->exec2);
}

// Time (variant, strip) candidates on the context's current state, restoring
// the state after each; returns the fastest.
int tune_context(gvh_context *c, int steps, int *variant, int *strip, float *best_ms) {
  const size_t bytes = (3 * c->qstride + 16 * (size_t)c->G) * sizeof(double);
  void *save = nullptr;
  HIP_TRY(hipMalloc(&save, bytes));
  // the tuning steps' failure counters are not the event's: kept aside and
  // put back with the state
  unsigned long long diag_save[DIAG_N];
  hipError_t e = hipMemcpyAsync(diag_save, c->diag, sizeof diag_save, hipMemcpyDeviceToHost,
                                c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(save, c->arena, bytes, hipMemcpyDeviceToDevice, c->stream);
  const int variants[2] = {0, 3};
  const int strips[5] = {4, 8, 16, 32, 64};
  const int v0 = c->variant, s0 = c->strip, iq = c->iq, iu = c->iu, iup = c->iup;
  const long long sd = c->steps_done;
  float best = 1e30f;
  int bv = v0, bs = s0, rc = GVH_OK;
  const double t0 = c->Hp.initialProperTimePoint, dt = c->L.latticeSpacingProperTime;
  for (int vi = 0; vi < 2 && e == hipSuccess && !rc; ++vi)
    for (int si = 0; si < 5 && e == hipSuccess && !rc; ++si) {
      if (strips[si] > c->ny && si > 0) continue;
      c->variant = variants[vi];
      c->strip = std::min(strips[si], c->ny);
      float total = 0;
      for (int n = 0; n < steps + 1 && !rc; ++n) {
        float a = 0, b2 = 0;
        rc = gvh_step_timed(c, t0 + n * dt, &a, &b2);
        if (n > 0) total += a + b2;  // the first step warms up
      }
      if (!rc && total < best) {
        best = total;
        bv = c->variant;
        bs = c->strip;
      }
      // restore the state and roles
      if (e == hipSuccess)
        e = hipMemcpyAsync(c->arena, save, bytes, hipMemcpyDeviceToDevice, c->stream);
      c->iq = iq;
      c->iu = iu;
      c->iup = iup;
      c->steps_done = sd;
    }
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->diag, diag_save, sizeof diag_save, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(save);
  c->variant = bv;
  c->strip = bs;
  if (rc) return rc;
  if (e != hipSuccess) return gvh_fail(GVH_EHIP, "autotune: %s", hipGetErrorString(e));
  if (variant) *variant = bv;
  if (strip) *strip = bs;
  if (best_ms) *best_ms = best / steps;
  return GVH_OK;
}

// An event whose admission failed: finished at once with the error code in
// its result, so that the events queued behind it and the running ones go on.
void fail_pending(gvh_batch *b, const Pending &P, int rc) {
  Finished F;
  F.r.id = P.id;
  F.r.steps = 0;
  F.r.frozen = 0;
  F.r.t_final = P.H.initialProperTimePoint;
  F.r.surface_rows = 0;
  F.r.t_submit = P.t_submit;
  F.r.t_admit = F.r.t_done = seconds(b->origin, Clock::now());
  F.r.status = rc;
  F.r.newton_not_converged = 0;
  b->done[P.id] = std::move(F);
  b->unpolled.push_back(P.id);
}

int admit(gvh_batch *b) {
  while (!b->queue.empty()) {
    Pending &P = b->queue.front();
    if (b->used + P.footprint > b->budget) break;
    int s = -1;
    for (int k = 0; k < b->max_slots; ++k)
      if (!b->live[k].ctx && b->h_slots[k].status != SLOT_RUN) {
        s = k;
        break;
      }
    if (s < 0) break;
    gvh_options o = b->opts;
    o.stream = b->stream;
    gvh_context *c = nullptr;
    int rc = gvh_create(&P.L, &P.H, &o, &c);
    if (rc) {
      fail_pending(b, P, rc);
      b->queue.pop_front();
      continue;
    }
    if (P.newton_budget > 0 && P.newton_budget < MAX_ITERS) c->max_iters = P.newton_budget;
    const size_t n = P.e.size();
    std::vector<double> p(n), ut(n, 1.0), zero(n, 0.0);
    for (size_t i = 0; i < n; ++i) p[i] = host_eos_p(c->eos, P.e[i]);
    const double t0 = P.H.initialProperTimePoint;
    rc = gvh_set_initial_state(c, t0, P.e.data(), p.data(), ut.data(), zero.data(), zero.data(),
                               zero.data(), nullptr, nullptr);
    if (!rc && b->autotune) {
      // one tuning per shape class: 2+1D or 3+1D, y extent in steps of 40,
      // eta extent in steps of 16
      const auto key = std::make_tuple(c->nz > 1 ? 1 : 0, c->ny / 40, c->nz / 16);
      auto it = b->tuned.find(key);
      if (it == b->tuned.end()) {
        int v = 0, st = 16;
        rc = tune_context(c, 3, &v, &st, nullptr);
        if (!rc) it = b->tuned.emplace(key, std::make_pair(v, st)).first;
      }
      if (!rc) {
        c->variant = it->second.first;
        c->strip = std::min(it->second.second, c->ny);
      }
    }
    double e_fo = host_e_of_T(c->eos, P.H.freezeoutTemperatureGeV / 0.197326938);
    if (!rc && P.surface) rc = fo_init(c, b->fofreq, e_fo, P.fo_flags, 2);
    if (rc) {
      gvh_destroy(c);
      fail_pending(b, P, rc);
      b->queue.pop_front();
      continue;
    }
    BatchSlot S{};
    S.A = make_stage_args(c, 1, t0);
    S.A.reg = nullptr;  // batched events keep no validity arrays
    S.B = context_buffers(c);
    S.R = context_roles(c);
    S.t0 = t0;
    S.steps = 0;
    S.status = SLOT_RUN;
    S.max_steps = P.L.numProperTimePoints + 1;
    S.fofreq = b->fofreq;
    S.fo_flags = P.fo_flags;
    S.bulk = c->bulk ? 1 : 0;
    S.e_fo = e_fo;
    S.ring = P.surface ? fo_ring(c->fo) : nullptr;
    S.halves = 2;
    S.nint = (long long)c->nx * c->ny * c->nz;
    Live &L = b->live[s];
    L = Live();
    L.id = P.id;
    L.ctx = c;
    L.surface = P.surface;
    L.keep = P.keep_state;
    L.t_submit = P.t_submit;
    L.t_admit = seconds(b->origin, Clock::now());
    L.footprint = P.footprint;
    L.max_steps = S.max_steps;
    L.has_z = c->nz > 1;
    L.lock = c->variant == 3;
    L.ntiles = tiles_of(c);
    b->h_slots[s] = S;
    HIP_TRY(hipMemcpyAsync(&b->d_slots[s], &b->h_slots[s], sizeof(BatchSlot), hipMemcpyHostToDevice,
                           b->stream));
    b->used += P.footprint;
    b->queue.pop_front();
    b->dirty = true;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  return GVH_OK;
}

int retire(gvh_batch *b, int s) {
  int rc0 = complete_search(b, s);
  if (rc0) return rc0;
  Live &L = b->live[s];
  const BatchSlot &S = b->h_slots[s];
  gvh_context *c = L.ctx;
  // hand the final roles back to the context so its state can be read
  c->iq = S.R.iq;
  c->iu = S.R.iu;
  c->iup = S.R.iup;
  c->ie = S.R.ie;
  c->steps_done = S.steps;
  Finished F;
  F.r.id = L.id;
  F.r.steps = S.steps;
  F.r.frozen = S.frozen_stop;
  F.r.t_final = S.t0 + S.steps * S.A.dt;
  F.r.surface_rows = L.rows.total;
  F.r.t_submit = L.t_submit;
  F.r.t_admit = L.t_admit;
  F.r.t_done = seconds(b->origin, Clock::now());
  F.r.status = S.error ? GVH_ENOCONV : GVH_OK;
  gvh_diagnostics dg{};
  int rc = gvh_get_diagnostics(c, &dg);
  if (rc) return rc;
  F.r.newton_not_converged = dg.newton_not_converged;
  F.rows = std::move(L.rows);
  if (L.keep) {
    F.ctx = c;
    F.footprint = L.footprint;
  } else {
    gvh_destroy(c);
    b->used -= L.footprint;
  }
  b->done[L.id] = std::move(F);
  b->unpolled.push_back(L.id);
  // free the slot on the device and the host
  b->h_slots[s].status = SLOT_FREE;
  HIP_TRY(hipMemcpyAsync(&b->d_slots[s].status, &b->h_slots[s].status, sizeof(int),
                         hipMemcpyHostToDevice, b->stream));
  L = Live();
  b->dirty = true;
  return GVH_OK;
}

}  // namespace

extern "C" void gvh_batch_options_default(gvh_batch_options *o) {
  if (!o) return;
  o->hbm_bytes = 0;
  o->max_events = 64;
  o->fofreq = 10;
  o->autotune = 1;
}

extern "C" int gvh_batch_create(const gvh_options *opts, const gvh_batch_options *bo,
                                gvh_batch **out) {
  if (!opts || !out) return gvh_fail(GVH_EINVAL, "gvh_batch_create: NULL argument");
  *out = nullptr;
  gvh_batch_options B;
  gvh_batch_options_default(&B);
  if (bo) B = *bo;
  if (B.max_events < 1 || B.max_events > 4096)
    return gvh_fail(GVH_EINVAL, "max_events must be in [1, 4096]");
  if (B.fofreq < 3 || B.fofreq > GVH_FO_MAX_FREQ)
    return gvh_fail(GVH_EINVAL, "fofreq must be in [3, %d]", GVH_FO_MAX_FREQ);
  {  // the options every event's context will be created with
    HostEos H;
    const int rc = eos_from_options(opts, H);
    if (rc) return rc;
    if (opts->precision != GVH_FP64 && opts->precision != GVH_MIXED)
      return gvh_fail(GVH_EINVAL, "a batch runs GVH_FP64 or GVH_MIXED, not precision %d",
                      opts->precision);
  }
  gvh_batch *b = new gvh_batch();
  b->opts = *opts;
  // keep the tabulated EoS alive for the contexts created later
  if (opts->eos == GVH_EOS_TABLE && opts->tab_n > 1) {
    b->tab_p.assign(opts->tab_p, opts->tab_p + opts->tab_n);
    b->tab_cs2.assign(opts->tab_cs2, opts->tab_cs2 + opts->tab_n);
    b->tab_T.assign(opts->tab_T, opts->tab_T + opts->tab_n);
    b->opts.tab_p = b->tab_p.data();
    b->opts.tab_cs2 = b->tab_cs2.data();
    b->opts.tab_T = b->tab_T.data();
  }
  b->max_slots = B.max_events;
  b->fofreq = B.fofreq;
  b->autotune = B.autotune;
  b->origin = Clock::now();
  hipError_t e = hipStreamCreate(&b->stream);
  for (hipStream_t &q : b->search)
    if (e == hipSuccess) e = hipStreamCreate(&q);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->recorded, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc((void **)&b->d_slots, b->max_slots * sizeof(BatchSlot));
  if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_slots, b->max_slots * sizeof(BatchSlot));
  if (e == hipSuccess) {
    std::memset(b->h_slots, 0, b->max_slots * sizeof(BatchSlot));
    e = hipMemcpy(b->d_slots, b->h_slots, b->max_slots * sizeof(BatchSlot), hipMemcpyHostToDevice);
  }
  size_t free_b = 0, total_b = 0;
  if (e == hipSuccess) e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) {
    gvh_batch_destroy(b);
    return gvh_fail(GVH_EHIP, "batch set-up: %s", hipGetErrorString(e));
  }
  b->budget = B.hbm_bytes ? (size_t)B.hbm_bytes : free_b / 10 * 8;
  b->live.resize(b->max_slots);
  b->srch.resize(b->max_slots);
  *out = b;
  return GVH_OK;
}

extern "C" int gvh_batch_destroy(gvh_batch *b) {
  if (!b) return GVH_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (hipStream_t q : b->search)  // a search left by an error path
    if (q) (void)hipStreamSynchronize(q);
  destroy_graph(b);
  for (Live &L : b->live)
    if (L.ctx) gvh_destroy(L.ctx);
  for (auto &kv : b->done)
    if (kv.second.ctx) gvh_destroy(kv.second.ctx);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  if (b->d_slots) (void)hipFree(b->d_slots);
  if (b->h_slots) (void)hipHostFree(b->h_slots);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (hipStream_t &q : b->search)
    if (q) (void)hipStreamDestroy(q);
  for (gvh_batch::Search &W : b->srch) {
    if (W.stage) (void)hipHostFree(W.stage);
    if (W.ev) (void)hipEventDestroy(W.ev);
  }
  if (b->recorded) (void)hipEventDestroy(b->recorded);
  delete b;
  return GVH_OK;
}

extern "C" int gvh_batch_submit(gvh_batch *b, const gvh_event *ev, long long *id) {
  if (!b || !ev || !ev->e) return gvh_fail(GVH_EINVAL, "gvh_batch_submit: NULL argument");
  const gvh_lattice_parameters &L = ev->lattice;
  int rc = check_event_params(&L, &ev->hydro);
  if (rc) return rc;
  if (L.numProperTimePoints < 0) return gvh_fail(GVH_EINVAL, "numProperTimePoints must be >= 0");
  if (ev->newton_budget < 0) return gvh_fail(GVH_EINVAL, "newton_budget must be >= 0");
  if (ev->surface && !(L.numLatticePointsX > 1 && L.numLatticePointsY > 1))
    return gvh_fail(GVH_EINVAL, "the surface needs a 2+1D or 3+1D lattice");
  Pending P;
  P.id = b->next_id++;
  P.L = L;
  P.L.numComputationalLatticePointsX = L.numLatticePointsX + 4;
  P.L.numComputationalLatticePointsY = L.numLatticePointsY + 4;
  P.L.numComputationalLatticePointsRapidity = L.numLatticePointsRapidity + 4;
  P.H = ev->hydro;
  const size_t n = (size_t)L.numLatticePointsX * L.numLatticePointsY * L.numLatticePointsRapidity;
  P.e.assign(ev->e, ev->e + n);
  P.surface = ev->surface;
  P.keep_state = ev->keep_state;
  P.fo_flags = ev->fo_flags;
  P.newton_budget = ev->newton_budget;
  P.t_submit = seconds(b->origin, Clock::now());
  P.footprint = event_footprint(P.L, b->opts.precision, P.surface, b->fofreq);
  if (P.footprint > b->budget)
    return gvh_fail(GVH_ENOMEM, "event needs %zu bytes, the batch budget is %zu", P.footprint,
                    b->budget);
  b->queue.push_back(std::move(P));
  if (id) *id = b->queue.back().id;
  return GVH_OK;
}

extern "C" int gvh_batch_step(gvh_batch *b, int *nrunning) {
  if (!b) return gvh_fail(GVH_EINVAL, "batch is NULL");
  int rc = admit(b);
  if (rc) return rc;
  if (b->dirty && (rc = rebuild(b))) return rc;
  int running = 0;
  for (int s = 0; s < b->max_slots; ++s) running += b->live[s].ctx != nullptr;
  if (running == 0) {
    if (nrunning) *nrunning = 0;
    return GVH_OK;
  }
  // the windows due at this step (loop index n = steps before it + 1): their
  // history is complete once the record graph has run.  A search still in
  // flight is completed first if this step records into its ring half again
  // (window m + 2) or its event has a new window due.
  std::vector<int> due;
  for (int s = 0; s < b->max_slots; ++s) {
    const Live &L = b->live[s];
    if (!L.ctx) continue;
    const int n = L.steps + 1;
    const bool is_due = L.surface && n <= L.max_steps && n % b->fofreq == b->fofreq - 1;
    if (is_due) due.push_back(s);
    if (b->srch[s].phase && (is_due || n / b->fofreq >= b->srch[s].m + 2) &&
        (rc = complete_search(b, s)))
      return rc;
  }
  HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  if (!due.empty()) HIP_TRY(hipEventRecord(b->recorded, b->stream));
  HIP_TRY(hipGraphLaunch(b->exec2, b->stream));
  HIP_TRY(hipMemcpyAsync(b->h_slots, b->d_slots, b->max_slots * sizeof(BatchSlot),
                         hipMemcpyDeviceToHost, b->stream));
  // searches in flight move on by a phase if the previous one has finished;
  // the due windows start theirs beside the stages
  for (int s = 0; s < b->max_slots; ++s)
    if (b->srch[s].phase && (rc = advance_search(b, s, false))) return rc;
  for (int s : due)
    if ((rc = start_search(b, s, b->live[s].steps + 1))) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  running = 0;
  for (int s = 0; s < b->max_slots; ++s) {
    Live &L = b->live[s];
    if (!L.ctx) continue;
    const BatchSlot &S = b->h_slots[s];
    L.steps = S.steps;
    if (S.status == SLOT_DONE) {
      if ((rc = retire(b, s))) return rc;
    } else {
      ++running;
    }
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (nrunning) *nrunning = running + (int)b->queue.size();
  return GVH_OK;
}

extern "C" int gvh_batch_run(gvh_batch *b) {
  if (!b) return gvh_fail(GVH_EINVAL, "batch is NULL");
  int n = 1;
  while (n > 0) {
    int rc = gvh_batch_step(b, &n);
    if (rc) return rc;
  }
  return GVH_OK;
}

extern "C" int gvh_batch_poll(gvh_batch *b, gvh_event_result *out, int max, int *n) {
  if (!b || !n || (max > 0 && !out)) return gvh_fail(GVH_EINVAL, "gvh_batch_poll: NULL argument");
  int k = 0;
  while (k < max && !b->unpolled.empty()) {
    out[k++] = b->done[b->unpolled.front()].r;
    b->unpolled.pop_front();
  }
  *n = k;
  return GVH_OK;
}

extern "C" int gvh_batch_rows(gvh_batch *b, long long id, double *rows, long long nrows) {
  if (!b) return gvh_fail(GVH_EINVAL, "batch is NULL");
  auto it = b->done.find(id);
  if (it == b->done.end()) return gvh_fail(GVH_EINVAL, "event %lld has not finished", id);
  const long long have = it->second.rows.total;
  if (nrows < 0 || nrows > have)
    return gvh_fail(GVH_EINVAL, "event %lld has %lld rows, asked for %lld", id, have, nrows);
  if (nrows && !rows) return gvh_fail(GVH_EINVAL, "rows is NULL");
  it->second.rows.copy_to(rows, nrows);
  return GVH_OK;
}

extern "C" int gvh_batch_field(gvh_batch *b, long long id, int field, double *host_out) {
  if (!b) return gvh_fail(GVH_EINVAL, "batch is NULL");
  auto it = b->done.find(id);
  if (it == b->done.end() || !it->second.ctx)
    return gvh_fail(GVH_EINVAL, "event %lld has not finished or its state was not kept", id);
  return gvh_get_field(it->second.ctx, field, 0, host_out);
}

extern "C" int gvh_batch_release(gvh_batch *b, long long id) {
  if (!b) return gvh_fail(GVH_EINVAL, "batch is NULL");
  auto it = b->done.find(id);
  if (it == b->done.end()) return gvh_fail(GVH_EINVAL, "event %lld has not finished", id);
  if (it->second.ctx) {
    gvh_destroy(it->second.ctx);
    b->used -= it->second.footprint;
  }
  b->done.erase(it);
  return GVH_OK;
}

extern "C" int gvh_batch_stats(gvh_batch *b, long long *queued, long long *running,
                               unsigned long long *hbm_used, unsigned long long *hbm_budget) {
  if (!b) return gvh_fail(GVH_EINVAL, "batch is NULL");
  long long r = 0;
  for (const Live &L : b->live) r += L.ctx != nullptr;
  if (queued) *queued = (long long)b->queue.size();
  if (running) *running = r;
  if (hbm_used) *hbm_used = b->used;
  if (hbm_budget) *hbm_budget = b->budget;
  return GVH_OK;
}

extern "C" int gvh_autotune(gvh_context *c, int steps, int *variant, int *strip, float *ms_per_step) {
  if (!c) return gvh_fail(GVH_EINVAL, "context is NULL");
  if (!c->initialized) return gvh_fail(GVH_EINVAL, "state not initialised");
  if (steps < 1) return gvh_fail(GVH_EINVAL, "steps must be >= 1");
  return tune_context(c, steps, variant, strip, ms_per_step);
}
