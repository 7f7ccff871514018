// SpMM launch-shape autotuner.
//
// The sliced copy of a sector Hamiltonian can give each matrix row 1, 2 or 4
// lanes of a wavefront (spmm_sell<V, NB, G>: slices of 64 / G rows).  Which
// is fastest depends on the sector dimension (one lane per row fills the 256
// CUs only for large sectors), on the mean row length (gathers per lane) and
// on the block width and element precision the solver will use.  With tuning
// on, every Hamiltonian build looks its shape up in a process-wide table
// keyed by (precision, block width, floor(log2 dim), mean row length rounded
// to a quarter octave); on a miss it builds the three layouts one after the
// other, times each on the build's stream (one warm-up and kRuns timed
// launches on zero vectors, each launch timed on its own and the fastest
// kept, since work of other streams only ever slows a launch down), records
// the timings and keeps the fastest layout.  Sectors of one scan share a key,
// so one build per scan pays for the timing: builds of a key that is being
// timed wait for that result instead of timing it again (so concurrent
// solves do not all time at once and disturb each other's measurements).
// The table keeps the first measurement of a key.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <map>
#include <mutex>
#include <set>

#include "objects.h"

namespace edx {
void spmm(const edx_hamiltonian& h, Ctx& ctx, int precision, int nb, const void* X, void* Y, int64_t ldy);

namespace {

constexpr int kRuns = 7;
constexpr int kLanes[3] = {1, 2, 4};

struct TuneKey {
  int prec, nb, dim_log2, row_q;
  bool operator<(const TuneKey& o) const {
    if (prec != o.prec) return prec < o.prec;
    if (nb != o.nb) return nb < o.nb;
    if (dim_log2 != o.dim_log2) return dim_log2 < o.dim_log2;
    return row_q < o.row_q;
  }
};

struct TuneEntry {
  int lanes;
  float ms[3];
};

struct TuneState {
  std::mutex mu;
  std::condition_variable cv;  // a key's timing finished
  bool enabled = false;
  int prec = EDX_PREC_MIXED;
  int nb = 6;
  std::map<TuneKey, TuneEntry> table;
  std::set<TuneKey> busy;  // keys being timed
};

TuneState& state() {
  static TuneState s;
  return s;
}

TuneKey key_for(const edx_hamiltonian& h, int prec, int nb) {
  const double row = h.dim ? (double)h.n_entries / (double)h.dim : 0.0;
  TuneKey k;
  k.prec = prec;
  k.nb = nb;
  k.dim_log2 = h.dim > 0 ? (int)std::floor(std::log2((double)h.dim)) : 0;
  k.row_q = row > 0 ? (int)std::lround(4.0 * std::log2(row)) : -1000;
  return k;
}

// Fastest of kRuns individually timed SpMM launches with the layout h
// currently has (after one warm-up launch).
float time_spmm(const edx_hamiltonian& h, Ctx& ctx, int prec, int nb, const void* X, void* Y) {
  spmm(h, ctx, prec, nb, X, Y, h.dim);  // warm-up
  hipEvent_t ev[kRuns + 1];
  for (auto& e : ev) EDX_HIP(hipEventCreate(&e));
  EDX_HIP(hipEventRecord(ev[0], ctx.stream));
  for (int r = 0; r < kRuns; ++r) {
    spmm(h, ctx, prec, nb, X, Y, h.dim);
    EDX_HIP(hipEventRecord(ev[r + 1], ctx.stream));
  }
  EDX_HIP(hipEventSynchronize(ev[kRuns]));
  float best = INFINITY;
  for (int r = 0; r < kRuns; ++r) {
    float ms = 0.0f;
    EDX_HIP(hipEventElapsedTime(&ms, ev[r], ev[r + 1]));
    best = std::min(best, ms);
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return best;
}

}  // namespace

// Times the three layouts of h at (prec, nb), leaves h built with the
// fastest and returns the entry.
TuneEntry tune_spmm(edx_hamiltonian& h, Ctx& ctx, int prec, int nb) {
  const size_t esz = prec == EDX_PREC_FP64 ? 16 : 8;
  DevBuf<char> X((size_t)h.dim * nb * esz), Y((size_t)h.dim * nb * esz);
  EDX_HIP(hipMemsetAsync(X.p, 0, X.bytes(), ctx.stream));
  TuneEntry e{1, {0.0f, 0.0f, 0.0f}};
  float best = INFINITY;
  for (int c = 0; c < 3; ++c) {
    if (h.sell_g != kLanes[c] || h.sell_slices == 0) build_sell(h, ctx, kLanes[c]);
    e.ms[c] = time_spmm(h, ctx, prec, nb, X.p, Y.p);
    if (e.ms[c] < best) {
      best = e.ms[c];
      e.lanes = kLanes[c];
    }
  }
  if (h.sell_g != e.lanes) build_sell(h, ctx, e.lanes);
  return e;
}

// Lane count for a new build: the context's fixed choice, else the tuned
// one, else the heuristic.  On a miss of the tuning table the plan says so
// (and marks the key as being timed); finish_sell then times the layouts
// once the table is complete.
SellPlan plan_sell(const edx_hamiltonian& h, Ctx& ctx) {
  SellPlan p;
  if (ctx.autotune_group > 0) {
    p.lanes = ctx.autotune_group;
    return p;
  }
  TuneState& s = state();
  bool on;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    on = s.enabled;
    p.prec = s.prec;
    p.nb = s.nb;
  }
  p.lanes = default_sell_lanes(h.dim, h.n_entries);
  if (!on || h.dim <= 500) return p;
  const TuneKey k = key_for(h, p.prec, p.nb);
  std::unique_lock<std::mutex> lk(s.mu);
  // another build is timing this key: wait for its result (bounded)
  s.cv.wait_for(lk, std::chrono::seconds(30), [&] { return !s.busy.count(k); });
  auto it = s.table.find(k);
  if (it != s.table.end()) {
    p.lanes = it->second.lanes;
  } else if (!s.busy.count(k)) {  // still being timed elsewhere: keep the heuristic
    s.busy.insert(k);
    p.tune = true;
  }
  return p;
}

void finish_sell(edx_hamiltonian& h, Ctx& ctx, SellPlan& p) {
  if (!p.tune) return;
  TuneState& s = state();
  const TuneKey k = key_for(h, p.prec, p.nb);
  TuneEntry e{0, {0.0f, 0.0f, 0.0f}};
  try {
    e = tune_spmm(h, ctx, p.prec, p.nb);
  } catch (...) {
    abandon_sell(h, p);
    throw;
  }
  std::lock_guard<std::mutex> lk(s.mu);
  s.table.emplace(k, e);  // the first measurement of a key is kept
  s.busy.erase(k);
  s.cv.notify_all();
  p.tune = false;
}

void abandon_sell(const edx_hamiltonian& h, SellPlan& p) {
  if (!p.tune) return;
  TuneState& s = state();
  std::lock_guard<std::mutex> lk(s.mu);
  s.busy.erase(key_for(h, p.prec, p.nb));
  s.cv.notify_all();
  p.tune = false;
}

bool tuning_enabled() {
  TuneState& s = state();
  std::lock_guard<std::mutex> lk(s.mu);
  return s.enabled;
}

namespace {
struct SchedState {
  std::mutex mu;
  std::condition_variable cv;
  std::map<SchedKey, SchedEntry> table;
  std::set<SchedKey> busy;
};
SchedState& sched_state() {
  static SchedState s;
  return s;
}
}  // namespace

SchedLookup schedule_lookup(const SchedKey& k, SchedEntry* out) {
  SchedState& s = sched_state();
  std::unique_lock<std::mutex> lk(s.mu);
  s.cv.wait_for(lk, std::chrono::seconds(30), [&] { return !s.busy.count(k); });
  auto it = s.table.find(k);
  if (it != s.table.end()) {
    *out = it->second;
    return SchedLookup::found;
  }
  if (s.busy.count(k)) return SchedLookup::unavailable;
  s.busy.insert(k);
  return SchedLookup::claimed;
}

void schedule_store(const SchedKey& k, const SchedEntry& e) {
  SchedState& s = sched_state();
  std::lock_guard<std::mutex> lk(s.mu);
  s.table.emplace(k, e);
  s.busy.erase(k);
  s.cv.notify_all();
}

void schedule_release(const SchedKey& k) {
  SchedState& s = sched_state();
  std::lock_guard<std::mutex> lk(s.mu);
  s.busy.erase(k);
  s.cv.notify_all();
}

void build_sell_for(edx_hamiltonian& h, Ctx& ctx) {
  SellPlan p = plan_sell(h, ctx);
  try {
    build_sell(h, ctx, p.lanes);
  } catch (...) {
    abandon_sell(h, p);
    throw;
  }
  finish_sell(h, ctx, p);
}

}  // namespace edx

extern "C" {

int edx_set_autotune(int32_t enable, int32_t precision, int32_t nb) {
  if (precision != EDX_PREC_FP64 && precision != EDX_PREC_MIXED) return EDX_ERR_INVALID;
  if (nb < 1 || nb > 16) return EDX_ERR_INVALID;
  auto& s = edx::state();
  std::lock_guard<std::mutex> lk(s.mu);
  s.enabled = enable != 0;
  s.prec = precision;
  s.nb = nb;
  return EDX_OK;
}

int edx_autotune_table(int32_t capacity, int32_t* keys, float* ms, int32_t* n) {
  if (!n || (capacity > 0 && (!keys || !ms))) return EDX_ERR_INVALID;
  auto& s = edx::state();
  std::lock_guard<std::mutex> lk(s.mu);
  int i = 0;
  for (const auto& kv : s.table) {
    if (i < capacity) {
      int32_t* r = keys + 5 * i;
      r[0] = kv.first.prec;
      r[1] = kv.first.nb;
      r[2] = kv.first.dim_log2;
      r[3] = kv.first.row_q;
      r[4] = kv.second.lanes;
      for (int c = 0; c < 3; ++c) ms[3 * i + c] = kv.second.ms[c];
    }
    ++i;
  }
  *n = i;
  return EDX_OK;
}

int edx_autotune_clear(void) {
  {
    auto& s = edx::state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.table.clear();
  }
  auto& s = edx::sched_state();
  std::lock_guard<std::mutex> lk(s.mu);
  s.table.clear();
  return EDX_OK;
}

int edx_schedule_table(int32_t capacity, int32_t* keys, float* ms, int32_t* n) {
  if (!n || (capacity > 0 && (!keys || !ms))) return EDX_ERR_INVALID;
  auto& s = edx::sched_state();
  std::lock_guard<std::mutex> lk(s.mu);
  int i = 0;
  for (const auto& kv : s.table) {
    if (i < capacity) {
      int32_t* r = keys + 6 * i;
      r[0] = kv.first.prec;
      r[1] = kv.first.b;
      r[2] = kv.first.dim_log2;
      r[3] = kv.first.basis;
      r[4] = kv.second.gram;
      r[5] = kv.second.update;
      float* m = ms + 5 * i;
      m[0] = kv.second.gram_ms[0];
      m[1] = kv.second.gram_ms[1];
      for (int c = 0; c < 3; ++c) m[2 + c] = kv.second.update_ms[c];
    }
    ++i;
  }
  *n = i;
  return EDX_OK;
}

int edx_set_solver_schedule(edx_ctx* ctx, int32_t gram, int32_t update) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(gram >= 0 && gram <= 2, EDX_ERR_INVALID, "Gram schedule must be 0 (automatic), 1 or 2");
    EDX_REQUIRE(update >= 0 && update <= 3, EDX_ERR_INVALID, "update schedule must be 0 (automatic) or 1..3");
    ctx->sched_gram = gram;
    ctx->sched_update = update;
  });
}

int edx_hamiltonian_autotune(edx_ctx* ctx, edx_hamiltonian* ham, int32_t precision, int32_t nb, int32_t* lanes,
                             float* ms) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(precision == EDX_PREC_FP64 || precision == EDX_PREC_MIXED, EDX_ERR_INVALID, "unknown precision");
    EDX_REQUIRE(nb >= 1 && nb <= 16, EDX_ERR_INVALID, "block width must be 1..16");
    EDX_REQUIRE(ham->dim > 0, EDX_ERR_INVALID, "empty Hamiltonian");
    EDX_REQUIRE(!ham->on_the_fly, EDX_ERR_INVALID, "an on-the-fly Hamiltonian has no sliced table to tune");
    edx::DeviceGuard dg(ctx->device);
    const edx::TuneEntry e = edx::tune_spmm(*ham, *ctx, precision, nb);
    if (lanes) *lanes = e.lanes;
    if (ms)
      for (int c = 0; c < 3; ++c) ms[c] = e.ms[c];
  });
}

}  // extern "C"
