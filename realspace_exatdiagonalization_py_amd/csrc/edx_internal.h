// Internal declarations shared by the edx HIP translation units.
//
// Masks are uint64 with site i (1-based) at bit i-1.  Ranks are colex ranks
// (sum_j C(b_j, j+1) over set bits b_0 < b_1 < ...), equal to the position of
// the mask in ascending numeric order among masks of the same popcount.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/edx.h"

#define EDX_MAX_SITES 63
#define EDX_MAX_GROUP 64

namespace edx {
// Largest palette of a packed sliced table: 32-bit words (complex128: 16 KiB
// of LDS per workgroup), and the wide layout (16-bit indices), whose palette
// is staged in LDS per 256-thread workgroup up to kPalLdsBytes (small or
// complex128 single-table products, grouped products), per 1024-thread
// workgroup up to kPalLdsBytesLarge (large complex64 products and larger
// palettes), and read through the caches above.
constexpr int kPalMax = 1024;
constexpr int kPalMaxWide = 16384;
constexpr int kPalLdsBytes = 24576;
constexpr int kPalLdsBytesLarge = 65536;
}  // namespace edx
#define EDX_MAX_TERMS 65535

namespace edx {

// ---------------------------------------------------------------------------
// Errors: every API call returns an int status and records text in the ctx.
// ---------------------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

#define EDX_HIP(call)                                                          \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess)                                                      \
      throw ::edx::Error{_e == hipErrorOutOfMemory ? EDX_ERR_OOM : EDX_ERR_HIP, \
                         std::string(#call) + ": " + hipGetErrorString(_e)};   \
  } while (0)

#define EDX_REQUIRE(cond, code, text)                                          \
  do {                                                                         \
    if (!(cond)) throw ::edx::Error{(code), (text)};                           \
  } while (0)

// Stream of the API call running on this thread (set by guarded() from the
// call's ctx).  Device buffers allocated during a call are stream-ordered on
// it: hipMallocAsync / hipFreeAsync, so that freeing memory does not
// synchronise the whole device (hipFree does) and streams of other contexts
// keep running.
inline hipStream_t& tls_stream() {
  thread_local hipStream_t s = nullptr;
  return s;
}

struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(tls_stream()) { tls_stream() = s; }
  ~StreamScope() { tls_stream() = prev; }
};

struct Ctx;
// Copies `rows` rows of `width` bytes from device memory (row pitch spitch)
// to host memory (row pitch dpitch) through the context's pinned staging
// buffers, on the context's stream; returns when the host copy is complete.
// Unlike a runtime copy into pageable memory it holds no runtime lock while
// the host side copies.
void copy_to_host(Ctx& ctx, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows);

// Device memory cache.  Freeing device memory through the runtime
// (hipFree, and also hipFreeAsync) takes a runtime-wide lock for
// milliseconds, during which no thread can launch or synchronise, so
// concurrent solves stall each other.  Released blocks are therefore kept:
// cache_free records, on the stream that used the block (or, for arrays that
// may be shared between contexts, on every live context stream of the
// device), the events after which it may be reused; cache_alloc hands out the
// smallest cached block that fits (at most 1/4 larger, or 1 MiB for small
// requests) and makes the allocating stream wait on those events (the host
// waits instead when there is no stream).  Only a miss calls hipMalloc; when
// that fails the cached blocks are returned to the runtime and it is retried.
void* cache_alloc(size_t bytes, int device, hipStream_t s);
void cache_free(void* p, int device, hipStream_t s, bool shared);

// Device allocation freed in the destructor, from the memory cache.  Work
// areas (Ordered, the default) are ordered on the stream of the API call that
// allocates them; arrays owned by the long-lived C-ABI objects (ObjBuf below)
// may be used by other contexts and freed from any thread, so their blocks
// become reusable only after the work of every live stream at release time.
template <class T, bool Ordered = true>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t st = nullptr;
  int dev = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), st(o.st), dev(o.dev) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; st = o.st; dev = o.dev; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    st = tls_stream();
    if (!count) return;
    EDX_HIP(hipGetDevice(&dev));
    p = static_cast<T*>(cache_alloc(count * sizeof(T), dev, st));
  }
  void release() {
    if (p) cache_free(p, dev, st, !Ordered);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

template <class T>
using ObjBuf = DevBuf<T, false>;

// Live per-class kernel timing with HIP events on the launch stream
// (enabled on request; used by benchmarks to time kernels inside a solve).
// K_GEMM: block updates (at most 16 outputs); K_COMBINE: combinations of
// more outputs (restarts, Ritz vectors) on FP64 vector FMAs, which bound
// them rather than HBM; K_COMBINE_MFMA: the same on the FP32 matrix cores
// (complex64).
enum KernelClass {
  K_SPMM = 0,
  K_SPMM32 = 1,
  K_GRAM = 2,
  K_GEMM = 3,
  K_COMBINE = 4,
  K_COMBINE_MFMA = 5,
  K_NCLASS = EDX_KERNEL_CLASSES
};

struct KernelTimer {
  bool on = false;
  struct Pending {
    hipEvent_t a, b;
    int cls;
    double bytes, flops;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double ms[K_NCLASS] = {};
  int64_t launches[K_NCLASS] = {};
  double bytes[K_NCLASS] = {};
  double flops[K_NCLASS] = {};  // FP64 (or FP32 matrix-core) arithmetic

  hipEvent_t take() {
    if (pool.empty()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  // Returns the start event (null when timing is off).
  hipEvent_t start(hipStream_t st) {
    if (!on) return nullptr;
    hipEvent_t e = take();
    if (e) (void)hipEventRecord(e, st);
    return e;
  }
  void stop(hipStream_t st, hipEvent_t a, int cls, double nbytes, double nflops = 0.0) {
    if (!on || !a) return;
    hipEvent_t b = take();
    if (!b) return;
    (void)hipEventRecord(b, st);
    pending.push_back(Pending{a, b, cls, nbytes, nflops});
    if (pending.size() > 4096) flush();
  }
  void flush() {
    for (auto& p : pending) {
      (void)hipEventSynchronize(p.b);
      float t = 0.f;
      if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
        ms[p.cls] += t;
        launches[p.cls] += 1;
        bytes[p.cls] += p.bytes;
        flops[p.cls] += p.flops;
      }
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void reset() {
    flush();
    for (int i = 0; i < K_NCLASS; ++i) ms[i] = 0, launches[i] = 0, bytes[i] = 0, flops[i] = 0;
  }
  ~KernelTimer() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  // Thread whose API call is using the context (0: none) and its nesting
  // depth: a context serves one thread at a time (its stream, staging
  // buffers, kernel timer and error text are not shared between threads).
  std::atomic<uintptr_t> owner{0};
  int depth = 0;
  int autotune_group = 0;  // lanes per row in the SpMV (0 = pick by heuristics)
  // 1: tables built on this context get the palette-packed sliced copy when
  // their values fit (edx_set_table_packing)
  int table_pack = 1;
  // batched solves driven from this context: round graphs captured, replayed
  int64_t graph_captures = 0, graph_replays = 0;
  // orthogonalisation schedule of the eigensolver (0 = tuned or default):
  // Gram 1 = vector FMAs, 2 = FP32 matrix cores (complex64, blocks <= 8);
  // block update 1..3 = stage/rows variants (kUpdateVariants); combination
  // of more than 16 outputs 1 = vector FMAs, 2 = FP32 matrix cores
  // (complex64)
  int sched_gram = 0;
  int sched_update = 0;
  int sched_combine = 0;
  KernelTimer timer;
  // pinned host staging owned by the context (allocated once, reused by
  // every solve on it): projection coefficients device -> host, and the
  // ring of update coefficients host -> device
  double2* pinned_in = nullptr;      // coherent: kernels write Gram results into it directly
  double2* pinned_in_dev = nullptr;  // its device address
  double2* pinned_out = nullptr;
  char* stage[2] = {nullptr, nullptr};  // pinned staging for device -> host copies (lazy)
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

// Orthogonalisation schedules of the eigensolver picked by timing (see
// tune.hip): a process-wide table keyed by (precision, block width,
// floor(log2 dim), basis size).
struct SchedKey {
  int prec, b, dim_log2, basis;
  bool operator<(const SchedKey& o) const {
    if (prec != o.prec) return prec < o.prec;
    if (b != o.b) return b < o.b;
    if (dim_log2 != o.dim_log2) return dim_log2 < o.dim_log2;
    return basis < o.basis;
  }
};
struct SchedEntry {
  int gram = 0, update = 0;            // kept choices (1-based)
  float gram_ms[2] = {0.f, 0.f};       // vector FMAs, matrix cores (0: not timed)
  float update_ms[3] = {0.f, 0.f, 0.f};
};
enum class SchedLookup { found, claimed, unavailable };
// true when the autotuner is on (edx_set_autotune)
bool tuning_enabled();
// found: *out holds the entry; claimed: the caller times the key and calls
// schedule_store (or schedule_release on failure); unavailable: another
// caller is still timing it, use the defaults
SchedLookup schedule_lookup(const SchedKey& k, SchedEntry* out);
void schedule_store(const SchedKey& k, const SchedEntry& e);
void schedule_release(const SchedKey& k);

// Largest Krylov basis and block width of the eigensolver.
constexpr int kMaxBasis = 224;
constexpr int kMaxBlock = 16;
// Sizes of the context's pinned staging buffers (double2 elements).
// pinned_in holds a Gram result ((kMaxBasis + kMaxBlock) x kMaxBlock), the
// factor written by the device Cholesky of a block (16 x 16) and its status
// word, then a ring of kLogSlots block-step logs: projection coefficients
// ((kMaxBasis + kMaxBlock) x kMaxBlock), the block coupling R (16 x 16) and
// four words of status.
constexpr size_t kPinnedGram = (size_t)(kMaxBasis + kMaxBlock) * kMaxBlock;
constexpr size_t kPinnedFactor = kMaxBlock * kMaxBlock + 1;
constexpr int kLogSlots = 16;
constexpr size_t kLogC = (size_t)(kMaxBasis + kMaxBlock) * kMaxBlock;
constexpr size_t kLogSlot = kLogC + kMaxBlock * kMaxBlock + 4;
constexpr size_t kPinnedIn = kPinnedGram + kPinnedFactor + kLogSlots * kLogSlot;
// pinned_out: a ring of four update-coefficient slots ((kMaxBasis +
// kMaxBlock) x kMaxBlock each), then the coefficients of a restart
// combination (up to (kMaxBasis + kMaxBlock) squared).
constexpr size_t kPinnedRing = 4 * (size_t)(kMaxBasis + kMaxBlock) * kMaxBlock;
constexpr size_t kPinnedCombine = (size_t)(kMaxBasis + kMaxBlock) * (kMaxBasis + kMaxBlock);
constexpr size_t kPinnedOut = kPinnedRing + kPinnedCombine;
// Device -> host copies into pageable memory go through two pinned staging
// buffers of this many bytes each (owned by the context, allocated on first
// use): DMA into one while the host copies the other out.
constexpr size_t kStageBytes = 8u << 20;

// Scoped device selection for API calls.
struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    EDX_HIP(hipGetDevice(&prev));
    if (prev != dev) EDX_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

// ---------------------------------------------------------------------------
// Group description shared by the catalog / phase / table kernels.
// ---------------------------------------------------------------------------
// One rank's catalog entry packed for a single 16-byte load (the on-the-fly
// product's gather): orbit index, group element, alpha (complex64).
struct __attribute__((aligned(16))) RankRec {
  int32_t orbit;
  int32_t g;
  float ax, ay;
};

struct GroupView {
  int n_site;
  int n_filled;
  int n_g;
  int fermionic;
  int product_phase;        // 1: alpha as running complex products (C == 1 case)
  const uint8_t* perm;      // [n_g][n_site] 0-based target bit
  const double* angle;      // [n_g][n_site] arg(eta_g(i))
  const double2* phase;     // [n_g][n_site] eta_g(i)
  const int32_t* inv;       // [n_g] index of the inverse element
  const uint32_t* binom;    // [(n_site+1)][(n_filled+1)] C(b, j) (entries < 2^31 used)
};

// Scan helpers (scan.hip).  Exclusive prefix sums, deterministic.
// counts of type uint32 -> offsets uint32/int64; returns the total.
uint64_t exclusive_scan_u32(Ctx& ctx, const uint32_t* in, uint32_t* out, size_t n);
uint64_t exclusive_scan_u32_to_i64(Ctx& ctx, const uint32_t* in, int64_t* out, size_t n);
// popcount of every word, scanned: out[w] = sum_{v<w} popcount(words[v]).
uint64_t exclusive_scan_popcount64(Ctx& ctx, const uint64_t* words, uint32_t* out, size_t n);

inline unsigned grid_for(size_t n, unsigned block) {
  size_t g = (n + block - 1) / block;
  return (unsigned)(g ? g : 1);
}

}  // namespace edx

// ---------------------------------------------------------------------------
// Device helpers (inlined into every kernel that uses them).
// ---------------------------------------------------------------------------
namespace edxd {

__device__ __forceinline__ int ctz64(uint64_t x) { return __builtin_ctzll(x); }
__device__ __forceinline__ int popc64(uint64_t x) { return __builtin_popcountll(x); }

// Colex rank with the binomial table (row b has stride nf1 = n_filled + 1).
__device__ __forceinline__ uint64_t rank_colex(uint64_t m, const uint32_t* binom, int nf1) {
  uint64_t r = 0;
  int j = 1;
  while (m) {
    int b = ctz64(m);
    r += binom[b * nf1 + j];
    ++j;
    m &= m - 1;
  }
  return r;
}

// Mask of a colex rank: for j = k..1 take the largest b with C(b, j) <= r.
__device__ __forceinline__ uint64_t unrank_colex(uint64_t r, int n_site, int k, const uint32_t* binom,
                                                 int nf1) {
  uint64_t m = 0;
  int b = n_site - 1;
  for (int j = k; j >= 1; --j) {
    while (b > 0 && (uint64_t)binom[b * nf1 + j] > r) --b;
    m |= 1ull << b;
    r -= binom[b * nf1 + j];
    --b;
  }
  return m;
}

// Image of m under group element g (permutation part only).
__device__ __forceinline__ uint64_t image(uint64_t m, const uint8_t* perm_g) {
  uint64_t img = 0;
  while (m) {
    img |= 1ull << perm_g[ctz64(m)];
    m &= m - 1;
  }
  return img;
}

// Parity (0/1) of the permutation restricted to the occupied sites of m.
__device__ __forceinline__ int perm_parity(uint64_t m, const uint8_t* perm_g) {
  uint64_t placed = 0;
  int par = 0;
  while (m) {
    int p = perm_g[ctz64(m)];
    par ^= popc64(placed >> p) & 1;  // targets above p already placed (p itself is free)
    placed |= 1ull << p;
    m &= m - 1;
  }
  return par;
}

// Phase of applying g to m in exponent form: exp(i * sum angle) with the angle
// sum formed per byte (bits ascending from 0.0), bytes added in ascending order
// from 0.0, then multiplied by the fermion sign as the complex (-1 + 0i).
__device__ __forceinline__ double2 alpha_exp(uint64_t m, int n_site, const double* angle_g,
                                             const uint8_t* perm_g, int fermionic) {
  const int n_bytes = (n_site + 7) >> 3;
  double tot = 0.0;
  for (int b = 0; b < n_bytes; ++b) {
    uint32_t byte = (uint32_t)(m >> (8 * b)) & 0xffu;
    double s = 0.0;
    while (byte) {
      int i = __builtin_ctz(byte);
      s = s + angle_g[8 * b + i];
      byte &= byte - 1;
    }
    tot = tot + s;
  }
  double sn, cs;
  sincos(tot, &sn, &cs);
  double2 a = make_double2(cs, sn);
  if (fermionic && perm_parity(m, perm_g)) {
    // (re + i im) * (-1 + 0i), formed term by term as in a complex product
    double re = a.x * -1.0 - a.y * 0.0;
    double im = a.x * 0.0 + a.y * -1.0;
    a = make_double2(re, im);
  }
  return a;
}

// Phase of applying g to m as a running complex product over occupied sites
// (ascending), starting at 1 + 0i, then times the integer parity (+1/-1).
__device__ __forceinline__ double2 alpha_prod(uint64_t m, const double2* phase_g,
                                              const uint8_t* perm_g, int fermionic) {
  double re = 1.0, im = 0.0;
  uint64_t t = m;
  while (t) {
    double2 e = phase_g[ctz64(t)];
    double nr = re * e.x - im * e.y;
    double ni = re * e.y + im * e.x;
    re = nr;
    im = ni;
    t &= t - 1;
  }
  double p = (fermionic && perm_parity(m, perm_g)) ? -1.0 : 1.0;
  double nr = re * p - im * 0.0;
  double ni = re * 0.0 + im * p;
  return make_double2(nr, ni);
}

}  // namespace edxd
