// Context management and device-memory helpers of the edx C ABI.
#include <cstring>

#include <atomic>
#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "dense.h"
#include "objects.h"

namespace edx {
std::string& tls_error() {
  thread_local std::string s;
  return s;
}
}  // namespace edx

namespace edx {
namespace {
// Live context streams per device (the memory cache orders the reuse of
// shared blocks after their work).
// Immortal (never destroyed): objects may release memory while the process
// exits, after static destructors have run.
std::mutex& g_streams_mu = *new std::mutex;
std::vector<std::pair<int, hipStream_t>>& g_streams = *new std::vector<std::pair<int, hipStream_t>>;

void register_stream(int device, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_streams_mu);
  g_streams.emplace_back(device, s);
}

void unregister_stream(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_streams_mu);
  for (size_t i = 0; i < g_streams.size(); ++i)
    if (g_streams[i].second == s) {
      g_streams.erase(g_streams.begin() + (long)i);
      break;
    }
}
}  // namespace

namespace {
struct CachedBlock {
  void* p;
  int device;
  std::vector<hipEvent_t> deps;  // to complete before the block is reused
};
std::mutex& g_cache_mu = *new std::mutex;
std::multimap<size_t, CachedBlock>& g_cache = *new std::multimap<size_t, CachedBlock>;  // free blocks by size
std::unordered_map<void*, size_t>& g_sizes = *new std::unordered_map<void*, size_t>;  // every block owned -> size
std::vector<hipEvent_t>& g_events = *new std::vector<hipEvent_t>;                       // spare events
// Set when the process starts exiting: blocks released afterwards are left to
// the process teardown (the runtime may already be shutting down).
std::atomic<bool> g_exiting{false};
void mark_exiting() { g_exiting.store(true); }

hipEvent_t take_event() {
  if (!g_events.empty()) {
    hipEvent_t e = g_events.back();
    g_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

size_t round_bytes(size_t b) {
  if (b <= (1u << 20)) return (b + 4095) / 4096 * 4096;
  return (b + (2u << 20) - 1) / (2u << 20) * (2u << 20);
}

// Returns every cached block of the device to the runtime (after its events).
void drain_cache(int device) {
  for (auto it = g_cache.begin(); it != g_cache.end();) {
    if (it->second.device != device) {
      ++it;
      continue;
    }
    for (hipEvent_t e : it->second.deps) {
      (void)hipEventSynchronize(e);
      g_events.push_back(e);
    }
    (void)hipFree(it->second.p);
    g_sizes.erase(it->second.p);
    it = g_cache.erase(it);
  }
}
}  // namespace

void* cache_alloc(size_t bytes, int device, hipStream_t s) {
  static const bool registered = [] { return std::atexit(mark_exiting) == 0; }();
  (void)registered;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  const size_t want = round_bytes(bytes);
  const size_t limit = std::max(want + want / 4, want + (1u << 20));
  for (auto it = g_cache.lower_bound(want); it != g_cache.end() && it->first <= limit; ++it) {
    if (it->second.device != device) continue;
    CachedBlock blk = std::move(it->second);
    g_cache.erase(it);
    for (hipEvent_t e : blk.deps) {
      if (s) EDX_HIP(hipStreamWaitEvent(s, e, 0));
      else EDX_HIP(hipEventSynchronize(e));
      g_events.push_back(e);
    }
    return blk.p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    drain_cache(device);
    e = hipMalloc(&p, want);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error{EDX_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed"};
  }
  g_sizes[p] = want;
  return p;
}

void cache_free(void* p, int device, hipStream_t s, bool shared) {
  if (!p || g_exiting.load()) return;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto sz = g_sizes.find(p);
  if (sz == g_sizes.end()) return;  // not a cache block
  CachedBlock blk{p, device, {}};
  int prev = 0;
  (void)hipGetDevice(&prev);
  if (prev != device) (void)hipSetDevice(device);
  auto record_on = [&](hipStream_t st) {
    hipEvent_t e = take_event();
    if (e && hipEventRecord(e, st) == hipSuccess) blk.deps.push_back(e);
    else if (e) g_events.push_back(e);
  };
  if (shared || !s) {
    // every live context stream (the allocating one may be gone already)
    std::lock_guard<std::mutex> lk2(g_streams_mu);
    for (auto& st : g_streams)
      if (st.first == device) record_on(st.second);
  } else {
    record_on(s);
  }
  (void)hipGetLastError();
  if (prev != device) (void)hipSetDevice(prev);
  g_cache.emplace(sz->second, std::move(blk));
}

void copy_to_host(Ctx& ctx, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows) {
  if (!rows || !width) return;
  for (int i = 0; i < 2; ++i) {
    if (!ctx.stage[i]) EDX_HIP(hipHostMalloc((void**)&ctx.stage[i], kStageBytes, hipHostMallocDefault));
    if (!ctx.stage_ev[i]) EDX_HIP(hipEventCreateWithFlags(&ctx.stage_ev[i], hipEventDisableTiming));
  }
  // pieces of at most kStageBytes: whole rows when a row fits, else row parts
  struct Piece {
    size_t row, off, bytes;
  };
  std::vector<Piece> pieces;
  if (width <= kStageBytes) {
    const size_t per = kStageBytes / width;
    for (size_t r = 0; r < rows; r += per) pieces.push_back({r, 0, std::min(per, rows - r) * width});
  } else {
    for (size_t r = 0; r < rows; ++r)
      for (size_t o = 0; o < width; o += kStageBytes) pieces.push_back({r, o, std::min(kStageBytes, width - o)});
  }
  auto issue = [&](size_t k) {
    const Piece& pc = pieces[k];
    char* buf = ctx.stage[k % 2];
    if (width <= kStageBytes) {
      const size_t nr = pc.bytes / width;
      EDX_HIP(hipMemcpy2DAsync(buf, width, static_cast<const char*>(src) + pc.row * spitch, spitch, width, nr,
                               hipMemcpyDeviceToHost, ctx.stream));
    } else {
      EDX_HIP(hipMemcpyAsync(buf, static_cast<const char*>(src) + pc.row * spitch + pc.off, pc.bytes,
                             hipMemcpyDeviceToHost, ctx.stream));
    }
    EDX_HIP(hipEventRecord(ctx.stage_ev[k % 2], ctx.stream));
  };
  issue(0);
  for (size_t k = 0; k < pieces.size(); ++k) {
    EDX_HIP(hipEventSynchronize(ctx.stage_ev[k % 2]));
    if (k + 1 < pieces.size()) issue(k + 1);  // the other buffer was emptied by the previous iteration
    const Piece& pc = pieces[k];
    const char* buf = ctx.stage[k % 2];
    if (width <= kStageBytes) {
      const size_t nr = pc.bytes / width;
      for (size_t r = 0; r < nr; ++r) std::memcpy(static_cast<char*>(dst) + (pc.row + r) * dpitch, buf + r * width, width);
    } else {
      std::memcpy(static_cast<char*>(dst) + pc.row * dpitch + pc.off, buf, pc.bytes);
    }
  }
}

}  // namespace edx

namespace {
// Streaming-read reference for the bandwidth figures: a grid-stride sum over
// n 16-byte words with four loads in flight per lane; the sum is written only
// if it equals an impossible value, so the loads cannot be dropped.
__global__ void __launch_bounds__(256) stream_read(const float4* __restrict__ a, size_t n, float* sink) {
  float s = 0.f;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const float4 x0 = a[i], x1 = a[i + stride], x2 = a[i + 2 * stride], x3 = a[i + 3 * stride];
    s += x0.x + x1.y + x2.z + x3.w;
  }
  for (; i < n; i += stride) s += a[i].x;
  if (s == -1.2345e-30f) sink[0] = s;
}
}  // namespace

extern "C" {

int32_t edx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int edx_ctx_create(int32_t device, edx_ctx** out) {
  return edx::guarded(nullptr, [&] {
    EDX_REQUIRE(out, EDX_ERR_INVALID, "null argument");
    *out = nullptr;
    const int n = edx_device_count();
    EDX_REQUIRE(n > 0, EDX_ERR_HIP, "no HIP device is visible");
    EDX_REQUIRE(device >= 0 && device < n, EDX_ERR_INVALID, "device index out of range");
    edx::DeviceGuard dg(device);
    // freed stream-ordered memory stays in the device pool for reuse
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
    auto ctx = new edx_ctx();
    ctx->device = device;
    // EDX_TABLE_PACKING=0: new contexts start with table packing off
    if (const char* pk = std::getenv("EDX_TABLE_PACKING")) ctx->table_pack = pk[0] == '0' ? 0 : 1;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    // Gram results are written by kernels straight into pinned_in: coherent
    // (fine-grained) host memory, read by the host after a stream sync
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&ctx->pinned_in, edx::kPinnedIn * sizeof(double2),
                        hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&ctx->pinned_in_dev, ctx->pinned_in, 0);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&ctx->pinned_out, edx::kPinnedOut * sizeof(double2), hipHostMallocDefault);
    if (e != hipSuccess) {
      if (ctx->pinned_in) (void)hipHostFree(ctx->pinned_in);
      if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
      delete ctx;
      throw edx::Error{EDX_ERR_HIP, std::string("context setup: ") + hipGetErrorString(e)};
    }
    edx::register_stream(device, ctx->stream);
    // device address of the no-alloc entry point's scratch array (once per device)
    edx::noalloc_init(device);
    *out = ctx;
  });
}

int edx_ctx_destroy(edx_ctx* ctx) {
  if (!ctx) return EDX_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    edx::unregister_stream(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  if (ctx->pinned_in) (void)hipHostFree(ctx->pinned_in);
  if (ctx->pinned_out) (void)hipHostFree(ctx->pinned_out);
  for (int i = 0; i < 2; ++i) {
    if (ctx->stage[i]) (void)hipHostFree(ctx->stage[i]);
    if (ctx->stage_ev[i]) (void)hipEventDestroy(ctx->stage_ev[i]);
  }
  delete ctx;
  (void)hipSetDevice(prev);
  return EDX_OK;
}

const char* edx_last_error(const edx_ctx* ctx) {
  if (ctx) return ctx->last_error.c_str();
  return edx::tls_error().c_str();
}

int edx_synchronize(edx_ctx* ctx) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_set_spmv_group(edx_ctx* ctx, int32_t lanes) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4, EDX_ERR_INVALID,
                "lanes per row must be 0 (automatic), 1, 2 or 4");
    ctx->autotune_group = lanes;
  });
}

int edx_hamiltonian_set_spmv_group(edx_ctx* ctx, edx_hamiltonian* ham, int32_t lanes) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4, EDX_ERR_INVALID, "lanes per row must be 1, 2 or 4");
    EDX_REQUIRE(!ham->on_the_fly, EDX_ERR_INVALID, "an on-the-fly Hamiltonian has no sliced table");
    edx::DeviceGuard dg(ctx->device);
    if (lanes != ham->sell_g) edx::build_sell(*ham, *ctx, lanes);
  });
}

int edx_hamiltonian_info(const edx_hamiltonian* ham, int64_t* dim, int64_t* n_entries, int32_t* spmv_group) {
  if (!ham) return EDX_ERR_INVALID;
  if (dim) *dim = ham->dim;
  if (n_entries) *n_entries = ham->n_entries;
  if (spmv_group) *spmv_group = ham->sell_g;
  return EDX_OK;
}

int edx_set_table_packing(edx_ctx* ctx, int32_t enable) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(enable == 0 || enable == 1, EDX_ERR_INVALID, "table packing must be 0 or 1");
    ctx->table_pack = enable;
  });
}

int edx_hamiltonian_layout(const edx_hamiltonian* ham, int32_t* packed, int32_t* n_palette, int64_t* stored_slots,
                           int64_t* table_bytes) {
  if (!ham) return EDX_ERR_INVALID;
  if (packed) *packed = ham->sell_packed;
  if (n_palette) *n_palette = ham->n_pal;
  if (stored_slots) *stored_slots = ham->sell_entries;
  if (table_bytes)
    *table_bytes = (int64_t)(ham->sell_off.bytes() + ham->sell_col.bytes() + ham->sell_idx.bytes() +
                             ham->sell_val.bytes() + ham->sell_val32.bytes() + ham->pal.bytes() + ham->pal32.bytes());
  return EDX_OK;
}

int edx_kernel_stats(edx_ctx* ctx, int32_t enable, int32_t reset, double* ms, int64_t* launches, double* bytes) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timer.flush();
    if (reset) ctx->timer.reset();
    if (enable >= 0) ctx->timer.on = enable != 0;
    for (int i = 0; i < edx::K_NCLASS; ++i) {
      if (ms) ms[i] = ctx->timer.ms[i];
      if (launches) launches[i] = ctx->timer.launches[i];
      if (bytes) bytes[i] = ctx->timer.bytes[i];
    }
  });
}

int edx_kernel_flops(edx_ctx* ctx, double* flops) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && flops, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timer.flush();
    for (int i = 0; i < edx::K_NCLASS; ++i) flops[i] = ctx->timer.flops[i];
  });
}

int edx_device_alloc(edx_ctx* ctx, uint64_t bytes, void** out) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && out, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    *out = nullptr;
    if (bytes) EDX_HIP(hipMalloc(out, bytes));
  });
}

int edx_device_free(edx_ctx* ctx, void* p) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    if (p) EDX_HIP(hipFree(p));
  });
}

int edx_memcpy_h2d(edx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && dst && src, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_memcpy_d2h(edx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && dst && src, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_mem_info(edx_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && free_bytes && total_bytes, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    size_t f = 0, t = 0;
    EDX_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
  });
}

int edx_device_copy(edx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && dst && src, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_time_stream_read(edx_ctx* ctx, const void* src, uint64_t bytes, int32_t n_rep, double* best_ms) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && src && best_ms && n_rep > 0 && bytes >= 16 && bytes % 16 == 0, EDX_ERR_INVALID,
                "bad argument (bytes: a positive multiple of 16)");
    edx::DeviceGuard dg(ctx->device);
    edx::DevBuf<float> sink;
    sink.alloc(1);
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>(65536, (n + 255) / 256);
    hipEvent_t a, b;
    EDX_HIP(hipEventCreate(&a));
    EDX_HIP(hipEventCreate(&b));
    float best = INFINITY;
    for (int r = 0; r <= n_rep; ++r) {  // the first launch warms up
      EDX_HIP(hipEventRecord(a, ctx->stream));
      hipLaunchKernelGGL(stream_read, dim3(grid), dim3(256), 0, ctx->stream, (const float4*)src, n, sink.p);
      EDX_HIP(hipEventRecord(b, ctx->stream));
      EDX_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      EDX_HIP(hipEventElapsedTime(&ms, a, b));
      if (r > 0) best = std::min(best, ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *best_ms = best;
  });
}

// Host-only: eigen-decomposition of a Hermitian n x n matrix given row-major as
// (re, im) pairs, using its upper (upper != 0) or lower triangle.  w: n
// ascending eigenvalues; z (optional): row-major eigenvectors as columns.
int edx_host_herm_eig(int32_t n, const double* a, int32_t upper, double* w, double* z) {
  return edx::guarded(nullptr, [&] {
    EDX_REQUIRE(n >= 0 && a && w, EDX_ERR_INVALID, "bad argument");
    edx::dense::Mat A(n, n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) A(i, j) = edx::dense::cplx(a[2 * (i * n + j)], a[2 * (i * n + j) + 1]);
    std::vector<double> ww;
    edx::dense::Mat Z;
    edx::dense::herm_eig(A, ww, Z, upper == 0);
    for (int i = 0; i < n; ++i) w[i] = ww[i];
    if (z)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          z[2 * (i * n + j)] = Z(i, j).real();
          z[2 * (i * n + j) + 1] = Z(i, j).imag();
        }
  });
}

}  // extern "C"
