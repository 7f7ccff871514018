// Context management and device-memory helpers of the edx C ABI.
#include <cstring>

#include "dense.h"
#include "objects.h"

namespace edx {
std::string& tls_error() {
  thread_local std::string s;
  return s;
}
}  // namespace edx

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
    *out = ctx;
  });
}

int edx_ctx_destroy(edx_ctx* ctx) {
  if (!ctx) return EDX_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  if (ctx->pinned_in) (void)hipHostFree(ctx->pinned_in);
  if (ctx->pinned_out) (void)hipHostFree(ctx->pinned_out);
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
