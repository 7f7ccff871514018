// Allocation-free variant of edx_matvec for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, hosts
// that forbid hipMalloc).  The three work areas of edx_matvec (the
// column-major copy of x, its interleaved copy and the column-major y) are
// carved out of one statically allocated device array per GPU instead of the
// device memory cache.
#include <mutex>

#include "objects.h"

namespace edx {
void interleave(hipStream_t st, int precision, int64_t dim, int nb, const void* src, int64_t ld, void* dst);
void apply_table(const edx_hamiltonian& h, Ctx& ctx, int precision, int nb, const void* X, void* Y, int64_t ldy);
}  // namespace edx

namespace {

// 96 MiB per GPU, held from the first context on that GPU to the end of the
// process: one batch of nb <= 16 complex128 vectors of length dim needs
// 3 * 16 * nb * dim bytes, so a single vector fits up to dim = 2 097 152.
constexpr size_t kScratchBytes = size_t(96) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex& g_mu = *new std::mutex;  // immortal: used while the process exits
void* g_ptr[kMaxDevices] = {};       // device address of `scratch` on each GPU
hipError_t g_err[kMaxDevices] = {};  // failure of its look-up, if any
std::mutex g_call_mu[kMaxDevices];   // serialises no-alloc calls per GPU

}  // namespace

namespace edx {

// Looks up the device address of the scratch array on `device` (which must be
// current) once per device; called from edx_ctx_create.  A failed look-up is
// recorded and reported by the entry point, it does not fail the context.
void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device] || g_err[device] != hipSuccess) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    g_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return;
  }
  g_ptr[device] = p;
}

}  // namespace edx

extern "C" {

int edx_matvec_noalloc(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t n_vec, const double* x, double* y) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham && x && y, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(n_vec >= 0, EDX_ERR_INVALID, "negative vector count");
    const int64_t D = ham->dim;
    if (D == 0 || n_vec == 0) return;
    const int dev = ctx->device;
    EDX_REQUIRE(dev >= 0 && dev < kMaxDevices, EDX_ERR_INVALID, "device index out of range");
    const size_t per_vec = (size_t)3 * 16 * (size_t)D;
    EDX_REQUIRE(per_vec <= kScratchBytes, EDX_ERR_INVALID,
                "edx_matvec_noalloc: sector dimension " + std::to_string(D) +
                    " exceeds the scratch limit (3 x 16 x dim bytes per vector must fit in " +
                    std::to_string(kScratchBytes) + " bytes, i.e. dim <= " +
                    std::to_string(kScratchBytes / 48) + ")");
    char* base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      EDX_REQUIRE(g_err[dev] == hipSuccess, EDX_ERR_HIP,
                  std::string("no-alloc scratch unavailable: hipGetSymbolAddress: ") + hipGetErrorString(g_err[dev]));
      base = static_cast<char*>(g_ptr[dev]);
    }
    EDX_REQUIRE(base, EDX_ERR_INTERNAL, "no-alloc scratch was not set up on this device");
    std::lock_guard<std::mutex> call_lk(g_call_mu[dev]);
    edx::DeviceGuard dg(dev);
    hipStream_t st = ctx->stream;
    // the live kernel timer may create events; this path creates nothing
    const bool timer_on = ctx->timer.on;
    ctx->timer.on = false;
    struct Restore {
      edx_ctx* c;
      bool on;
      ~Restore() { c->timer.on = on; }
    } restore{ctx, timer_on};
    const int nb_max = (int)std::min<size_t>(16, kScratchBytes / per_vec);
    // every element of the three areas that a batch reads is written by that
    // batch first (rows < dim of nb vectors), so nothing needs clearing
    for (int32_t v0 = 0; v0 < n_vec; v0 += nb_max) {
      const int nb = std::min<int32_t>(nb_max, n_vec - v0);
      const size_t area = (size_t)nb * (size_t)D * 16;
      double2* xc = reinterpret_cast<double2*>(base);
      double2* xi = reinterpret_cast<double2*>(base + area);
      double2* yc = reinterpret_cast<double2*>(base + 2 * area);
      EDX_HIP(hipMemcpyAsync(xc, x + (size_t)v0 * D * 2, area, hipMemcpyHostToDevice, st));
      edx::interleave(st, EDX_PREC_FP64, D, nb, xc, D, xi);
      edx::apply_table(*ham, *ctx, EDX_PREC_FP64, nb, xi, yc, D);
      EDX_HIP(hipMemcpyAsync(y + (size_t)v0 * D * 2, yc, area, hipMemcpyDeviceToHost, st));
      EDX_HIP(hipStreamSynchronize(st));
    }
  });
}

}  // extern "C"
