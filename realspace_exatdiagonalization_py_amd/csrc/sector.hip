// Irrep filter: an orbit belongs to the sector of the 1-D irrep chi iff
// |chi(h) - alpha_h(rep)| <= atol for every stabilizer element h
// (reference _is_orbit_compatible / build_symmetry_sector_basis,
// SRE:1484-1546).  Kept orbits are compacted in catalog (ascending rep) order.
#include "objects.h"

using edx::DevBuf;

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock)
sector_flags(int64_t n_orbits, int n_g, const double2* chi, double atol, const uint32_t* stab_off,
             const uint8_t* stab_g, const double2* stab_phase, uint32_t* flags) {
  __shared__ double2 chi_s[EDX_MAX_GROUP];
  for (int i = threadIdx.x; i < n_g; i += blockDim.x) chi_s[i] = chi[i];
  __syncthreads();
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n_orbits) return;
  uint32_t ok = 1;
  for (uint32_t j = stab_off[o]; j < stab_off[o + 1]; ++j) {
    const double2 c = chi_s[stab_g[j]];
    const double2 a = stab_phase[j];
    // "not |chi - alpha| <= atol" rejects, so a NaN difference rejects too
    if (!(hypot(c.x - a.x, c.y - a.y) <= atol)) {
      ok = 0;
      break;
    }
  }
  flags[o] = ok;
}

__global__ void __launch_bounds__(kBlock)
sector_compact(int64_t n_orbits, const uint32_t* flags, const uint32_t* prefix, const uint64_t* reps,
               const int32_t* stab_order, uint64_t* basis_reps, int32_t* basis_stab, int32_t* orbit_to_basis) {
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n_orbits) return;
  if (flags[o]) {
    const uint32_t b = prefix[o];
    basis_reps[b] = reps[o];
    basis_stab[b] = stab_order[o];
    orbit_to_basis[o] = (int32_t)b;
  } else {
    orbit_to_basis[o] = -1;
  }
}

}  // namespace

extern "C" {

int edx_sector_build(edx_ctx* ctx, const edx_catalog* cat, const double* chi, double atol, edx_sector** out,
                     int64_t* dim) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && cat && chi && out, EDX_ERR_INVALID, "null argument");
    *out = nullptr;
    edx::DeviceGuard dg(ctx->device);
    const int64_t n_orb = cat->n_orbits;
    DevBuf<double2> chi_d(cat->n_g);
    EDX_HIP(hipMemcpyAsync(chi_d.p, chi, (size_t)cat->n_g * 16, hipMemcpyHostToDevice, ctx->stream));
    DevBuf<uint32_t> flags(n_orb), prefix(n_orb);
    hipLaunchKernelGGL(sector_flags, dim3(edx::grid_for(n_orb, kBlock)), dim3(kBlock), 0, ctx->stream, n_orb,
                       cat->n_g, (const double2*)chi_d.p, atol, (const uint32_t*)cat->stab_off.p,
                       (const uint8_t*)cat->stab_g.p, (const double2*)cat->stab_phase.p, flags.p);
    EDX_HIP(hipGetLastError());
    const uint64_t D = edx::exclusive_scan_u32(*ctx, flags.p, prefix.p, (size_t)n_orb);
    auto sec = new edx_sector();
    try {
      sec->device = ctx->device;
      sec->cat = cat;
      sec->dim = (int64_t)D;
      sec->reps.alloc(D);
      sec->stab_order.alloc(D);
      sec->orbit_to_basis.alloc(n_orb);
      hipLaunchKernelGGL(sector_compact, dim3(edx::grid_for(n_orb, kBlock)), dim3(kBlock), 0, ctx->stream, n_orb,
                         (const uint32_t*)flags.p, (const uint32_t*)prefix.p, (const uint64_t*)cat->reps.p,
                         (const int32_t*)cat->stab_order.p, sec->reps.p, sec->stab_order.p,
                         sec->orbit_to_basis.p);
      EDX_HIP(hipGetLastError());
      EDX_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      delete sec;
      throw;
    }
    *out = sec;
    if (dim) *dim = (int64_t)D;
  });
}

int edx_sector_fetch(edx_ctx* ctx, const edx_sector* sec, uint64_t* reps, int32_t* stab_order,
                     int32_t* orbit_to_basis) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && sec, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    if (reps && sec->reps.n)
      EDX_HIP(hipMemcpyAsync(reps, sec->reps.p, sec->reps.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    if (stab_order && sec->stab_order.n)
      EDX_HIP(hipMemcpyAsync(stab_order, sec->stab_order.p, sec->stab_order.bytes(), hipMemcpyDeviceToHost,
                             ctx->stream));
    if (orbit_to_basis && sec->orbit_to_basis.n)
      EDX_HIP(hipMemcpyAsync(orbit_to_basis, sec->orbit_to_basis.p, sec->orbit_to_basis.bytes(),
                             hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_sector_destroy(edx_sector* sec) {
  if (!sec) return EDX_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(sec->device);
  delete sec;
  (void)hipSetDevice(prev);
  return EDX_OK;
}

}  // extern "C"
