// Consumers of sector eigenvectors that reuse the representative tables:
// expansion to Fock amplitudes (reference
// observables/entanglement_spectrum.py:82-144), the projected position
// operator (observables/charge_pump.py:81-162) and the one-body density matrix
// (observables/off_diagonal_long_range_order.py:43-140).  Group actions use
// the product-form phase of apply_operation_to_mask (SRE:977-1008) and
// complex products are formed term by term as numpy forms them
// (re = a.re b.re - a.im b.im, im = a.re b.im + a.im b.re; no fused ops).
#include <cmath>

#include "objects.h"

using edx::DevBuf;

namespace {

constexpr int kBlock = 256;
constexpr int kSumBlocks = 1024;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// Amplitude of the Fock state of colex rank r:
//   sum over g ascending with g.s == m of (c/sqrt(|G||Stab|)) * conj(chi(g)) * alpha_g(s)
// (convention 0, the reference's formula), or with chi(g) * conj(alpha_g(s))
// (convention 1, the expansion under which the block's eigenvectors are
// eigenvectors of the Fock-space Hamiltonian).  Entries with |amp| <= atol
// are set to zero, as the reference prunes them.
__global__ void __launch_bounds__(kBlock)
fock_amp(edx::GroupView gv, int64_t C, const int32_t* __restrict__ rep_rank, const int32_t* __restrict__ o2b,
         const uint64_t* __restrict__ sec_reps, const int32_t* __restrict__ sec_stab, const double2* __restrict__ chi,
         const double2* __restrict__ coeff, double atol, int convention, double2* __restrict__ amp) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= C) return;
  double2 acc = make_double2(0.0, 0.0);
  const int o = rep_rank[r];
  const int col = o >= 0 ? o2b[o] : -1;
  if (col >= 0) {
    const double2 c = coeff[col];
    if (hypot(c.x, c.y) > atol) {
      const uint64_t s = sec_reps[col];
      const uint64_t m = edxd::unrank_colex((uint64_t)r, gv.n_site, gv.n_filled, gv.binom, gv.n_filled + 1);
      const double d = sqrt((double)gv.n_g * (double)sec_stab[col]);
      const double2 pref = make_double2(c.x / d, c.y / d);
      for (int g = 0; g < gv.n_g; ++g) {
        const uint8_t* pg = gv.perm + (size_t)g * gv.n_site;
        if (edxd::image(s, pg) != m) continue;
        const double2 a = edxd::alpha_prod(s, gv.phase + (size_t)g * gv.n_site, pg, gv.fermionic);
        const double2 t = convention == 0 ? cmul(cmul(pref, conj2(chi[g])), a) : cmul(cmul(pref, chi[g]), conj2(a));
        acc = make_double2(acc.x + t.x, acc.y + t.y);
      }
    }
  }
  if (!(hypot(acc.x, acc.y) > atol)) acc = make_double2(0.0, 0.0);
  amp[r] = acc;
}

// partial[b] = sum of |amp|^2 over a grid-stride range (fixed grid, so the
// total is summed in a fixed order).
__global__ void __launch_bounds__(kBlock) sumsq(int64_t n, const double2* amp, double* partial) {
  __shared__ double red[kBlock];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double2 a = amp[i];
    const double h = hypot(a.x, a.y);
    s += h * h;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kBlock) nonzero_flags(int64_t n, const double2* amp, uint32_t* flags) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double2 a = amp[i];
  flags[i] = (a.x != 0.0 || a.y != 0.0) ? 1u : 0u;
}

__global__ void __launch_bounds__(kBlock)
fock_compact(edx::GroupView gv, int64_t n, const double2* amp, const uint32_t* flags, const uint32_t* pos,
             double inv_norm, uint64_t* masks, double2* out) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= n || !flags[r]) return;
  const double2 a = amp[r];
  const uint32_t p = pos[r];
  masks[p] = edxd::unrank_colex((uint64_t)r, gv.n_site, gv.n_filled, gv.binom, gv.n_filled + 1);
  out[p] = make_double2(a.x * inv_norm, a.y * inv_norm);
}

// prod over occupied sites (ascending) of site_phases, from 1 + 0i.
__device__ __forceinline__ double2 position_phase(uint64_t m, const double2* ph) {
  double2 z = make_double2(1.0, 0.0);
  while (m) {
    z = cmul(z, ph[edxd::ctz64(m)]);
    m &= m - 1;
  }
  return z;
}

// One column of <chi_to| U |chi_from>: the source representative s pairs with
// its row in the target sector; elem = (1/|G|) sum_g chi_to(g) conj(chi_from(g)) U(g.s).
__global__ void __launch_bounds__(kBlock)
position_elems(edx::GroupView gv, int64_t D_from, const uint64_t* __restrict__ reps_from,
               const int32_t* __restrict__ rep_rank, const int32_t* __restrict__ o2b_to,
               const double2* __restrict__ chi_to, const double2* __restrict__ chi_from,
               const double2* __restrict__ site_ph, int64_t* rows, double2* vals) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= D_from) return;
  const uint64_t s = reps_from[col];
  const uint64_t idx = edxd::rank_colex(s, gv.binom, gv.n_filled + 1);
  const int row = o2b_to[rep_rank[idx]];
  if (row < 0) {
    rows[col] = -1;
    return;
  }
  double2 elem = make_double2(0.0, 0.0);
  for (int g = 0; g < gv.n_g; ++g) {
    const uint64_t m = edxd::image(s, gv.perm + (size_t)g * gv.n_site);
    const double2 t = cmul(cmul(chi_to[g], conj2(chi_from[g])), position_phase(m, site_ph));
    elem = make_double2(elem.x + t.x, elem.y + t.y);
  }
  const double inv_ng = 1.0 / (double)gv.n_g;
  elem = make_double2(elem.x * inv_ng, elem.y * inv_ng);
  if (hypot(elem.x, elem.y) > 1e-13) {
    rows[col] = row;
    vals[col] = elem;
  } else {
    rows[col] = -1;
  }
}

// One (column, g) pair of the one-body density matrix, accumulated into LDS
// (n_site <= kLdsSites) and then into rho with atomics; larger lattices add to
// rho directly.  rho[i][j] = <psi| a_i^dag a_j |psi>, 0-based, row-major.
constexpr int kLdsSites = 40;

__global__ void __launch_bounds__(kBlock)
one_body_density(edx::GroupView gv, int64_t D, const uint64_t* __restrict__ reps, const int32_t* __restrict__ stab,
                 const int32_t* __restrict__ rep_rank, const int32_t* __restrict__ rep_sym,
                 const float2* __restrict__ rep_amp, const int32_t* __restrict__ o2b,
                 const double2* __restrict__ chi, const double2* __restrict__ coeff, double* rho) {
  __shared__ double acc[kLdsSites * kLdsSites * 2];
  const int N = gv.n_site;
  const bool lds = N <= kLdsSites;
  if (lds) {
    for (int i = threadIdx.x; i < N * N * 2; i += kBlock) acc[i] = 0.0;
    __syncthreads();
  }
  double* dst = lds ? acc : rho;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t col = t / gv.n_g;
  const int g = (int)(t % gv.n_g);
  if (col < D) {
    const double2 c = coeff[col];
    if (c.x != 0.0 || c.y != 0.0) {
      const double inv_ng = 1.0 / (double)gv.n_g;
      const double inv_sqrt_ng = 1.0 / sqrt((double)gv.n_g);
      const double hc = hypot(c.x, c.y);
      const double diag_w = hc * hc * inv_ng;
      const double ket_norm = inv_sqrt_ng / sqrt((double)stab[col]);
      const uint64_t s = reps[col];
      const uint8_t* pg = gv.perm + (size_t)g * gv.n_site;
      const uint64_t km = edxd::image(s, pg);
      const double2 kph = edxd::alpha_prod(s, gv.phase + (size_t)g * gv.n_site, pg, gv.fermionic);
      const double2 cg = conj2(chi[g]);
      const double2 ket_basis = cmul(make_double2(ket_norm * cg.x, ket_norm * cg.y), kph);
      const double2 ket_amp = cmul(c, ket_basis);
      uint64_t occ = km;
      while (occ) {
        const int j = edxd::ctz64(occ);  // 0-based
        occ &= occ - 1;
        atomicAdd(&dst[2 * (j * N + j)], diag_w);
        for (int i = 0; i < N; ++i) {
          if (i == j || ((km >> i) & 1ull)) continue;
          const uint64_t nm = (km & ~(1ull << j)) | (1ull << i);
          const uint64_t idx = edxd::rank_colex(nm, gv.binom, gv.n_filled + 1);
          const int o = rep_rank[idx];
          const int row = o >= 0 ? o2b[o] : -1;
          if (row < 0) continue;
          const float2 a32 = rep_amp[idx];
          const double2 coef = cmul(make_double2((double)a32.x, (double)a32.y), conj2(chi[rep_sym[idx]]));
          const double sq = sqrt((double)stab[row]) * inv_sqrt_ng;
          const double2 cc = conj2(coef);
          const double2 bra_basis = make_double2(sq * cc.x, sq * cc.y);
          const double2 bra_amp = cmul(coeff[row], bra_basis);
          double2 term = cmul(conj2(bra_amp), ket_amp);
          if (gv.fermionic) {
            const int lo = min(i, j), hi = max(i, j);
            const uint64_t between = hi - lo > 1 ? (((1ull << (hi - lo - 1)) - 1ull) << (lo + 1)) : 0ull;
            if (edxd::popc64(km & between) & 1) term = cmul(term, make_double2(-1.0, 0.0));
          }
          atomicAdd(&dst[2 * (i * N + j)], term.x);
          atomicAdd(&dst[2 * (i * N + j) + 1], term.y);
        }
      }
    }
  }
  if (lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < N * N * 2; i += kBlock)
      if (acc[i] != 0.0) atomicAdd(&rho[i], acc[i]);
  }
}

void require_same_group(const edx_catalog* cat, const edx_sector* sec) {
  EDX_REQUIRE(sec->cat == cat, EDX_ERR_INVALID, "sector was built from a different catalog");
}

}  // namespace

extern "C" {

int edx_fock_expand(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec, const double* chi,
                    const double* coeff, double atol, int32_t convention, int64_t capacity, uint64_t* masks,
                    double* amps, int64_t* n_out) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && cat && sec && chi && coeff && n_out, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(convention == 0 || convention == 1, EDX_ERR_INVALID, "convention must be 0 or 1");
    require_same_group(cat, sec);
    edx::DeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    const int64_t C = cat->n_total, D = sec->dim;
    *n_out = 0;
    DevBuf<double2> d_chi(cat->n_g), d_c(std::max<int64_t>(D, 1)), d_amp(std::max<int64_t>(C, 1));
    EDX_HIP(hipMemcpyAsync(d_chi.p, chi, (size_t)cat->n_g * 16, hipMemcpyHostToDevice, st));
    if (D) EDX_HIP(hipMemcpyAsync(d_c.p, coeff, (size_t)D * 16, hipMemcpyHostToDevice, st));
    double norm2 = 0.0;
    if (C > 0 && D > 0) {
      hipLaunchKernelGGL(fock_amp, dim3(edx::grid_for(C, kBlock)), dim3(kBlock), 0, st, cat->view(), C,
                         (const int32_t*)cat->rep_rank.p, (const int32_t*)sec->orbit_to_basis.p,
                         (const uint64_t*)sec->reps.p, (const int32_t*)sec->stab_order.p, (const double2*)d_chi.p,
                         (const double2*)d_c.p, atol, convention, d_amp.p);
      EDX_HIP(hipGetLastError());
      DevBuf<double> partial(kSumBlocks);
      hipLaunchKernelGGL(sumsq, dim3(kSumBlocks), dim3(kBlock), 0, st, C, (const double2*)d_amp.p, partial.p);
      EDX_HIP(hipGetLastError());
      std::vector<double> hp(kSumBlocks);
      EDX_HIP(hipMemcpyAsync(hp.data(), partial.p, kSumBlocks * 8, hipMemcpyDeviceToHost, st));
      EDX_HIP(hipStreamSynchronize(st));
      for (double x : hp) norm2 += x;
    }
    EDX_REQUIRE(norm2 > atol, EDX_ERR_INVALID, "Expanded state has near-zero norm.");
    DevBuf<uint32_t> flags(C), pos(C);
    hipLaunchKernelGGL(nonzero_flags, dim3(edx::grid_for(C, kBlock)), dim3(kBlock), 0, st, C,
                       (const double2*)d_amp.p, flags.p);
    EDX_HIP(hipGetLastError());
    const int64_t nnz = (int64_t)edx::exclusive_scan_u32(*ctx, flags.p, pos.p, (size_t)C);
    *n_out = nnz;
    if (!masks || !amps) return;  // size query
    EDX_REQUIRE(capacity >= nnz, EDX_ERR_INVALID, "output capacity below the number of amplitudes");
    DevBuf<uint64_t> d_m(std::max<int64_t>(nnz, 1));
    DevBuf<double2> d_a(std::max<int64_t>(nnz, 1));
    hipLaunchKernelGGL(fock_compact, dim3(edx::grid_for(C, kBlock)), dim3(kBlock), 0, st, cat->view(), C,
                       (const double2*)d_amp.p, (const uint32_t*)flags.p, (const uint32_t*)pos.p,
                       1.0 / std::sqrt(norm2), d_m.p, d_a.p);
    EDX_HIP(hipGetLastError());
    if (nnz) {
      EDX_HIP(hipMemcpyAsync(masks, d_m.p, (size_t)nnz * 8, hipMemcpyDeviceToHost, st));
      EDX_HIP(hipMemcpyAsync(amps, d_a.p, (size_t)nnz * 16, hipMemcpyDeviceToHost, st));
    }
    EDX_HIP(hipStreamSynchronize(st));
  });
}

int edx_position_operator(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec_to, const edx_sector* sec_from,
                          const double* chi_to, const double* chi_from, const double* site_phases, int64_t* rows,
                          double* vals) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && cat && sec_to && sec_from && chi_to && chi_from && site_phases && rows && vals,
                EDX_ERR_INVALID, "null argument");
    require_same_group(cat, sec_to);
    require_same_group(cat, sec_from);
    edx::DeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    const int64_t D = sec_from->dim;
    if (D == 0) return;
    DevBuf<double2> d_ct(cat->n_g), d_cf(cat->n_g), d_ph(cat->n_site), d_vals(D);
    DevBuf<int64_t> d_rows(D);
    EDX_HIP(hipMemcpyAsync(d_ct.p, chi_to, (size_t)cat->n_g * 16, hipMemcpyHostToDevice, st));
    EDX_HIP(hipMemcpyAsync(d_cf.p, chi_from, (size_t)cat->n_g * 16, hipMemcpyHostToDevice, st));
    EDX_HIP(hipMemcpyAsync(d_ph.p, site_phases, (size_t)cat->n_site * 16, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(position_elems, dim3(edx::grid_for(D, kBlock)), dim3(kBlock), 0, st, cat->view(), D,
                       (const uint64_t*)sec_from->reps.p, (const int32_t*)cat->rep_rank.p,
                       (const int32_t*)sec_to->orbit_to_basis.p, (const double2*)d_ct.p, (const double2*)d_cf.p,
                       (const double2*)d_ph.p, d_rows.p, d_vals.p);
    EDX_HIP(hipGetLastError());
    EDX_HIP(hipMemcpyAsync(rows, d_rows.p, (size_t)D * 8, hipMemcpyDeviceToHost, st));
    EDX_HIP(hipMemcpyAsync(vals, d_vals.p, (size_t)D * 16, hipMemcpyDeviceToHost, st));
    EDX_HIP(hipStreamSynchronize(st));
  });
}

int edx_one_body_density(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec, const double* chi,
                         const double* coeff, double* rho) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && cat && sec && chi && coeff && rho, EDX_ERR_INVALID, "null argument");
    require_same_group(cat, sec);
    edx::DeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    const int N = cat->n_site;
    const int64_t D = sec->dim;
    DevBuf<double> d_rho((size_t)N * N * 2);
    EDX_HIP(hipMemsetAsync(d_rho.p, 0, d_rho.bytes(), st));
    if (D > 0) {
      DevBuf<double2> d_chi(cat->n_g), d_c(D);
      EDX_HIP(hipMemcpyAsync(d_chi.p, chi, (size_t)cat->n_g * 16, hipMemcpyHostToDevice, st));
      EDX_HIP(hipMemcpyAsync(d_c.p, coeff, (size_t)D * 16, hipMemcpyHostToDevice, st));
      const int64_t work = D * cat->n_g;
      hipLaunchKernelGGL(one_body_density, dim3(edx::grid_for(work, kBlock)), dim3(kBlock), 0, st, cat->view(), D,
                         (const uint64_t*)sec->reps.p, (const int32_t*)sec->stab_order.p,
                         (const int32_t*)cat->rep_rank.p, (const int32_t*)cat->rep_sym.p,
                         (const float2*)cat->rep_amp.p, (const int32_t*)sec->orbit_to_basis.p,
                         (const double2*)d_chi.p, (const double2*)d_c.p, d_rho.p);
      EDX_HIP(hipGetLastError());
    }
    EDX_HIP(hipMemcpyAsync(rho, d_rho.p, d_rho.bytes(), hipMemcpyDeviceToHost, st));
    EDX_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
