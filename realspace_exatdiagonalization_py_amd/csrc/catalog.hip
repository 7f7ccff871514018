// Orbit catalog: representative tables (rep, g, alpha) over the colex-ranked
// fixed-filling basis, orbit list and stabilizers.
//
// Reference: _build_orbit_catalog_kernel (SRE:433-575) scans the basis in rank
// order, marks masks that are the minimum of their orbit, then scatters
// (orbit, g, alpha_g(rep)) to rank(g.rep) for g ascending, so the largest g
// with g.rep == mask wins.  Here every mask PULLS its own entry instead:
// with H(m) = {h : h.m = min_h h.m} the winning g is max{inv(h) : h in H(m)},
// so each thread writes only its own rank (coalesced), and the orbit index is
// resolved afterwards from a rep bitmap and its prefix popcounts.
//
//   K1 catalog_scan     one lane per rank: unrank, |G| images, min, g, alpha;
//                       rank(rep) parked in rep_rank; rep flags -> bitmap via ballot
//   K2 scan             prefix popcount of the bitmap words  -> orbit numbering
//   K3 catalog_finalize rank(rep) -> orbit index              (rank/select)
//   K4 catalog_reps     per bitmap word: rep masks, stabilizer bit sets
//   K5 scan             stabilizer offsets
//   K6 catalog_stab     stabilizer element lists and phases
#include <cmath>
#include <cstring>

#include "objects.h"

using edx::DevBuf;
using edx::Error;

namespace {

constexpr int kBlock = 256;

// Copies the small group tables into LDS; returns pointers into shared memory.
struct SharedGroup {
  uint8_t* perm;
  uint32_t* binom;
  int32_t* inv;
};

__device__ SharedGroup load_group(const edx::GroupView& gv, unsigned char* smem) {
  const int np = gv.n_g * gv.n_site;
  const int nb = (gv.n_site + 1) * (gv.n_filled + 1);
  uint32_t* binom = reinterpret_cast<uint32_t*>(smem);
  int32_t* inv = reinterpret_cast<int32_t*>(binom + nb);
  uint8_t* perm = reinterpret_cast<uint8_t*>(inv + gv.n_g);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) binom[i] = gv.binom[i];
  for (int i = threadIdx.x; i < gv.n_g; i += blockDim.x) inv[i] = gv.inv[i];
  for (int i = threadIdx.x; i < np; i += blockDim.x) perm[i] = gv.perm[i];
  __syncthreads();
  return SharedGroup{perm, binom, inv};
}

size_t group_smem_bytes(int n_site, int n_filled, int n_g) {
  return (size_t)(n_site + 1) * (n_filled + 1) * 4 + (size_t)n_g * 4 + (size_t)n_g * n_site + 16;
}

__device__ __forceinline__ double2 alpha_of(const edx::GroupView& gv, uint64_t m, int g,
                                            const uint8_t* perm_s) {
  const uint8_t* pg = perm_s + g * gv.n_site;
  if (gv.product_phase) return edxd::alpha_prod(m, gv.phase + (size_t)g * gv.n_site, pg, gv.fermionic);
  return edxd::alpha_exp(m, gv.n_site, gv.angle + (size_t)g * gv.n_site, pg, gv.fermionic);
}

__global__ void __launch_bounds__(kBlock)
catalog_scan(edx::GroupView gv, int64_t n_total, int32_t* rep_rank, int32_t* rep_sym, float2* rep_amp,
             uint64_t* words) {
  extern __shared__ unsigned char smem[];
  SharedGroup sg = load_group(gv, smem);
  const int nf1 = gv.n_filled + 1;
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = r < n_total;
  bool is_rep = false;
  if (valid) {
    const uint64_t m = edxd::unrank_colex((uint64_t)r, gv.n_site, gv.n_filled, sg.binom, nf1);
    uint64_t mn = ~0ull;
    int gmax = -1;
    for (int g = 0; g < gv.n_g; ++g) {
      const uint64_t img = edxd::image(m, sg.perm + g * gv.n_site);
      const int gi = sg.inv[g];
      if (img < mn) {
        mn = img;
        gmax = gi;
      } else if (img == mn && gi > gmax) {
        gmax = gi;
      }
    }
    is_rep = (mn == m);
    const double2 a = alpha_of(gv, mn, gmax, sg.perm);
    rep_rank[r] = (int32_t)edxd::rank_colex(mn, sg.binom, nf1);
    rep_sym[r] = gmax;
    rep_amp[r] = make_float2((float)a.x, (float)a.y);
  }
  const uint64_t bal = __ballot(is_rep);
  if ((threadIdx.x & 63) == 0) {
    const int64_t w = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (w * 64 < n_total) words[w] = bal;
  }
}

__global__ void __launch_bounds__(kBlock)
catalog_finalize(int64_t n_total, int32_t* rep_rank, const uint64_t* words, const uint32_t* word_prefix) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= n_total) return;
  const uint32_t rr = (uint32_t)rep_rank[r];
  const uint32_t w = rr >> 6;
  const uint64_t below = (rr & 63) ? (words[w] & ((1ull << (rr & 63)) - 1)) : 0ull;
  rep_rank[r] = (int32_t)(word_prefix[w] + (uint32_t)edxd::popc64(below));
}

__global__ void __launch_bounds__(kBlock)
catalog_reps(edx::GroupView gv, int64_t n_words, const uint64_t* words, const uint32_t* word_prefix,
             uint64_t* reps, int32_t* stab_order, uint64_t* stab_bits) {
  extern __shared__ unsigned char smem[];
  SharedGroup sg = load_group(gv, smem);
  const int nf1 = gv.n_filled + 1;
  const int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (w >= n_words) return;
  uint64_t bits = words[w];
  uint32_t orbit = word_prefix[w];
  while (bits) {
    const int j = edxd::ctz64(bits);
    const uint64_t m = edxd::unrank_colex((uint64_t)(w * 64 + j), gv.n_site, gv.n_filled, sg.binom, nf1);
    uint64_t sb = 0;
    for (int g = 0; g < gv.n_g; ++g)
      if (edxd::image(m, sg.perm + g * gv.n_site) == m) sb |= 1ull << g;
    reps[orbit] = m;
    stab_order[orbit] = edxd::popc64(sb);
    stab_bits[orbit] = sb;
    ++orbit;
    bits &= bits - 1;
  }
}

__global__ void __launch_bounds__(kBlock)
catalog_stab(edx::GroupView gv, int64_t n_orbits, const uint64_t* reps, const uint64_t* stab_bits,
             const uint32_t* stab_off, uint8_t* stab_g, double2* stab_phase) {
  extern __shared__ unsigned char smem[];
  SharedGroup sg = load_group(gv, smem);
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n_orbits) return;
  const uint64_t m = reps[o];
  uint64_t sb = stab_bits[o];
  uint32_t j = stab_off[o];
  while (sb) {
    const int g = edxd::ctz64(sb);
    stab_g[j] = (uint8_t)g;
    stab_phase[j] = alpha_of(gv, m, g, sg.perm);
    ++j;
    sb &= sb - 1;
  }
}

// Phase refresh for the stored (orbit, g) of every rank and for every
// stabilizer element.  fresh = 0: product-form alpha, as
// update_orbit_stabilizer_phases computes it (SRE:1220-1269); fresh = 1: the
// formulas of a fresh catalog build (alpha_of), so that the tables equal
// those of edx_catalog_build at the new phases bit for bit.
__device__ __forceinline__ double2 refresh_alpha(const edx::GroupView& gv, uint64_t m, int g, const uint8_t* perm_s,
                                                 int fresh) {
  if (fresh) return alpha_of(gv, m, g, perm_s);
  return edxd::alpha_prod(m, gv.phase + (size_t)g * gv.n_site, perm_s + g * gv.n_site, gv.fermionic);
}

__global__ void __launch_bounds__(kBlock)
catalog_update_table(edx::GroupView gv, int64_t n_total, const int32_t* rep_rank, const int32_t* rep_sym,
                     const uint64_t* reps, float2* rep_amp, int fresh) {
  extern __shared__ unsigned char smem[];
  SharedGroup sg = load_group(gv, smem);
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= n_total) return;
  const int g = rep_sym[r];
  const uint64_t m = reps[rep_rank[r]];
  const double2 a = refresh_alpha(gv, m, g, sg.perm, fresh);
  rep_amp[r] = make_float2((float)a.x, (float)a.y);
}

__global__ void __launch_bounds__(kBlock)
catalog_update_stab(edx::GroupView gv, int64_t n_orbits, const uint64_t* reps, const uint32_t* stab_off,
                    const uint8_t* stab_g, double2* stab_phase, int fresh) {
  extern __shared__ unsigned char smem[];
  SharedGroup sg = load_group(gv, smem);
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n_orbits) return;
  const uint64_t m = reps[o];
  for (uint32_t j = stab_off[o]; j < stab_off[o + 1]; ++j) stab_phase[j] = refresh_alpha(gv, m, stab_g[j], sg.perm, fresh);
}

__global__ void __launch_bounds__(kBlock)
catalog_lookup(edx::GroupView gv, int64_t n, const uint64_t* masks, const int32_t* rep_rank, const int32_t* rep_sym,
               const float2* rep_amp, const uint64_t* reps, uint64_t* rep_out, int32_t* g_out, double2* a_out) {
  extern __shared__ unsigned char smem[];
  SharedGroup sg = load_group(gv, smem);
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t idx = edxd::rank_colex(masks[i], sg.binom, gv.n_filled + 1);
  rep_out[i] = reps[rep_rank[idx]];
  g_out[i] = rep_sym[idx];
  const float2 a = rep_amp[idx];
  a_out[i] = make_double2((double)a.x, (double)a.y);
}

uint64_t binom_u64(int n, int k) {
  if (k < 0 || k > n) return 0;
  uint64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
  return r;
}

// Validates the group and uploads perm / phases / angles / inverse / binomials.
// Every site phase +1 or -1 (no twist): alpha is a sign, exact in complex64.
bool exact_site_phases(const double* phases, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const double re = phases[2 * i], im = phases[2 * i + 1];
    if (im != 0.0 || (re != 1.0 && re != -1.0)) return false;
  }
  return true;
}

void upload_group(edx_catalog& cat, edx_ctx& ctx, const int32_t* perm, const double* phases,
                  const double* angles) {
  const int N = cat.n_site, G = cat.n_g;
  std::vector<uint8_t> p0((size_t)G * N);
  cat.h_perm.assign(perm, perm + (size_t)G * N);
  cat.exact_alpha = exact_site_phases(phases, (size_t)G * N);
  for (int g = 0; g < G; ++g) {
    std::vector<int> seen(N, 0);
    for (int i = 0; i < N; ++i) {
      const int t = perm[(size_t)g * N + i];
      EDX_REQUIRE(t >= 1 && t <= N, EDX_ERR_INVALID, "permutation entries must lie in 1..n_site");
      EDX_REQUIRE(!seen[t - 1], EDX_ERR_INVALID, "operation " + std::to_string(g) + " is not a permutation");
      seen[t - 1] = 1;
      p0[(size_t)g * N + i] = (uint8_t)(t - 1);
    }
  }
  std::vector<int32_t> inv(G, -1);
  for (int g = 0; g < G; ++g) {
    for (int h = 0; h < G && inv[g] < 0; ++h) {
      bool ok = true;
      for (int i = 0; i < N && ok; ++i) ok = p0[(size_t)h * N + p0[(size_t)g * N + i]] == i;
      if (ok) inv[g] = h;
    }
    EDX_REQUIRE(inv[g] >= 0, EDX_ERR_INVALID,
                "symmetry operations must form a group: operation " + std::to_string(g) +
                    " has no inverse in the list");
  }
  const int nf1 = cat.n_filled + 1;
  std::vector<uint32_t> bt((size_t)(N + 1) * nf1, 0);
  for (int b = 0; b <= N; ++b)
    for (int j = 0; j <= cat.n_filled && j <= b; ++j) {
      const uint64_t v = binom_u64(b, j);
      bt[(size_t)b * nf1 + j] = v > 0xffffffffull ? 0xffffffffu : (uint32_t)v;
    }
  cat.perm.alloc(p0.size());
  cat.angle.alloc((size_t)G * N);
  cat.phase.alloc((size_t)G * N);
  cat.inv.alloc(G);
  cat.binom.alloc(bt.size());
  EDX_HIP(hipMemcpyAsync(cat.perm.p, p0.data(), p0.size(), hipMemcpyHostToDevice, ctx.stream));
  EDX_HIP(hipMemcpyAsync(cat.angle.p, angles, (size_t)G * N * 8, hipMemcpyHostToDevice, ctx.stream));
  EDX_HIP(hipMemcpyAsync(cat.phase.p, phases, (size_t)G * N * 16, hipMemcpyHostToDevice, ctx.stream));
  EDX_HIP(hipMemcpyAsync(cat.inv.p, inv.data(), (size_t)G * 4, hipMemcpyHostToDevice, ctx.stream));
  EDX_HIP(hipMemcpyAsync(cat.binom.p, bt.data(), bt.size() * 4, hipMemcpyHostToDevice, ctx.stream));
  EDX_HIP(hipStreamSynchronize(ctx.stream));
}

void build_catalog(edx_ctx& ctx, edx_catalog& cat) {
  const int64_t C = cat.n_total;
  const edx::GroupView gv = cat.view();
  const size_t smem = group_smem_bytes(cat.n_site, cat.n_filled, cat.n_g);
  const int64_t n_words = (C + 63) / 64;

  cat.rep_rank.alloc(C);
  cat.rep_sym.alloc(C);
  cat.rep_amp.alloc(C);
  DevBuf<uint64_t> words(n_words);
  DevBuf<uint32_t> word_prefix(n_words);

  hipLaunchKernelGGL(catalog_scan, dim3(edx::grid_for(C, kBlock)), dim3(kBlock), smem, ctx.stream, gv, C,
                     cat.rep_rank.p, cat.rep_sym.p, cat.rep_amp.p, words.p);
  EDX_HIP(hipGetLastError());
  cat.n_orbits = (int64_t)edx::exclusive_scan_popcount64(ctx, words.p, word_prefix.p, n_words);
  hipLaunchKernelGGL(catalog_finalize, dim3(edx::grid_for(C, kBlock)), dim3(kBlock), 0, ctx.stream, C,
                     cat.rep_rank.p, (const uint64_t*)words.p, (const uint32_t*)word_prefix.p);
  EDX_HIP(hipGetLastError());

  const int64_t n_orb = cat.n_orbits;
  cat.reps.alloc(n_orb);
  cat.stab_order.alloc(n_orb);
  cat.stab_off.alloc(n_orb + 1);
  DevBuf<uint64_t> stab_bits(n_orb);
  hipLaunchKernelGGL(catalog_reps, dim3(edx::grid_for(n_words, kBlock)), dim3(kBlock), smem, ctx.stream, gv,
                     n_words, (const uint64_t*)words.p, (const uint32_t*)word_prefix.p, cat.reps.p,
                     cat.stab_order.p, stab_bits.p);
  EDX_HIP(hipGetLastError());
  const uint64_t n_stab = edx::exclusive_scan_u32(ctx, reinterpret_cast<const uint32_t*>(cat.stab_order.p),
                                                  cat.stab_off.p, (size_t)n_orb);
  const uint32_t tail = (uint32_t)n_stab;
  EDX_HIP(hipMemcpyAsync(cat.stab_off.p + n_orb, &tail, 4, hipMemcpyHostToDevice, ctx.stream));
  cat.n_stab = (int64_t)n_stab;
  cat.stab_g.alloc(n_stab);
  cat.stab_phase.alloc(n_stab);
  hipLaunchKernelGGL(catalog_stab, dim3(edx::grid_for(n_orb, kBlock)), dim3(kBlock), smem, ctx.stream, gv,
                     n_orb, (const uint64_t*)cat.reps.p, (const uint64_t*)stab_bits.p,
                     (const uint32_t*)cat.stab_off.p, cat.stab_g.p, cat.stab_phase.p);
  EDX_HIP(hipGetLastError());
  EDX_HIP(hipStreamSynchronize(ctx.stream));
}

}  // namespace

extern "C" {

int edx_catalog_build(edx_ctx* ctx, int32_t n_site, int32_t n_filled, int32_t n_g, const int32_t* perm,
                      const double* perm_phases, const double* perm_angles, int32_t fermionic,
                      edx_catalog** out) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && out && perm && perm_phases && perm_angles, EDX_ERR_INVALID, "null argument");
    *out = nullptr;
    EDX_REQUIRE(n_site >= 1, EDX_ERR_INVALID, "n_site must be >= 1");
    EDX_REQUIRE(n_site <= EDX_MAX_SITES, EDX_ERR_INVALID,
                "n_site = " + std::to_string(n_site) + " exceeds the 64-bit kernel limit (63); use a smaller system.");
    EDX_REQUIRE(n_filled >= 0 && n_filled <= n_site, EDX_ERR_ASSERT, "0 <= n_filled <= n_site violated");
    EDX_REQUIRE(n_g >= 1 && n_g <= EDX_MAX_GROUP, EDX_ERR_INVALID, "group order must be in 1..64");
    const uint64_t C = binom_u64(n_site, n_filled);
    EDX_REQUIRE(C > 0 && C < (1ull << 31), EDX_ERR_INVALID,
                "C(n_site, n_filled) must be below 2^31 (int32 representative tables)");
    edx::DeviceGuard dg(ctx->device);
    auto cat = new edx_catalog();
    try {
      cat->device = ctx->device;
      cat->n_site = n_site;
      cat->n_filled = n_filled;
      cat->n_g = n_g;
      cat->fermionic = fermionic ? 1 : 0;
      cat->product_phase = (n_filled == 0 || n_filled == n_site) ? 1 : 0;
      cat->n_total = (int64_t)C;
      upload_group(*cat, *ctx, perm, perm_phases, perm_angles);
      build_catalog(*ctx, *cat);
    } catch (...) {
      delete cat;
      throw;
    }
    *out = cat;
  });
}

int edx_catalog_upload(edx_ctx* ctx, int32_t n_site, int32_t n_filled, int32_t n_g, const int32_t* perm,
                       const double* perm_phases, const double* perm_angles, int32_t fermionic, int64_t n_orbits,
                       int64_t n_stab, const int32_t* rep_rank, const int32_t* rep_sym, const float* rep_amp,
                       const uint64_t* reps, const int32_t* stab_order, const int64_t* stab_offsets,
                       const int32_t* stab_g, const double* stab_phase, edx_catalog** out) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && out && perm && perm_phases && perm_angles && rep_rank && rep_sym && rep_amp && reps &&
                    stab_order && stab_offsets && (n_stab == 0 || (stab_g && stab_phase)),
                EDX_ERR_INVALID, "null argument");
    *out = nullptr;
    EDX_REQUIRE(n_site >= 1 && n_site <= EDX_MAX_SITES, EDX_ERR_INVALID, "n_site must be in 1..63");
    EDX_REQUIRE(n_filled >= 0 && n_filled <= n_site, EDX_ERR_ASSERT, "0 <= n_filled <= n_site violated");
    EDX_REQUIRE(n_g >= 1 && n_g <= EDX_MAX_GROUP, EDX_ERR_INVALID, "group order must be in 1..64");
    const uint64_t C = binom_u64(n_site, n_filled);
    EDX_REQUIRE(C > 0 && C < (1ull << 31), EDX_ERR_INVALID,
                "C(n_site, n_filled) must be below 2^31 (int32 representative tables)");
    EDX_REQUIRE(n_orbits >= 1 && (uint64_t)n_orbits <= C && n_stab >= n_orbits, EDX_ERR_INVALID,
                "catalog sizes out of range");
    // the tables must describe a catalog of these sizes: offsets, group
    // indices, orbit indices and representatives are checked before use
    EDX_REQUIRE(stab_offsets[0] == 0 && stab_offsets[n_orbits] == n_stab, EDX_ERR_INVALID,
                "stabilizer offsets must start at 0 and end at n_stab");
    const uint64_t lim = n_site >= 64 ? ~0ull : (1ull << n_site);
    for (int64_t o = 0; o < n_orbits; ++o) {
      EDX_REQUIRE(stab_offsets[o + 1] - stab_offsets[o] == stab_order[o] && stab_order[o] >= 1, EDX_ERR_INVALID,
                  "stabilizer offsets disagree with the stabilizer orders");
      EDX_REQUIRE(reps[o] < lim && __builtin_popcountll(reps[o]) == n_filled && (o == 0 || reps[o - 1] < reps[o]),
                  EDX_ERR_INVALID, "representatives must be ascending masks of the catalog's filling");
    }
    for (int64_t e = 0; e < n_stab; ++e)
      EDX_REQUIRE(stab_g[e] >= 0 && stab_g[e] < n_g, EDX_ERR_INVALID, "stabilizer element out of range");
    for (uint64_t r = 0; r < C; ++r)
      EDX_REQUIRE(rep_rank[r] >= 0 && rep_rank[r] < n_orbits && rep_sym[r] >= 0 && rep_sym[r] < n_g,
                  EDX_ERR_INVALID, "representative table entry out of range");
    edx::DeviceGuard dg(ctx->device);
    auto cat = new edx_catalog();
    try {
      cat->device = ctx->device;
      cat->n_site = n_site;
      cat->n_filled = n_filled;
      cat->n_g = n_g;
      cat->fermionic = fermionic ? 1 : 0;
      cat->product_phase = (n_filled == 0 || n_filled == n_site) ? 1 : 0;
      cat->n_total = (int64_t)C;
      cat->n_orbits = n_orbits;
      cat->n_stab = n_stab;
      upload_group(*cat, *ctx, perm, perm_phases, perm_angles);
      std::vector<uint32_t> off((size_t)n_orbits + 1);
      for (int64_t o = 0; o <= n_orbits; ++o) off[o] = (uint32_t)stab_offsets[o];
      std::vector<uint8_t> g8((size_t)n_stab);
      for (int64_t e = 0; e < n_stab; ++e) g8[e] = (uint8_t)stab_g[e];
      cat->rep_rank.alloc(C);
      cat->rep_sym.alloc(C);
      cat->rep_amp.alloc(C);
      cat->reps.alloc(n_orbits);
      cat->stab_order.alloc(n_orbits);
      cat->stab_off.alloc(n_orbits + 1);
      cat->stab_g.alloc(n_stab);
      cat->stab_phase.alloc(n_stab);
      hipStream_t st = ctx->stream;
      EDX_HIP(hipMemcpyAsync(cat->rep_rank.p, rep_rank, C * 4, hipMemcpyHostToDevice, st));
      EDX_HIP(hipMemcpyAsync(cat->rep_sym.p, rep_sym, C * 4, hipMemcpyHostToDevice, st));
      EDX_HIP(hipMemcpyAsync(cat->rep_amp.p, rep_amp, C * 8, hipMemcpyHostToDevice, st));
      EDX_HIP(hipMemcpyAsync(cat->reps.p, reps, n_orbits * 8, hipMemcpyHostToDevice, st));
      EDX_HIP(hipMemcpyAsync(cat->stab_order.p, stab_order, n_orbits * 4, hipMemcpyHostToDevice, st));
      EDX_HIP(hipMemcpyAsync(cat->stab_off.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, st));
      if (n_stab) {
        EDX_HIP(hipMemcpyAsync(cat->stab_g.p, g8.data(), g8.size(), hipMemcpyHostToDevice, st));
        EDX_HIP(hipMemcpyAsync(cat->stab_phase.p, stab_phase, n_stab * 16, hipMemcpyHostToDevice, st));
      }
      EDX_HIP(hipStreamSynchronize(st));
    } catch (...) {
      delete cat;
      throw;
    }
    *out = cat;
  });
}

int edx_catalog_clone(edx_ctx* ctx, const edx_catalog* src, edx_catalog** out) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && src && out, EDX_ERR_INVALID, "null argument");
    *out = nullptr;
    edx::DeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    auto cat = new edx_catalog();
    try {
      cat->device = ctx->device;
      cat->n_site = src->n_site;
      cat->n_filled = src->n_filled;
      cat->n_g = src->n_g;
      cat->fermionic = src->fermionic;
      cat->product_phase = src->product_phase;
      cat->exact_alpha = src->exact_alpha;
      cat->n_total = src->n_total;
      cat->n_orbits = src->n_orbits;
      cat->n_stab = src->n_stab;
      cat->h_perm = src->h_perm;
      auto copy = [&](auto& dst, const auto& from) {
        dst.alloc(from.n);
        if (from.n) EDX_HIP(hipMemcpyAsync(dst.p, from.p, from.bytes(), hipMemcpyDeviceToDevice, st));
      };
      copy(cat->perm, src->perm);
      copy(cat->angle, src->angle);
      copy(cat->phase, src->phase);
      copy(cat->inv, src->inv);
      copy(cat->binom, src->binom);
      copy(cat->rep_rank, src->rep_rank);
      copy(cat->rep_sym, src->rep_sym);
      copy(cat->rep_amp, src->rep_amp);
      copy(cat->reps, src->reps);
      copy(cat->stab_order, src->stab_order);
      copy(cat->stab_off, src->stab_off);
      copy(cat->stab_g, src->stab_g);
      copy(cat->stab_phase, src->stab_phase);
      EDX_HIP(hipStreamSynchronize(st));
    } catch (...) {
      delete cat;
      throw;
    }
    *out = cat;
  });
}

int edx_catalog_info(const edx_catalog* cat, int64_t* n_total, int64_t* n_orbits, int64_t* n_stab) {
  if (!cat) return EDX_ERR_INVALID;
  if (n_total) *n_total = cat->n_total;
  if (n_orbits) *n_orbits = cat->n_orbits;
  if (n_stab) *n_stab = cat->n_stab;
  return EDX_OK;
}

int edx_catalog_fetch(edx_ctx* ctx, const edx_catalog* cat, int32_t* rep_rank, int32_t* rep_sym, float* rep_amp,
                      uint64_t* reps, int32_t* stab_order, int64_t* stab_offsets, int32_t* stab_g,
                      double* stab_phase) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && cat, EDX_ERR_INVALID, "null argument");
    edx::DeviceGuard dg(ctx->device);
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
      if (dst && bytes) EDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    };
    d2h(rep_rank, cat->rep_rank.p, cat->rep_rank.bytes());
    d2h(rep_sym, cat->rep_sym.p, cat->rep_sym.bytes());
    d2h(rep_amp, cat->rep_amp.p, cat->rep_amp.bytes());
    d2h(reps, cat->reps.p, cat->reps.bytes());
    d2h(stab_order, cat->stab_order.p, cat->stab_order.bytes());
    d2h(stab_phase, cat->stab_phase.p, cat->stab_phase.bytes());
    std::vector<uint32_t> off;
    std::vector<uint8_t> g8;
    if (stab_offsets) {
      off.resize(cat->n_orbits + 1);
      d2h(off.data(), cat->stab_off.p, off.size() * 4);
    }
    if (stab_g) {
      g8.resize(cat->n_stab);
      d2h(g8.data(), cat->stab_g.p, g8.size());
    }
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    if (stab_offsets)
      for (size_t i = 0; i < off.size(); ++i) stab_offsets[i] = off[i];
    if (stab_g)
      for (size_t i = 0; i < g8.size(); ++i) stab_g[i] = g8[i];
  });
}

}  // extern "C"

namespace {
__global__ void rank_records(int64_t n, const int32_t* rep_rank, const int32_t* rep_sym, const float2* rep_amp,
                             edx::RankRec* out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float2 a = rep_amp[i];
  out[i] = edx::RankRec{rep_rank[i], rep_sym[i], a.x, a.y};
}

void fill_rank_records(edx_catalog& cat, hipStream_t st) {
  if (cat.n_total == 0) return;
  hipLaunchKernelGGL(rank_records, dim3(edx::grid_for(cat.n_total, kBlock)), dim3(kBlock), 0, st, cat.n_total,
                     (const int32_t*)cat.rep_rank.p, (const int32_t*)cat.rep_sym.p, (const float2*)cat.rep_amp.p,
                     cat.rank_rec.p);
  EDX_HIP(hipGetLastError());
}
}  // namespace

namespace edx {
void ensure_rank_records(edx_catalog& cat, Ctx& ctx) {
  std::lock_guard<std::mutex> lk(cat.rank_rec_mu);
  if (cat.rank_rec.p || cat.n_total == 0) return;
  cat.rank_rec.alloc((size_t)cat.n_total);
  fill_rank_records(cat, ctx.stream);
  EDX_HIP(hipStreamSynchronize(ctx.stream));
}
}  // namespace edx

namespace {
void refresh_phases(edx_ctx* ctx, edx_catalog* cat, const int32_t* perm, const double* perm_phases,
                    const double* perm_angles, int fresh) {
    EDX_REQUIRE(ctx && cat && perm && perm_phases && perm_angles, EDX_ERR_INVALID, "null argument");
    const size_t n = (size_t)cat->n_g * cat->n_site;
    EDX_REQUIRE(std::memcmp(perm, cat->h_perm.data(), n * 4) == 0, EDX_ERR_INVALID,
                "update_orbit_stabilizer_phases: the new group must have the same permutations");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipMemcpyAsync(cat->phase.p, perm_phases, n * 16, hipMemcpyHostToDevice, ctx->stream));
    EDX_HIP(hipMemcpyAsync(cat->angle.p, perm_angles, n * 8, hipMemcpyHostToDevice, ctx->stream));
    cat->exact_alpha = exact_site_phases(perm_phases, n);
    const edx::GroupView gv = cat->view();
    const size_t smem = group_smem_bytes(cat->n_site, cat->n_filled, cat->n_g);
    hipLaunchKernelGGL(catalog_update_stab, dim3(edx::grid_for(cat->n_orbits, kBlock)), dim3(kBlock), smem,
                       ctx->stream, gv, cat->n_orbits, (const uint64_t*)cat->reps.p,
                       (const uint32_t*)cat->stab_off.p, (const uint8_t*)cat->stab_g.p, cat->stab_phase.p, fresh);
    EDX_HIP(hipGetLastError());
    hipLaunchKernelGGL(catalog_update_table, dim3(edx::grid_for(cat->n_total, kBlock)), dim3(kBlock), smem,
                       ctx->stream, gv, cat->n_total, (const int32_t*)cat->rep_rank.p,
                       (const int32_t*)cat->rep_sym.p, (const uint64_t*)cat->reps.p, cat->rep_amp.p, fresh);
    EDX_HIP(hipGetLastError());
    {
      std::lock_guard<std::mutex> lk(cat->rank_rec_mu);
      if (cat->rank_rec.p) fill_rank_records(*cat, ctx->stream);
    }
    EDX_HIP(hipStreamSynchronize(ctx->stream));
}
}  // namespace

extern "C" {

int edx_catalog_update_phases(edx_ctx* ctx, edx_catalog* cat, const int32_t* perm, const double* perm_phases,
                              const double* perm_angles) {
  return edx::guarded(ctx, [&] { refresh_phases(ctx, cat, perm, perm_phases, perm_angles, 0); });
}

int edx_catalog_rephase(edx_ctx* ctx, edx_catalog* cat, const int32_t* perm, const double* perm_phases,
                        const double* perm_angles) {
  return edx::guarded(ctx, [&] { refresh_phases(ctx, cat, perm, perm_phases, perm_angles, 1); });
}

int edx_catalog_lookup(edx_ctx* ctx, const edx_catalog* cat, int64_t n, const uint64_t* masks, uint64_t* rep,
                       int32_t* g, double* alpha) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && cat && masks && rep && g && alpha, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(n >= 0, EDX_ERR_INVALID, "negative count");
    const uint64_t lim = cat->n_site >= 64 ? ~0ull : (1ull << cat->n_site);
    for (int64_t i = 0; i < n; ++i)
      EDX_REQUIRE(masks[i] < lim && __builtin_popcountll(masks[i]) == cat->n_filled, EDX_ERR_INVALID,
                  "mask outside the fixed-filling basis of the catalog");
    if (n == 0) return;
    edx::DeviceGuard dg(ctx->device);
    DevBuf<uint64_t> dm(n), dr(n);
    DevBuf<int32_t> dg_(n);
    DevBuf<double2> da(n);
    EDX_HIP(hipMemcpyAsync(dm.p, masks, n * 8, hipMemcpyHostToDevice, ctx->stream));
    const size_t smem = group_smem_bytes(cat->n_site, cat->n_filled, cat->n_g);
    hipLaunchKernelGGL(catalog_lookup, dim3(edx::grid_for(n, kBlock)), dim3(kBlock), smem, ctx->stream, cat->view(), n,
                       (const uint64_t*)dm.p, (const int32_t*)cat->rep_rank.p, (const int32_t*)cat->rep_sym.p,
                       (const float2*)cat->rep_amp.p, (const uint64_t*)cat->reps.p, dr.p, dg_.p, da.p);
    EDX_HIP(hipGetLastError());
    EDX_HIP(hipMemcpyAsync(rep, dr.p, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipMemcpyAsync(g, dg_.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipMemcpyAsync(alpha, da.p, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_catalog_destroy(edx_catalog* cat) {
  if (!cat) return EDX_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(cat->device);
  delete cat;
  (void)hipSetDevice(prev);
  return EDX_OK;
}

}  // extern "C"
