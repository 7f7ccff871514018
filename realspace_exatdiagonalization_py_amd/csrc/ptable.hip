// Sector Hamiltonian: projection table as a row-grouped sparse matrix, and
// the complex SpMV/SpMM that applies it.
//
// Reference: _vectorized_block_triplets (SRE:595-693) generates, per column
// representative m and bilinear term (i -> j) with i occupied and j empty,
// the entry at row = sector index of the orbit of m ^ (b_i | b_j):
//     t * s_JW(m) * alpha * conj(chi(g)) * sqrt(stab_row / stab_col)
// with (orbit, g, alpha) read from the representative tables at
// rank(m ^ hop).  MatrixFreeHamiltonian turns the triplets into a CSR
// (SRE:1958-1963) and every H x is one CSR product.
//
// Device build:
//   P1 count   lane per column: candidate hops, row lookup, atomic row counts
//   P2 scan    row counts -> row_ptr (int64)
//   P3 fill    same walk; key (col << 16 | term) at row_ptr[row] + atomic slot
//   P4 sort    per row: keys ascending, so the row order is (col, term) and the
//              result never depends on atomic ordering (LDS sized by the
//              longest row)
//   P5 values  4 lanes per row: element recomputed from (col, term) in the
//              reference's multiplication order, written to the CSR arrays
//              and (untwisted tables) straight into the sliced copy
//   P6 diag    per row: density terms in term order
// Hops are enumerated per occupied site over terms grouped by source site,
// so every lane walks the same number of candidates.
#include <algorithm>
#include <type_traits>
#include <cmath>

#include "objects.h"

using edx::DevBuf;

namespace {

constexpr int kBlock = 256;

struct TermView {
  int n_site;
  int n_filled;
  int n_terms;
  int fermionic;
  const int32_t* toff;      // [n_site+1] terms grouped by source bit
  const uint16_t* tk;       // [n_terms] original index, grouped order
  const uint8_t* tto;       // [n_terms] target bit, grouped order
  const uint8_t* from_bit;  // [n_terms] by original index
  const uint8_t* to_bit;    // [n_terms] by original index
  const double2* amp;       // [n_terms] by original index
  const uint64_t* between;  // [n_terms] JW string mask by original index
  const uint32_t* binom;    // [(n_site+1)][(n_filled+1)]
};

struct SectorView {
  int64_t dim;
  const uint64_t* reps;
  const int32_t* stab;
  const int32_t* o2b;
  const int32_t* rep_rank;
  const int32_t* rep_sym;
  const float2* rep_amp;
};

struct SharedTerms {
  int32_t* toff;
  uint16_t* tk;
  uint8_t* tto;
  uint32_t* binom;
};

__device__ SharedTerms load_terms(const TermView& tv, unsigned char* smem) {
  const int nb = (tv.n_site + 1) * (tv.n_filled + 1);
  uint32_t* binom = reinterpret_cast<uint32_t*>(smem);
  int32_t* toff = reinterpret_cast<int32_t*>(binom + nb);
  uint16_t* tk = reinterpret_cast<uint16_t*>(toff + tv.n_site + 1);
  uint8_t* tto = reinterpret_cast<uint8_t*>(tk + tv.n_terms);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) binom[i] = tv.binom[i];
  for (int i = threadIdx.x; i <= tv.n_site; i += blockDim.x) toff[i] = tv.toff[i];
  for (int i = threadIdx.x; i < tv.n_terms; i += blockDim.x) {
    tk[i] = tv.tk[i];
    tto[i] = tv.tto[i];
  }
  __syncthreads();
  return SharedTerms{toff, tk, tto, binom};
}

size_t terms_smem_bytes(int n_site, int n_filled, int n_terms) {
  return (size_t)(n_site + 1) * (n_filled + 1) * 4 + (size_t)(n_site + 1) * 4 + (size_t)n_terms * 3 + 16;
}

// Lanes that share one row of the value pass: each takes every kItemLanes-th
// entry, so more catalog lookups are in flight (the count and fill passes
// keep one lane per column, which measured faster).
constexpr int kItemLanes = 4;

// Visits every valid hop of column rep m: f(row, term).
template <class F>
__device__ __forceinline__ void for_each_hop(const TermView& tv, const SectorView& sv, const SharedTerms& st,
                                             uint64_t m, F&& f) {
  const int nf1 = tv.n_filled + 1;
  uint64_t occ = m;
  while (occ) {
    const int s = edxd::ctz64(occ);
    occ &= occ - 1;
    const int q1 = st.toff[s + 1];
    for (int q = st.toff[s]; q < q1; ++q) {
      const int to = st.tto[q];
      if ((m >> to) & 1ull) continue;
      const uint64_t nm = m ^ (1ull << s) ^ (1ull << to);
      const uint64_t idx = edxd::rank_colex(nm, st.binom, nf1);
      const int32_t row = sv.o2b[sv.rep_rank[idx]];
      if (row >= 0) f(row, (int)st.tk[q]);
    }
  }
}

__global__ void __launch_bounds__(kBlock) ptab_count(TermView tv, SectorView sv, uint32_t* cnt) {
  extern __shared__ unsigned char smem[];
  SharedTerms st = load_terms(tv, smem);
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= sv.dim) return;
  for_each_hop(tv, sv, st, sv.reps[c], [&](int32_t row, int) { atomicAdd(&cnt[row], 1u); });
}

__global__ void __launch_bounds__(kBlock)
ptab_fill(TermView tv, SectorView sv, const int64_t* row_ptr, uint32_t* cursor, uint64_t* keys) {
  extern __shared__ unsigned char smem[];
  SharedTerms st = load_terms(tv, smem);
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= sv.dim) return;
  for_each_hop(tv, sv, st, sv.reps[c], [&](int32_t row, int k) {
    const uint32_t slot = atomicAdd(&cursor[row], 1u);
    keys[row_ptr[row] + slot] = ((uint64_t)c << 16) | (uint64_t)k;
  });
}

// Longest row (entry count) of the table: one atomic max per workgroup.
__global__ void __launch_bounds__(kBlock) row_len_max(int64_t dim, const uint32_t* cnt, uint32_t* out) {
  __shared__ uint32_t wmax[kBlock / 64];
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  uint32_t m = r < dim ? cnt[r] : 0u;
  for (int sh = 32; sh >= 1; sh >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, sh, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) m = max(m, wmax[w]);
    atomicMax(out, m);
  }
}

__device__ __forceinline__ void insertion_sort_global(uint64_t* a, int64_t n) {
  for (int64_t i = 1; i < n; ++i) {
    const uint64_t v = a[i];
    int64_t j = i - 1;
    while (j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = v;
  }
}

// Per row: keys ascending.  NT threads per block, each sorting its row in
// CAP slots of LDS (rows longer than CAP are sorted in place in global
// memory); CAP and NT are chosen from the longest row so that blocks stay
// small in LDS and many of them run per CU.
template <int CAP, int NT>
__global__ void __launch_bounds__(NT) ptab_sort(int64_t dim, const int64_t* row_ptr, uint64_t* keys) {
  __shared__ uint64_t buf[CAP * NT];
  const int64_t r = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (r >= dim) return;
  const int64_t b = row_ptr[r];
  const int64_t n = row_ptr[r + 1] - b;
  if (n <= 1) return;
  if (n > CAP) {
    insertion_sort_global(keys + b, n);
    return;
  }
  uint64_t* s = buf + threadIdx.x;  // element i at s[i * NT]
  for (int i = 0; i < n; ++i) s[i * NT] = keys[b + i];
  for (int i = 1; i < n; ++i) {
    const uint64_t v = s[i * NT];
    int j = i - 1;
    while (j >= 0 && s[j * NT] > v) {
      s[(j + 1) * NT] = s[j * NT];
      --j;
    }
    s[(j + 1) * NT] = v;
  }
  for (int i = 0; i < n; ++i) keys[b + i] = s[i * NT];
}

// Rows of at most kRankCap entries: one wavefront per row (four per
// workgroup).  The row is staged in LDS and every lane places its keys at
// their rank, the number of smaller keys plus the equal ones before it, read
// as LDS broadcasts: the stable order of the insertion sort above, without
// its one-thread-per-row serial shifts and its LDS-bound occupancy.
constexpr int kRankCap = 128;
__global__ void __launch_bounds__(256) ptab_sort_rank(int64_t dim, const int64_t* row_ptr, uint64_t* keys) {
  __shared__ uint64_t buf[4][kRankCap];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + w;
  int64_t b = 0, n = 0;
  if (r < dim) {
    b = row_ptr[r];
    n = row_ptr[r + 1] - b;
  }
  for (int i = lane; i < n; i += 64) buf[w][i] = keys[b + i];
  __syncthreads();
  for (int i = lane; i < n; i += 64) {
    const uint64_t v = buf[w][i];
    int pos = 0;
    for (int j = 0; j < n; ++j) {
      const uint64_t u = buf[w][j];
      pos += (u < v) || (u == v && j < i);
    }
    keys[b + pos] = v;
  }
}

// Rank sort up to rows of kRankCap entries; above, 32 KB of LDS per block up
// to rows of 64 entries, 64 KB above (only the longest rows can reach it).
void launch_sort(hipStream_t st, int64_t dim, uint32_t max_row, const int64_t* row_ptr, uint64_t* keys) {
  if (max_row <= (uint32_t)kRankCap)
    hipLaunchKernelGGL(ptab_sort_rank, dim3(edx::grid_for(dim, 4)), dim3(256), 0, st, dim, row_ptr, keys);
  else if (max_row <= 32)
    hipLaunchKernelGGL((ptab_sort<32, 128>), dim3(edx::grid_for(dim, 128)), dim3(128), 0, st, dim, row_ptr, keys);
  else if (max_row <= 64)
    hipLaunchKernelGGL((ptab_sort<64, 64>), dim3(edx::grid_for(dim, 64)), dim3(64), 0, st, dim, row_ptr, keys);
  else
    hipLaunchKernelGGL((ptab_sort<128, 64>), dim3(edx::grid_for(dim, 64)), dim3(64), 0, st, dim, row_ptr, keys);
  EDX_HIP(hipGetLastError());
}

// Complex product formed as (a.re b.re - a.im b.im) + (a.re b.im + a.im b.re) i
// with separate roundings (the file is compiled without FP contraction).
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Sliced copy written by ptab_values itself (col == nullptr: not written):
// G lanes per row, slices of 64 / G rows, `rows` = slices * (64 / G).
struct SellOut {
  int G;
  int64_t rows;
  const int64_t* off;
  int32_t* col;
  double2* val;
  float2* val32;
};

__global__ void __launch_bounds__(kBlock)
ptab_values(TermView tv, SectorView sv, const double2* chi, const int64_t* row_ptr, const uint64_t* keys,
            int32_t* col, uint16_t* term, double2* val, SellOut so) {
  __shared__ double2 chi_s[EDX_MAX_GROUP];
  __shared__ uint32_t binom_s[(EDX_MAX_SITES + 1) * (EDX_MAX_SITES + 1)];
  const int nb = (tv.n_site + 1) * (tv.n_filled + 1);
  for (int i = threadIdx.x; i < nb; i += blockDim.x) binom_s[i] = tv.binom[i];
  for (int i = threadIdx.x; i < EDX_MAX_GROUP; i += blockDim.x) chi_s[i] = chi[i];
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t r = t / kItemLanes;
  const int sub = (int)(t % kItemLanes);  // this lane takes entries k with k % kItemLanes == sub
  // slots of this row in the sliced copy: entry k at base + (k / G) * 64 + k % G
  int64_t sbase = 0, slots = 0;
  if (so.col && r < so.rows) {
    const int R = 64 / so.G;
    const int64_t sl = r / R;
    sbase = so.off[sl] + (r % R) * so.G;
    slots = (so.off[sl + 1] - so.off[sl]) / 64 * so.G;
  }
  int64_t n_row = 0;
  if (r < sv.dim) {
    const int nf1 = tv.n_filled + 1;
    const double stab_row = (double)sv.stab[r];
    const int64_t e0 = row_ptr[r], e1 = row_ptr[r + 1];
    n_row = e1 - e0;
    for (int64_t e = e0 + sub; e < e1; e += kItemLanes) {
      const int64_t k = e - e0;
      const uint64_t key = keys[e];
      const int64_t c = (int64_t)(key >> 16);
      const int kt = (int)(key & 0xffffu);
      const uint64_t m = sv.reps[c];
      const uint64_t nm = m ^ (1ull << tv.from_bit[kt]) ^ (1ull << tv.to_bit[kt]);
      const uint64_t idx = edxd::rank_colex(nm, binom_s, nf1);
      const int g = sv.rep_sym[idx];
      const float2 a32 = sv.rep_amp[idx];
      const double2 a = make_double2((double)a32.x, (double)a32.y);
      const double s = (tv.fermionic && (edxd::popc64(m & tv.between[kt]) & 1)) ? -1.0 : 1.0;
      const double2 amp = tv.amp[kt];
      // t * hop_phase, with hop_phase the complex (s + 0i)
      double2 h = make_double2(amp.x * s - amp.y * 0.0, amp.x * 0.0 + amp.y * s);
      h = cmul(h, a);
      const double2 cc = make_double2(chi_s[g].x, -chi_s[g].y);
      h = cmul(h, cc);
      const double q = sqrt(stab_row / (double)sv.stab[c]);
      h = make_double2(h.x * q - h.y * 0.0, h.x * 0.0 + h.y * q);
      col[e] = (int32_t)c;
      term[e] = (uint16_t)kt;
      val[e] = h;
      if (slots) {
        const int64_t pos = sbase + (k / so.G) * 64 + (k % so.G);
        so.col[pos] = (int32_t)c;
        so.val[pos] = h;
        so.val32[pos] = make_float2((float)h.x, (float)h.y);
      }
    }
  }
  for (int64_t k = sub; k < slots; k += kItemLanes) {  // padding, and the rows past the last one
    if (k < n_row) continue;
    const int64_t pos = sbase + (k / so.G) * 64 + (k % so.G);
    so.col[pos] = -1;
    so.val[pos] = make_double2(0.0, 0.0);
    so.val32[pos] = make_float2(0.0f, 0.0f);
  }
}

__global__ void __launch_bounds__(kBlock)
ptab_diag(int64_t dim, const uint64_t* reps, int n_density, const uint8_t* vi, const uint8_t* vj,
          const double2* v, double2* hdiag, float2* hdiag32) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= dim) return;
  const uint64_t m = reps[r];
  double re = 0.0, im = 0.0;
  for (int k = 0; k < n_density; ++k) {
    if (((m >> vi[k]) & 1ull) && ((m >> vj[k]) & 1ull)) {
      re = re + v[k].x;
      im = im + v[k].y;
    }
  }
  hdiag[r] = make_double2(re, im);
  hdiag32[r] = make_float2((float)re, (float)im);
}

}  // namespace

// ---------------------------------------------------------------------------
// SpMM: Y[j][r] = hdiag[r] X[r][j] + sum_e val[e] X[col[e]][j],  j < nb.
// X interleaved (row-major [dim][nb]) so one gather serves all nb vectors;
// Y column-major with leading dimension ldy.  G lanes share one row.
// ---------------------------------------------------------------------------
namespace edx {

template <typename V> struct Real;
template <> struct Real<double2> { using T = double; };
template <> struct Real<float2> { using T = float; };

// ---------------------------------------------------------------------------
// SpMM on the sliced-ELL copy of the table (SELL-64).  Slice s holds R = 64/G
// consecutive rows; lane l of its wavefront serves row s*R + l/G and takes
// every G-th entry of that row.  Entry t of lane l sits at off[s] + 64 t + l,
// so the column and element loads of a wavefront are contiguous; padding has
// column -1.  X is interleaved (nb values per row, one contiguous gather per
// entry); Y is column-major with leading dimension ldy.
// ---------------------------------------------------------------------------
// Slice s of one table (wavefront `lane` lanes); shared by the single-table
// and the grouped kernels, so both give identical results.  LAYOUT 0: `col`
// and `val` are the columns (-1 = padding) and elements.  LAYOUT 1: `col`
// holds the palette-packed words (column | index << bits, 0xFFFFFFFF =
// padding) and `val` is the palette (in LDS).  LAYOUT 2 (wide): `col` holds
// the columns, `idx` the 16-bit palette indices and `val` the palette (in
// LDS).  The elements multiplied carry the bits of the unpacked copy's, so all
// layouts give identical results.
template <typename V, int NB, int G, int LAYOUT>
__device__ __forceinline__ void sell_slice(int64_t s, int lane, int64_t dim, const int64_t* __restrict__ off,
                                           const int32_t* __restrict__ col, const uint16_t* __restrict__ idx,
                                           const V* __restrict__ val, int bits, const V* __restrict__ hdiag,
                                           const V* __restrict__ X, V* __restrict__ Y, int64_t ldy) {
  constexpr bool PACKED = LAYOUT == 1;
  using T = typename Real<V>::T;
  const int64_t o0 = off[s];
  const int64_t L = (off[s + 1] - o0) >> 6;
  const int64_t row = s * (64 / G) + lane / G;
  T ar[NB], ai[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) ar[q] = ai[q] = (T)0;
  const int32_t* cp = col + o0 + lane;
  const uint16_t* ip = idx + o0 + lane;
  const V* vp = val + o0 + lane;
  const uint32_t cmask = (1u << bits) - 1u;
  // the next entry's column (and palette index) is loaded while this entry's
  // element and x values are fetched, so each entry costs one memory round
  // trip
  int cn = L > 0 ? __builtin_nontemporal_load(cp) : -1;
  int in = LAYOUT == 2 && L > 0 ? (int)__builtin_nontemporal_load(ip) : 0;
  for (int64_t t = 0; t < L; ++t) {
    const int w = cn;
    const int wi = in;
    if (t + 1 < L) {
      cn = __builtin_nontemporal_load(cp + (t + 1) * 64);
      if (LAYOUT == 2) in = (int)__builtin_nontemporal_load(ip + (t + 1) * 64);
    }
    if (PACKED ? (uint32_t)w != 0xFFFFFFFFu : w >= 0) {
      const int c = PACKED ? (int)((uint32_t)w & cmask) : w;
      const V v = PACKED ? val[(uint32_t)w >> bits] : LAYOUT == 2 ? val[wi] : vp[t * 64];
      const V* x = X + (int64_t)c * NB;
      V xv[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) xv[q] = x[q];
      // every gather of the entry is issued before its products (otherwise
      // the compiler may hold back the last ones until the first arrive,
      // which costs a second round trip: complex128 at 6 vectors)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        ar[q] = fma(v.x, xv[q].x, ar[q]);
        ar[q] = fma(-v.y, xv[q].y, ar[q]);
        ai[q] = fma(v.x, xv[q].y, ai[q]);
        ai[q] = fma(v.y, xv[q].x, ai[q]);
      }
    }
  }
  if (G > 1) {
#pragma unroll
    for (int sh = G / 2; sh >= 1; sh >>= 1)
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        ar[q] += __shfl_xor(ar[q], sh, G);
        ai[q] += __shfl_xor(ai[q], sh, G);
      }
  }
  if (lane % G == 0 && row < dim) {
    const V d = hdiag[row];
    const V* xr = X + row * NB;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const V xv = xr[q];
      V out;
      out.x = fma(-d.y, xv.y, fma(d.x, xv.x, ar[q]));
      out.y = fma(d.y, xv.x, fma(d.x, xv.y, ai[q]));
      Y[(int64_t)q * ldy + row] = out;
    }
  }
}

// Copies a palette of n values into LDS (all threads of the workgroup take
// part, before any of them returns).
template <typename V>
__device__ __forceinline__ const V* stage_palette(const V* __restrict__ pal, int n) {
  extern __shared__ __attribute__((aligned(16))) char smem[];  // holds double2 palettes
  V* p = reinterpret_cast<V*>(smem);
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = pal[i];
  __syncthreads();
  return p;
}

template <typename V, int NB, int G, bool PACKED, int BLK = 256>
__global__ void __launch_bounds__(BLK)
spmm_sell(int64_t dim, int64_t n_slices, const int64_t* __restrict__ off, const int32_t* __restrict__ col,
          const V* __restrict__ val, int n_pal, int bits, const V* __restrict__ hdiag, const V* __restrict__ X,
          V* __restrict__ Y, int64_t ldy) {
  const V* v = PACKED ? stage_palette(val, n_pal) : val;
  const int64_t s = (int64_t)blockIdx.x * (BLK / 64) + (threadIdx.x >> 6);
  if (s >= n_slices) return;
  sell_slice<V, NB, G, PACKED ? 1 : 0>(s, threadIdx.x & 63, dim, off, col, nullptr, v, bits, hdiag, X, Y, ldy);
}

// The same on the wide layout (columns and 16-bit palette indices); LDS:
// the palette is staged in LDS, else read through the caches.  BLK threads
// per workgroup (BLK / 64 slices): large palettes are staged once per 1024
// threads, so that LDS does not limit the wavefronts per CU.
template <typename V, int NB, int G, bool LDS, int BLK>
__global__ void __launch_bounds__(BLK)
spmm_sell_wide(int64_t dim, int64_t n_slices, const int64_t* __restrict__ off, const int32_t* __restrict__ col,
               const uint16_t* __restrict__ idx, const V* __restrict__ pal, int n_pal, const V* __restrict__ hdiag,
               const V* __restrict__ X, V* __restrict__ Y, int64_t ldy) {
  const V* v = LDS ? stage_palette(pal, n_pal) : pal;
  const int64_t s = (int64_t)blockIdx.x * (BLK / 64) + (threadIdx.x >> 6);
  if (s >= n_slices) return;
  sell_slice<V, NB, G, 2>(s, threadIdx.x & 63, dim, off, col, idx, v, 0, hdiag, X, Y, ldy);
}

// Grouped SpMM: the sliced products of up to kGroupMax tables (same element
// precision, block width, lanes per row and layout) in one launch.
// Workgroups [first[i], first[i + 1]) serve table i, four slices each, as in
// spmm_sell.
constexpr int kGroupMax = 32;
template <typename V>
struct SellTable {
  int64_t dim, n_slices, ldy;
  const int64_t* off;
  const int32_t* col;
  const uint16_t* idx;  // palette indices (wide layout)
  const V* val;  // elements, or the palette of a packed table
  int n_pal, bits;
  const V* hdiag;
  const V* X;
  V* Y;
};
template <typename V>
struct SellGroup {
  int n;
  int64_t first[kGroupMax + 1];
  SellTable<V> t[kGroupMax];
};

template <typename V, int NB, int G, bool PACKED>
__global__ void __launch_bounds__(256) spmm_sell_group(SellGroup<V> g) {
  const int64_t blk = blockIdx.x;
  int i = 0;
  while (i + 1 < g.n && blk >= g.first[i + 1]) ++i;
  const SellTable<V>& t = g.t[i];
  const V* v = PACKED ? stage_palette(t.val, t.n_pal) : t.val;
  const int64_t s = (blk - g.first[i]) * 4 + (threadIdx.x >> 6);
  if (s >= t.n_slices) return;
  sell_slice<V, NB, G, PACKED ? 1 : 0>(s, threadIdx.x & 63, t.dim, t.off, t.col, nullptr, v, t.bits, t.hdiag, t.X,
                                       t.Y, t.ldy);
}

template <typename V, int NB, int G, bool LDS>
__global__ void __launch_bounds__(256) spmm_sell_group_wide(SellGroup<V> g) {
  const int64_t blk = blockIdx.x;
  int i = 0;
  while (i + 1 < g.n && blk >= g.first[i + 1]) ++i;
  const SellTable<V>& t = g.t[i];
  const V* v = LDS ? stage_palette(t.val, t.n_pal) : t.val;
  const int64_t s = (blk - g.first[i]) * 4 + (threadIdx.x >> 6);
  if (s >= t.n_slices) return;
  sell_slice<V, NB, G, 2>(s, threadIdx.x & 63, t.dim, t.off, t.col, t.idx, v, 0, t.hdiag, t.X, t.Y, t.ldy);
}

template <typename V, int NB, bool LDS, int BLK>
void launch_wide(const edx_hamiltonian& h, hipStream_t st, size_t lds, const V* val, int n_pal, const V* hd,
                 const V* X, V* Y, int64_t ldy) {
  const int64_t* off = h.sell_off.p;
  const int32_t* col = h.sell_col.p;
  const uint16_t* idx = h.sell_idx.p;
  constexpr int S = BLK / 64;
  const dim3 grid((unsigned)((h.sell_slices + S - 1) / S));
  switch (h.sell_g) {
    case 1:
      hipLaunchKernelGGL((spmm_sell_wide<V, NB, 1, LDS, BLK>), grid, dim3(BLK), lds, st, h.dim, h.sell_slices, off,
                         col, idx, val, n_pal, hd, X, Y, ldy);
      break;
    case 2:
      hipLaunchKernelGGL((spmm_sell_wide<V, NB, 2, LDS, BLK>), grid, dim3(BLK), lds, st, h.dim, h.sell_slices, off,
                         col, idx, val, n_pal, hd, X, Y, ldy);
      break;
    default:
      hipLaunchKernelGGL((spmm_sell_wide<V, NB, 4, LDS, BLK>), grid, dim3(BLK), lds, st, h.dim, h.sell_slices, off,
                         col, idx, val, n_pal, hd, X, Y, ldy);
  }
  EDX_HIP(hipGetLastError());
}

template <typename V, int NB, bool PACKED>
void launch_sell_nb(const edx_hamiltonian& h, hipStream_t st, const V* val, const V* hd, const V* X, V* Y,
                    int64_t ldy) {
  const dim3 grid((unsigned)((h.sell_slices + 3) / 4));
  const int64_t* off = h.sell_off.p;
  const int32_t* col = h.sell_col.p;
  const int n_pal = PACKED ? h.n_pal : 0;
  const size_t lds = PACKED ? (size_t)n_pal * sizeof(V) : 0;
  // Packed complex64 products of large tables run on 1024-thread workgroups
  // (16 slices each, the palette staged once per 16 wavefronts): 48-site
  // sector 6.52 -> 6.31 ms, twisted 36-site sector 1523 -> 1473 us.
  // Complex128 products (48-site, 10 vectors: 27.1 -> 30.7 ms) and tables of
  // fewer than 8192 such workgroups (Hubbard 3x4, 6 vectors: 47.6 -> 52.6 us)
  // keep 256 threads.
  const bool big = PACKED && std::is_same<V, float2>::value && (h.sell_slices + 15) / 16 >= 8192;
  if (PACKED && h.sell_packed == 2) {
    if (lds <= (size_t)kPalLdsBytes && !big) launch_wide<V, NB, true, 256>(h, st, lds, val, n_pal, hd, X, Y, ldy);
    else if (lds <= (size_t)kPalLdsBytesLarge) launch_wide<V, NB, true, 1024>(h, st, lds, val, n_pal, hd, X, Y, ldy);
    else launch_wide<V, NB, false, 256>(h, st, 0, val, n_pal, hd, X, Y, ldy);
    return;
  }
  if (PACKED && big) {
    const dim3 g2((unsigned)((h.sell_slices + 15) / 16));
    switch (h.sell_g) {
      case 1:
        hipLaunchKernelGGL((spmm_sell<V, NB, 1, PACKED, 1024>), g2, dim3(1024), lds, st, h.dim, h.sell_slices, off,
                           col, val, n_pal, h.pack_bits, hd, X, Y, ldy);
        break;
      case 2:
        hipLaunchKernelGGL((spmm_sell<V, NB, 2, PACKED, 1024>), g2, dim3(1024), lds, st, h.dim, h.sell_slices, off,
                           col, val, n_pal, h.pack_bits, hd, X, Y, ldy);
        break;
      default:
        hipLaunchKernelGGL((spmm_sell<V, NB, 4, PACKED, 1024>), g2, dim3(1024), lds, st, h.dim, h.sell_slices, off,
                           col, val, n_pal, h.pack_bits, hd, X, Y, ldy);
    }
    EDX_HIP(hipGetLastError());
    return;
  }
  switch (h.sell_g) {
    case 1:
      hipLaunchKernelGGL((spmm_sell<V, NB, 1, PACKED>), grid, dim3(256), lds, st, h.dim, h.sell_slices, off, col,
                         val, n_pal, h.pack_bits, hd, X, Y, ldy);
      break;
    case 2:
      hipLaunchKernelGGL((spmm_sell<V, NB, 2, PACKED>), grid, dim3(256), lds, st, h.dim, h.sell_slices, off, col,
                         val, n_pal, h.pack_bits, hd, X, Y, ldy);
      break;
    default:
      hipLaunchKernelGGL((spmm_sell<V, NB, 4, PACKED>), grid, dim3(256), lds, st, h.dim, h.sell_slices, off, col,
                         val, n_pal, h.pack_bits, hd, X, Y, ldy);
  }
  EDX_HIP(hipGetLastError());
}

template <typename V>
void launch_sell(const edx_hamiltonian& h, hipStream_t st, int nb, const V* val, const V* hd, const V* X, V* Y,
                 int64_t ldy) {
  switch (nb) {
#define EDX_SELL_CASE(N) \
  case N:                \
    if (h.sell_packed) launch_sell_nb<V, N, true>(h, st, val, hd, X, Y, ldy); \
    else launch_sell_nb<V, N, false>(h, st, val, hd, X, Y, ldy); \
    break;
    EDX_SELL_CASE(1) EDX_SELL_CASE(2) EDX_SELL_CASE(3) EDX_SELL_CASE(4) EDX_SELL_CASE(5) EDX_SELL_CASE(6)
    EDX_SELL_CASE(7) EDX_SELL_CASE(8) EDX_SELL_CASE(9) EDX_SELL_CASE(10) EDX_SELL_CASE(11) EDX_SELL_CASE(12)
    EDX_SELL_CASE(13) EDX_SELL_CASE(14) EDX_SELL_CASE(15) EDX_SELL_CASE(16)
#undef EDX_SELL_CASE
    default:
      throw Error{EDX_ERR_INTERNAL, "SpMM block width out of range"};
  }
}

// Algorithmic bytes of one SpMM launch: per row the diagonal, the row's own x
// entries and the y writes; per stored entry the column index, the element
// and one gather of nb x values (padding of the sliced layout not counted).
double spmm_bytes(const edx_hamiltonian& h, int precision, int nb) {
  const double s = precision == EDX_PREC_FP64 ? 16.0 : 8.0;
  // a packed entry is one 4-byte word, a wide one its column and a 2-byte
  // index (the palette stays in LDS)
  const double elem = h.sell_packed == 1 ? 0.0 : h.sell_packed == 2 ? 2.0 : s;
  return (double)h.dim * (s + 2.0 * nb * s) + (double)h.n_entries * (4.0 + elem + nb * s);
}

// ---------------------------------------------------------------------------
// On-the-fly product (no stored table).  Reference: _apply_hamiltonian_kernel
// (SRE:826-923) recomputes H x from the terms and the representative tables
// at every call.  Here lane group of row r (G lanes, each taking every G-th
// occupied site) walks the valid hops of r's representative m, as the table
// build does for column r: target mask nm = m ^ hop, O(N_e) colex rank of
// nm, (orbit, g, alpha) gathered from the catalog, row x = orbit_to_basis,
// element v(x, r) = t s_JW(m) alpha conj(chi(g)) sqrt(stab_x / stab_r) in the
// reference's multiplication order.  That is entry H[x, r]; for catalogs with
// site phases +-1 (alpha = +-1, the block Hermitian to FP64 rounding)
// H[r, x] = conj(H[x, r]) summed over the hops, so y[r] = sum conj(v) X[x]
// without atomics.  Bytes per entry: the three catalog arrays at rank(nm),
// the row index and stabilizer order of x, and the gather of X[x].
// ---------------------------------------------------------------------------
struct FlyView {
  int64_t dim;
  int n_site, n_filled, n_terms, fermionic;
  const uint64_t* reps;
  const int32_t* stab;
  const int32_t* o2b;
  const RankRec* recs;
  const uint32_t* binom;
  const int32_t* toff;
  const uint8_t* to;
  const double2* amp;
  const uint64_t* between;
  const double2* chi;
};

size_t fly_smem_bytes(int n_site, int n_filled, int n_terms) {
  return (size_t)n_terms * 16 + (size_t)n_terms * 8 + (size_t)EDX_MAX_GROUP * 16 + (size_t)(n_site + 1) * 4 +
         (size_t)(n_site + 1) * (n_filled + 1) * 4 + (size_t)n_terms + 16;
}

template <typename V, int NB, int G>
__global__ void __launch_bounds__(256)
spmm_fly(FlyView fv, const V* __restrict__ hdiag, const V* __restrict__ X, V* __restrict__ Y, int64_t ldy) {
  using T = typename Real<V>::T;
  extern __shared__ unsigned char fly_smem[];
  double2* chi = reinterpret_cast<double2*>(fly_smem);
  double2* amp = chi + EDX_MAX_GROUP;
  uint64_t* betw = reinterpret_cast<uint64_t*>(amp + fv.n_terms);
  int32_t* toff = reinterpret_cast<int32_t*>(betw + fv.n_terms);
  const int nb = (fv.n_site + 1) * (fv.n_filled + 1);
  uint32_t* binom = reinterpret_cast<uint32_t*>(toff + fv.n_site + 1);
  uint8_t* to = reinterpret_cast<uint8_t*>(binom + nb);
  for (int i = threadIdx.x; i < fv.n_terms; i += 256) {
    amp[i] = fv.amp[i];
    betw[i] = fv.between[i];
    to[i] = fv.to[i];
  }
  for (int i = threadIdx.x; i < EDX_MAX_GROUP; i += 256) chi[i] = fv.chi[i];
  for (int i = threadIdx.x; i <= fv.n_site; i += 256) toff[i] = fv.toff[i];
  for (int i = threadIdx.x; i < nb; i += 256) binom[i] = fv.binom[i];
  __syncthreads();
  const int64_t r = ((int64_t)blockIdx.x * 256 + threadIdx.x) / G;
  const int sub = (int)(threadIdx.x % G);
  T ar[NB], ai[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) ar[q] = ai[q] = (T)0;
  if (r < fv.dim) {
    const int nf1 = fv.n_filled + 1;
    const uint64_t m = fv.reps[r];
    const double stab_r = (double)fv.stab[r];
    uint64_t occ = m;
    int site = 0;
    while (occ) {
      const int s = edxd::ctz64(occ);
      occ &= occ - 1;
      if ((site++ % G) != sub) continue;
      const int q1 = toff[s + 1];
      for (int q = toff[s]; q < q1; ++q) {
        const int tb = to[q];
        if ((m >> tb) & 1ull) continue;
        const uint64_t nm = m ^ (1ull << s) ^ (1ull << tb);
        const uint64_t idx = edxd::rank_colex(nm, binom, nf1);
        const RankRec rec = fv.recs[idx];  // one 16-byte load
        const int32_t x = fv.o2b[rec.orbit];
        if (x < 0) continue;
        const int g = rec.g;
        const double2 a = make_double2((double)rec.ax, (double)rec.ay);
        const double sj = (fv.fermionic && (edxd::popc64(m & betw[q]) & 1)) ? -1.0 : 1.0;
        const double2 t = amp[q];
        double2 h = make_double2(t.x * sj - t.y * 0.0, t.x * 0.0 + t.y * sj);
        h = cmul(h, a);
        h = cmul(h, make_double2(chi[g].x, -chi[g].y));
        const double qs = sqrt((double)fv.stab[x] / stab_r);
        h = make_double2(h.x * qs - h.y * 0.0, h.x * 0.0 + h.y * qs);
        const T vr = (T)h.x, vi = (T)-h.y;  // conj(v)
        const V* xp = X + (int64_t)x * NB;
#pragma unroll
        for (int q2 = 0; q2 < NB; ++q2) {
          const V xv = xp[q2];
          ar[q2] = fma(vr, xv.x, ar[q2]);
          ar[q2] = fma(-vi, xv.y, ar[q2]);
          ai[q2] = fma(vr, xv.y, ai[q2]);
          ai[q2] = fma(vi, xv.x, ai[q2]);
        }
      }
    }
  }
  if (G > 1) {
#pragma unroll
    for (int sh = G / 2; sh >= 1; sh >>= 1)
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        ar[q] += __shfl_xor(ar[q], sh, G);
        ai[q] += __shfl_xor(ai[q], sh, G);
      }
  }
  if (sub == 0 && r < fv.dim) {
    const V d = hdiag[r];
    const V* xr = X + r * NB;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const V xv = xr[q];
      V out;
      out.x = fma(-d.y, xv.y, fma(d.x, xv.x, ar[q]));
      out.y = fma(d.y, xv.x, fma(d.x, xv.y, ai[q]));
      Y[(int64_t)q * ldy + r] = out;
    }
  }
}

constexpr int kFlyLanes = 4;

FlyView fly_view(const edx_hamiltonian& h) {
  const edx_sector& sec = *h.sec;
  const edx_catalog& cat = *sec.cat;
  EDX_REQUIRE(cat.exact_alpha, EDX_ERR_INVALID,
              "the catalog's phases were refreshed to a twist after this on-the-fly Hamiltonian was built: its "
              "block is no longer Hermitian; build a stored table");
  EDX_REQUIRE(cat.rank_rec.p, EDX_ERR_INTERNAL, "catalog rank records missing");
  return FlyView{h.dim, cat.n_site, cat.n_filled, h.n_bilinear, cat.fermionic, sec.reps.p, sec.stab_order.p,
                 sec.orbit_to_basis.p, cat.rank_rec.p, cat.binom.p, h.fly_toff.p, h.fly_to.p, h.fly_amp.p,
                 h.fly_between.p, h.fly_chi.p};
}

template <typename V, int NB>
void launch_fly_nb(const edx_hamiltonian& h, hipStream_t st, const V* hd, const V* X, V* Y, int64_t ldy) {
  const FlyView fv = fly_view(h);
  const size_t smem = fly_smem_bytes(fv.n_site, fv.n_filled, fv.n_terms);
  hipLaunchKernelGGL((spmm_fly<V, NB, kFlyLanes>), dim3(grid_for((size_t)h.dim * kFlyLanes, 256)), dim3(256), smem,
                     st, fv, hd, X, Y, ldy);
  EDX_HIP(hipGetLastError());
}

template <typename V>
void launch_fly(const edx_hamiltonian& h, hipStream_t st, int nb, const V* hd, const V* X, V* Y, int64_t ldy) {
  switch (nb) {
#define EDX_FLY_CASE(N) \
  case N:               \
    launch_fly_nb<V, N>(h, st, hd, X, Y, ldy); \
    break;
    EDX_FLY_CASE(1) EDX_FLY_CASE(2) EDX_FLY_CASE(3) EDX_FLY_CASE(4) EDX_FLY_CASE(5) EDX_FLY_CASE(6)
    EDX_FLY_CASE(7) EDX_FLY_CASE(8) EDX_FLY_CASE(9) EDX_FLY_CASE(10) EDX_FLY_CASE(11) EDX_FLY_CASE(12)
    EDX_FLY_CASE(13) EDX_FLY_CASE(14) EDX_FLY_CASE(15) EDX_FLY_CASE(16)
#undef EDX_FLY_CASE
    default:
      throw Error{EDX_ERR_INTERNAL, "SpMM block width out of range"};
  }
}

// Algorithmic bytes of one on-the-fly launch: per row its mask, stabilizer
// order, diagonal, own x entries and y writes; per entry the three catalog
// arrays at the target's rank (16 B), its orbit's sector row and stabilizer
// order (8 B) and the gather of nb x values.
double fly_bytes(const edx_hamiltonian& h, int precision, int nb) {
  const double s = precision == EDX_PREC_FP64 ? 16.0 : 8.0;
  return (double)h.dim * (8.0 + 4.0 + s + 2.0 * nb * s) + (double)h.n_entries * (16.0 + 8.0 + nb * s);
}

// Applies H to nb <= 16 interleaved vectors.  precision selects the
// complex128 or complex64 element copy (on-the-fly tables: the vectors'
// precision, elements formed in FP64).
void spmm(const edx_hamiltonian& h, Ctx& ctx, int precision, int nb, const void* X, void* Y, int64_t ldy) {
  if (h.dim == 0) return;
  EDX_REQUIRE(nb >= 1 && nb <= 16, EDX_ERR_INVALID, "at most 16 vectors per SpMM");
  hipStream_t st = ctx.stream;
  hipEvent_t t0 = ctx.timer.start(st);
  if (h.on_the_fly) {
    if (precision == EDX_PREC_FP64)
      launch_fly<double2>(h, st, nb, h.hdiag.p, (const double2*)X, (double2*)Y, ldy);
    else
      launch_fly<float2>(h, st, nb, h.hdiag32.p, (const float2*)X, (float2*)Y, ldy);
    ctx.timer.stop(st, t0, precision == EDX_PREC_FP64 ? K_SPMM : K_SPMM32, fly_bytes(h, precision, nb));
    return;
  }
  if (precision == EDX_PREC_FP64)
    launch_sell<double2>(h, st, nb, h.sell_packed ? h.pal.p : h.sell_val.p, h.hdiag.p, (const double2*)X,
                         (double2*)Y, ldy);
  else
    launch_sell<float2>(h, st, nb, h.sell_packed ? h.pal32.p : h.sell_val32.p, h.hdiag32.p, (const float2*)X,
                        (float2*)Y, ldy);
  ctx.timer.stop(st, t0, precision == EDX_PREC_FP64 ? K_SPMM : K_SPMM32, spmm_bytes(h, precision, nb));
}

template <typename V, int NB, bool PACKED>
void launch_group_packed(const SellGroup<V>& g, int lanes, unsigned blocks, size_t lds, hipStream_t st) {
  switch (lanes) {
    case 1: hipLaunchKernelGGL((spmm_sell_group<V, NB, 1, PACKED>), dim3(blocks), dim3(256), lds, st, g); break;
    case 2: hipLaunchKernelGGL((spmm_sell_group<V, NB, 2, PACKED>), dim3(blocks), dim3(256), lds, st, g); break;
    default: hipLaunchKernelGGL((spmm_sell_group<V, NB, 4, PACKED>), dim3(blocks), dim3(256), lds, st, g);
  }
  EDX_HIP(hipGetLastError());
}

template <typename V, int NB>
void launch_group_nb(const SellGroup<V>& g, int lanes, int packed, unsigned blocks, hipStream_t st) {
  int n_pal = 0;
  for (int i = 0; i < g.n; ++i) n_pal = std::max(n_pal, g.t[i].n_pal);
  const size_t lds = (size_t)n_pal * sizeof(V);
  if (packed == 2) {
    if (lds <= (size_t)kPalLdsBytes) {
      switch (lanes) {
        case 1: hipLaunchKernelGGL((spmm_sell_group_wide<V, NB, 1, true>), dim3(blocks), dim3(256), lds, st, g); break;
        case 2: hipLaunchKernelGGL((spmm_sell_group_wide<V, NB, 2, true>), dim3(blocks), dim3(256), lds, st, g); break;
        default: hipLaunchKernelGGL((spmm_sell_group_wide<V, NB, 4, true>), dim3(blocks), dim3(256), lds, st, g);
      }
    } else {
      switch (lanes) {
        case 1: hipLaunchKernelGGL((spmm_sell_group_wide<V, NB, 1, false>), dim3(blocks), dim3(256), 0, st, g); break;
        case 2: hipLaunchKernelGGL((spmm_sell_group_wide<V, NB, 2, false>), dim3(blocks), dim3(256), 0, st, g); break;
        default: hipLaunchKernelGGL((spmm_sell_group_wide<V, NB, 4, false>), dim3(blocks), dim3(256), 0, st, g);
      }
    }
    EDX_HIP(hipGetLastError());
  } else if (packed == 1) {
    launch_group_packed<V, NB, true>(g, lanes, blocks, lds, st);
  } else {
    launch_group_packed<V, NB, false>(g, lanes, blocks, 0, st);
  }
}

template <typename V>
void launch_group(const SpmmJob* jobs, const std::vector<int>& idx, int precision, hipStream_t st) {
  SellGroup<V> g{};
  g.n = (int)idx.size();
  g.first[0] = 0;
  for (int i = 0; i < g.n; ++i) {
    const SpmmJob& j = jobs[idx[i]];
    const edx_hamiltonian& h = *j.h;
    SellTable<V>& t = g.t[i];
    t.dim = h.dim;
    t.n_slices = h.sell_slices;
    t.ldy = j.ldy;
    t.off = h.sell_off.p;
    t.col = h.sell_col.p;
    if constexpr (std::is_same<V, double2>::value) {
      t.val = h.sell_packed ? h.pal.p : h.sell_val.p;
      t.hdiag = h.hdiag.p;
    } else {
      t.val = h.sell_packed ? h.pal32.p : h.sell_val32.p;
      t.hdiag = h.hdiag32.p;
    }
    t.n_pal = h.sell_packed ? h.n_pal : 0;
    t.bits = h.pack_bits;
    t.idx = h.sell_idx.p;
    t.X = (const V*)j.X;
    t.Y = (V*)j.Y;
    g.first[i + 1] = g.first[i] + (h.sell_slices + 3) / 4;
  }
  EDX_REQUIRE(g.first[g.n] < (int64_t(1) << 31), EDX_ERR_INTERNAL, "grouped SpMM grid");
  const unsigned blocks = (unsigned)g.first[g.n];
  const int lanes = jobs[idx[0]].h->sell_g;
  const int packed = jobs[idx[0]].h->sell_packed;
  switch (jobs[idx[0]].nb) {
#define EDX_GROUP_CASE(N) \
  case N:                 \
    launch_group_nb<V, N>(g, lanes, packed, blocks, st); \
    break;
    EDX_GROUP_CASE(1) EDX_GROUP_CASE(2) EDX_GROUP_CASE(3) EDX_GROUP_CASE(4) EDX_GROUP_CASE(5) EDX_GROUP_CASE(6)
    EDX_GROUP_CASE(7) EDX_GROUP_CASE(8) EDX_GROUP_CASE(9) EDX_GROUP_CASE(10) EDX_GROUP_CASE(11) EDX_GROUP_CASE(12)
    EDX_GROUP_CASE(13) EDX_GROUP_CASE(14) EDX_GROUP_CASE(15) EDX_GROUP_CASE(16)
#undef EDX_GROUP_CASE
    default:
      throw Error{EDX_ERR_INTERNAL, "SpMM block width out of range"};
  }
  (void)precision;
}

double spmm_group_bytes(const SpmmJob* jobs, int n, int precision) {
  double b = 0.0;
  for (int i = 0; i < n; ++i)
    if (jobs[i].h->dim > 0)
      b += jobs[i].h->on_the_fly ? fly_bytes(*jobs[i].h, precision, jobs[i].nb)
                                 : spmm_bytes(*jobs[i].h, precision, jobs[i].nb);
  return b;
}

// Applies each job's H to its block (all of one precision) on ctx's stream.
// Sliced tables of equal block width and lanes per row share one launch (up
// to kGroupMax tables each); on-the-fly tables run one by one.  Every table
// gets exactly the arithmetic of spmm().
void spmm_group(const SpmmJob* jobs, int n, Ctx& ctx, int precision) {
  hipStream_t st = ctx.stream;
  std::vector<char> left(n, 1);
  for (int i = 0; i < n; ++i) {
    const edx_hamiltonian& h = *jobs[i].h;
    EDX_REQUIRE(jobs[i].nb >= 1 && jobs[i].nb <= 16, EDX_ERR_INVALID, "at most 16 vectors per SpMM");
    if (h.dim == 0) {
      left[i] = 0;
    } else if (h.on_the_fly) {
      spmm(h, ctx, precision, jobs[i].nb, jobs[i].X, jobs[i].Y, jobs[i].ldy);
      left[i] = 0;
    }
  }
  for (int i0 = 0; i0 < n; ++i0) {
    if (!left[i0]) continue;
    std::vector<int> idx;
    double bytes = 0.0;
    for (int i = i0; i < n && (int)idx.size() < kGroupMax; ++i)
      if (left[i] && jobs[i].nb == jobs[i0].nb && jobs[i].h->sell_g == jobs[i0].h->sell_g &&
          jobs[i].h->sell_packed == jobs[i0].h->sell_packed) {
        idx.push_back(i);
        left[i] = 0;
        bytes += spmm_bytes(*jobs[i].h, precision, jobs[i].nb);
      }
    hipEvent_t t0 = ctx.timer.start(st);
    if (precision == EDX_PREC_FP64) launch_group<double2>(jobs, idx, precision, st);
    else launch_group<float2>(jobs, idx, precision, st);
    ctx.timer.stop(st, t0, precision == EDX_PREC_FP64 ? K_SPMM : K_SPMM32, bytes);
  }
}

// ---- construction of the sliced copy from the CSR table
__global__ void sell_slice_len(int64_t dim, int64_t n_slices, int G, const int64_t* row_ptr, uint32_t* words) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n_slices) return;
  const int R = 64 / G;
  int64_t L = 0;
  for (int64_t r = s * R; r < min(dim, (s + 1) * R); ++r) {
    const int64_t n = row_ptr[r + 1] - row_ptr[r];
    L = max(L, (n + G - 1) / G);
  }
  words[s] = (uint32_t)(L * 64);
}

// Sum of the table entries of row c in column r (columns ascend within a row).
__device__ double2 entry_sum(const int64_t* row_ptr, const int32_t* col, const double2* val, int64_t c, int32_t r) {
  int64_t lo = row_ptr[c], hi = row_ptr[c + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (col[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  double2 s = make_double2(0.0, 0.0);
  for (int64_t e = lo; e < row_ptr[c + 1] && col[e] == r; ++e) {
    s.x += val[e].x;
    s.y += val[e].y;
  }
  return s;
}

// herm = 1 stores the Hermitian part: every entry of the run (r, c) is halved
// and the run's first entry also carries conj(sum of the run (c, r)) / 2, so the
// run sums to (H_rc + conj(H_cr)) / 2.  kItemLanes lanes share a row, each
// taking every kItemLanes-th entry: the transposed look-ups of a row are in
// flight together, and with G = 4 the lanes of a slice store contiguous slots.
__global__ void sell_fill(int64_t dim, int G, int herm, const int64_t* row_ptr, const int32_t* col,
                          const double2* val, const int64_t* off, int32_t* scol, double2* sval, float2* sval32) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t r = t / kItemLanes;
  const int sub = (int)(t % kItemLanes);
  if (r >= dim) return;
  const int R = 64 / G;
  const int64_t s = r / R;
  const int64_t base = off[s] + (r % R) * G;
  const int64_t e0 = row_ptr[r], e1 = row_ptr[r + 1];
  for (int64_t e = e0 + sub; e < e1; e += kItemLanes) {
    const int64_t k = e - e0;
    const int64_t pos = base + (k / G) * 64 + (k % G);
    const int32_t c = col[e];
    double2 v = val[e];
    if (herm) {
      v = make_double2(0.5 * v.x, 0.5 * v.y);
      if (e == e0 || col[e - 1] != c) {
        const double2 t = entry_sum(row_ptr, col, val, c, (int32_t)r);
        v.x += 0.5 * t.x;
        v.y -= 0.5 * t.y;
      }
    }
    scol[pos] = c;
    sval[pos] = v;
    sval32[pos] = make_float2((float)v.x, (float)v.y);
  }
}

// Y[j][r] = hdiag[r] X[r][j] + sum_e val[e] X[col[e]][j] straight from the CSR
// table (X interleaved [dim][nb], Y column-major with leading dimension ldy):
// the table's own operator where the sliced copy holds its Hermitian part.
template <typename V>
__global__ void csr_apply(int64_t dim, int nb, const int64_t* row_ptr, const int32_t* col, const double2* val,
                          const double2* hd, const V* X, V* Y, int64_t ldy) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= dim) return;
  for (int j = 0; j < nb; ++j) {
    const V x = X[r * nb + j];
    double yr = hd[r].x * x.x - hd[r].y * x.y, yi = hd[r].x * x.y + hd[r].y * x.x;
    for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
      const V xc = X[(int64_t)col[e] * nb + j];
      const double2 a = val[e];
      yr += a.x * xc.x - a.y * xc.y;
      yi += a.x * xc.y + a.y * xc.x;
    }
    V y;
    y.x = (typename Real<V>::T)yr;
    y.y = (typename Real<V>::T)yi;
    Y[(int64_t)j * ldy + r] = y;
  }
}

// The table's operator: the sliced SpMM, or the CSR arrays when the sliced
// copy holds the Hermitian part.
void apply_table(const edx_hamiltonian& h, Ctx& ctx, int precision, int nb, const void* X, void* Y, int64_t ldy) {
  if (!h.herm_part) {
    spmm(h, ctx, precision, nb, X, Y, ldy);
    return;
  }
  if (precision == EDX_PREC_FP64)
    hipLaunchKernelGGL(csr_apply<double2>, dim3(grid_for(h.dim, 256)), dim3(256), 0, ctx.stream, h.dim, nb,
                       (const int64_t*)h.row_ptr.p, (const int32_t*)h.col.p, (const double2*)h.val.p,
                       (const double2*)h.hdiag.p, (const double2*)X, (double2*)Y, ldy);
  else
    hipLaunchKernelGGL(csr_apply<float2>, dim3(grid_for(h.dim, 256)), dim3(256), 0, ctx.stream, h.dim, nb,
                       (const int64_t*)h.row_ptr.p, (const int32_t*)h.col.p, (const double2*)h.val.p,
                       (const double2*)h.hdiag.p, (const float2*)X, (float2*)Y, ldy);
  EDX_HIP(hipGetLastError());
}

int default_sell_lanes(int64_t dim, int64_t n_entries) {
  // rows of 16 or more entries split over 4 lanes, of 6 or more over 2: more
  // gathers in flight per row.  Measured on MI355X for rows of 25 entries
  // (D = 2.3e5) and 82 entries (D = 1.6e7): 4 lanes are fastest at every
  // block width in both precisions, by up to 4.4x over 1 lane for complex128
  // at 8 vectors, where one lane per row keeps too few loads in flight.
  const double mean_row = dim ? (double)n_entries / (double)dim : 0.0;
  if (mean_row >= 16) return 4;
  if (mean_row >= 6) return 2;
  return 1;
}

// Slice offsets and storage of the sliced copy with G lanes per row.
int64_t prepare_sell(edx_hamiltonian& h, Ctx& ctx, int G) {
  EDX_REQUIRE(G == 1 || G == 2 || G == 4, EDX_ERR_INVALID, "lanes per row must be 1, 2 or 4");
  hipStream_t st = ctx.stream;
  const int64_t D = h.dim;
  const int64_t ns = (D + 64 / G - 1) / (64 / G);
  h.sell_g = G;
  h.sell_slices = ns;
  h.sell_entries = 0;
  h.sell_packed = 0;
  h.n_pal = 0;
  h.pack_bits = 0;
  h.pal.release();
  h.pal32.release();
  h.sell_idx.release();
  h.sell_off.alloc((size_t)ns + 1);
  if (ns == 0) {
    const int64_t zero = 0;
    EDX_HIP(hipMemcpyAsync(h.sell_off.p, &zero, 8, hipMemcpyHostToDevice, st));
    EDX_HIP(hipStreamSynchronize(st));
    return 0;
  }
  DevBuf<uint32_t> words((size_t)ns);
  hipLaunchKernelGGL(sell_slice_len, dim3(grid_for(ns, 256)), dim3(256), 0, st, D, ns, G,
                     (const int64_t*)h.row_ptr.p, words.p);
  EDX_HIP(hipGetLastError());
  const int64_t total = (int64_t)exclusive_scan_u32_to_i64(ctx, words.p, h.sell_off.p, (size_t)ns);
  EDX_HIP(hipMemcpyAsync(h.sell_off.p + ns, &total, 8, hipMemcpyHostToDevice, st));
  h.sell_col.alloc((size_t)std::max<int64_t>(total, 1));
  h.sell_val.alloc((size_t)std::max<int64_t>(total, 1));
  h.sell_val32.alloc((size_t)std::max<int64_t>(total, 1));
  h.sell_entries = total;
  return total;
}

// Builds (or rebuilds with another lane count) the sliced copy of the table
// from its CSR arrays.
void build_sell(edx_hamiltonian& h, Ctx& ctx, int G) {
  prepare_sell(h, ctx, G);
  if (h.sell_slices == 0) return;
  hipStream_t st = ctx.stream;
  const int64_t D = h.dim;
  EDX_HIP(hipMemsetAsync(h.sell_col.p, 0xff, h.sell_col.bytes(), st));
  EDX_HIP(hipMemsetAsync(h.sell_val.p, 0, h.sell_val.bytes(), st));
  EDX_HIP(hipMemsetAsync(h.sell_val32.p, 0, h.sell_val32.bytes(), st));
  hipLaunchKernelGGL(sell_fill, dim3(grid_for((size_t)D * kItemLanes, 256)), dim3(256), 0, st, D, G, h.herm_part,
                     (const int64_t*)h.row_ptr.p, (const int32_t*)h.col.p, (const double2*)h.val.p,
                     (const int64_t*)h.sell_off.p, h.sell_col.p, h.sell_val.p, h.sell_val32.p);
  EDX_HIP(hipGetLastError());
  EDX_HIP(hipStreamSynchronize(st));
  pack_sell(h, ctx);
}

// Column-major [nb][ld] -> interleaved [dim][nb].
template <typename V>
__global__ void interleave_kernel(int64_t dim, int nb, const V* src, int64_t ld, V* dst) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= dim) return;
  for (int j = 0; j < nb; ++j) dst[r * nb + j] = src[(int64_t)j * ld + r];
}

void interleave(hipStream_t st, int precision, int64_t dim, int nb, const void* src, int64_t ld, void* dst) {
  if (dim == 0) return;
  if (precision == EDX_PREC_FP64)
    hipLaunchKernelGGL(interleave_kernel<double2>, dim3(grid_for(dim, 256)), dim3(256), 0, st, dim, nb,
                       (const double2*)src, ld, (double2*)dst);
  else
    hipLaunchKernelGGL(interleave_kernel<float2>, dim3(grid_for(dim, 256)), dim3(256), 0, st, dim, nb,
                       (const float2*)src, ld, (float2*)dst);
  EDX_HIP(hipGetLastError());
}

// Dense row-major complex128 copy of the block (entries summed per row in
// stored order, diagonal added last).
__global__ void dense_kernel(int64_t dim, const int64_t* row_ptr, const int32_t* col, const double2* val,
                             const double2* hdiag, double2* out) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= dim) return;
  double2* o = out + r * dim;
  for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
    double2 a = o[col[e]];
    a.x = a.x + val[e].x;
    a.y = a.y + val[e].y;
    o[col[e]] = a;
  }
  o[r].x = o[r].x + hdiag[r].x;
  o[r].y = o[r].y + hdiag[r].y;
}

__global__ void unit_block(int64_t dim, int nb, int64_t j0, double2* X) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= dim * nb) return;
  const int64_t r = i / nb, q = i % nb;
  X[i] = make_double2(r == j0 + q ? 1.0 : 0.0, 0.0);
}

__global__ void conj_inplace(int64_t n, double2* a) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i].y = -a[i].y;
}

// Dense block of an on-the-fly Hamiltonian: H applied to 16 unit vectors at a
// time gives columns c of H, written as rows c of out; conjugated, out[c][r]
// = conj(H[r][c]) = H[c][r] (the block is Hermitian).
void dense_matrix_fly(const edx_hamiltonian& h, Ctx& ctx, double2* out) {
  const int64_t D = h.dim;
  DevBuf<double2> X((size_t)D * 16);
  for (int64_t j0 = 0; j0 < D; j0 += 16) {
    const int nb = (int)std::min<int64_t>(16, D - j0);
    hipLaunchKernelGGL(unit_block, dim3(grid_for((size_t)D * nb, 256)), dim3(256), 0, ctx.stream, D, nb, j0, X.p);
    EDX_HIP(hipGetLastError());
    spmm(h, ctx, EDX_PREC_FP64, nb, X.p, out + j0 * D, D);
  }
  hipLaunchKernelGGL(conj_inplace, dim3(grid_for((size_t)D * D, 256)), dim3(256), 0, ctx.stream, D * D, out);
  EDX_HIP(hipGetLastError());
}

void dense_matrix(const edx_hamiltonian& h, Ctx& ctx, double2* out) {
  if (h.on_the_fly) {
    dense_matrix_fly(h, ctx, out);
    return;
  }
  hipStream_t st = ctx.stream;
  EDX_HIP(hipMemsetAsync(out, 0, (size_t)h.dim * h.dim * sizeof(double2), st));
  hipLaunchKernelGGL(dense_kernel, dim3(grid_for(h.dim, 256)), dim3(256), 0, st, h.dim,
                     (const int64_t*)h.row_ptr.p, (const int32_t*)h.col.p, (const double2*)h.val.p,
                     (const double2*)h.hdiag.p, out);
  EDX_HIP(hipGetLastError());
}

}  // namespace edx

namespace {

// Shared body of the stored-table and on-the-fly builds.  fly = 1 keeps the
// grouped term arrays, the characters and the diagonal, counts the entries
// and stores no table.
int build_hamiltonian(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec, const double* chi,
                      int32_t n_bilinear, const int32_t* i_from, const int32_t* i_to, const double* t,
                      int32_t n_density, const int32_t* v_i, const int32_t* v_j, const double* v,
                      edx_hamiltonian** out, int64_t* n_entries, int fly) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && cat && sec && chi && out, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(!fly || cat->exact_alpha, EDX_ERR_INVALID,
                "on-the-fly products need a catalog whose site phases are all +1 or -1 (the block is then "
                "Hermitian); build a stored table for twisted catalogs");
    EDX_REQUIRE(sec->cat == cat, EDX_ERR_INVALID, "sector was built from a different catalog");
    EDX_REQUIRE(n_bilinear >= 0 && n_bilinear <= EDX_MAX_TERMS, EDX_ERR_INVALID, "too many bilinear terms");
    EDX_REQUIRE(n_density >= 0, EDX_ERR_INVALID, "negative density term count");
    EDX_REQUIRE(n_bilinear == 0 || (i_from && i_to && t), EDX_ERR_INVALID, "null bilinear term arrays");
    EDX_REQUIRE(n_density == 0 || (v_i && v_j && v), EDX_ERR_INVALID, "null density term arrays");
    *out = nullptr;
    const int N = cat->n_site;
    // ---- host-side term preparation
    std::vector<uint8_t> fb(n_bilinear), tb(n_bilinear);
    std::vector<uint64_t> betw(n_bilinear);
    std::vector<std::vector<int>> by_src(N);
    for (int k = 0; k < n_bilinear; ++k) {
      const int a = i_from[k], b = i_to[k];
      EDX_REQUIRE(a >= 1 && a <= N && b >= 1 && b <= N, EDX_ERR_INVALID, "bilinear term site out of range");
      EDX_REQUIRE(a != b, EDX_ERR_INVALID,
                  "bilinear term with i == j: the hop gate would remove a particle; pass on-site energies as "
                  "density terms (i, i, v)");
      fb[k] = (uint8_t)(a - 1);
      tb[k] = (uint8_t)(b - 1);
      const int lo = std::min(a, b), hi = std::max(a, b);
      betw[k] = (hi - lo > 1) ? (((1ull << (hi - lo - 1)) - 1ull) << lo) : 0ull;
      by_src[a - 1].push_back(k);
    }
    std::vector<int32_t> toff(N + 1, 0);
    std::vector<uint16_t> tk;
    std::vector<uint8_t> tto;
    for (int s = 0; s < N; ++s) {
      toff[s] = (int32_t)tk.size();
      for (int k : by_src[s]) {
        tk.push_back((uint16_t)k);
        tto.push_back(tb[k]);
      }
    }
    toff[N] = (int32_t)tk.size();
    std::vector<uint8_t> dvi(n_density), dvj(n_density);
    for (int k = 0; k < n_density; ++k) {
      EDX_REQUIRE(v_i[k] >= 1 && v_i[k] <= N && v_j[k] >= 1 && v_j[k] <= N, EDX_ERR_INVALID,
                  "density term site out of range");
      dvi[k] = (uint8_t)(v_i[k] - 1);
      dvj[k] = (uint8_t)(v_j[k] - 1);
    }
    edx::DeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    const int T = n_bilinear;
    DevBuf<int32_t> d_toff(N + 1);
    DevBuf<uint16_t> d_tk(std::max(T, 1));
    DevBuf<uint8_t> d_tto(std::max(T, 1)), d_fb(std::max(T, 1)), d_tb(std::max(T, 1));
    DevBuf<double2> d_amp(std::max(T, 1));
    DevBuf<uint64_t> d_betw(std::max(T, 1));
    DevBuf<double2> d_chi(EDX_MAX_GROUP);
    DevBuf<uint8_t> d_vi(std::max(n_density, 1)), d_vj(std::max(n_density, 1));
    DevBuf<double2> d_v(std::max(n_density, 1));
    std::vector<double> chi_pad(2 * EDX_MAX_GROUP, 0.0);
    std::copy(chi, chi + 2 * cat->n_g, chi_pad.begin());
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
      if (bytes) EDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    };
    h2d(d_toff.p, toff.data(), toff.size() * 4);
    h2d(d_tk.p, tk.data(), tk.size() * 2);
    h2d(d_tto.p, tto.data(), tto.size());
    h2d(d_fb.p, fb.data(), fb.size());
    h2d(d_tb.p, tb.data(), tb.size());
    h2d(d_amp.p, t, (size_t)T * 16);
    h2d(d_betw.p, betw.data(), betw.size() * 8);
    h2d(d_chi.p, chi_pad.data(), chi_pad.size() * 8);
    h2d(d_vi.p, dvi.data(), dvi.size());
    h2d(d_vj.p, dvj.data(), dvj.size());
    h2d(d_v.p, v, (size_t)n_density * 16);

    TermView tv{N, cat->n_filled, T, cat->fermionic, d_toff.p, d_tk.p, d_tto.p, d_fb.p, d_tb.p, d_amp.p,
                d_betw.p, cat->binom.p};
    const int64_t D = sec->dim;
    SectorView sv{D, sec->reps.p, sec->stab_order.p, sec->orbit_to_basis.p, cat->rep_rank.p, cat->rep_sym.p,
                  cat->rep_amp.p};
    auto ham = new edx_hamiltonian();
    edx::SellPlan plan;  // sliced copy written by ptab_values (not for twisted tables)
    bool sell_written = false;
    try {
      ham->device = ctx->device;
      ham->sec = sec;
      ham->dim = D;
      ham->n_bilinear = T;
      ham->herm_part = cat->exact_alpha ? 0 : 1;
      ham->hdiag.alloc(D);
      ham->hdiag32.alloc(D);
      const size_t smem = terms_smem_bytes(N, cat->n_filled, T);
      int64_t E = 0;
      if (fly) {
        ham->on_the_fly = 1;
        ham->sell_g = 0;
        edx::ensure_rank_records(*const_cast<edx_catalog*>(cat), *ctx);
        EDX_REQUIRE(edx::fly_smem_bytes(N, cat->n_filled, T) <= 65536, EDX_ERR_INVALID,
                    "too many bilinear terms for the on-the-fly product's shared-memory term list");
        std::vector<double2> amp_g(std::max(T, 1));
        std::vector<uint64_t> betw_g(std::max(T, 1));
        for (size_t q = 0; q < tk.size(); ++q) {
          amp_g[q] = make_double2(t[2 * tk[q]], t[2 * tk[q] + 1]);
          betw_g[q] = betw[tk[q]];
        }
        ham->fly_toff.alloc(N + 1);
        ham->fly_to.alloc(std::max(T, 1));
        ham->fly_amp.alloc(std::max(T, 1));
        ham->fly_between.alloc(std::max(T, 1));
        ham->fly_chi.alloc(EDX_MAX_GROUP);
        h2d(ham->fly_toff.p, toff.data(), toff.size() * 4);
        h2d(ham->fly_to.p, tto.data(), tto.size());
        h2d(ham->fly_amp.p, amp_g.data(), (size_t)T * 16);
        h2d(ham->fly_between.p, betw_g.data(), (size_t)T * 8);
        h2d(ham->fly_chi.p, chi_pad.data(), chi_pad.size() * 8);
        if (D > 0) {
          DevBuf<uint32_t> cnt(D);
          DevBuf<int64_t> rp(D + 1);
          EDX_HIP(hipMemsetAsync(cnt.p, 0, D * 4, st));
          hipLaunchKernelGGL(ptab_count, dim3(edx::grid_for(D, kBlock)), dim3(kBlock), smem, st, tv, sv, cnt.p);
          EDX_HIP(hipGetLastError());
          E = (int64_t)edx::exclusive_scan_u32_to_i64(*ctx, cnt.p, rp.p, (size_t)D);
          hipLaunchKernelGGL(ptab_diag, dim3(edx::grid_for(D, kBlock)), dim3(kBlock), 0, st, D,
                             (const uint64_t*)sec->reps.p, n_density, (const uint8_t*)d_vi.p,
                             (const uint8_t*)d_vj.p, (const double2*)d_v.p, ham->hdiag.p, ham->hdiag32.p);
          EDX_HIP(hipGetLastError());
        }
        EDX_HIP(hipStreamSynchronize(st));
        ham->n_entries = E;
        *out = ham;
        if (n_entries) *n_entries = E;
        return;
      }
      ham->row_ptr.alloc(D + 1);
      if (D > 0) {
        DevBuf<uint32_t> cnt(D), rmax(1);
        EDX_HIP(hipMemsetAsync(cnt.p, 0, D * 4, st));
        EDX_HIP(hipMemsetAsync(rmax.p, 0, 4, st));
        hipLaunchKernelGGL(ptab_count, dim3(edx::grid_for(D, kBlock)), dim3(kBlock), smem, st, tv, sv, cnt.p);
        EDX_HIP(hipGetLastError());
        hipLaunchKernelGGL(row_len_max, dim3(edx::grid_for(D, kBlock)), dim3(kBlock), 0, st, D,
                           (const uint32_t*)cnt.p, rmax.p);
        EDX_HIP(hipGetLastError());
        E = (int64_t)edx::exclusive_scan_u32_to_i64(*ctx, cnt.p, ham->row_ptr.p, (size_t)D);
        uint32_t max_row = 0;
        EDX_HIP(hipMemcpyAsync(&max_row, rmax.p, 4, hipMemcpyDeviceToHost, st));
        EDX_HIP(hipStreamSynchronize(st));
        EDX_HIP(hipMemcpyAsync(ham->row_ptr.p + D, &E, 8, hipMemcpyHostToDevice, st));
        DevBuf<uint64_t> keys(std::max<int64_t>(E, 1));
        EDX_HIP(hipMemsetAsync(cnt.p, 0, D * 4, st));
        hipLaunchKernelGGL(ptab_fill, dim3(edx::grid_for(D, kBlock)), dim3(kBlock), smem, st, tv, sv,
                           (const int64_t*)ham->row_ptr.p, cnt.p, keys.p);
        EDX_HIP(hipGetLastError());
        launch_sort(st, D, max_row, (const int64_t*)ham->row_ptr.p, keys.p);
        ham->col.alloc(std::max<int64_t>(E, 1));
        ham->term.alloc(std::max<int64_t>(E, 1));
        ham->val.alloc(std::max<int64_t>(E, 1));
        ham->n_entries = E;
        // the sliced copy is written in the same pass, except for twisted
        // tables (their copy needs the finished table's transposed entries)
        SellOut so{1, 0, nullptr, nullptr, nullptr, nullptr};
        if (!ham->herm_part) {
          plan = edx::plan_sell(*ham, *ctx);
          edx::prepare_sell(*ham, *ctx, plan.lanes);
          sell_written = true;
          so = SellOut{plan.lanes, ham->sell_slices * (64 / plan.lanes), (const int64_t*)ham->sell_off.p,
                       ham->sell_col.p, ham->sell_val.p, ham->sell_val32.p};
        }
        const int64_t value_rows = so.col ? std::max<int64_t>(so.rows, D) : D;
        hipLaunchKernelGGL(ptab_values, dim3(edx::grid_for(value_rows * kItemLanes, kBlock)), dim3(kBlock), 0, st,
                           tv, sv,
                           (const double2*)d_chi.p, (const int64_t*)ham->row_ptr.p, (const uint64_t*)keys.p,
                           ham->col.p, ham->term.p, ham->val.p, so);
        EDX_HIP(hipGetLastError());
        hipLaunchKernelGGL(ptab_diag, dim3(edx::grid_for(D, kBlock)), dim3(kBlock), 0, st, D,
                           (const uint64_t*)sec->reps.p, n_density, (const uint8_t*)d_vi.p,
                           (const uint8_t*)d_vj.p, (const double2*)d_v.p, ham->hdiag.p, ham->hdiag32.p);
        EDX_HIP(hipGetLastError());
        EDX_HIP(hipStreamSynchronize(st));
      } else {
        const int64_t zero = 0;
        EDX_HIP(hipMemcpyAsync(ham->row_ptr.p, &zero, 8, hipMemcpyHostToDevice, st));
        EDX_HIP(hipStreamSynchronize(st));
      }
      ham->n_entries = E;
      if (sell_written) {
        edx::pack_sell(*ham, *ctx);
        edx::finish_sell(*ham, *ctx, plan);
      }
      else edx::build_sell_for(*ham, *ctx);
    } catch (...) {
      edx::abandon_sell(*ham, plan);
      delete ham;
      throw;
    }
    *out = ham;
    if (n_entries) *n_entries = ham->n_entries;
  });
}

}  // namespace

extern "C" {

int edx_hamiltonian_build(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec, const double* chi,
                          int32_t n_bilinear, const int32_t* i_from, const int32_t* i_to, const double* t,
                          int32_t n_density, const int32_t* v_i, const int32_t* v_j, const double* v,
                          edx_hamiltonian** out, int64_t* n_entries) {
  return build_hamiltonian(ctx, cat, sec, chi, n_bilinear, i_from, i_to, t, n_density, v_i, v_j, v, out,
                           n_entries, 0);
}

int edx_hamiltonian_build_on_the_fly(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec,
                                     const double* chi, int32_t n_bilinear, const int32_t* i_from,
                                     const int32_t* i_to, const double* t, int32_t n_density, const int32_t* v_i,
                                     const int32_t* v_j, const double* v, edx_hamiltonian** out,
                                     int64_t* n_entries) {
  return build_hamiltonian(ctx, cat, sec, chi, n_bilinear, i_from, i_to, t, n_density, v_i, v_j, v, out,
                           n_entries, 1);
}

int edx_hamiltonian_fetch_table(edx_ctx* ctx, const edx_hamiltonian* ham, int64_t* row, int64_t* col,
                                int32_t* term, double* vals, double* h_diag) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(!ham->on_the_fly, EDX_ERR_INVALID, "an on-the-fly Hamiltonian stores no table");
    edx::DeviceGuard dg(ctx->device);
    const int64_t E = ham->n_entries, D = ham->dim;
    std::vector<int64_t> rp(D + 1);
    std::vector<int32_t> c32(E);
    EDX_HIP(hipMemcpyAsync(rp.data(), ham->row_ptr.p, (D + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (E) EDX_HIP(hipMemcpyAsync(c32.data(), ham->col.p, E * 4, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<uint16_t> t16(term ? E : 0);
    if (term && E) {
      EDX_REQUIRE(ham->term.n >= (size_t)E, EDX_ERR_INVALID, "this Hamiltonian carries no term indices");
      EDX_HIP(hipMemcpyAsync(t16.data(), ham->term.p, E * 2, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (vals && E) EDX_HIP(hipMemcpyAsync(vals, ham->val.p, E * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (h_diag && D) EDX_HIP(hipMemcpyAsync(h_diag, ham->hdiag.p, D * 16, hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    if (row)
      for (int64_t r = 0; r < D; ++r)
        for (int64_t e = rp[r]; e < rp[r + 1]; ++e) row[e] = r;
    if (col)
      for (int64_t e = 0; e < E; ++e) col[e] = c32[e];
    if (term)
      for (int64_t e = 0; e < E; ++e) term[e] = t16[e];
  });
}

int edx_hamiltonian_fetch_csr(edx_ctx* ctx, const edx_hamiltonian* ham, int64_t row_begin, int64_t row_end,
                              int64_t* row_ptr, int32_t* col, double* vals, double* h_diag) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham && row_ptr, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(!ham->on_the_fly, EDX_ERR_INVALID, "an on-the-fly Hamiltonian stores no table");
    EDX_REQUIRE(0 <= row_begin && row_begin <= row_end && row_end <= ham->dim, EDX_ERR_INVALID,
                "row range outside the sector");
    edx::DeviceGuard dg(ctx->device);
    const int64_t n = row_end - row_begin;
    EDX_HIP(hipMemcpyAsync(row_ptr, ham->row_ptr.p + row_begin, (n + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t e0 = row_ptr[0], e1 = row_ptr[n];
    for (int64_t r = 0; r <= n; ++r) row_ptr[r] -= e0;
    if (col && e1 > e0)
      EDX_HIP(hipMemcpyAsync(col, ham->col.p + e0, (e1 - e0) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (vals && e1 > e0)
      EDX_HIP(hipMemcpyAsync(vals, ham->val.p + e0, (e1 - e0) * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (h_diag && n)
      EDX_HIP(hipMemcpyAsync(h_diag, ham->hdiag.p + row_begin, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

namespace {
__global__ void to_c64(int64_t n, const double2* a, float2* b) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
}  // namespace

int edx_hamiltonian_from_csr(edx_ctx* ctx, int64_t dim, int64_t nnz, const int64_t* row_ptr, const int32_t* col,
                             const double* vals, const double* h_diag, edx_hamiltonian** out) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && row_ptr && h_diag && out && (nnz == 0 || (col && vals)), EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(dim >= 0 && nnz >= 0 && row_ptr[0] == 0 && row_ptr[dim] == nnz, EDX_ERR_INVALID,
                "row_ptr must start at 0 and end at nnz");
    for (int64_t r = 0; r < dim; ++r)
      EDX_REQUIRE(row_ptr[r] <= row_ptr[r + 1], EDX_ERR_INVALID, "row_ptr must be non-decreasing");
    for (int64_t e = 0; e < nnz; ++e)
      EDX_REQUIRE(col[e] >= 0 && col[e] < dim, EDX_ERR_INVALID, "column index out of range");
    *out = nullptr;
    edx::DeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    auto ham = new edx_hamiltonian();
    try {
      ham->device = ctx->device;
      ham->dim = dim;
      ham->n_entries = nnz;
      ham->row_ptr.alloc(dim + 1);
      ham->hdiag.alloc(dim);
      ham->hdiag32.alloc(dim);
      ham->col.alloc(std::max<int64_t>(nnz, 1));
      ham->val.alloc(std::max<int64_t>(nnz, 1));
      EDX_HIP(hipMemcpyAsync(ham->row_ptr.p, row_ptr, (dim + 1) * 8, hipMemcpyHostToDevice, st));
      if (dim) EDX_HIP(hipMemcpyAsync(ham->hdiag.p, h_diag, dim * 16, hipMemcpyHostToDevice, st));
      if (nnz) {
        EDX_HIP(hipMemcpyAsync(ham->col.p, col, nnz * 4, hipMemcpyHostToDevice, st));
        EDX_HIP(hipMemcpyAsync(ham->val.p, vals, nnz * 16, hipMemcpyHostToDevice, st));
      }
      if (dim) {
        hipLaunchKernelGGL(to_c64, dim3(edx::grid_for(dim, 256)), dim3(256), 0, st, dim, (const double2*)ham->hdiag.p,
                           ham->hdiag32.p);
        EDX_HIP(hipGetLastError());
      }
      EDX_HIP(hipStreamSynchronize(st));
      edx::build_sell_for(*ham, *ctx);
    } catch (...) {
      delete ham;
      throw;
    }
    *out = ham;
  });
}

int edx_matvec(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t n_vec, const double* x, double* y) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham && x && y, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(n_vec >= 0, EDX_ERR_INVALID, "negative vector count");
    const int64_t D = ham->dim;
    if (D == 0 || n_vec == 0) return;
    edx::DeviceGuard dg(ctx->device);
    hipStream_t st = ctx->stream;
    for (int32_t v0 = 0; v0 < n_vec; v0 += 16) {
      const int nb = std::min<int32_t>(16, n_vec - v0);
      DevBuf<double2> xc((size_t)nb * D), xi((size_t)nb * D), yc((size_t)nb * D);
      EDX_HIP(hipMemcpyAsync(xc.p, x + (size_t)v0 * D * 2, (size_t)nb * D * 16, hipMemcpyHostToDevice, st));
      edx::interleave(st, EDX_PREC_FP64, D, nb, xc.p, D, xi.p);
      edx::apply_table(*ham, *ctx, EDX_PREC_FP64, nb, xi.p, yc.p, D);
      EDX_HIP(hipMemcpyAsync(y + (size_t)v0 * D * 2, yc.p, (size_t)nb * D * 16, hipMemcpyDeviceToHost, st));
      EDX_HIP(hipStreamSynchronize(st));
    }
  });
}

int edx_matvec_device(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t precision, int32_t n_vec,
                      const void* x_dev, void* y_dev) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham && x_dev && y_dev, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(precision == EDX_PREC_FP64 || precision == EDX_PREC_MIXED, EDX_ERR_INVALID, "bad precision");
    edx::DeviceGuard dg(ctx->device);
    edx::apply_table(*ham, *ctx, precision, n_vec, x_dev, y_dev, ham->dim);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"

namespace {

// Checked job list of a grouped product: Y_i = H_i X_i on device buffers
// (interleaved X [dim][n_vec], column-major Y [n_vec][dim]).
std::vector<edx::SpmmJob> group_jobs(edx_ctx* ctx, int32_t n, const edx_hamiltonian* const* hams, int32_t precision,
                                     int32_t n_vec, const void* const* x_dev, void* const* y_dev) {
  EDX_REQUIRE(ctx && hams && x_dev && y_dev && n >= 1, EDX_ERR_INVALID, "null argument or n < 1");
  EDX_REQUIRE(precision == EDX_PREC_FP64 || precision == EDX_PREC_MIXED, EDX_ERR_INVALID, "bad precision");
  EDX_REQUIRE(n_vec >= 1 && n_vec <= 16, EDX_ERR_INVALID, "1 to 16 vectors per product");
  std::vector<edx::SpmmJob> jobs(n);
  for (int i = 0; i < n; ++i) {
    EDX_REQUIRE(hams[i] && x_dev[i] && y_dev[i], EDX_ERR_INVALID, "null Hamiltonian or buffer");
    EDX_REQUIRE(hams[i]->device == ctx->device, EDX_ERR_INVALID, "Hamiltonian on another device");
    EDX_REQUIRE(hams[i]->on_the_fly || hams[i]->dim == 0 || hams[i]->sell_slices > 0, EDX_ERR_INVALID,
                "Hamiltonian without a sliced table");
    EDX_REQUIRE((int64_t)(hams[i]->sell_slices + 3) / 4 < (int64_t(1) << 31), EDX_ERR_INVALID, "table too large");
    jobs[i] = edx::SpmmJob{hams[i], n_vec, x_dev[i], y_dev[i], hams[i]->dim};
  }
  return jobs;
}

}  // namespace

extern "C" {

int edx_matvec_group(edx_ctx* ctx, int32_t n, const edx_hamiltonian* const* hams, int32_t precision, int32_t n_vec,
                     const void* const* x_dev, void* const* y_dev) {
  return edx::guarded(ctx, [&] {
    const std::vector<edx::SpmmJob> jobs = group_jobs(ctx, n, hams, precision, n_vec, x_dev, y_dev);
    edx::DeviceGuard dg(ctx->device);
    edx::spmm_group(jobs.data(), n, *ctx, precision);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_matvec_group_graph(edx_ctx* ctx, int32_t n, const edx_hamiltonian* const* hams, int32_t precision,
                           int32_t n_vec, const void* const* x_dev, void* const* y_dev, edx_graph** out) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(out, EDX_ERR_INVALID, "null argument");
    *out = nullptr;
    const std::vector<edx::SpmmJob> jobs = group_jobs(ctx, n, hams, precision, n_vec, x_dev, y_dev);
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    auto g = new edx_graph();
    g->device = ctx->device;
    // the launches only: the kernel timer's events stay out of the graph
    const bool timer_on = ctx->timer.on;
    ctx->timer.on = false;
    hipError_t cap = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    try {
      EDX_HIP(cap);
      edx::spmm_group(jobs.data(), n, *ctx, precision);
      EDX_HIP(hipStreamEndCapture(ctx->stream, &g->graph));
      cap = hipErrorUnknown;  // capture ended
      EDX_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      ctx->timer.on = timer_on;
      if (cap == hipSuccess) {  // still capturing: end it so the stream is usable again
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &partial);
        if (partial) (void)hipGraphDestroy(partial);
      }
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      delete g;
      throw;
    }
    ctx->timer.on = timer_on;
    *out = g;
  });
}

int edx_graph_launch(edx_ctx* ctx, const edx_graph* g) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && g && g->exec, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(g->device == ctx->device, EDX_ERR_INVALID, "graph captured on another device");
    edx::DeviceGuard dg(ctx->device);
    EDX_HIP(hipGraphLaunch(g->exec, ctx->stream));
  });
}

int edx_graph_destroy(edx_graph* g) {
  if (!g) return EDX_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(g->device);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  (void)hipSetDevice(prev);
  delete g;
  return EDX_OK;
}

int edx_time_matvec(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t precision, int32_t n_vec, int32_t n_rep,
                    const void* x_dev, void* y_dev, double* avg_ms) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham && x_dev && y_dev && avg_ms && n_rep > 0, EDX_ERR_INVALID, "bad argument");
    edx::DeviceGuard dg(ctx->device);
    hipEvent_t a, b;
    EDX_HIP(hipEventCreate(&a));
    EDX_HIP(hipEventCreate(&b));
    EDX_HIP(hipEventRecord(a, ctx->stream));
    for (int i = 0; i < n_rep; ++i) edx::spmm(*ham, *ctx, precision, n_vec, x_dev, y_dev, ham->dim);
    EDX_HIP(hipEventRecord(b, ctx->stream));
    EDX_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    EDX_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = (double)ms / n_rep;
  });
}

int edx_hamiltonian_destroy(edx_hamiltonian* ham) {
  if (!ham) return EDX_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(ham->device);
  delete ham;
  (void)hipSetDevice(prev);
  return EDX_OK;
}

}  // extern "C"
