// Palette packing of the sliced (SELL-64) table copy.
//
// A sector table holds few distinct elements: each is a hopping amplitude
// times a Jordan-Wigner or permutation sign, a character of the irrep and a
// stabilizer-norm ratio (SRE:595-693), so a checkerboard, Hubbard or Haldane
// sector has tens to a few hundred distinct complex128 values among its 10^6
// - 10^9 entries.  The packed copy stores one 32-bit word per slot,
//
//     word = column | (palette index << pack_bits),   0xFFFFFFFF = padding,
//
// with pack_bits = bit length of the dimension (so an all-ones column field is
// never a valid column), and the distinct values once, in complex128 (pal)
// and rounded to complex64 (pal32).  The SpMM streams 4 bytes per entry
// instead of 4 + 16 (complex128) or 4 + 8 (complex64) and reads the element
// from the palette held in LDS; the elements it multiplies carry the same bits
// as the unpacked copy's, so both layouts give identical results.
//
// A table with more distinct values than the index bits left above the
// column allow (the Hermitian parts of twisted blocks: their complex64-rounded
// alpha give hundreds to a few thousand values) gets the wide layout instead:
// the int32 columns stay and a 16-bit palette index per slot replaces the
// elements (6 bytes per entry), for at most kPalMaxWide values (palettes
// beyond kPalLdsBytes are read through the caches instead of LDS).
//
// Build: a hash pass collects the distinct values (64-bit fingerprint of the
// value's bits in an open-addressing table of kSlots slots, one compare-and-
// swap per new value), the host numbers the occupied slots, and an encode
// pass writes the words or indices after checking, per entry, that the slot's
// value has exactly the entry's complex128 and complex64 bits.  A table whose
// values do not fit (more than kPalMaxWide distinct values) or whose check
// fails keeps the unpacked copy.
#include "objects.h"

namespace edx {

namespace {

constexpr int kSlots = 65536;  // hash slots (power of two, >= 4 kPalMaxWide)
// Longest probe sequence: at most kPalMaxWide keys in kSlots slots (load <=
// 1/4) keep linear-probing runs far shorter; a longer one ends the packing.
constexpr int kMaxProbe = 64;

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

__device__ __forceinline__ unsigned long long value_key(double2 v) {
  const uint64_t a = (uint64_t)__double_as_longlong(v.x);
  const uint64_t b = (uint64_t)__double_as_longlong(v.y);
  const uint64_t k = mix64(a ^ mix64(b + 0x9e3779b97f4a7c15ull));
  return k ? k : 1ull;  // 0 marks an empty slot
}

__device__ __forceinline__ bool same_bits(double2 a, double2 b) {
  return __double_as_longlong(a.x) == __double_as_longlong(b.x) &&
         __double_as_longlong(a.y) == __double_as_longlong(b.y);
}

// Inserts the value of every stored slot (column >= 0) into the hash table.
// flag[0] bit 0: packing is impossible (more than kPalMaxWide distinct
// values, or a probe run longer than kMaxProbe); every thread then stops.
// flag[1] counts the distinct values inserted.
__global__ void __launch_bounds__(256)
pal_insert(int64_t n, const int32_t* __restrict__ col, const double2* __restrict__ val, unsigned long long* keys,
           double2* rep, int* flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (__atomic_load_n(flag, __ATOMIC_RELAXED)) return;
    if (col[i] < 0) continue;
    const double2 v = val[i];
    const unsigned long long k = value_key(v);
    unsigned s = (unsigned)k & (kSlots - 1);
    int probe = 0;
    for (; probe < kMaxProbe; ++probe) {
      // a plain read settles almost every entry; a stale empty slot only costs
      // the compare-and-swap, which returns the key that is there
      unsigned long long cur = __atomic_load_n(&keys[s], __ATOMIC_RELAXED);
      if (cur == k) break;
      if (cur == 0ull) {
        cur = atomicCAS(&keys[s], 0ull, k);
        if (cur == 0ull) {
          rep[s] = v;
          if (atomicAdd(flag + 1, 1) >= kPalMaxWide) atomicOr(flag, 1);
          break;
        }
        if (cur == k) break;
      }
      s = (s + 1) & (kSlots - 1);
    }
    if (probe == kMaxProbe) atomicOr(flag, 1);
  }
}

// Palette in both precisions from the numbered slots (slot_of[i] = slot of
// palette entry i); the complex64 rounding is the table build's.
__global__ void pal_gather(int n_pal, const int* slot_of, const double2* rep, double2* pal, float2* pal32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pal) return;
  const double2 v = rep[slot_of[i]];
  pal[i] = v;
  pal32[i] = make_float2((float)v.x, (float)v.y);
}

// Packed words (words != null: column | index << bits, padding all ones) or
// 16-bit indices (idx != null, padding 0); flag bit 1: an entry whose slot is
// missing or holds other bits (a fingerprint collision), bit 2: a complex64
// element that is not the rounding of its complex128 one.
__global__ void __launch_bounds__(256)
pal_encode(int64_t n, const int32_t* __restrict__ col, const double2* __restrict__ val,
           const float2* __restrict__ val32, const unsigned long long* __restrict__ keys,
           const double2* __restrict__ rep, const int* __restrict__ index_of_slot, int bits,
           uint32_t* __restrict__ words, uint16_t* __restrict__ idx, int* flag) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t c = col[i];
    if (c < 0) {
      if (words) words[i] = 0xFFFFFFFFu;
      else idx[i] = 0;
      continue;
    }
    const double2 v = val[i];
    const unsigned long long k = value_key(v);
    unsigned s = (unsigned)k & (kSlots - 1);
    int probe = 0;
    while (probe < kMaxProbe && keys[s] != k) {
      s = (s + 1) & (kSlots - 1);
      ++probe;
    }
    if (probe == kMaxProbe || !same_bits(rep[s], v)) {
      atomicOr(flag, 2);
      continue;
    }
    const float2 w = val32[i];
    const float2 r = make_float2((float)v.x, (float)v.y);
    if (__float_as_uint(w.x) != __float_as_uint(r.x) || __float_as_uint(w.y) != __float_as_uint(r.y))
      atomicOr(flag, 4);
    if (words) words[i] = (uint32_t)c | ((uint32_t)index_of_slot[s] << bits);
    else idx[i] = (uint16_t)index_of_slot[s];
  }
}

unsigned grid_stride_blocks(int64_t n) {
  const int64_t want = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, 256 * 64));
}

}  // namespace

int pack_bits_for(int64_t dim) {
  int b = 0;
  while (b < 63 && (int64_t(1) << b) <= dim) ++b;
  return std::max(b, 1);
}

// Replaces h's unpacked sliced copy by a packed one when its values allow:
// the 32-bit words when the palette fits the bits above the column, else the
// wide layout; returns true when packed.  Called once the copy is complete.
bool pack_sell(edx_hamiltonian& h, Ctx& ctx) {
  if (!ctx.table_pack || h.sell_packed || h.sell_entries == 0 || h.on_the_fly) return false;
  const int bits = pack_bits_for(h.dim);
  const int cap = bits < 31 ? std::min(kPalMax, 1 << (32 - bits)) : 0;
  hipStream_t st = ctx.stream;
  const int64_t n = h.sell_entries;
  DevBuf<unsigned long long> keys(kSlots);
  DevBuf<double2> rep(kSlots);
  DevBuf<int> flag(2);
  EDX_HIP(hipMemsetAsync(keys.p, 0, keys.bytes(), st));
  EDX_HIP(hipMemsetAsync(flag.p, 0, flag.bytes(), st));
  hipLaunchKernelGGL(pal_insert, dim3(grid_stride_blocks(n)), dim3(256), 0, st, n, (const int32_t*)h.sell_col.p,
                     (const double2*)h.sell_val.p, keys.p, rep.p, flag.p);
  EDX_HIP(hipGetLastError());
  std::vector<unsigned long long> hk(kSlots);
  int hflag = 0;
  EDX_HIP(hipMemcpyAsync(hk.data(), keys.p, keys.bytes(), hipMemcpyDeviceToHost, st));
  EDX_HIP(hipMemcpyAsync(&hflag, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  EDX_HIP(hipStreamSynchronize(st));
  if (hflag) return false;
  std::vector<int> index_of_slot(kSlots, -1), slot_of;
  for (int s = 0; s < kSlots; ++s)
    if (hk[s]) {
      index_of_slot[s] = (int)slot_of.size();
      slot_of.push_back(s);
    }
  const int n_pal = (int)slot_of.size();
  if (n_pal == 0 || n_pal > kPalMaxWide) return false;
  const bool wide = n_pal > cap;
  DevBuf<int> d_index(kSlots), d_slot(n_pal);
  ObjBuf<double2> pal(n_pal);
  ObjBuf<float2> pal32(n_pal);
  ObjBuf<int32_t> words(wide ? 0 : n);
  ObjBuf<uint16_t> idx(wide ? n : 0);
  EDX_HIP(hipMemcpyAsync(d_index.p, index_of_slot.data(), d_index.bytes(), hipMemcpyHostToDevice, st));
  EDX_HIP(hipMemcpyAsync(d_slot.p, slot_of.data(), d_slot.bytes(), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(pal_gather, dim3((n_pal + 255) / 256), dim3(256), 0, st, n_pal, (const int*)d_slot.p,
                     (const double2*)rep.p, pal.p, pal32.p);
  EDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(pal_encode, dim3(grid_stride_blocks(n)), dim3(256), 0, st, n, (const int32_t*)h.sell_col.p,
                     (const double2*)h.sell_val.p, (const float2*)h.sell_val32.p,
                     (const unsigned long long*)keys.p, (const double2*)rep.p, (const int*)d_index.p, bits,
                     wide ? nullptr : (uint32_t*)words.p, wide ? idx.p : nullptr, flag.p);
  EDX_HIP(hipGetLastError());
  EDX_HIP(hipMemcpyAsync(&hflag, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  EDX_HIP(hipStreamSynchronize(st));
  if (hflag) return false;
  if (wide) {
    h.sell_idx = std::move(idx);
  } else {
    h.sell_col = std::move(words);
    h.pack_bits = bits;
  }
  h.sell_val.release();
  h.sell_val32.release();
  h.pal = std::move(pal);
  h.pal32 = std::move(pal32);
  h.n_pal = n_pal;
  h.sell_packed = wide ? 2 : 1;
  return true;
}

}  // namespace edx
