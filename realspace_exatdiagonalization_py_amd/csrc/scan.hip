// Deterministic device-wide exclusive prefix sums (three-phase tile scan).
//
// Phase 1: each 256-thread block reduces a tile of 256*ITEMS inputs.
// Phase 2: one block scans the tile sums in order.
// Phase 3: each block rescans its tile and adds the tile offset.
// The summation order is fixed by the tile layout, so results never depend
// on scheduling.
#include "edx_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;

struct LoadU32 {
  const uint32_t* p;
  __device__ uint64_t operator()(size_t i) const { return p[i]; }
};
struct LoadPopc {
  const uint64_t* p;
  __device__ uint64_t operator()(size_t i) const { return (uint64_t)__builtin_popcountll(p[i]); }
};

// Inclusive scan of one value per thread across the block; returns the
// inclusive prefix and writes the block total to *total.
__device__ uint64_t block_inclusive_scan(uint64_t v, uint64_t* lds_waves, uint64_t* total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int off = 1; off < 64; off <<= 1) {
    uint64_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) lds_waves[wave] = v;
  __syncthreads();
  uint64_t base = 0, tot = 0;
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) base += lds_waves[w];
    tot += lds_waves[w];
  }
  __syncthreads();
  *total = tot;
  return v + base;
}

template <class Load>
__global__ void __launch_bounds__(kBlock) tile_reduce(Load ld, size_t n, uint64_t* tile_sums) {
  __shared__ uint64_t w[kBlock / 64];
  size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k)
    if (base + k < n) s += ld(base + k);
  uint64_t tot;
  block_inclusive_scan(s, w, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock) scan_tile_sums(uint64_t* sums, size_t n_tiles, uint64_t* grand) {
  __shared__ uint64_t w[kBlock / 64];
  uint64_t carry = 0;
  for (size_t start = 0; start < n_tiles; start += kBlock) {
    size_t i = start + threadIdx.x;
    uint64_t v = i < n_tiles ? sums[i] : 0;
    uint64_t tot;
    uint64_t inc = block_inclusive_scan(v, w, &tot);
    if (i < n_tiles) sums[i] = carry + inc - v;
    carry += tot;
  }
  if (threadIdx.x == 0) *grand = carry;
}

template <class Load, class Out>
__global__ void __launch_bounds__(kBlock) tile_scan(Load ld, size_t n, const uint64_t* tile_off, Out* out) {
  __shared__ uint64_t w[kBlock / 64];
  size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;
  uint64_t vals[kItems];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    vals[k] = (base + k < n) ? ld(base + k) : 0;
    s += vals[k];
  }
  uint64_t tot;
  uint64_t inc = block_inclusive_scan(s, w, &tot);
  uint64_t run = tile_off[blockIdx.x] + inc - s;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (base + k < n) out[base + k] = (Out)run;
    run += vals[k];
  }
}

template <class Load, class Out>
uint64_t run_scan(edx::Ctx& ctx, Load ld, size_t n, Out* out) {
  if (n == 0) return 0;
  size_t n_tiles = (n + kTile - 1) / kTile;
  edx::DevBuf<uint64_t> sums(n_tiles + 1);
  hipLaunchKernelGGL(tile_reduce<Load>, dim3((unsigned)n_tiles), dim3(kBlock), 0, ctx.stream, ld, n, sums.p);
  EDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(scan_tile_sums, dim3(1), dim3(kBlock), 0, ctx.stream, sums.p, n_tiles, sums.p + n_tiles);
  EDX_HIP(hipGetLastError());
  hipLaunchKernelGGL((tile_scan<Load, Out>), dim3((unsigned)n_tiles), dim3(kBlock), 0, ctx.stream, ld, n,
                     (const uint64_t*)sums.p, out);
  EDX_HIP(hipGetLastError());
  uint64_t total = 0;
  EDX_HIP(hipMemcpyAsync(&total, sums.p + n_tiles, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx.stream));
  EDX_HIP(hipStreamSynchronize(ctx.stream));
  return total;
}

}  // namespace

namespace edx {

uint64_t exclusive_scan_u32(Ctx& ctx, const uint32_t* in, uint32_t* out, size_t n) {
  return run_scan(ctx, LoadU32{in}, n, out);
}

uint64_t exclusive_scan_u32_to_i64(Ctx& ctx, const uint32_t* in, int64_t* out, size_t n) {
  return run_scan(ctx, LoadU32{in}, n, out);
}

uint64_t exclusive_scan_popcount64(Ctx& ctx, const uint64_t* words, uint32_t* out, size_t n) {
  return run_scan(ctx, LoadPopc{words}, n, out);
}

}  // namespace edx
