// Sector eigensolver: thick-restarted block Lanczos with full (two-pass
// classical Gram-Schmidt) reorthogonalisation.
//
// Reference: diagonalize_block_matrixfree (SRE:2248-2299) forms the dense
// block for n <= 500 and calls LAPACK eigh on its upper triangle, and for
// larger blocks runs ARPACK eigs(which="SR", tol=1e-10) on the CSR operator,
// returning the real parts in ascending order.  With twisted boundaries the
// blocks are Hermitian only up to the complex64 rounding of the stored alpha;
// the sliced operator this solver applies then holds the Hermitian part
// (H + H^H) / 2, whose eigenvalues are ARPACK's real parts to second order in
// that rounding, except inside multiplets the rounding splits, where the
// real parts themselves are fixed only to the rounding (~1e-8).
//
// Krylov state: basis Q (nq vectors whose H-images are known), residual block
// Vr (b vectors) and coupling B with H Q = Q M + Vr B.  Each step applies H to
// Vr (one SpMM over b interleaved vectors), projects the result against
// [Q, Vr] twice, extends M, and orthonormalises the remainder (Cholesky QR,
// twice) into the next Vr.  Ritz residuals are ||B y||.  When the basis is
// full, the k lowest Ritz vectors are kept (thick restart).  The block size
// bounds the multiplicity of eigenvalues that are resolved.
//
// EDX_PREC_MIXED runs the sweeps on complex64 vectors and matrix elements with
// FP64 projections, small-matrix algebra and Ritz problems, then restarts in
// FP64 from the complex64 Ritz vectors to reach the requested tolerance.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <thread>
#include <type_traits>

#include "dense.h"
#include "objects.h"

using edx::DevBuf;
using edx::dense::cplx;
using edx::dense::Mat;

namespace edx {
void spmm(const edx_hamiltonian& h, Ctx& ctx, int precision, int nb, const void* X, void* Y, int64_t ldy);
void interleave(hipStream_t st, int precision, int64_t dim, int nb, const void* src, int64_t ld, void* dst);
void dense_matrix(const edx_hamiltonian& h, Ctx& ctx, double2* out);
}  // namespace edx

namespace {

constexpr int kMaxVec = edx::kMaxBasis;  // basis vectors the Gram/update kernels accept
constexpr int kMaxW = edx::kMaxBlock;    // block width
// P = [Q, W]^H W of a block step is staged in LDS by step_coeffs: the basis
// is capped so that (nbasis + b) * b <= kStepLds (48 KiB of complex128)
constexpr int kStepLds = 3072;
// Gram of complex64 vectors against at most 8 on the FP32 matrix cores from
// this many rows (below, its grid is too small to fill the device).
constexpr int64_t kMfmaGramRows = int64_t(1) << 20;
// Rows per lane of the update kernel (each coefficient read serves them all).
constexpr int kUpdateRows = 2;
// Block-update schedules (blocks of at most 8 vectors): vectors per prefetch
// stage and rows per lane.  Every variant sums each output in the same order
// (input vectors ascending, one lane per row), so they give identical
// results.  Defaults: 2/2 for complex64, 2/1 for complex128 (two rows per
// lane hold too many FP64 sums: 2/1 was 1.0-1.6x faster at D = 2^17 and
// 2^23, blocks 6-8); the tuner may pick another.
struct UpdateVariant {
  int stage, rows;
};
constexpr UpdateVariant kUpdateVariants[3] = {{2, 2}, {2, 1}, {4, 1}};
// Gram schedules: vector FMAs, or the FP32 matrix cores (complex64 Grams
// against at most 8 vectors).
constexpr int kGramVector = 1, kGramMatrix = 2;
// Sectors from this many rows have their schedule tuned (smaller ones are
// bound by launch latency, where timings say little).
constexpr int64_t kScheduleTuneRows = 65536;
// A tuned schedule must beat the default by this fraction to replace it.
constexpr float kScheduleMargin = 0.03f;
// Largest LDS staging of an update's interleaved output copy (bytes).
constexpr size_t kStageLdsBytes = 48 * 1024;
constexpr size_t kLogC = edx::kLogC;

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <typename V> __device__ __forceinline__ V from_d2(double re, double im);
template <> __device__ __forceinline__ double2 from_d2<double2>(double re, double im) { return make_double2(re, im); }
template <> __device__ __forceinline__ float2 from_d2<float2>(double re, double im) {
  return make_float2((float)re, (float)im);
}

// Column vectors taken from two arrays: vector i is A0 + i*ld0 for i < n0 and
// A1 + (i - n0)*ld1 otherwise.  Lets one launch cover [Q, W] without a copy.
template <typename V>
struct VecPair {
  const V* a0;
  int64_t ld0;
  int n0;
  const V* a1;
  int64_t ld1;
  int n1;
  __device__ __forceinline__ const V* vec(int i) const {
    return i < n0 ? a0 + (int64_t)i * ld0 : a1 + (int64_t)(i - n0) * ld1;
  }
};

// Launch predicate of the kernels of a device block step.  ctl[0] is the
// step's status (0: one Cholesky QR pass, 1: a second pass follows on the
// device, 2: the host takes over the step), ctl[1] is set once a step has
// been handed to the host: every later predicated kernel then returns at once
// until the host clears it.  mode 0 always runs; mode 1 runs unless halted;
// mode 2 (second Cholesky QR pass) runs when not halted and ctl[0] == 1.
struct StepPred {
  const int* ctl = nullptr;
  int mode = 0;
  __device__ __forceinline__ bool run() const {
    if (!ctl || mode == 0) return true;
    if (ctl[1]) return false;
    return mode == 1 || ctl[0] == 1;
  }
  // the first update writes the interleaved copy only when it is final
  __device__ __forceinline__ bool final_update() const { return !ctl || mode != 1 || ctl[0] == 0; }
};

// Sum of NV per-lane values over the 64 lanes of a wavefront by halving: at
// each lane distance O (32, 16, ..., 1) a lane pair splits its current set of
// values in two, each lane sends the half it drops and adds the partner's copy
// of the half it keeps (N values take N/2 shuffles); a set of odd size is
// summed whole at that distance instead.  Afterwards a lane holds NF sums, for
// value indices base .. base + NF; lanes that differ only in the bits of `dup`
// hold the same sums.  The order of additions is fixed, so results do not
// depend on scheduling.
template <int N, int O>
struct WaveSplitSum {
  static constexpr int NF = (N % 2 == 0) ? WaveSplitSum<N / 2, O / 2>::NF : WaveSplitSum<N, O / 2>::NF;
  __device__ __forceinline__ static void run(double* v, int lane, int& base, int& dup) {
    if constexpr (N % 2 == 0) {
      constexpr int H = N / 2;
      const bool up = (lane & O) != 0;
#pragma unroll
      for (int i = 0; i < H; ++i) {
        const double send = up ? v[i] : v[H + i];
        const double keep = up ? v[H + i] : v[i];
        v[i] = keep + __shfl_xor(send, O, 64);
      }
      if (up) base += H;
      WaveSplitSum<H, O / 2>::run(v, lane, base, dup);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] += __shfl_xor(v[i], O, 64);
      dup |= O;
      WaveSplitSum<N, O / 2>::run(v, lane, base, dup);
    }
  }
};
template <int N>
struct WaveSplitSum<N, 0> {
  static constexpr int NF = N;
  __device__ __forceinline__ static void run(double*, int, int&, int&) {}
};

// Keeps the loads issued just before it in place: no code motion moves them
// past this point (towards their uses), and the scheduler does not start the
// following arithmetic (which waits on earlier loads) before them.  A loop
// that loads the next stage and then computes the current one thereby keeps
// the next stage's loads in flight during the arithmetic.
#define EDX_KEEP_PREFETCH()          \
  do {                               \
    asm volatile("" ::: "memory");   \
    __builtin_amdgcn_sched_barrier(0); \
  } while (0)

// partial[rb][i*NW + j] = sum over row block rb of conj(A_i[r]) W_j[r].
// One workgroup per (tile of TI vectors of A, row block); the 1-D grid is
// mapped so that all tiles of a row block run on one XCD (workgroups are
// dispatched to the 8 XCDs round-robin), whose L2 then serves that block of W
// to every tile.  Lanes take consecutive rows (coalesced loads).  The loop
// keeps two register stages: the loads of the next row are in flight while the
// current one is multiplied.  Loads are unconditional: rows past the block
// read row ld - 1 and vector indices past nv read vector nv - 1; every vector
// is zero in rows [D, ld) with ld > D, so the extra rows add nothing, and the
// sums of padded vector slots are never written.  Each thread keeps a TI x NW
// complex tile, summed over the wavefront by WaveSplitSum and over the four
// wavefronts in a fixed order.
template <typename V, int TI, int NW>
__global__ void __launch_bounds__(256)
gram_tiles(int64_t D, int64_t ld, VecPair<V> A, const V* __restrict__ Wb, int64_t ldw, int64_t rows_per_block,
           int n_itiles, int n_tiles, double2* __restrict__ partial, StepPred pred) {
  constexpr int NV = TI * NW * 2;
  __shared__ double red[4][NV];
  if (!pred.run()) return;
  const int per_xcd = (n_tiles + 7) / 8;
  const int logical = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (logical >= n_tiles) return;
  const int itile = logical % n_itiles, rblk = logical / n_itiles;
  const int nv = A.n0 + A.n1;
  const int i0 = itile * TI;
  const int64_t rb = (int64_t)rblk * rows_per_block;
  const int64_t re = min(D, rb + rows_per_block);
  const V* ap[TI];
#pragma unroll
  for (int t = 0; t < TI; ++t) ap[t] = A.vec(min(i0 + t, nv - 1));
  const V* wp[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) wp[j] = Wb + (int64_t)j * ldw;
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  const int64_t nstage = (re - rb + 255) / 256;
  V w0[NW], a0[TI], w1[NW], a1[TI];
#define EDX_GRAM_LOAD(WB, AB, K)                                     \
  {                                                                  \
    int64_t rr = rb + (K) * 256 + (int64_t)threadIdx.x;              \
    rr = rr < re ? rr : ld - 1;                                      \
    _Pragma("unroll") for (int j = 0; j < NW; ++j) WB[j] = wp[j][rr]; \
    _Pragma("unroll") for (int t = 0; t < TI; ++t) AB[t] = ap[t][rr]; \
    EDX_KEEP_PREFETCH();                                             \
  }
#define EDX_GRAM_FMA(WB, AB)                         \
  {                                                  \
    _Pragma("unroll") for (int j = 0; j < NW; ++j) { \
      const double2 w = to_d2(WB[j]);                \
      _Pragma("unroll") for (int t = 0; t < TI; ++t) { \
        const double2 v = to_d2(AB[t]);              \
        double& xr = acc[(t * NW + j) * 2];          \
        double& xi = acc[(t * NW + j) * 2 + 1];      \
        xr = fma(v.x, w.x, xr);                      \
        xr = fma(v.y, w.y, xr);                      \
        xi = fma(v.x, w.y, xi);                      \
        xi = fma(-v.y, w.x, xi);                     \
      }                                              \
    }                                                \
  }
  EDX_GRAM_LOAD(w0, a0, 0)
  for (int64_t k = 0; k < nstage; k += 2) {
    EDX_GRAM_LOAD(w1, a1, k + 1)
    EDX_GRAM_FMA(w0, a0)
    EDX_GRAM_LOAD(w0, a0, k + 2)
    EDX_GRAM_FMA(w1, a1)
  }
#undef EDX_GRAM_LOAD
#undef EDX_GRAM_FMA
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int base = 0, dup = 0;
  WaveSplitSum<NV, 32>::run(acc, lane, base, dup);
  if ((lane & dup) == 0) {
#pragma unroll
    for (int i = 0; i < WaveSplitSum<NV, 32>::NF; ++i) red[wave][base + i] = acc[i];
  }
  __syncthreads();
  const int nout = nv * NW;
  for (int o = threadIdx.x; o < TI * NW; o += 256) {
    const int t = o / NW, j = o % NW;
    if (i0 + t < nv) {
      double xr = 0.0, xi = 0.0;
      for (int wv = 0; wv < 4; ++wv) {
        xr += red[wv][o * 2];
        xi += red[wv][o * 2 + 1];
      }
      partial[(int64_t)rblk * nout + (int64_t)(i0 + t) * NW + j] = make_double2(xr, xi);
    }
  }
}

// out[o] = sum over blocks of partial[blk][o]; one wavefront per output,
// lanes stride the blocks and the sum is finished by a fixed butterfly, so
// the result does not depend on scheduling.
__global__ void __launch_bounds__(256) gram_reduce(int nblk, int nout, const double2* partial, double2* out,
                                                   StepPred pred) {
  if (!pred.run()) return;
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double xr = 0.0, xi = 0.0;
  if (o < nout)
    for (int b = lane; b < nblk; b += 64) {
      const double2 p = partial[(int64_t)b * nout + o];
      xr += p.x;
      xi += p.y;
    }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    xr += __shfl_xor(xr, off, 64);
    xi += __shfl_xor(xi, off, 64);
  }
  if (o < nout && lane == 0) out[o] = make_double2(xr, xi);
}

// Vectors loaded per stage of gemm_tiles (two stages in flight): 2 for the
// block updates (nk <= 16 outputs, one tile; 1-7 % faster than 4 in
// complex64 at D = 2.3e5 and 3.8e6, equal in complex128) and for the wide
// restart combinations (tiles of 12 outputs).
constexpr int kGemmStage = 2;
constexpr int kCombineStage = 2;
constexpr int kCombineTile = 12;

// Out_k[r] = sum_i A_i[r] S[i][k] for k < nk (S row-major, leading dimension
// ldS), FP64 accumulation.  One workgroup per (256 RPT rows, tile of NK
// outputs); the 1-D grid is mapped so that the tiles of a row block run on one
// XCD, whose L2 then serves the block's inputs to all of them.  Each lane owns
// RPT rows (r, r + 256, ...), so each coefficient read from LDS serves RPT
// rows.  The tile's coefficients are staged in LDS (every lane reads the same
// entry, a broadcast), with zero rows up to a multiple of two stages, so the
// input loop needs no bounds test: it loads vector min(i, nv - 1) and the zero
// coefficient cancels it.  Loads of the next U vectors are in flight while the
// current ones are multiplied.  Lanes past D read row D (zero
// padding, never written) and store nothing.  With one tile (nk <= NK) every
// input of row r is read by the lane of row r before it writes any output of
// row r, so Out may be one of the A arrays (same leading dimension); Xi
// (optional, one tile only) receives the interleaved copy [r][k].
template <typename V, int NK, int U, int RPT = 1>
__global__ void __launch_bounds__(256)
gemm_tiles(int64_t D, VecPair<V> A, int nk, V* Ob, int64_t ldo, const double2* __restrict__ S, int ldS, V* Xi,
           int n_kt, int n_tiles, int stage_xi, StepPred pred) {
  extern __shared__ double2 s_tile[];  // [nvp][NK]
  if (!pred.run()) return;
  if (!pred.final_update()) Xi = nullptr;
  const int per_xcd = (n_tiles + 7) / 8;
  const int logical = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (logical >= n_tiles) return;  // whole workgroup
  const int kt = logical % n_kt, rb = logical / n_kt;
  const int k0 = kt * NK;
  const int nv = A.n0 + A.n1;
  const int nvp = (nv + 2 * U - 1) / (2 * U) * (2 * U);
  for (int e = threadIdx.x; e < nvp * NK; e += 256) {
    const int i = e / NK, k = e % NK;
    s_tile[e] = (i < nv && k0 + k < nk) ? S[(int64_t)i * ldS + k0 + k] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  int64_t rr[RPT], rc[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    rr[q] = (int64_t)rb * 256 * RPT + 256 * q + threadIdx.x;
    rc[q] = rr[q] < D ? rr[q] : D;
  }
  double pr[RPT][NK], pi[RPT][NK];
#pragma unroll
  for (int q = 0; q < RPT; ++q)
#pragma unroll
    for (int k = 0; k < NK; ++k) pr[q][k] = pi[q][k] = 0.0;
  V b0[U][RPT], b1[U][RPT];
#define EDX_GEMM_LOAD(B, I0)                                                  \
  {                                                                           \
    _Pragma("unroll") for (int u = 0; u < U; ++u) {                           \
      const V* vp = A.vec(min((I0) + u, nv - 1));                             \
      _Pragma("unroll") for (int q = 0; q < RPT; ++q) B[u][q] = vp[rc[q]];    \
    }                                                                         \
    EDX_KEEP_PREFETCH();                                                      \
  }
#define EDX_GEMM_FMA(B, I0)                                \
  {                                                        \
    _Pragma("unroll") for (int u = 0; u < U; ++u) {        \
      double2 x[RPT];                                      \
      _Pragma("unroll") for (int q = 0; q < RPT; ++q) x[q] = to_d2(B[u][q]); \
      _Pragma("unroll") for (int k = 0; k < NK; ++k) {     \
        const double2 sv = s_tile[((I0) + u) * NK + k];    \
        _Pragma("unroll") for (int q = 0; q < RPT; ++q) {  \
          pr[q][k] = fma(x[q].x, sv.x, pr[q][k]);          \
          pr[q][k] = fma(-x[q].y, sv.y, pr[q][k]);         \
          pi[q][k] = fma(x[q].x, sv.y, pi[q][k]);          \
          pi[q][k] = fma(x[q].y, sv.x, pi[q][k]);          \
        }                                                  \
      }                                                    \
    }                                                      \
  }
  EDX_GEMM_LOAD(b0, 0)
  for (int i0 = 0; i0 < nvp; i0 += 2 * U) {
    EDX_GEMM_LOAD(b1, i0 + U)
    EDX_GEMM_FMA(b0, i0)
    EDX_GEMM_LOAD(b0, i0 + 2 * U)
    EDX_GEMM_FMA(b1, i0 + U)
  }
#undef EDX_GEMM_LOAD
#undef EDX_GEMM_FMA
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    if (rr[q] >= D) continue;
#pragma unroll
    for (int k = 0; k < NK; ++k)
      if (k0 + k < nk) Ob[(int64_t)(k0 + k) * ldo + rr[q]] = from_d2<V>(pr[q][k], pi[q][k]);
  }
  if (!Xi) return;
  // interleaved copy [r][k] for the next SpMM (one tile, k0 = 0): staged in
  // LDS (when the launch gave room for it) so that the rows of the workgroup
  // are written as one contiguous run
  if (!stage_xi) {
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      if (rr[q] >= D) continue;
#pragma unroll
      for (int k = 0; k < NK; ++k)
        if (k < nk) Xi[rr[q] * nk + k] = from_d2<V>(pr[q][k], pi[q][k]);
    }
    return;
  }
  __syncthreads();  // the coefficients are no longer read
  V* stage = reinterpret_cast<V*>(s_tile);
  const int64_t row0 = (int64_t)rb * 256 * RPT;
#pragma unroll
  for (int q = 0; q < RPT; ++q)
#pragma unroll
    for (int k = 0; k < NK; ++k)
      if (k < nk) stage[(q * 256 + (int)threadIdx.x) * nk + k] = from_d2<V>(pr[q][k], pi[q][k]);
  __syncthreads();
  const int64_t n = (min(D, row0 + 256 * RPT) - row0) * nk;
  for (int64_t e = threadIdx.x; e < n; e += 256) Xi[row0 * nk + e] = stage[e];
}

// ---------------------------------------------------------------------------
// Matrix-core Gram.  Operand layout of v_mfma_f32_16x16x4f32 on gfx950
// (checked with a one-wave probe): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; the result tile holds D[4 (l / 16) + v][l % 16],
// v = 0..3.  The K index of a product may be mapped to any row as long as
// both operands use the same map.
// ---------------------------------------------------------------------------
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// Gram of complex64 vectors on the FP32 matrix cores: partial[rb][i*nw + j] =
// sum over row block rb of conj(A_i[r]) W_j[r], nw <= 8.  A wave computes, for
// MT tiles of 16 vectors of A, the real products C1 = Ar^T [Wr | Wi] and
// C2 = Ai^T [Wr | Wi] (column j < 8: Wr_j, j >= 8: Wi_(j-8)), so that
// Re G = C1[., j] + C2[., j + 8] and Im G = C1[., j + 8] - C2[., j].  K runs
// over rows: in a chunk of 16 rows, lane group g takes rows 4g .. 4g + 3,
// one per MFMA step, and loads them with two 16-byte loads per vector (the
// four lane groups together read 128 contiguous bytes of each vector).  The
// FP32 tiles are added into FP64 accumulators every 32 rows; the products of
// complex64 values and these short FP32 sums err by about 2^-24 of their
// size, so a Gram entry of unit vectors errs by ~1e-9 at D = 1.6e7, far below
// the complex64 rounding of the vectors themselves.  Each wave takes every
// fourth chunk of the row block; the four waves' sums are added in a fixed
// order, so results do not depend on scheduling.  Chunks never read past
// row ld - 1 (row blocks and chunks start at multiples of 16, ld is a
// multiple of 64 above D), and rows [D, ld) are zero.  The 1-D grid maps all
// tiles of a row block onto one XCD, whose L2 then serves W to all of them.
template <int MT>
__global__ void __launch_bounds__(256)
gram_mfma32(int64_t D, VecPair<float2> A, const float2* __restrict__ Wb, int64_t ldw, int nw,
            int64_t rows_per_block, int n_itiles, int n_tiles, double2* __restrict__ partial, StepPred pred) {
  __shared__ double2 red[4][MT * 16][8];
  if (!pred.run()) return;
  const int per_xcd = (n_tiles + 7) / 8;
  const int logical = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (logical >= n_tiles) return;
  const int itile = logical % n_itiles, rblk = logical / n_itiles;
  const int nv = A.n0 + A.n1;
  const int i0 = itile * MT * 16;
  const int64_t rb = (int64_t)rblk * rows_per_block;
  const int64_t re = min(D, rb + rows_per_block);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const float2* ap[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) ap[t] = A.vec(min(i0 + t * 16 + col, nv - 1)) + 4 * g;
  const int wm = col & 7;
  const bool wlive = wm < nw, wim = col >= 8;
  const float2* wp = Wb + (int64_t)(wlive ? wm : 0) * ldw + 4 * g;
  f32x4_t c1[MT], c2[MT];
  double s1[MT][4], s2[MT][4];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    c1[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    c2[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int v = 0; v < 4; ++v) s1[t][v] = s2[t][v] = 0.0;
  }
  // chunk starts past the block's end are replaced by the block's first rows
  // (always inside the array); their loads are discarded
  auto safe = [&](int64_t r) { return r < re ? r : rb; };
  float4 a0[MT][2], a1[MT][2], w0[2], w1[2];
#define EDX_MG_LOAD(AB, WB, R)                                                        \
  {                                                                                    \
    const int64_t r_ = (R);                                                            \
    _Pragma("unroll") for (int t = 0; t < MT; ++t) {                                   \
      const float4* q = reinterpret_cast<const float4*>(ap[t] + r_);                   \
      AB[t][0] = q[0];                                                                 \
      AB[t][1] = q[1];                                                                 \
    }                                                                                  \
    const float4* qw = reinterpret_cast<const float4*>(wp + r_);                       \
    WB[0] = qw[0];                                                                     \
    WB[1] = qw[1];                                                                     \
    EDX_KEEP_PREFETCH();                                                               \
  }
#define EDX_MG_STEP(AB, WB)                                                           \
  {                                                                                    \
    const float wv[4] = {wim ? WB[0].y : WB[0].x, wim ? WB[0].w : WB[0].z,            \
                         wim ? WB[1].y : WB[1].x, wim ? WB[1].w : WB[1].z};           \
    _Pragma("unroll") for (int t = 0; t < MT; ++t) {                                   \
      const float ar[4] = {AB[t][0].x, AB[t][0].z, AB[t][1].x, AB[t][1].z};           \
      const float ai[4] = {AB[t][0].y, AB[t][0].w, AB[t][1].y, AB[t][1].w};           \
      _Pragma("unroll") for (int k = 0; k < 4; ++k) {                                  \
        const float b = wlive ? wv[k] : 0.f;                                           \
        c1[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[k], b, c1[t], 0, 0, 0);         \
        c2[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[k], b, c2[t], 0, 0, 0);         \
      }                                                                                \
    }                                                                                  \
  }
  int64_t r = rb + wave * 16;
  EDX_MG_LOAD(a0, w0, safe(r))
  for (; r < re; r += 128) {
    EDX_MG_LOAD(a1, w1, safe(r + 64))
    EDX_MG_STEP(a0, w0)
    EDX_MG_LOAD(a0, w0, safe(r + 128))
    if (r + 64 < re) EDX_MG_STEP(a1, w1)
#pragma unroll
    for (int t = 0; t < MT; ++t) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        s1[t][v] += (double)c1[t][v];
        s2[t][v] += (double)c2[t][v];
      }
      c1[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      c2[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    }
  }
#undef EDX_MG_LOAD
#undef EDX_MG_STEP
  // lane (g, col) holds C1[4g + v][col] and C2[4g + v][col]; columns j and
  // j + 8 sit in lanes col and col ^ 8
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double o1 = __shfl_xor(s1[t][v], 8, 64);
      const double o2 = __shfl_xor(s2[t][v], 8, 64);
      if (col < 8) red[wave][t * 16 + 4 * g + v][col] = make_double2(s1[t][v] + o2, o1 - s2[t][v]);
    }
  __syncthreads();
  const int nout = nv * nw;
  for (int o = threadIdx.x; o < MT * 16 * 8; o += 256) {
    const int i = o / 8, j = o % 8;
    if (i0 + i < nv && j < nw) {
      double xr = 0.0, xi = 0.0;
      for (int wv = 0; wv < 4; ++wv) {
        xr += red[wv][i][j].x;
        xi += red[wv][i][j].y;
      }
      partial[(int64_t)rblk * nout + (int64_t)(i0 + i) * nw + j] = make_double2(xr, xi);
    }
  }
}

// Combination of complex64 vectors on the FP32 matrix cores (restart
// combinations): Out_k = sum_j A_j S[j][k] for the NT * 8 outputs k of
// column group blockIdx.y (S row-major, FP64).  Operand layout as for the
// Gram above; K runs over the vectors of A, four per step (lane group g takes
// vector j + g).  In real form Re Out_k = sum Ar Sr - Ai Si and
// Im Out_k = sum Ar Si + Ai Sr: a 16 x 16 result tile holds 8 outputs
// (columns n < 8 their real parts, n >= 8 their imaginary parts) and takes
// two products per step, Ar times B1 = [Sr | Si] and Ai times
// B2 = [-Si | Sr].  A wave works on strips of 32 rows: lane (g, c) loads rows
// 2c and 2c + 1 of its vector with one 16-byte load, which feed two row tiles
// (the even and the odd rows of the strip); loads run two steps ahead.  The
// FP32 tiles (8 products per entry and step) are added into FP64 sums every
// kCombineFlush steps, so an output errs by a few FP32 roundings of its
// terms before its one rounding to complex64 (every step: 48-site restarts,
// 104 -> 57 vectors, took 10.8 ms against 9.3 ms every second step and 15.6
// ms on FP64 vector FMAs, and their FP64 refinement needed the matvecs it
// needs after vector-FMA restarts on 2 of 3 sectors, against 0 of 3 every
// second step).  The group's coefficients are staged in LDS as complex64 (rows
// padded so that the four lane groups read distinct banks), zero past nv and
// nk; vector indices past nv load vector nv - 1, which the zero coefficients
// cancel.  Strips never read past row ld - 1 (the launch requires ld >= D
// rounded up to 32); rows past D are not written.  Lane (g, c) ends with the
// real (c < 8) or imaginary part of output c % 8 of each tile at rows
// 8g .. 8g + 7 of the strip; it swaps half of them with lane c ^ 8 and writes
// four rows of that output as two 16-byte stores.
constexpr int kCombineFlush = 1;
constexpr int kCombineWorkgroups = 512;  // two per CU
// Row stride (complex64 entries) of the coefficient staging: 2 * ldc = 16
// (mod 64) dwords, so that rows j .. j + 3 start 16 LDS banks apart.
constexpr int combine_ldc(int nt) { return (nt * 8 + 31) / 32 * 32 + 8; }

template <int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
combine_mfma32(int64_t D, VecPair<float2> A, int nk, float2* __restrict__ Ob, int64_t ldo,
               const double2* __restrict__ S, int ldS, int64_t n_strips, StepPred pred) {
  constexpr int NC = NT * 8;              // outputs of a column group
  constexpr int LDC = combine_ldc(NT);    // LDS row stride in complex64 entries
  extern __shared__ float2 s_coef[];  // [nvp][LDC]
  if (!pred.run()) return;
  const int nv = A.n0 + A.n1;
  const int steps = (nv + 3) / 4;
  const int k0 = (int)blockIdx.y * NC;
  for (int e = threadIdx.x; e < steps * 4 * NC; e += 256) {
    const int i = e / NC, k = e % NC;
    float2 v = make_float2(0.f, 0.f);
    if (i < nv && k0 + k < nk) {
      const double2 s = S[(int64_t)i * ldS + k0 + k];
      v = make_float2((float)s.x, (float)s.y);
    }
    s_coef[i * LDC + k] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane & 15, g = lane >> 4;
  const bool re_lane = c < 8;
  const float2* crow = s_coef + g * LDC + (c & 7);
  for (int64_t strip = (int64_t)blockIdx.x * 4 + wave; strip < n_strips; strip += (int64_t)gridDim.x * 4) {
    const int64_t row0 = strip * 32;
    const int64_t lr = row0 + 2 * c;
    f32x4_t acc[2][NT];
    double sum[2][NT][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        acc[h][t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int v = 0; v < 4; ++v) sum[h][t][v] = 0.0;
      }
    auto load = [&](int s) -> float4 {
      return *reinterpret_cast<const float4*>(A.vec(min(4 * s + g, nv - 1)) + lr);
    };
    auto step = [&](const float4& a, int s) {
      const float2* cp = crow + 4 * s * LDC;
      float b1[NT], b2[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const float2 sc = cp[t * 8];
        b1[t] = re_lane ? sc.x : sc.y;
        b2[t] = re_lane ? -sc.y : sc.x;
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b1[t], acc[0][t], 0, 0, 0);
        acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b1[t], acc[1][t], 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b2[t], acc[0][t], 0, 0, 0);
        acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b2[t], acc[1][t], 0, 0, 0);
      }
      if ((s + 1) % kCombineFlush == 0 || s + 1 == steps) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int t = 0; t < NT; ++t) {
#pragma unroll
            for (int v = 0; v < 4; ++v) sum[h][t][v] += (double)acc[h][t][v];
            acc[h][t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
          }
      }
    };
    float4 a0 = load(0), a1 = load(1), a2;
    for (int s = 0; s < steps; s += 3) {
      a2 = load(s + 2);
      EDX_KEEP_PREFETCH();
      step(a0, s);
      a0 = load(s + 3);
      EDX_KEEP_PREFETCH();
      if (s + 1 < steps) step(a1, s + 1);
      a1 = load(s + 4);
      EDX_KEEP_PREFETCH();
      if (s + 2 < steps) step(a2, s + 2);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      float x[4][2], y[2][2];
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int h = 0; h < 2; ++h) x[v][h] = (float)sum[h][t][v];
      // the real lane keeps rows 8g .. 8g + 3 (v = 0, 1), the imaginary lane
      // rows 8g + 4 .. 8g + 7 (v = 2, 3)
#pragma unroll
      for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int h = 0; h < 2; ++h) y[w][h] = __shfl_xor(re_lane ? x[2 + w][h] : x[w][h], 8, 64);
      const int k = k0 + t * 8 + (c & 7);
      if (k >= nk) continue;
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        const int64_t row = row0 + 8 * g + 2 * (re_lane ? w : 2 + w);
        const float re0 = re_lane ? x[w][0] : y[w][0], im0 = re_lane ? y[w][0] : x[2 + w][0];
        const float re1 = re_lane ? x[w][1] : y[w][1], im1 = re_lane ? y[w][1] : x[2 + w][1];
        float2* op = Ob + (int64_t)k * ldo + row;
        if (row + 1 < D) *reinterpret_cast<float4*>(op) = make_float4(re0, im0, re1, im1);
        else if (row < D) *op = make_float2(re0, im0);
      }
    }
  }
}

// Launch of gemm_tiles: Out[:, 0..nk) = [A0 | A1] S over vectors of D rows
// (leading dimension ld), output tiles of NK (see the kernel for Xi and
// aliasing).
template <typename VT, int NK, int U = kGemmStage, int RPT = 1>
void launch_gemm_tiles(hipStream_t st, int64_t D, int64_t ld, const void* A0, int n0, const void* A1, int n1,
                       void* Out, int nk, const double2* S, int ldS, void* Xi, StepPred pred) {
  VecPair<VT> A{(const VT*)A0, ld, n0, (const VT*)A1, ld, n1};
  const int nv = n0 + n1;
  const int n_kt = (nk + NK - 1) / NK;
  const int64_t n_tiles = (int64_t)n_kt * ((D + 256 * RPT - 1) / (256 * RPT));
  EDX_REQUIRE(n_tiles < (1ll << 30), EDX_ERR_INTERNAL, "update grid");
  const int nvp = (nv + 2 * U - 1) / (2 * U) * (2 * U);
  // LDS: the coefficients, and (with an interleaved copy, room permitting)
  // the staged outputs
  size_t lds = (size_t)nvp * NK * sizeof(double2);
  const size_t stage = (size_t)256 * RPT * nk * sizeof(VT);
  const int stage_xi = Xi && stage <= kStageLdsBytes ? 1 : 0;
  if (stage_xi) lds = std::max(lds, stage);
  hipLaunchKernelGGL((gemm_tiles<VT, NK, U, RPT>), dim3((unsigned)((n_tiles + 7) / 8 * 8)), dim3(256), lds, st, D, A,
                     nk, (VT*)Out, ld, S, ldS, (VT*)Xi, n_kt, (int)n_tiles, stage_xi, pred);
}

// Combination schedules (more than 16 outputs): FP64 vector FMAs in tiles
// of 9 to 12 outputs, or (complex64) the FP32 matrix cores.
constexpr int kCombineVector = 1, kCombineMatrix = 2;
// Largest coefficient staging of a matrix-core combination (bytes).
constexpr size_t kCombineLdsBytes = 64 * 1024;
// Automatic schedule: complex64 combinations of vectors of at least this
// many rows run on the matrix cores.
constexpr int64_t kMfmaCombineRows = int64_t(1) << 16;

int combine_tiles(int nk) { return std::min(8, (nk + 7) / 8); }

// Whether combine_mfma32 takes nv inputs and nk outputs of D rows with
// leading dimension ld (coefficients within its LDS, whole 32-row strips
// inside every vector, 16-byte aligned rows).
bool combine_mfma_fits(int nv, int nk, int64_t D, int64_t ld) {
  const size_t lds = (size_t)((nv + 3) / 4 * 4) * combine_ldc(combine_tiles(nk)) * sizeof(float2);
  return lds <= kCombineLdsBytes && ld >= (D + 31) / 32 * 32 && ld % 2 == 0;
}

template <int NT>
void launch_combine_nt(hipStream_t st, int64_t D, const VecPair<float2>& A, int nk, float2* Out, int64_t ldo,
                       const double2* S, int ldS, StepPred pd) {
  const int nv = A.n0 + A.n1;
  const size_t lds = (size_t)((nv + 3) / 4 * 4) * combine_ldc(NT) * sizeof(float2);
  const int64_t n_strips = (D + 31) / 32;
  const int n_groups = (nk + NT * 8 - 1) / (NT * 8);
  // each workgroup stages the coefficients once for at least 8 strips
  const int64_t wgs = std::min<int64_t>(kCombineWorkgroups, std::max<int64_t>(1, n_strips / 8));
  hipLaunchKernelGGL((combine_mfma32<NT>), dim3((unsigned)wgs, (unsigned)n_groups), dim3(256), lds, st, D, A, nk, Out,
                     ldo, S, ldS, n_strips, pd);
}

void launch_combine_mfma(hipStream_t st, int64_t D, const VecPair<float2>& A, int nk, float2* Out, int64_t ldo,
                         const double2* S, int ldS, StepPred pd) {
  switch (combine_tiles(nk)) {
    case 1: launch_combine_nt<1>(st, D, A, nk, Out, ldo, S, ldS, pd); break;
    case 2: launch_combine_nt<2>(st, D, A, nk, Out, ldo, S, ldS, pd); break;
    case 3: launch_combine_nt<3>(st, D, A, nk, Out, ldo, S, ldS, pd); break;
    case 4: launch_combine_nt<4>(st, D, A, nk, Out, ldo, S, ldS, pd); break;
    case 5: launch_combine_nt<5>(st, D, A, nk, Out, ldo, S, ldS, pd); break;
    case 6: launch_combine_nt<6>(st, D, A, nk, Out, ldo, S, ldS, pd); break;
    case 7: launch_combine_nt<7>(st, D, A, nk, Out, ldo, S, ldS, pd); break;
    default: launch_combine_nt<8>(st, D, A, nk, Out, ldo, S, ldS, pd);
  }
}

// Out[:, 0..nk) = [A0 | A1] S with the given schedule (0: automatic).  Out
// must not overlap the inputs.  A matrix-core schedule that does not apply
// to the shape is an error when `strict`, else the vector schedule runs.
// Returns the kernel-timer class of the launch.
template <typename VT>
int launch_combine(hipStream_t st, int64_t D, int64_t ld, const void* A0, int n0, const void* A1, int n1, void* Out,
                    int nk, const double2* S, int ldS, int schedule, bool strict, StepPred pd) {
  if (std::is_same<VT, float2>::value && schedule != kCombineVector) {
    auto aligned = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
    const bool fits = combine_mfma_fits(n0 + n1, nk, D, ld) && aligned(A0) && aligned(A1) && aligned(Out);
    EDX_REQUIRE(fits || !strict || schedule != kCombineMatrix, EDX_ERR_INVALID,
                "the matrix-core combination needs ld >= dim rounded up to 32, an even ld, 16-byte aligned "
                "buffers and a smaller basis");
    if (fits && (schedule == kCombineMatrix || D >= kMfmaCombineRows)) {
      VecPair<float2> A{(const float2*)A0, ld, n0, (const float2*)A1, ld, n1};
      launch_combine_mfma(st, D, A, nk, (float2*)Out, ld, S, ldS, pd);
      return edx::K_COMBINE_MFMA;
    }
  }
  // as many tiles as with kCombineTile outputs each, but of equal width
  // (9 to 12 outputs), so the last one carries no padding: the combination
  // is bound by FP64 FMAs (60 -> 33: 84.8 us in tiles of 11 against 91.4 us
  // in tiles of 12, complex64, D = 225k)
  constexpr int R = kUpdateRows;
  const int n_kt = (nk + kCombineTile - 1) / kCombineTile;
  switch ((nk + n_kt - 1) / n_kt) {
    case 9: launch_gemm_tiles<VT, 9, kCombineStage, R>(st, D, ld, A0, n0, A1, n1, Out, nk, S, ldS, nullptr, pd); break;
    case 10: launch_gemm_tiles<VT, 10, kCombineStage, R>(st, D, ld, A0, n0, A1, n1, Out, nk, S, ldS, nullptr, pd); break;
    case 11: launch_gemm_tiles<VT, 11, kCombineStage, R>(st, D, ld, A0, n0, A1, n1, Out, nk, S, ldS, nullptr, pd); break;
    default:
      launch_gemm_tiles<VT, kCombineTile, kCombineStage, R>(st, D, ld, A0, n0, A1, n1, Out, nk, S, ldS, nullptr, pd);
  }
  return edx::K_COMBINE;
}

// Second Cholesky QR pass of a block, on the device: G (b x b, row-major,
// Hermitian) = R^H R with R upper triangular, S = R^-1 (row-major, the
// coefficients of the update dst <- dst S) and R copied to Rout (host-visible)
// for the projected matrix.  One wavefront; column j of R is formed after
// column j - 1 (lanes take the entries of a row), then lane j forms column j
// of R^-1.  The formulas are those of the host factorisation
// (edx::dense::cholesky_upper, upper_inverse).  A pivot at or below
// tiny * max diag sets status[0] = 1 and S = I (the update then leaves dst
// unchanged); the caller only takes this path when its error estimate makes
// the factorisation safe, and treats status 1 as an internal error.
__global__ void __launch_bounds__(64) block_cholinv(int b, const double2* __restrict__ G, double tiny,
                                                   double2* __restrict__ S, double2* __restrict__ Rout,
                                                   int* __restrict__ status, const double2* __restrict__ R1,
                                                   StepPred pred) {
  __shared__ double2 A[kMaxW][kMaxW], R[kMaxW][kMaxW], X[kMaxW][kMaxW], P1[kMaxW][kMaxW];
  if (!pred.run()) return;
  const int lane = threadIdx.x;
  for (int e = lane; e < kMaxW * kMaxW; e += 64) {
    const int i = e / kMaxW, j = e % kMaxW;
    A[i][j] = (i < b && j < b) ? G[i * b + j] : make_double2(0.0, 0.0);
    P1[i][j] = (R1 && i < b && j < b) ? R1[i * b + j] : make_double2(0.0, 0.0);
    R[i][j] = X[i][j] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  double dmax = 0.0;
  for (int i = 0; i < b; ++i) dmax = fmax(dmax, A[i][i].x);
  bool ok = dmax > 0.0;
  for (int j = 0; j < b; ++j) {
    double s = A[j][j].x;
    for (int k = 0; k < j; ++k) s -= R[k][j].x * R[k][j].x + R[k][j].y * R[k][j].y;
    if (!(s > tiny * dmax)) ok = false;
    const double rjj = sqrt(fmax(s, tiny * dmax > 0.0 ? tiny * dmax : 1.0));
    const int i = lane;
    if (i > j && i < b) {
      double tr = A[j][i].x, ti = A[j][i].y;
      for (int k = 0; k < j; ++k) {  // t -= conj(R(k,j)) R(k,i)
        const double2 a = R[k][j], c = R[k][i];
        tr -= a.x * c.x + a.y * c.y;
        ti -= a.x * c.y - a.y * c.x;
      }
      R[j][i] = make_double2(tr / rjj, ti / rjj);
    }
    if (lane == 0) R[j][j] = make_double2(rjj, 0.0);
    __syncthreads();
  }
  if (lane < b) {
    const int j = lane;
    X[j][j] = make_double2(1.0 / R[j][j].x, 0.0);
    for (int i = j - 1; i >= 0; --i) {
      double sr = 0.0, si = 0.0;
      for (int k = i + 1; k <= j; ++k) {
        const double2 a = R[i][k], c = X[k][j];
        sr += a.x * c.x - a.y * c.y;
        si += a.x * c.y + a.y * c.x;
      }
      X[i][j] = make_double2(-sr / R[i][i].x, -si / R[i][i].x);
    }
  }
  __syncthreads();
  for (int e = lane; e < b * b; e += 64) {
    const int i = e / b, j = e % b;
    S[e] = ok ? X[i][j] : make_double2(i == j ? 1.0 : 0.0, 0.0);
    if (R1) {  // Rout = R R1 (both upper triangular)
      double sr = 0.0, si = 0.0;
      for (int k = i; k <= j; ++k) {
        const double2 a = R[i][k], c = P1[k][j];
        sr += a.x * c.x - a.y * c.y;
        si += a.x * c.y + a.y * c.x;
      }
      Rout[e] = make_double2(sr, si);
    } else {
      Rout[e] = R[i][j];
    }
  }
  if (lane == 0) status[0] = ok ? 0 : 1;
}

// Coefficients of a block step from its Gram P = [Q, W]^H W ((nbasis + b) x b,
// row-major; Q the nbasis basis vectors, W = H Vr), as the host computes them
// (extend): C = P[0, nbasis), G = W^H W - C^H C, drop = max_j (W^H W)_jj /
// G_jj, G = R1^H R1, and S = [-C R1^-1; R1^-1] so that [Q, W] S =
// (W - Q C) R1^-1.  The status (ctl[0]) is 0 when the error estimate
// unit_roundoff * max(drop, 1) * cond(R1)^2 is within single_tol (one pass
// suffices), 1 when it is within 1/4 (a second pass on the device is safe),
// and 2 otherwise (also instead of 1 when no device second pass follows,
// second_pass = 0), or when drop >= kappa2 (a second projection is needed)
// or G is not positive definite; status 2 also sets the halt word ctl[1], and
// the host finishes the step.  The log receives C, R1 and the status.
__global__ void __launch_bounds__(256)
step_coeffs(const double2* __restrict__ P, int nbasis, int b, double kappa2, double tiny, double uround,
            double single_tol, int second_pass, double2* __restrict__ S, double2* __restrict__ R1out,
            double2* __restrict__ log, int* __restrict__ ctl) {
  __shared__ double2 G[kMaxW][kMaxW], R[kMaxW][kMaxW], X[kMaxW][kMaxW];
  __shared__ double2 Ps[kStepLds];  // P, read once with all lanes
  __shared__ int status_s;
  if (ctl[1]) return;  // halted: the host finishes an earlier step first
  const int tid = threadIdx.x;
  for (int e = tid; e < (nbasis + b) * b; e += 256) {
    const double2 v = P[e];
    Ps[e] = v;
    if (e < nbasis * b) log[e] = v;
  }
  __syncthreads();
  for (int e = tid; e < kMaxW * kMaxW; e += 256) {
    const int i = e / kMaxW, j = e % kMaxW;
    double2 g = make_double2(0.0, 0.0);
    if (i < b && j < b) {
      g = Ps[(nbasis + i) * b + j];
      for (int k = 0; k < nbasis; ++k) {  // g -= conj(P(k,i)) P(k,j)
        const double2 a = Ps[k * b + i], c = Ps[k * b + j];
        g.x -= a.x * c.x + a.y * c.y;
        g.y -= a.x * c.y - a.y * c.x;
      }
    }
    G[i][j] = g;
    R[i][j] = X[i][j] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  double drop = 0.0, dmax = 0.0;
  for (int j = 0; j < b; ++j) {
    const double g = G[j][j].x;
    drop = fmax(drop, g > 0.0 ? Ps[(nbasis + j) * b + j].x / g : INFINITY);
    dmax = fmax(dmax, g);
  }
  bool ok = dmax > 0.0;
  for (int j = 0; j < b; ++j) {
    double sd = G[j][j].x;
    for (int k = 0; k < j; ++k) sd -= R[k][j].x * R[k][j].x + R[k][j].y * R[k][j].y;
    if (!(sd > tiny * dmax)) ok = false;
    const double rjj = sqrt(fmax(sd, dmax > 0.0 ? tiny * dmax : 1.0));
    const int i = tid;
    if (i > j && i < b) {
      double tr = G[j][i].x, ti = G[j][i].y;
      for (int k = 0; k < j; ++k) {
        const double2 a = R[k][j], c = R[k][i];
        tr -= a.x * c.x + a.y * c.y;
        ti -= a.x * c.y - a.y * c.x;
      }
      R[j][i] = make_double2(tr / rjj, ti / rjj);
    }
    if (tid == 0) R[j][j] = make_double2(rjj, 0.0);
    __syncthreads();
  }
  if (tid < b) {
    const int j = tid;
    X[j][j] = make_double2(1.0 / R[j][j].x, 0.0);
    for (int i = j - 1; i >= 0; --i) {
      double sr = 0.0, si = 0.0;
      for (int k = i + 1; k <= j; ++k) {
        const double2 a = R[i][k], c = X[k][j];
        sr += a.x * c.x - a.y * c.y;
        si += a.x * c.y + a.y * c.x;
      }
      X[i][j] = make_double2(-sr / R[i][i].x, -si / R[i][i].x);
    }
  }
  __syncthreads();
  if (tid == 0) {
    double fr = 0.0, fx = 0.0;
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < b; ++j) {
        fr += R[i][j].x * R[i][j].x + R[i][j].y * R[i][j].y;
        fx += X[i][j].x * X[i][j].x + X[i][j].y * X[i][j].y;
      }
    const double cond = sqrt(fr) * sqrt(fx);
    const double est = uround * fmax(drop, 1.0) * cond * cond;
    int st = 2;
    if (ok && drop < kappa2) st = est <= single_tol ? 0 : (est <= 0.25 && second_pass ? 1 : 2);
    status_s = st;
    double2* tail = log + kLogC + kMaxW * kMaxW;
    tail[0] = make_double2((double)st, 0.0);
    tail[1] = make_double2(drop, cond);
    ctl[0] = st;
    if (st == 2) ctl[1] = 1;
  }
  __syncthreads();
  if (status_s == 2) return;
  for (int e = tid; e < b * b; e += 256) {
    R1out[e] = R[e / b][e % b];
    log[kLogC + e] = R[e / b][e % b];
  }
  for (int e = tid; e < (nbasis + b) * b; e += 256) {
    const int i = e / b, j = e % b;
    double2 v;
    if (i < nbasis) {  // -sum_k C(i,k) R1^-1(k,j)
      double sr = 0.0, si = 0.0;
      for (int k = 0; k <= j; ++k) {
        const double2 a = Ps[i * b + k], c = X[k][j];
        sr += a.x * c.x - a.y * c.y;
        si += a.x * c.y + a.y * c.x;
      }
      v = make_double2(-sr, -si);
    } else {
      v = X[i - nbasis][j];
    }
    S[e] = v;
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Complex standard normal entries (Box-Muller on a counter-based hash).
template <typename V>
__global__ void fill_gaussian(int64_t D, int nvec, V* out, int64_t ld, uint64_t seed) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= D) return;
  for (int j = 0; j < nvec; ++j) {
    const uint64_t h1 = mix64(seed ^ mix64((uint64_t)j * 0x100000001b3ull + (uint64_t)r * 2 + 1));
    const uint64_t h2 = mix64(h1 ^ 0x632be59bd9b4e019ull);
    const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double rad = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(6.283185307179586 * u2, &sn, &cs);
    out[(int64_t)j * ld + r] = from_d2<V>(rad * cs, rad * sn);
  }
}

// dst[:, j] = src[:, j] for j < gridDim.y, rows < D (one element per thread).
template <typename V>
__global__ void copy_vecs(int64_t D, const V* __restrict__ src, int64_t lds, V* __restrict__ dst, int64_t ldd) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= D) return;
  dst[(int64_t)blockIdx.y * ldd + r] = src[(int64_t)blockIdx.y * lds + r];
}

template <typename A, typename B>
__global__ void convert_vec(int64_t D, int nvec, const A* src, int64_t lds, B* dst, int64_t ldd) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= D) return;
  for (int j = 0; j < nvec; ++j) {
    const double2 v = to_d2(src[(int64_t)j * lds + r]);
    dst[(int64_t)j * ldd + r] = from_d2<B>(v.x, v.y);
  }
}

struct Options {
  int nev = 1;
  int block = 1;
  int basis = 0;  // max Ritz basis size (0: automatic)
  double tol = 1e-10;     // relative residual ||H x - theta x|| / max(1, |theta|)
  double eig_tol = 0.0;   // > 0: a pair is also done when its cluster's sum r^2 / gap <= eig_tol
  // Ritz values of an earlier, larger basis of the same operator (or of a
  // lower-precision copy): the gaps of the eig_tol rule are measured to these
  // too, so a small basis after a restart cannot overestimate them
  std::vector<double> gap_ref, gap_ref_res;  // values and residual norms
  // return, past the nev-th pair, the Ritz vectors of its cluster too (values
  // within 1e-4 relative of it), so that a restart from them holds the cluster
  bool out_cluster = false;
  // > 0: also stop once the worst relative residual of the nev pairs is
  // within stall_ceiling and has not dropped by 10 % over this many checks
  // (a lower-precision phase that has reached its rounding floor)
  int stall_checks = 0;
  double stall_ceiling = 0.0;
  int max_restarts = 5000;
  int check_every = 4;    // Ritz check every this many block steps (and when the basis is full)
  uint64_t seed = 1;
};

// Automatic Krylov basis size: 5 (nev + b) vectors; for complex64 sweeps
// after many levels (nev >= 8) up to ratio / 4.5 vectors, with ratio = SpMM
// bytes / bytes of one vector.  A block step's Gram and update read the basis
// about twice, so that size keeps the per-step vector work below about half
// of the SpMM's, and a larger basis restarts less often.  Measured (mixed,
// eigenvalue bound 1e-12): the 48-site checkerboard sector (nev = 10, block
// 8, ratio ~790) takes 6.7 s with 176 vectors against 7.9 s with 96 and 7.4 s
// with 208; the 36-site Haldane sector (nev = 5, ratio ~560) is faster with
// 50 vectors (0.75 s) than with 125 (0.85 s), and so is the FP64 phase of the
// checkerboard solve with 100 rather than 200, so both keep 5 (nev + b).
int auto_basis(int64_t D, int64_t E, int prec, int b, int nev) {
  const int base = std::max({5 * (nev + b), 6 * b, 24});
  if (prec == EDX_PREC_FP64 || nev < 8 || D <= 0) return base;
  // the count keeps the element bytes of an unpacked table: on the
  // palette-packed 48-site checkerboard sectors (nev = 10, b = 4) its 104
  // vectors stay the fastest basis (84, 92, 116 and 128 took 2-12 % longer)
  const double esz = 8.0;
  const double spmm = (double)D * (esz + 2.0 * b * esz) + (double)E * (4.0 + esz + (double)b * esz);
  return std::max(base, (int)(spmm / ((double)D * esz) / 4.5));
}

// Convergence of the nev lowest Ritz pairs.  A pair passes only when its
// cluster (the Ritz values within 1e-4 relative of it, which may reach past
// the nev-th) is resolved as a whole: until then the error of each of its Ritz
// values is first order in the residuals, not residual^2 / gap.  The cluster
// then passes when all its members meet the relative residual tol, or
// (eig_tol > 0) when its eigenvalue error bound, the sum of its squared
// residuals divided by the distance to the nearest Ritz value outside it, is
// below eig_tol (the residual-squared-over-gap bound for Hermitian matrices).
// `ref` / `ref_res` hold the Ritz values and residual norms of an earlier,
// larger basis.  Those whose residual places an eigenvalue within half the
// window of the pair are copies of the cluster: the current basis must hold
// as many (else a copy has not been captured yet).  The distance is also
// taken to the values outside the window, since a basis of a few vectors has
// its higher Ritz values far up the spectrum.
bool converged(const std::vector<double>& th, const std::vector<double>& res, int nev, double tol, double eig_tol,
               const std::vector<double>& ref, const std::vector<double>& ref_res, double* worst) {
  bool ok = true;
  double w = 0.0;
  const int n = (int)th.size();
  for (int i = 0; i < nev; ++i) {
    const double sc = std::max(1.0, std::fabs(th[i]));
    w = std::max(w, res[i] / sc);
    int members = 0;
    bool all_tol = true;
    double rc2 = 0.0, gap = INFINITY;
    for (int j = 0; j < n; ++j) {
      const double d = std::fabs(th[j] - th[i]);
      if (d <= 1e-4 * sc) {
        ++members;
        rc2 += res[j] * res[j];
        if (res[j] / std::max(1.0, std::fabs(th[j])) > tol) all_tol = false;
      } else {
        gap = std::min(gap, d);
      }
    }
    int ref_members = 0;
    for (size_t j = 0; j < ref.size(); ++j) {
      const double d = std::fabs(ref[j] - th[i]);
      if (d > 1e-4 * sc) gap = std::min(gap, d);
      else if (d + ref_res[j] <= 0.5e-4 * sc) ++ref_members;
    }
    if (ref_members > members) {
      ok = false;
      continue;
    }
    if (all_tol) continue;
    if (!(eig_tol > 0.0 && rc2 / gap <= eig_tol)) ok = false;
  }
  *worst = w;
  return ok;
}

struct Result {
  std::vector<double> vals;
  std::vector<double> ritz;  // every Ritz value of the final basis, ascending
  std::vector<double> ritz_res;  // and their residual norms
  int nvec = 0;              // Ritz vectors left in vec(0..nvec): nev, or more with out_cluster
  int64_t matvecs = 0;
  int restarts = 0;
  double max_rel_res = 0.0;
  // host time (ms): whole run, waiting on the stream, Ritz problems
  double ms_total = 0.0, ms_wait = 0.0, ms_ritz = 0.0;
  int syncs = 0, checks = 0;
};

// Wall-clock milliseconds since an arbitrary origin.
inline double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Block Lanczos on one precision.  Vectors live in two banks (ping-pong for
// the thick restart), each holding `cap` vectors of leading dimension ld.
// Every step makes one pass over the basis for the projections of H Vr (a
// second one only when that pass cancels most of the block) and one for the
// update that also forms the next block; the host receives the projections
// (pinned buffers) to extend the projected matrix.
class BlockLanczos {
 public:
  BlockLanczos(edx_ctx& ctx, const edx_hamiltonian& h, int prec, const Options& opt)
      : ctx_(ctx), h_(h), prec_(prec), opt_(opt) {
    D_ = h.dim;
    b_ = opt.block;
    esz_ = prec == EDX_PREC_FP64 ? 16 : 8;
    // leading dimension: a multiple of 64 rows with at least one row past D;
    // rows [D, ld) of every vector are zero (cleared below, never written:
    // kernels store rows < D only), which the Gram kernel's unconditional
    // loads rely on
    ld_ = (D_ + 64) / 64 * 64;
    int base = opt.basis > 0 ? opt.basis : auto_basis(h.dim, h.n_entries, prec, b_, opt.nev);
    base = (base + b_ - 1) / b_ * b_;
    nq_max_ = std::min({base, kMaxVec - 2 * b_, kStepLds / b_ - b_});
    nq_max_ = std::max(nq_max_ / b_ * b_, opt.nev + b_);
    cap_ = nq_max_ + b_;
    EDX_REQUIRE(cap_ + b_ <= kMaxVec && (nq_max_ + b_) * b_ <= kStepLds, EDX_ERR_INVALID,
                "Krylov basis too large for nev/block");
    // a single projection pass is kept while no column loses more than this
    // factor (squared) of its norm to it: the orthogonality it leaves is
    // about the unit roundoff times that factor
    kappa2_ = prec == EDX_PREC_FP64 ? 1e4 : 1e2;
    // one Cholesky QR pass suffices when its orthogonality error estimate is
    // below this (FP64 only; complex64 blocks always take two passes)
    unit_roundoff_ = prec == EDX_PREC_FP64 ? 1.1e-16 : 6e-8;
    single_qr_tol_ = prec == EDX_PREC_FP64 ? 1e-13 : 0.0;
    bank_[0].alloc((size_t)cap_ * ld_ * esz_);
    bank_[1].alloc((size_t)cap_ * ld_ * esz_);
    W_.alloc((size_t)b_ * ld_ * esz_);
    X_.alloc((size_t)b_ * D_ * esz_);
    const size_t pad = (size_t)(ld_ - D_) * esz_;
    EDX_HIP(hipMemset2DAsync(static_cast<char*>((void*)bank_[0].p) + D_ * esz_, ld_ * esz_, 0, pad, cap_,
                             ctx_.stream));
    EDX_HIP(hipMemset2DAsync(static_cast<char*>((void*)bank_[1].p) + D_ * esz_, ld_ * esz_, 0, pad, cap_,
                             ctx_.stream));
    EDX_HIP(hipMemset2DAsync(static_cast<char*>((void*)W_.p) + D_ * esz_, ld_ * esz_, 0, pad, b_, ctx_.stream));
    // Gram row blocks of about 8192 rows (32 per lane), a multiple of 256
    // rows each, at most 1024 of them
    n_rowblocks_ = (int)std::min<int64_t>(std::max<int64_t>((D_ + 8191) / 8192, 1), 1024);
    rows_per_block_ = ((D_ + n_rowblocks_ - 1) / n_rowblocks_ + 255) / 256 * 256;
    n_rowblocks_ = (int)((D_ + rows_per_block_ - 1) / rows_per_block_);
    partial_.alloc((size_t)n_rowblocks_ * kMaxOut);
    sdev_.alloc((size_t)kSlots * kSlot);
    static_assert(kMaxOut <= edx::kPinnedGram && kSlots * kSlot <= edx::kPinnedRing, "pinned staging too small");
    static_assert(kMaxW * kMaxW + 1 <= edx::kPinnedFactor, "pinned factor area too small");
    static_assert((size_t)(kMaxVec + kMaxW) * kMaxW <= edx::kLogC && edx::kLogSlot >= edx::kLogC + kMaxW * kMaxW + 4,
                  "block step log slot too small");
    hc_ = ctx.pinned_in;
    hc_dev_ = ctx.pinned_in_dev;
    hf_ = ctx.pinned_in + edx::kPinnedGram;
    hf_dev_ = ctx.pinned_in_dev + edx::kPinnedGram;
    hs_ = ctx.pinned_out;
    fdev_.alloc((size_t)2 * kMaxW * kMaxW);
    pbuf_.alloc((size_t)kMaxOut);
    sbuf_.alloc((size_t)kMaxOut);
    r1buf_.alloc((size_t)kMaxW * kMaxW);
    g2buf_.alloc((size_t)kMaxW * kMaxW);
    s2buf_.alloc((size_t)kMaxW * kMaxW);
    ctl_.alloc(4);
    EDX_HIP(hipMemsetAsync(ctl_.p, 0, 4 * sizeof(int), ctx_.stream));
    choose_schedule();
  }
  BlockLanczos(const BlockLanczos&) = delete;
  BlockLanczos& operator=(const BlockLanczos&) = delete;

  // Runs to convergence.  `start` (optional, device, same precision,
  // column-major with leading dimension start_ld) seeds the first block with
  // n_start vectors (rest random).  Ritz vectors are left in vec(0..nev).
  Result run(const void* start, int n_start, int64_t start_ld) {
    begin(start, n_start, start_ld);
    for (;;)
      if (step(false) && sync_point()) return res_;
  }

  // The same run in pieces, for a driver that advances several solves
  // together: begin(), then step() until it returns true, then
  // sync_point(), which returns true once the run has converged (result()).
  void begin(const void* start, int n_start, int64_t start_ld) {
    t_start_ = now_ms();
    res_ = Result();
    cur_ = 0;
    nq_ = 0;
    random_block(vec(cur_, 0), b_, opt_.seed);
    if (start && n_start > 0) {
      const int ns = std::min(n_start, b_);
      copy_block(vec(cur_, 0), ld_, start, start_ld, ns);
    }
    start_block();
    M_ = Mat(cap_, cap_);
    B_ = Mat(b_, cap_);
    theta_.clear();
    since_check_ = 0;
    best_worst_ = INFINITY;  // stall detection
    since_best_ = 0;
    pending_.clear();
  }

  // One block step on the device: W = H Vr (Vr at vec(nq), interleaved copy
  // in X_; enqueued here unless the caller has enqueued it, spmm_enqueued),
  // projection, coefficients, update; the host does not wait.  Returns true
  // when the host must now wait for the stream and call sync_point().
  bool step(bool spmm_enqueued) {
    if (!spmm_enqueued) edx::spmm(h_, ctx_, prec_, b_, X_.p, W_.p, ld_);
    enqueue_step(nq_, (int)pending_.size());
    pending_.push_back(nq_);
    nq_ += b_;
    full_ = nq_ + b_ > nq_max_;
    due_ = nq_ >= opt_.nev && (full_ || ++since_check_ >= opt_.check_every);
    return due_ || (int)pending_.size() >= edx::kLogSlots;
  }

  // Waits for the stream, folds the enqueued steps into the projected matrix
  // and, when a Ritz check is due, checks convergence, then finishes (returns
  // true) or restarts when the basis is full.
  bool sync_point() {
    wait();
    settle(true);
    const int n_enq = (int)pending_.size();
    int n_acc = 0;
    nq_ = absorb(pending_, M_, res_, n_acc);
    pending_.clear();
    if (n_acc < n_enq) {  // a step was finished on the host; the later ones were skipped
      since_check_ = std::max(0, since_check_ - (n_enq - n_acc));
      full_ = nq_ + b_ > nq_max_;
      due_ = nq_ >= opt_.nev && (full_ || since_check_ >= opt_.check_every);
    }
    if (!due_) return false;
    since_check_ = 0;
    const int nq = nq_;
    const bool full = full_;
    // coupling of the newest block: H Q = Q M + Vr B with B = [0, R]
    settle(false);
    B_ = Mat(b_, cap_);
    for (int i = 0; i < b_; ++i)
      for (int j = 0; j < b_; ++j) B_(i, nq - b_ + j) = Rstep_(i, j);
    const int nev = opt_.nev;
    // vectors kept by a thick restart
    int k = nev + (nq - nev) / 2;
    k = std::max(k, nev + std::min(b_, nq - nev));
    k = std::min(k, nq - b_);
    k = std::max(k, nev);
    // Ritz values, rows nq-b .. nq-1 of every Ritz vector (the residual
    // norms need only those) and, when restarting, the k leading vectors
    Mat Yl, Yk;
    ritz(M_, nq, theta_, full ? k : 0, full ? &Yk : nullptr, &Yl);
    std::vector<double> rn(nq);
    for (int c = 0; c < nq; ++c) {
      double rn2 = 0.0;
      for (int i = 0; i < b_; ++i) {
        cplx sum(0.0, 0.0);
        for (int j = 0; j < b_; ++j) sum += B_(i, nq - b_ + j) * Yl(j, c);
        rn2 += std::norm(sum);
      }
      rn[c] = std::sqrt(rn2);
    }
    double worst = 0.0;
    bool conv = converged(theta_, rn, nev, opt_.tol, opt_.eig_tol, opt_.gap_ref, opt_.gap_ref_res, &worst);
    res_.max_rel_res = worst;
    if (!conv && opt_.stall_checks > 0) {
      if (worst < 0.9 * best_worst_) {
        best_worst_ = worst;
        since_best_ = 0;
      } else if (++since_best_ >= opt_.stall_checks && worst <= opt_.stall_ceiling) {
        conv = true;
      }
    }
    if (conv) {
      int nout = nev;
      if (opt_.out_cluster) {
        const double sc = std::max(1.0, std::fabs(theta_[nev - 1]));
        while (nout < nq && nout < kMaxW && std::fabs(theta_[nout] - theta_[nev - 1]) <= 1e-4 * sc) ++nout;
      }
      Mat Yn;
      if (full && nout <= k) {
        Yn = Mat(nq, nout);
        for (int i = 0; i < nq; ++i)
          for (int c = 0; c < nout; ++c) Yn(i, c) = Yk(i, c);
      } else {
        std::vector<double> th2;
        ritz(M_, nq, th2, nout, &Yn, nullptr);
      }
      combine(vec(cur_, 0), nq, vec(1 - cur_, 0), nout, Yn);
      cur_ = 1 - cur_;
      res_.vals.assign(theta_.begin(), theta_.begin() + nev);
      res_.nvec = nout;
      res_.ritz = theta_;
      res_.ritz_res = rn;
      res_.restarts = restarts_;
      finish(res_, t_start_);
      return true;
    }
    if (full) {
      EDX_REQUIRE(restarts_ < opt_.max_restarts, EDX_ERR_NOCONV,
                  "block Lanczos did not converge within " + std::to_string(opt_.max_restarts) +
                      " restarts (max relative residual " + std::to_string(worst) + ")");
      combine(vec(cur_, 0), nq, vec(1 - cur_, 0), k, Yk);
      copy_block(vec(1 - cur_, k), ld_, vec(cur_, nq), ld_, b_);
      Mat Bn(b_, cap_);
      for (int i = 0; i < b_; ++i)
        for (int c = 0; c < k; ++c) {
          cplx sum(0.0, 0.0);
          for (int j = 0; j < b_; ++j) sum += B_(i, nq - b_ + j) * Yl(j, c);
          Bn(i, c) = sum;
        }
      M_ = Mat(cap_, cap_);
      for (int c = 0; c < k; ++c) M_(c, c) = theta_[c];
      B_ = Bn;
      nq_ = k;
      cur_ = 1 - cur_;
      ++restarts_;
    }
    return false;
  }

  const Result& result() const { return res_; }
  // What a grouped SpMM needs of this solve: its operator, precision, block
  // width, input (interleaved) and output (column-major, leading dimension
  // ld) blocks, and its stream.
  const edx_hamiltonian& op() const { return h_; }
  int precision() const { return prec_; }
  int block() const { return b_; }
  const void* spmm_in() const { return X_.p; }
  void* spmm_out() { return W_.p; }
  hipStream_t stream() const { return ctx_.stream; }

  // n vectors (rows < D) from src to dst, both of this run's precision.
  void copy_block(void* dst, int64_t ldd, const void* src, int64_t lds, int n) {
    if (n <= 0) return;
    const dim3 grid((unsigned)edx::grid_for(D_, 256), (unsigned)n);
    if (prec_ == EDX_PREC_FP64)
      hipLaunchKernelGGL(copy_vecs<double2>, grid, dim3(256), 0, ctx_.stream, D_, (const double2*)src, lds,
                         (double2*)dst, ldd);
    else
      hipLaunchKernelGGL(copy_vecs<float2>, grid, dim3(256), 0, ctx_.stream, D_, (const float2*)src, lds,
                         (float2*)dst, ldd);
    EDX_HIP(hipGetLastError());
  }

  // Device address of basis vector i of a bank.
  void* vec(int bank, int i) { return static_cast<char*>((void*)bank_[bank].p) + (size_t)i * ld_ * esz_; }
  int current() const { return cur_; }
  int64_t ld() const { return ld_; }

 private:
  static constexpr int kMaxOut = (kMaxVec + kMaxW) * kMaxW;  // Gram outputs of [Q, W] against W
  static constexpr int kSlots = 4;
  static constexpr size_t kSlot = (size_t)kMaxOut;

  template <typename VT, int TI, int NW>
  void launch_tiles(const void* A0, int n0, const void* A1, int n1, const void* W, StepPred pred) {
    const int nv = n0 + n1;
    const int n_itiles = (nv + TI - 1) / TI;
    const int n_tiles = n_itiles * n_rowblocks_;
    const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
    VecPair<VT> A{(const VT*)A0, ld_, n0, (const VT*)A1, ld_, n1};
    hipLaunchKernelGGL((gram_tiles<VT, TI, NW>), grid, dim3(256), 0, ctx_.stream, D_, ld_, A, (const VT*)W, ld_,
                       rows_per_block_, n_itiles, n_tiles, partial_.p, pred);
  }

  // Gram of complex64 vectors against at most 8 on the FP32 matrix cores.
  // Tiles of 32 vectors per wave, or 16 for Grams of at most 16 vectors (the
  // second Cholesky QR pass), whose second tile would only repeat loads.
  void launch_gram_mfma(const void* A0, int n0, const void* A1, int n1, const void* W, int nw, StepPred pred) {
    const int nv = n0 + n1;
    const int MT = nv <= 16 ? 1 : 2;
    const int n_itiles = (nv + 16 * MT - 1) / (16 * MT);
    const int n_tiles = n_itiles * n_rowblocks_;
    const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
    VecPair<float2> A{(const float2*)A0, ld_, n0, (const float2*)A1, ld_, n1};
    if (MT == 1)
      hipLaunchKernelGGL((gram_mfma32<1>), grid, dim3(256), 0, ctx_.stream, D_, A, (const float2*)W, ld_, nw,
                         rows_per_block_, n_itiles, n_tiles, partial_.p, pred);
    else
      hipLaunchKernelGGL((gram_mfma32<2>), grid, dim3(256), 0, ctx_.stream, D_, A, (const float2*)W, ld_, nw,
                         rows_per_block_, n_itiles, n_tiles, partial_.p, pred);
  }

  template <typename VT>
  void launch_tiles_nw(const void* A0, int n0, const void* A1, int n1, const void* W, int nw, StepPred pred) {
    // the Gram of a block of at most 4 vectors with itself (second Cholesky
    // QR pass): one tile of exactly that size on vector FMAs (on the matrix
    // cores it would fill 16-wide tiles with 4 live columns: 274 us against
    // the 63 us that its 0.5 GB take at the HBM peak, complex64, D = 1.6e7)
    if (n0 + n1 == nw && nw <= 4) {
      switch (nw) {
        case 1: launch_tiles<VT, 1, 1>(A0, n0, A1, n1, W, pred); return;
        case 2: launch_tiles<VT, 2, 2>(A0, n0, A1, n1, W, pred); return;
        case 3: launch_tiles<VT, 3, 3>(A0, n0, A1, n1, W, pred); return;
        default: launch_tiles<VT, 4, 4>(A0, n0, A1, n1, W, pred); return;
      }
    }
    if (std::is_same<VT, float2>::value && nw <= 8 && gram_mode_ == kGramMatrix) {
      launch_gram_mfma(A0, n0, A1, n1, W, nw, pred);
      return;
    }
    switch (nw) {
#define EDX_GRAM_CASE(NWV, TIV) \
  case NWV:                     \
    launch_tiles<VT, TIV, NWV>(A0, n0, A1, n1, W, pred); \
    break;
      // TI x NW tiles of 16 to 32 complex sums per thread.  Measured at
      // D = 3.8e6 and 2.3e5, 48 and 88 basis vectors, complex64 and
      // complex128: TI = 4 beats 3 at NW = 7, 8 (10-22 %), TI = 3 beats 2 at
      // NW = 9, 10 (17-21 %), TI = 2 beats 1 at NW = 16 (about 40 %)
      EDX_GRAM_CASE(1, 8) EDX_GRAM_CASE(2, 8) EDX_GRAM_CASE(3, 6)
      EDX_GRAM_CASE(5, 4) EDX_GRAM_CASE(7, 4) EDX_GRAM_CASE(8, 4)
      // from 2^20 rows the larger tile wins (fewer re-reads of W; D = 3.8e6:
      // 6-9 % at NW = 6, 3-11 % at NW = 4); below, the smaller one keeps more
      // workgroups per row block (D = 2.3e5: TI = 5 10-15 % slower at NW = 6)
      case 4:
        if (D_ >= (int64_t(1) << 20)) launch_tiles<VT, 8, 4>(A0, n0, A1, n1, W, pred);
        else launch_tiles<VT, 6, 4>(A0, n0, A1, n1, W, pred);
        break;
      case 6:
        if (D_ >= (int64_t(1) << 20)) launch_tiles<VT, 5, 6>(A0, n0, A1, n1, W, pred);
        else launch_tiles<VT, 4, 6>(A0, n0, A1, n1, W, pred);
        break;
      EDX_GRAM_CASE(9, 3) EDX_GRAM_CASE(10, 3) EDX_GRAM_CASE(11, 2) EDX_GRAM_CASE(12, 2)
      EDX_GRAM_CASE(13, 2) EDX_GRAM_CASE(14, 2) EDX_GRAM_CASE(15, 2) EDX_GRAM_CASE(16, 2)
#undef EDX_GRAM_CASE
      default:
        throw edx::Error{EDX_ERR_INTERNAL, "Gram block width out of range"};
    }
  }

  // Host copy of [A0 (n0 vectors), A1 (n1 vectors)]^H W as (n0 + n1) x nw.
  Mat gram_host(const void* A0, int n0, const void* A1, int n1, const void* W, int nw) {
    const int nv = n0 + n1;
    EDX_REQUIRE(nv <= kMaxVec + kMaxW && nw <= kMaxW, EDX_ERR_INTERNAL, "gram size");
    hipEvent_t t0 = ctx_.timer.start(ctx_.stream);
    if (prec_ == EDX_PREC_FP64) launch_tiles_nw<double2>(A0, n0, A1, n1, W, nw, StepPred{});
    else launch_tiles_nw<float2>(A0, n0, A1, n1, W, nw, StepPred{});
    EDX_HIP(hipGetLastError());
    const int nout = nv * nw;
    // the reduction writes the result straight into the pinned host buffer
    hipLaunchKernelGGL(gram_reduce, dim3((nout + 3) / 4), dim3(256), 0, ctx_.stream, n_rowblocks_, nout,
                       (const double2*)partial_.p, hc_dev_, StepPred{});
    EDX_HIP(hipGetLastError());
    ctx_.timer.stop(ctx_.stream, t0, edx::K_GRAM, (double)(nv + nw) * D_ * esz_, 8.0 * nv * nw * D_);
    wait();
    settle(true);
    Mat C(nv, nw);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nw; ++j) C(i, j) = cplx(hc_[i * nw + j].x, hc_[i * nw + j].y);
    return C;
  }

  // Device copy of [A0, A1]^H W ((n0 + n1) x nw, row-major) at out; no sync.
  void gram_device(const void* A0, int n0, const void* A1, int n1, const void* W, int nw, double2* out,
                   StepPred pred = StepPred{}) {
    const int nv = n0 + n1;
    EDX_REQUIRE(nv <= kMaxVec + kMaxW && nw <= kMaxW, EDX_ERR_INTERNAL, "gram size");
    hipEvent_t t0 = ctx_.timer.start(ctx_.stream);
    if (prec_ == EDX_PREC_FP64) launch_tiles_nw<double2>(A0, n0, A1, n1, W, nw, pred);
    else launch_tiles_nw<float2>(A0, n0, A1, n1, W, nw, pred);
    EDX_HIP(hipGetLastError());
    const int nout = nv * nw;
    hipLaunchKernelGGL(gram_reduce, dim3((nout + 3) / 4), dim3(256), 0, ctx_.stream, n_rowblocks_, nout,
                       (const double2*)partial_.p, out, pred);
    EDX_HIP(hipGetLastError());
    ctx_.timer.stop(ctx_.stream, t0, edx::K_GRAM, (double)(nv + nw) * D_ * esz_, 8.0 * nv * nw * D_);
  }

  // Second Cholesky QR pass of the block at dst without a host round trip:
  // Gram, factorisation (block_cholinv) and update dst <- dst R2^-1 with its
  // interleaved copy, all enqueued; R2 lands in the pinned factor area and
  // R = R2 R1 is formed by settle() once the stream has passed it.
  void cholqr_pass_device(void* dst, const Mat& R1, Mat* Rtarget) {
    double2* g2 = fdev_.p;
    double2* s2 = fdev_.p + kMaxW * kMaxW;
    gram_device(dst, b_, nullptr, 0, dst, b_, g2);
    hipLaunchKernelGGL(block_cholinv, dim3(1), dim3(64), 0, ctx_.stream, b_, (const double2*)g2, 1e-8, s2, hf_dev_,
                       reinterpret_cast<int*>(hf_dev_ + kMaxW * kMaxW), (const double2*)nullptr, StepPred{});
    EDX_HIP(hipGetLastError());
    launch_gemm(nullptr, 0, dst, b_, dst, b_, s2, b_, X_.p);
    r_pending_ = true;
    r1_pending_ = R1;
    r_target_ = Rtarget;
  }

 public:
  // Completes a pending device Cholesky pass (stream synchronised first
  // unless the caller just did): R = R2 R1 into the step's coupling.
  void settle(bool synced) {
    if (!r_pending_) return;
    if (!synced) wait();
    r_pending_ = false;
    const int st = *reinterpret_cast<volatile int*>(hf_ + kMaxW * kMaxW);
    EDX_REQUIRE(st == 0, EDX_ERR_INTERNAL, "device Cholesky QR pass met a non-positive pivot");
    Mat R2(b_, b_);
    for (int i = 0; i < b_; ++i)
      for (int j = 0; j < b_; ++j) R2(i, j) = cplx(hf_[i * b_ + j].x, hf_[i * b_ + j].y);
    *r_target_ = mul(R2, r1_pending_);
  }

 private:
  // Out[:, 0..nk) = [A0, A1] S with S ((n0 + n1) x nk) given on the host.  S
  // goes through a ring of pinned staging slots; the stream is synchronised
  // (by a Gram) between any four consecutive uploads, so a slot is never
  // rewritten while its copy is pending.
  void gemm_host(const void* A0, int n0, const void* A1, int n1, void* Out, int nk, const Mat& S, void* Xi) {
    const int nv = n0 + n1;
    EDX_REQUIRE((size_t)nv * nk <= kSlot && nk <= kMaxW, EDX_ERR_INTERNAL, "staging slot too small");
    double2* hs = hs_ + (size_t)slot_ * kSlot;
    double2* ds = sdev_.p + (size_t)slot_ * kSlot;
    slot_ = (slot_ + 1) % kSlots;
    for (int i = 0; i < nv; ++i)
      for (int k = 0; k < nk; ++k) hs[(size_t)i * nk + k] = make_double2(S(i, k).real(), S(i, k).imag());
    EDX_HIP(hipMemcpyAsync(ds, hs, (size_t)nv * nk * sizeof(double2), hipMemcpyHostToDevice, ctx_.stream));
    launch_gemm(A0, n0, A1, n1, Out, nk, ds, nk, Xi);
  }

  template <typename VT, int NK, int U = kGemmStage, int RPT = 1>
  void launch_gemm_t(const void* A0, int n0, const void* A1, int n1, void* Out, int nk, const double2* S, int ldS,
                     void* Xi, StepPred pred) {
    launch_gemm_tiles<VT, NK, U, RPT>(ctx_.stream, D_, ld_, A0, n0, A1, n1, Out, nk, S, ldS, Xi, pred);
  }

  // A block update of at most 8 outputs with the schedule of this solve.
  template <typename VT, int NK>
  void launch_update(const void* A0, int n0, const void* A1, int n1, void* Out, int nk, const double2* S, int ldS,
                     void* Xi, StepPred pd) {
    static_assert(kUpdateVariants[0].stage == 2 && kUpdateVariants[0].rows == 2 && kUpdateVariants[1].stage == 2 &&
                      kUpdateVariants[1].rows == 1 && kUpdateVariants[2].stage == 4 && kUpdateVariants[2].rows == 1,
                  "update variants and their instantiations below must agree");
    switch (upd_var_) {
      case 1: launch_gemm_t<VT, NK, 2, 1>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd); break;
      case 2: launch_gemm_t<VT, NK, 4, 1>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd); break;
      default: launch_gemm_t<VT, NK, 2, 2>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd);
    }
  }

  // Fastest of three individually timed runs of `f` (after one warm-up) on
  // the solve's stream, in milliseconds.
  template <class F>
  float time_fastest(F&& f) {
    f();
    hipEvent_t ev[4];
    for (auto& e : ev) EDX_HIP(hipEventCreate(&e));
    EDX_HIP(hipEventRecord(ev[0], ctx_.stream));
    for (int r = 0; r < 3; ++r) {
      f();
      EDX_HIP(hipEventRecord(ev[r + 1], ctx_.stream));
    }
    EDX_HIP(hipEventSynchronize(ev[3]));
    float best = INFINITY;
    for (int r = 0; r < 3; ++r) {
      float ms = 0.0f;
      EDX_HIP(hipEventElapsedTime(&ms, ev[r], ev[r + 1]));
      best = std::min(best, ms);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return best;
  }

  // Orthogonalisation schedule of this solve: the context's fixed choice,
  // else (tuner on, D >= kScheduleTuneRows, b <= 8) the process-wide table
  // entry of (precision, b, floor(log2 D), basis size), timed on a miss with
  // this solve's own banks at the shape of a block step three quarters into
  // the basis (a Gram of [Q, W] against W and the update [Q, W] S into the
  // other bank and X); else the defaults (matrix-core Gram for complex64
  // from kMfmaGramRows rows; update 2/2 for complex64, 2/1 for complex128).
  // The banks hold no
  // vectors yet: the timed kernels read them and write only areas that the
  // run overwrites before use.
  void choose_schedule() {
    gram_mode_ = (esz_ == 8 && D_ >= kMfmaGramRows) ? kGramMatrix : kGramVector;
    upd_var_ = esz_ == 8 ? 0 : 1;
    const bool gram_fixed = ctx_.sched_gram != 0, upd_fixed = ctx_.sched_update != 0;
    if (gram_fixed) gram_mode_ = ctx_.sched_gram;
    if (upd_fixed) upd_var_ = ctx_.sched_update - 1;
    if ((gram_fixed && upd_fixed) || b_ > 8 || D_ < kScheduleTuneRows || !edx::tuning_enabled()) return;
    const edx::SchedKey key{prec_, b_, (int)std::floor(std::log2((double)D_)), nq_max_};
    edx::SchedEntry e;
    const edx::SchedLookup look = edx::schedule_lookup(key, &e);
    if (look == edx::SchedLookup::unavailable) return;
    if (look == edx::SchedLookup::claimed) {
      const bool timer_on = ctx_.timer.on;
      ctx_.timer.on = false;
      try {
        const int nq = std::max(b_, nq_max_ * 3 / 4 / b_ * b_);
        void* Q = vec(0, 0);
        void* Out = vec(1, 0);
        EDX_HIP(hipMemsetAsync(sbuf_.p, 0, (size_t)(nq + b_) * b_ * sizeof(double2), ctx_.stream));
        const int gram_modes = esz_ == 8 ? 2 : 1;
        for (int g = 0; g < gram_modes; ++g) {
          gram_mode_ = g + 1;
          e.gram_ms[g] = time_fastest([&] { gram_device(Q, nq, W_.p, b_, W_.p, b_, pbuf_.p); });
        }
        // another schedule replaces the default only when it is faster by
        // more than kScheduleMargin (differences below it are timing noise
        // at this synthetic shape and did not carry over to solves)
        const int gram_default = (esz_ == 8 && D_ >= kMfmaGramRows) ? kGramMatrix : kGramVector;
        e.gram = gram_default;
        if (gram_modes == 2) {
          const int other = gram_default == kGramMatrix ? kGramVector : kGramMatrix;
          if (e.gram_ms[other - 1] < (1.0f - kScheduleMargin) * e.gram_ms[gram_default - 1]) e.gram = other;
        }
        for (int v = 0; v < 3; ++v) {
          upd_var_ = v;
          e.update_ms[v] = time_fastest([&] { launch_gemm(Q, nq, W_.p, b_, Out, b_, sbuf_.p, b_, X_.p); });
        }
        const int upd_default = esz_ == 8 ? 0 : 1;
        int best = upd_default;
        for (int v = 0; v < 3; ++v)
          if (e.update_ms[v] < (1.0f - kScheduleMargin) * e.update_ms[best]) best = v;
        e.update = best + 1;
      } catch (...) {
        ctx_.timer.on = timer_on;
        edx::schedule_release(key);
        throw;
      }
      ctx_.timer.on = timer_on;
      edx::schedule_store(key, e);
    }
    if (!gram_fixed) gram_mode_ = e.gram;
    if (!upd_fixed) upd_var_ = e.update - 1;
  }

  // One tile of the smallest width that holds nk (nk <= 16), otherwise the
  // context's combination schedule (vector tiles of up to 12 outputs: 97 us
  // against 129 us with tiles of 6 at 60 -> 33 vectors, complex64, D = 225k).
  // Returns the kernel-timer class of the launch.
  template <typename VT>
  int launch_gemm_v(const void* A0, int n0, const void* A1, int n1, void* Out, int nk, const double2* S, int ldS,
                     void* Xi, StepPred pd) {
    if (nk > 16)
      return launch_combine<VT>(ctx_.stream, D_, ld_, A0, n0, A1, n1, Out, nk, S, ldS, ctx_.sched_combine, false, pd);
    if (nk <= 2) launch_update<VT, 2>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd);
    else if (nk <= 4) launch_update<VT, 4>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd);
    else if (nk <= 6) launch_update<VT, 6>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd);
    else if (nk <= 8) launch_update<VT, 8>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd);
    else if (nk <= 12) launch_gemm_t<VT, 12>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd);
    else launch_gemm_t<VT, 16>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pd);
    return edx::K_GEMM;
  }

  // True when the nk output vectors at Out overlap the n vectors at A.
  bool overlaps(const void* A, int n, const void* Out, int nk) const {
    if (!A || n == 0) return false;
    const char* a = static_cast<const char*>(A);
    const char* o = static_cast<const char*>(Out);
    const size_t vb = (size_t)ld_ * esz_;
    return o < a + n * vb && a < o + nk * vb;
  }

  void launch_gemm(const void* A0, int n0, const void* A1, int n1, void* Out, int nk, const double2* S, int ldS,
                   void* Xi, StepPred pred = StepPred{}) {
    EDX_REQUIRE(nk >= 1 && nk <= kMaxVec + kMaxW && n0 + n1 >= 1 && n0 + n1 <= kMaxVec + kMaxW, EDX_ERR_INTERNAL,
                "update width");
    // several output tiles run in separate workgroups: the outputs must not
    // be inputs, and the interleaved copy is written by single-tile updates
    EDX_REQUIRE(nk <= 16 || (!Xi && !overlaps(A0, n0, Out, nk) && !overlaps(A1, n1, Out, nk)), EDX_ERR_INTERNAL,
                "multi-tile update aliases its inputs");
    hipEvent_t t0 = ctx_.timer.start(ctx_.stream);
    const int cls = prec_ == EDX_PREC_FP64 ? launch_gemm_v<double2>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pred)
                                           : launch_gemm_v<float2>(A0, n0, A1, n1, Out, nk, S, ldS, Xi, pred);
    EDX_HIP(hipGetLastError());
    ctx_.timer.stop(ctx_.stream, t0, cls,
                    (double)(n0 + n1 + (Xi ? 2 : 1) * nk) * D_ * esz_, 8.0 * (n0 + n1) * nk * D_);
  }

  // Out[:, 0..nk) = V[:, 0..nv) Y (Out must not alias V), one launch over
  // all output tiles, so the basis is read from HBM about once.
  // The coefficients go through the context's pinned combine area (the
  // stream is synchronised at the end, so the area is free again).
  void combine(const void* V, int nv, void* Out, int nk, const Mat& Y) {
    EDX_REQUIRE((size_t)nv * nk <= edx::kPinnedCombine, EDX_ERR_INTERNAL, "restart combination too large");
    double2* hs = ctx_.pinned_out + edx::kPinnedRing;
    for (int i = 0; i < nv; ++i)
      for (int k = 0; k < nk; ++k) hs[(size_t)i * nk + k] = make_double2(Y(i, k).real(), Y(i, k).imag());
    DevBuf<double2> S((size_t)nv * nk);
    EDX_HIP(hipMemcpyAsync(S.p, hs, (size_t)nv * nk * sizeof(double2), hipMemcpyHostToDevice, ctx_.stream));
    launch_gemm(V, nv, nullptr, 0, Out, nk, S.p, nk, nullptr);
    wait();
  }

  void random_block(void* out, int nvec, uint64_t seed) {
    if (prec_ == EDX_PREC_FP64)
      hipLaunchKernelGGL(fill_gaussian<double2>, dim3(edx::grid_for(D_, 256)), dim3(256), 0, ctx_.stream, D_, nvec,
                         (double2*)out, ld_, seed);
    else
      hipLaunchKernelGGL(fill_gaussian<float2>, dim3(edx::grid_for(D_, 256)), dim3(256), 0, ctx_.stream, D_, nvec,
                         (float2*)out, ld_, seed);
    EDX_HIP(hipGetLastError());
  }

  // Orthonormalises the block at V in place (Cholesky QR twice) and writes
  // its interleaved copy to X_.  False if it is numerically rank deficient.
  bool cholqr2(void* V, double tiny, Mat* R) {
    Mat Rt;
    for (int pass = 0; pass < 2; ++pass) {
      Mat G = gram_host(V, b_, nullptr, 0, V, b_);
      Mat Rp;
      if (!edx::dense::cholesky_upper(G, Rp, pass == 0 ? tiny : 1e-8)) return false;
      gemm_host(nullptr, 0, V, b_, V, b_, edx::dense::upper_inverse(Rp), pass == 1 ? X_.p : nullptr);
      Rt = pass == 0 ? Rp : mul(Rp, Rt);
    }
    if (R) *R = Rt;
    return true;
  }

  void start_block() {
    EDX_REQUIRE(cholqr2(vec(cur_, 0), 1e-24, nullptr), EDX_ERR_INTERNAL, "start block is rank deficient");
  }

  static Mat mul(const Mat& A, const Mat& B) {
    Mat C(A.n, B.m);
    for (int i = 0; i < A.n; ++i)
      for (int k = 0; k < A.m; ++k) {
        const cplx a = A(i, k);
        for (int j = 0; j < B.m; ++j) C(i, j) += a * B(k, j);
      }
    return C;
  }

  // Gram of W - Q C from the Gram of [Q, W] against W (Q orthonormal):
  // W^H W - C^H C.
  Mat remainder_gram(const Mat& P, int nbasis) const {
    Mat G(b_, b_);
    for (int i = 0; i < b_; ++i)
      for (int j = 0; j < b_; ++j) {
        cplx g = P(nbasis + i, j);
        for (int k = 0; k < nbasis; ++k) g -= std::conj(P(k, i)) * P(k, j);
        G(i, j) = g;
      }
    return G;
  }

  // S = [-C R^-1; R^-1] so that [Q, W] S = (W - Q C) R^-1.
  Mat update_coeffs(const Mat& C, int nbasis, const Mat* Rinv) const {
    Mat S(nbasis + b_, b_);
    for (int i = 0; i < nbasis; ++i)
      for (int j = 0; j < b_; ++j) {
        cplx sum(0.0, 0.0);
        if (Rinv)
          for (int k = 0; k < b_; ++k) sum += C(i, k) * (*Rinv)(k, j);
        else
          sum = C(i, j);
        S(i, j) = -sum;
      }
    for (int i = 0; i < b_; ++i)
      for (int j = 0; j < b_; ++j) S(nbasis + i, j) = Rinv ? (*Rinv)(i, j) : cplx(i == j ? 1.0 : 0.0, 0.0);
    return S;
  }

  // Log slot s of the pinned ring (host and device addresses).
  const double2* log_host(int s) const { return hc_ + edx::kPinnedGram + edx::kPinnedFactor + (size_t)s * edx::kLogSlot; }
  double2* log_dev(int s) const { return hc_dev_ + edx::kPinnedGram + edx::kPinnedFactor + (size_t)s * edx::kLogSlot; }

  // Enqueues the block step that follows a basis of q vectors (the SpMM has
  // been enqueued): Gram of [Q, W] against W, the coefficients (step_coeffs,
  // log slot `slot`), the update into vec(q + b) and, when the step's status
  // asks for it, the second Cholesky QR pass; nothing waits for the host.
  void enqueue_step(int q, int slot) {
    const int nbasis = q + b_;
    void* Q = vec(cur_, 0);
    void* dst = vec(cur_, nbasis);
    const StepPred p1{ctl_.p, 1}, p2{ctl_.p, 2};
    const double tiny = prec_ == EDX_PREC_FP64 ? 1e-24 : 1e-12;
    double2* lg = log_dev(slot);
    gram_device(Q, nbasis, W_.p, b_, W_.p, b_, pbuf_.p, p1);
    // complex64 steps always take a second Cholesky QR pass; FP64 steps rarely
    // do, so theirs is not enqueued and such a step goes to the host
    const int second = prec_ == EDX_PREC_FP64 ? 0 : 1;
    hipLaunchKernelGGL(step_coeffs, dim3(1), dim3(256), 0, ctx_.stream, (const double2*)pbuf_.p, nbasis, b_, kappa2_,
                       tiny, unit_roundoff_, single_qr_tol_, second, sbuf_.p, r1buf_.p, lg, ctl_.p);
    EDX_HIP(hipGetLastError());
    launch_gemm(Q, nbasis, W_.p, b_, dst, b_, sbuf_.p, b_, X_.p, p1);
    if (!second) return;
    gram_device(dst, b_, nullptr, 0, dst, b_, g2buf_.p, p2);
    hipLaunchKernelGGL(block_cholinv, dim3(1), dim3(64), 0, ctx_.stream, b_, (const double2*)g2buf_.p, 1e-8,
                       s2buf_.p, lg + kLogC, reinterpret_cast<int*>(lg + kLogC + kMaxW * kMaxW + 2),
                       (const double2*)r1buf_.p, p2);
    EDX_HIP(hipGetLastError());
    launch_gemm(nullptr, 0, dst, b_, dst, b_, s2buf_.p, b_, X_.p, p2);
  }

  // Adds block step q's projections C ((q + b) x b) to the projected matrix.
  void add_block(Mat& M, const Mat& C, int q) const {
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < b_; ++j) {
        M(i, q + j) = C(i, j);
        M(q + j, i) = std::conj(C(i, j));
      }
    for (int i = 0; i < b_; ++i)
      for (int j = 0; j < b_; ++j) M(q + i, q + j) = 0.5 * (C(q + i, j) + std::conj(C(q + j, i)));
  }

  // Folds the logs of the enqueued steps (stream synchronised) into M and the
  // newest coupling; a step the device handed over (status 2) is finished on
  // the host and ends the list (the steps after it were skipped).  Returns the
  // basis size; n_acc = steps accepted.
  int absorb(const std::vector<int>& pending, Mat& M, Result& res, int& n_acc) {
    int nq = pending.empty() ? 0 : pending[0];
    n_acc = 0;
    for (size_t s = 0; s < pending.size(); ++s) {
      const int q = pending[s];
      const int nbasis = q + b_;
      const double2* lg = log_host((int)s);
      const double2* tail = lg + kLogC + kMaxW * kMaxW;
      const int st = (int)tail[0].x;
      Mat C(nbasis, b_);
      if (st == 2) {
        resolve(q, C, Rstep_);
      } else {
        EDX_REQUIRE(st == 0 || st == 1, EDX_ERR_INTERNAL, "block step log without a status");
        if (st == 1)
          EDX_REQUIRE(*reinterpret_cast<const volatile int*>(tail + 2) == 0, EDX_ERR_INTERNAL,
                      "device Cholesky QR pass met a non-positive pivot");
        for (int i = 0; i < nbasis; ++i)
          for (int j = 0; j < b_; ++j) C(i, j) = cplx(lg[i * b_ + j].x, lg[i * b_ + j].y);
        Rstep_ = Mat(b_, b_);
        for (int i = 0; i < b_; ++i)
          for (int j = 0; j < b_; ++j) Rstep_(i, j) = cplx(lg[kLogC + i * b_ + j].x, lg[kLogC + i * b_ + j].y);
      }
      add_block(M, C, q);
      res.matvecs += b_;
      nq = nbasis;
      ++n_acc;
      if (st == 2) break;
    }
    return nq;
  }

  // Finishes on the host the block step after a basis of q vectors that the
  // device handed over: its Gram is still in pbuf_ (the predicated kernels
  // after it did nothing, and W_ = H Vr is unchanged), then the device steps
  // are released.
  void resolve(int q, Mat& C, Mat& R) {
    const int nbasis = q + b_;
    const int nv = nbasis + b_;
    std::vector<double2> h((size_t)nv * b_);
    EDX_HIP(hipMemcpyAsync(h.data(), pbuf_.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx_.stream));
    wait();
    Mat P(nv, b_);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < b_; ++j) P(i, j) = cplx(h[(size_t)i * b_ + j].x, h[(size_t)i * b_ + j].y);
    extend(P, vec(cur_, 0), nbasis, vec(cur_, nbasis), C, R);
    EDX_HIP(hipMemsetAsync(ctl_.p, 0, 4 * sizeof(int), ctx_.stream));
  }

  // Projects W_ = H Vr against the basis Q (nbasis vectors, Vr included),
  // given the Gram P = [Q, W]^H W: C = Q^H W (summed over passes), and writes
  // the next block dst with W - Q C = dst R (dst orthonormal, orthogonal to
  // Q) and its interleaved copy to X_.  Host-driven (waits on the stream).
  void extend(const Mat& P, void* Q, int nbasis, void* dst, Mat& C, Mat& R) {
    const double tiny = prec_ == EDX_PREC_FP64 ? 1e-24 : 1e-12;
    C = Mat(nbasis, b_);
    for (int i = 0; i < nbasis; ++i)
      for (int j = 0; j < b_; ++j) C(i, j) = P(i, j);
    Mat G = remainder_gram(P, nbasis);
    double drop = norm_drop(P, G, nbasis);
    Mat Cl = C;  // coefficients of the pass still to be applied to W_
    Mat R1;
    if (!(drop < kappa2_) || !edx::dense::cholesky_upper(G, R1, tiny)) {
      // apply the first pass explicitly, then project again
      gemm_host(Q, nbasis, W_.p, b_, W_.p, b_, update_coeffs(C, nbasis, nullptr), nullptr);
      Mat P2 = gram_host(Q, nbasis, W_.p, b_, W_.p, b_);
      for (int i = 0; i < nbasis; ++i)
        for (int j = 0; j < b_; ++j) {
          Cl(i, j) = P2(i, j);
          C(i, j) += P2(i, j);
        }
      G = remainder_gram(P2, nbasis);
      drop = norm_drop(P2, G, nbasis);
      if (!edx::dense::cholesky_upper(G, R1, tiny)) {
        // W - Q C is numerically rank deficient
        gemm_host(Q, nbasis, W_.p, b_, W_.p, b_, update_coeffs(Cl, nbasis, nullptr), nullptr);
        deflate(Q, nbasis, dst, R);
        return;
      }
    }
    // dst = (W - Q Cl) R1^-1.  Its Gram differs from the identity by about
    // eps * drop * cond(R1)^2 (the error of the Gram G, formed by
    // subtraction, carried through R1^-1); when that is below single_qr_tol_
    // dst is final, otherwise a second Cholesky QR pass is made on it.
    const Mat R1i = edx::dense::upper_inverse(R1);
    const double cond = frobenius(R1) * frobenius(R1i);
    const bool single = unit_roundoff_ * std::max(drop, 1.0) * cond * cond <= single_qr_tol_;
    gemm_host(Q, nbasis, W_.p, b_, dst, b_, update_coeffs(Cl, nbasis, &R1i), single ? X_.p : nullptr);
    if (single) {
      R = R1;
      return;
    }
    // With the Gram of dst within 1/4 of the identity (by the estimate) its
    // Cholesky factorisation cannot fail: the second pass runs on the device
    // without a host round trip
    if (unit_roundoff_ * std::max(drop, 1.0) * cond * cond <= 0.25) {
      cholqr_pass_device(dst, R1, &R);
      return;
    }
    Mat G2 = gram_host(dst, b_, nullptr, 0, dst, b_);
    Mat R2;
    if (edx::dense::cholesky_upper(G2, R2, 1e-8)) {
      gemm_host(nullptr, 0, dst, b_, dst, b_, edx::dense::upper_inverse(R2), X_.p);
      R = mul(R2, R1);
      return;
    }
    gemm_host(Q, nbasis, W_.p, b_, W_.p, b_, update_coeffs(Cl, nbasis, nullptr), nullptr);
    deflate(Q, nbasis, dst, R);
  }

  // Largest ratio ||w_j||^2 / ||w_j - Q c_j||^2 over the block's columns.
  double norm_drop(const Mat& P, const Mat& G, int nbasis) const {
    double d = 0.0;
    for (int j = 0; j < b_; ++j) {
      const double g = G(j, j).real();
      d = std::max(d, g > 0.0 ? P(nbasis + j, j).real() / g : INFINITY);
    }
    return d;
  }

  static double frobenius(const Mat& A) {
    double s = 0.0;
    for (const cplx& x : A.a) s += std::norm(x);
    return std::sqrt(s);
  }

  // W_ (orthogonal to Q) is numerically rank deficient: the next block is
  // made of fresh random directions orthogonal to Q, and R is the projection
  // of W_ on them.
  void deflate(void* Q, int nbasis, void* dst, Mat& R) {
    const double tiny = prec_ == EDX_PREC_FP64 ? 1e-24 : 1e-12;
    for (int attempt = 0; attempt < 3; ++attempt) {
      random_block(dst, b_, opt_.seed * 7919 + 104729 * (uint64_t)(++deflations_));
      for (int pass = 0; pass < 2; ++pass) {
        Mat P = gram_host(Q, nbasis, nullptr, 0, dst, b_);
        gemm_host(Q, nbasis, dst, b_, dst, b_, update_coeffs(P, nbasis, nullptr), nullptr);
      }
      if (!cholqr2(dst, tiny, nullptr)) continue;
      R = gram_host(dst, b_, nullptr, 0, W_.p, b_);
      return;
    }
    throw edx::Error{EDX_ERR_NOCONV, "could not extend the Krylov basis (sector dimension exhausted)"};
  }

  // Ritz values of the projected matrix M (nq x nq), the ncol leading Ritz
  // vectors (Yc, nq x ncol) and rows nq-b .. nq-1 of every Ritz vector (Yl).
  void ritz(const Mat& M, int nq, std::vector<double>& theta, int ncol, Mat* Yc, Mat* Yl) {
    const double t0 = now_ms();
    Mat A(nq, nq);
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j) A(i, j) = M(i, j);
    edx::dense::herm_eig_sel(A, true, theta, ncol, Yc, Yl ? nq - b_ : -1, Yl);
    ms_ritz_ += now_ms() - t0;
    ++checks_;
  }

  // Blocks until the stream has finished its work (time and count kept).
  void wait() {
    const double t0 = now_ms();
    EDX_HIP(hipStreamSynchronize(ctx_.stream));
    ms_wait_ += now_ms() - t0;
    ++syncs_;
  }

  void finish(Result& res, double t_start) {
    res.ms_total = now_ms() - t_start;
    res.ms_wait = ms_wait_;
    res.ms_ritz = ms_ritz_;
    res.syncs = syncs_;
    res.checks = checks_;
  }

  edx_ctx& ctx_;
  const edx_hamiltonian& h_;
  int prec_;
  Options opt_;
  int64_t D_ = 0, ld_ = 0;
  int b_ = 1, esz_ = 16, nq_max_ = 0, cap_ = 0, cur_ = 0, restarts_ = 0;
  // state of a run between begin() and the end of sync_point()
  Result res_;
  Mat M_, B_;
  std::vector<double> theta_;
  std::vector<int> pending_;  // basis size before each block step enqueued since the last sync
  int nq_ = 0, since_check_ = 0, since_best_ = 0;
  double best_worst_ = INFINITY, t_start_ = 0.0;
  bool full_ = false, due_ = false;
  int gram_mode_ = kGramVector;  // Gram schedule (kGramVector / kGramMatrix)
  int upd_var_ = 0;              // block-update schedule (index into kUpdateVariants)
  double kappa2_ = 1e4;
  double unit_roundoff_ = 1.1e-16;
  double single_qr_tol_ = 0.0;
  uint64_t deflations_ = 0;
  int n_rowblocks_ = 1;
  int64_t rows_per_block_ = 0;
  int slot_ = 0;
  DevBuf<char> bank_[2];
  DevBuf<char> W_, X_;
  DevBuf<double2> partial_, sdev_;
  double2* hc_ = nullptr;      // pinned Gram results (host address)
  double2* hc_dev_ = nullptr;  // the same buffer's device address
  double2* hf_ = nullptr;      // pinned factor of the device Cholesky pass (host address)
  double2* hf_dev_ = nullptr;  // its device address
  DevBuf<double2> fdev_;       // device Gram (b x b) and coefficients of the device Cholesky pass
  // device block steps: Gram, update coefficients, R1, second-pass Gram and
  // coefficients, and the status / halt words
  DevBuf<double2> pbuf_, sbuf_, r1buf_, g2buf_, s2buf_;
  DevBuf<int> ctl_;
  // a block step whose second Cholesky QR pass runs on the device leaves its
  // coupling R = R2 R1 pending until the stream has passed the factorisation
  bool r_pending_ = false;
  Mat r1_pending_;
  Mat* r_target_ = nullptr;
  Mat Rstep_;  // R of the newest block step
  double ms_wait_ = 0.0, ms_ritz_ = 0.0;
  int syncs_ = 0, checks_ = 0;
  double2* hs_ = nullptr;
};

// Dense path for small sectors: explicit block, eigen-decomposition of the
// Hermitian matrix formed from its upper triangle (the reference's
// eigh(UPLO="U") for n <= 500).
void dense_solve(edx_ctx& ctx, const edx_hamiltonian& h, int nev, double* vals, double* vecs) {
  const int64_t D = h.dim;
  DevBuf<double2> dmat((size_t)D * D);
  edx::dense_matrix(h, ctx, dmat.p);
  std::vector<double2> hm((size_t)D * D);
  EDX_HIP(hipMemcpyAsync(hm.data(), dmat.p, hm.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx.stream));
  EDX_HIP(hipStreamSynchronize(ctx.stream));
  Mat A((int)D, (int)D);
  for (int64_t i = 0; i < D; ++i)
    for (int64_t j = 0; j < D; ++j) A((int)i, (int)j) = cplx(hm[i * D + j].x, hm[i * D + j].y);
  std::vector<double> w;
  Mat Z;
  edx::dense::herm_eig(A, w, Z, false);
  const int k = (int)std::min<int64_t>(nev, D);
  for (int c = 0; c < k; ++c) {
    vals[c] = w[c];
    if (vecs)
      for (int64_t r = 0; r < D; ++r) {
        vecs[2 * ((size_t)c * D + r)] = Z((int)r, c).real();
        vecs[2 * ((size_t)c * D + r) + 1] = Z((int)r, c).imag();
      }
  }
}

void eigs_impl(edx_ctx* ctx, const edx_hamiltonian* ham, const edx_eigs_options& o, double* vals, double* vecs,
               double* info);
void eigs_batch_impl(edx_ctx* bctx, int n, edx_ctx* const* ctxs, const edx_hamiltonian* const* hams,
                     const edx_eigs_options& o, double* vals, double* const* vecs, double* info);

}  // namespace

extern "C" {

int edx_eigs_batch(edx_ctx* ctx, int32_t n, edx_ctx* const* ctxs, const edx_hamiltonian* const* hams,
                   const edx_eigs_options* o, double* vals, double* const* vecs, double* info) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && o && vals && n >= 1 && ctxs && hams, EDX_ERR_INVALID, "null argument or n < 1");
    // every solve drives its own context from this thread
    std::vector<std::unique_ptr<edx::CtxClaim>> claims;
    for (int i = 0; i < n; ++i) {
      EDX_REQUIRE(ctxs[i] && hams[i], EDX_ERR_INVALID, "null context or Hamiltonian");
      EDX_REQUIRE(ctxs[i] != ctx && ctxs[i]->device == ctx->device, EDX_ERR_INVALID,
                  "each solve needs its own context on the batch context's device");
      for (int j = 0; j < i; ++j) EDX_REQUIRE(ctxs[j] != ctxs[i], EDX_ERR_INVALID, "contexts must be distinct");
      claims.emplace_back(new edx::CtxClaim(ctxs[i]));
      if (!claims.back()->ok) throw edx::Error{EDX_ERR_BUSY, "a solve's context is in use by another thread"};
    }
    eigs_batch_impl(ctx, n, ctxs, hams, *o, vals, vecs, info);
  });
}

int edx_combine_vectors(edx_ctx* ctx, int32_t precision, const void* v_dev, int64_t ld, int64_t dim, int32_t nv,
                        const double* y, int32_t nk, void* out_dev, int32_t schedule) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && v_dev && y && out_dev, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(precision == EDX_PREC_FP64 || precision == EDX_PREC_MIXED, EDX_ERR_INVALID, "bad precision");
    EDX_REQUIRE(schedule >= 0 && schedule <= 2, EDX_ERR_INVALID, "combination schedule must be 0 (automatic), 1 or 2");
    // the vector schedule's lanes past dim read (and discard) row dim
    EDX_REQUIRE(dim >= 1 && ld > dim, EDX_ERR_INVALID, "need dim >= 1 and ld > dim");
    EDX_REQUIRE(nv >= 1 && nv <= kMaxVec + kMaxW && nk >= 1 && nk <= kMaxVec + kMaxW, EDX_ERR_INVALID,
                "nv and nk must be 1..240");
    EDX_REQUIRE(precision == EDX_PREC_MIXED || schedule != kCombineMatrix, EDX_ERR_INVALID,
                "the matrix-core combination takes complex64 vectors");
    const size_t esz = precision == EDX_PREC_FP64 ? sizeof(double2) : sizeof(float2);
    const char* vb = static_cast<const char*>(v_dev);
    const char* ob = static_cast<const char*>(out_dev);
    EDX_REQUIRE(ob + (size_t)nk * ld * esz <= vb || vb + (size_t)nv * ld * esz <= ob, EDX_ERR_INVALID,
                "the output overlaps the input vectors");
    edx::DeviceGuard dg(ctx->device);
    DevBuf<double2> S((size_t)nv * nk);
    EDX_HIP(hipMemcpyAsync(S.p, y, (size_t)nv * nk * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
    hipEvent_t t0 = ctx->timer.start(ctx->stream);
    const int cls =
        precision == EDX_PREC_FP64
            ? launch_combine<double2>(ctx->stream, dim, ld, v_dev, nv, nullptr, 0, out_dev, nk, S.p, nk, schedule,
                                      true, StepPred{})
            : launch_combine<float2>(ctx->stream, dim, ld, v_dev, nv, nullptr, 0, out_dev, nk, S.p, nk, schedule,
                                     true, StepPred{});
    EDX_HIP(hipGetLastError());
    ctx->timer.stop(ctx->stream, t0, cls, (double)(nv + nk) * dim * esz, 8.0 * nv * nk * dim);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int edx_set_combine_schedule(edx_ctx* ctx, int32_t schedule) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(schedule >= 0 && schedule <= 2, EDX_ERR_INVALID, "combination schedule must be 0 (automatic), 1 or 2");
    ctx->sched_combine = schedule;
  });
}

int edx_batch_graph_stats(edx_ctx* ctx, int64_t* captures, int64_t* replays) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx, EDX_ERR_INVALID, "null argument");
    if (captures) *captures = ctx->graph_captures;
    if (replays) *replays = ctx->graph_replays;
  });
}

int edx_eigs(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t nev, int32_t block, double tol, uint64_t seed,
             int32_t precision, double* vals, double* vecs, double* info) {
  edx_eigs_options o;
  edx_eigs_default_options(&o);
  o.nev = nev;
  o.block = block;
  o.tol = tol;
  o.seed = seed;
  o.precision = precision;
  return edx_eigs_ex(ctx, ham, &o, vals, vecs, info);
}

void edx_eigs_default_options(edx_eigs_options* o) {
  if (!o) return;
  o->nev = 1;
  o->block = 0;
  o->basis = 0;
  o->precision = EDX_PREC_FP64;
  o->max_restarts = 5000;
  o->tol = 1e-10;
  o->eig_tol = 0.0;
  o->mixed_tol = 3e-7;
  o->seed = 12345;
}

int edx_eigs_ex(edx_ctx* ctx, const edx_hamiltonian* ham, const edx_eigs_options* o, double* vals, double* vecs,
                double* info) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham && o && vals, EDX_ERR_INVALID, "null argument");
    eigs_impl(ctx, ham, *o, vals, vecs, info);
  });
}

}  // extern "C"

namespace {

void eigs_impl(edx_ctx* ctx, const edx_hamiltonian* ham, const edx_eigs_options& o, double* vals, double* vecs,
               double* info) {
  {
    const int nev = o.nev, block = o.block, precision = o.precision;
    const double tol = o.tol;
    const uint64_t seed = o.seed;
    EDX_REQUIRE(nev >= 1, EDX_ERR_INVALID, "nev must be >= 1");
    EDX_REQUIRE(precision == EDX_PREC_FP64 || precision == EDX_PREC_MIXED, EDX_ERR_INVALID, "bad precision");
    EDX_REQUIRE(tol > 0.0, EDX_ERR_INVALID, "tol must be positive");
    EDX_REQUIRE(o.max_restarts >= 1, EDX_ERR_INVALID, "max_restarts must be >= 1");
    edx::DeviceGuard dg(ctx->device);
    const int64_t D = ham->dim;
    if (info) std::fill(info, info + 4, 0.0);
    if (D == 0) return;
    const double t_all = now_ms();
    if (D <= 500) {
      dense_solve(*ctx, *ham, nev, vals, vecs);
      if (info) info[3] = 0;
      return;
    }
    EDX_REQUIRE(nev < D - 1, EDX_ERR_INVALID, "nev must be < sector dimension - 1");
    int b = block > 0 ? block : std::min(nev, 8);
    b = std::max(1, std::min(b, kMaxW));
    Options opt;
    opt.nev = nev;
    opt.block = b;
    opt.basis = o.basis;
    opt.tol = tol;
    opt.eig_tol = o.eig_tol;
    opt.max_restarts = o.max_restarts;
    opt.seed = seed;
    Result res;
    int64_t mv_total = 0;
    // EDX_TRACE=1 prints the work of each phase to stderr
    static const bool trace = [] {
      const char* e = getenv("EDX_TRACE");
      return e && *e && *e != '0';
    }();
    std::unique_ptr<BlockLanczos> fin;
    int fin_block = b;
    if (precision == EDX_PREC_MIXED) {
      // complex64 sweeps resolve the nev pairs to mixed_tol, then an FP64
      // block Lanczos starts from those Ritz vectors and the rest of the
      // nev-th pair's cluster (all in its first block), checks convergence
      // every step and judges clusters and gaps against the complex64 basis
      const int n32 = (int)std::min<int64_t>(std::min(nev, kMaxW), D - 2);
      Options o32 = opt;
      o32.nev = n32;
      o32.tol = std::max(tol, o.mixed_tol > 0 ? o.mixed_tol : 3e-7);
      o32.eig_tol = 0.0;
      o32.out_cluster = true;
      // a target below the complex64 rounding floor still hands over: the
      // sweeps stop once they stall anywhere below 1e-4
      o32.stall_checks = 8;
      o32.stall_ceiling = std::max(30.0 * o32.tol, 1e-4);
      BlockLanczos lz32(*ctx, *ham, EDX_PREC_MIXED, o32);
      Result r32 = lz32.run(nullptr, 0, 0);
      mv_total += r32.matvecs;
      if (trace)
        fprintf(stderr,
                "[edx eigs] D=%lld complex64 phase: nev=%d block=%d matvecs=%lld restarts=%d res=%.2e | host %.1f ms: "
                "wait %.1f (%d syncs), ritz %.1f (%d checks)\n",
                (long long)D, n32, b, (long long)r32.matvecs, r32.restarts, r32.max_rel_res, r32.ms_total, r32.ms_wait,
                r32.syncs, r32.ms_ritz, r32.checks);
      const int n_start = std::max(n32, std::min(r32.nvec, kMaxW));
      Options o64 = opt;
      o64.block = std::min(kMaxW, std::max(b, n_start));
      o64.basis = 0;
      o64.check_every = 1;
      o64.gap_ref = r32.ritz;
      o64.gap_ref_res = r32.ritz_res;
      fin_block = o64.block;
      fin.reset(new BlockLanczos(*ctx, *ham, EDX_PREC_FP64, o64));
      DevBuf<double2> start((size_t)n_start * lz32.ld());
      hipLaunchKernelGGL((convert_vec<float2, double2>), dim3(edx::grid_for(D, 256)), dim3(256), 0, ctx->stream, D,
                         n_start, (const float2*)lz32.vec(lz32.current(), 0), lz32.ld(), start.p, lz32.ld());
      EDX_HIP(hipGetLastError());
      res = fin->run(start.p, n_start, lz32.ld());
      res.restarts += r32.restarts;
    } else {
      opt.check_every = 2;
      fin.reset(new BlockLanczos(*ctx, *ham, EDX_PREC_FP64, opt));
      res = fin->run(nullptr, 0, 0);
    }
    mv_total += res.matvecs;
    if (trace)
      fprintf(stderr,
              "[edx eigs] D=%lld FP64 phase: nev=%d block=%d matvecs=%lld restarts=%d res=%.2e | host %.1f ms: "
              "wait %.1f (%d syncs), ritz %.1f (%d checks)\n",
              (long long)D, nev, fin_block, (long long)res.matvecs, res.restarts, res.max_rel_res, res.ms_total,
              res.ms_wait,
              res.syncs, res.ms_ritz, res.checks);
    for (int c = 0; c < nev; ++c) vals[c] = res.vals[c];
    const double t_copy = now_ms();
    if (vecs) edx::copy_to_host(*ctx, vecs, D * 16, fin->vec(fin->current(), 0), fin->ld() * 16, D * 16, nev);
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    if (trace)
      fprintf(stderr, "[edx eigs] D=%lld total %.1f ms, eigenvector copy %.1f ms\n", (long long)D, now_ms() - t_all,
              now_ms() - t_copy);
    if (info) {
      info[0] = (double)mv_total;
      info[1] = (double)res.restarts;
      info[2] = res.max_rel_res;
      info[3] = (double)b;
    }
  }
}

// Host threads that run the checks of a batch's solves.
constexpr int kBatchHostThreads = 8;

// One solve of a batch.
struct BatchSolve {
  int idx;
  edx_ctx* ctx;
  const edx_hamiltonian* h;
  std::unique_ptr<BlockLanczos> lz;
  hipEvent_t ev = nullptr;
  int64_t matvecs = 0;
  int restarts = 0;
};

// A captured HIP graph of one round's grouped SpMM launch(es).
struct RoundGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double bytes = 0.0;  // algorithmic bytes, for the kernel timer
};

// The round graphs of a batch phase, one per set of stepping solves.  Within
// a phase the launch arguments of a set (tables, interleaved inputs, outputs,
// leading dimensions) do not change, so a set's graph is captured once and
// replayed by every later round with the same set.
class RoundGraphs {
 public:
  explicit RoundGraphs(edx_ctx& bctx) : bctx_(bctx) {}
  RoundGraphs(const RoundGraphs&) = delete;
  RoundGraphs& operator=(const RoundGraphs&) = delete;
  ~RoundGraphs() { clear(); }

  // Enqueues the grouped SpMM of `jobs` (solve ids `key`) on the batch
  // stream as a replay of the set's graph, capturing it first if needed.
  void launch(const std::vector<int>& key, const std::vector<edx::SpmmJob>& jobs, int precision) {
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
      if (graphs_.size() >= kMaxGraphs) clear();
      it = graphs_.emplace(key, capture(jobs, precision)).first;
      ++captures_;
    }
    hipEvent_t t0 = bctx_.timer.start(bctx_.stream);
    EDX_HIP(hipGraphLaunch(it->second.exec, bctx_.stream));
    bctx_.timer.stop(bctx_.stream, t0, precision == EDX_PREC_FP64 ? edx::K_SPMM : edx::K_SPMM32, it->second.bytes);
    ++replays_;
  }
  int captures() const { return captures_; }
  int replays() const { return replays_; }

 private:
  static constexpr size_t kMaxGraphs = 64;

  RoundGraph capture(const std::vector<edx::SpmmJob>& jobs, int precision) {
    RoundGraph r;
    r.bytes = edx::spmm_group_bytes(jobs.data(), (int)jobs.size(), precision);
    // the launches only: the kernel timer's events stay out of the graph
    const bool timer_on = bctx_.timer.on;
    bctx_.timer.on = false;
    hipError_t cap = hipStreamBeginCapture(bctx_.stream, hipStreamCaptureModeThreadLocal);
    try {
      EDX_HIP(cap);
      edx::spmm_group(jobs.data(), (int)jobs.size(), bctx_, precision);
      EDX_HIP(hipStreamEndCapture(bctx_.stream, &r.graph));
      cap = hipErrorUnknown;  // capture ended
      EDX_HIP(hipGraphInstantiate(&r.exec, r.graph, nullptr, nullptr, 0));
    } catch (...) {
      bctx_.timer.on = timer_on;
      if (cap == hipSuccess) {  // still capturing: end it so the stream is usable again
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(bctx_.stream, &partial);
        if (partial) (void)hipGraphDestroy(partial);
      }
      if (r.exec) (void)hipGraphExecDestroy(r.exec);
      if (r.graph) (void)hipGraphDestroy(r.graph);
      throw;
    }
    bctx_.timer.on = timer_on;
    return r;
  }
  // Replays already enqueued stay valid: HIP keeps a graph's resources until
  // its launches have completed.
  void clear() {
    for (auto& kv : graphs_) {
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    graphs_.clear();
  }

  edx_ctx& bctx_;
  std::map<std::vector<int>, RoundGraph> graphs_;
  int captures_ = 0, replays_ = 0;
};

// Runs the BlockLanczos of every solve to convergence together.  Each round,
// the solves that are stepping hand their SpMM to one grouped launch on the
// batch context's stream (which first waits for each of their
// streams), replayed from the round graph of that set of solves, then enqueue
// the rest of their block step on their own streams (which wait for the
// grouped launch); a solve whose step asks for the host leaves the round,
// and once none is stepping, each waits for its own stream and runs its
// check (on a few host threads).  Solves that converge leave the batch.
// `starts[i]` (optional) seeds solve i as BlockLanczos::run does.
void run_batch_phase(edx_ctx& bctx, std::vector<BatchSolve*>& solves, int precision,
                     const std::vector<std::pair<const void*, int>>& starts, int* graph_counts) {
  for (size_t i = 0; i < solves.size(); ++i) {
    BatchSolve* s = solves[i];
    edx::StreamScope scope(s->ctx->stream);
    s->lz->begin(starts[i].first, starts[i].second, s->lz->ld());
  }
  hipEvent_t done_ev;
  EDX_HIP(hipEventCreateWithFlags(&done_ev, hipEventDisableTiming));
  RoundGraphs graphs(bctx);
  try {
    std::vector<BatchSolve*> active = solves;
    std::vector<edx::SpmmJob> jobs;
    std::vector<int> key;
    while (!active.empty()) {
      std::vector<BatchSolve*> stepping = active;
      while (!stepping.empty()) {
        jobs.clear();
        key.clear();
        for (BatchSolve* s : stepping) {
          EDX_HIP(hipEventRecord(s->ev, s->ctx->stream));
          EDX_HIP(hipStreamWaitEvent(bctx.stream, s->ev, 0));
          jobs.push_back(edx::SpmmJob{&s->lz->op(), s->lz->block(), s->lz->spmm_in(), s->lz->spmm_out(), s->lz->ld()});
          key.push_back(s->idx);
        }
        graphs.launch(key, jobs, precision);
        EDX_HIP(hipEventRecord(done_ev, bctx.stream));
        std::vector<BatchSolve*> next;
        for (BatchSolve* s : stepping) {
          edx::StreamScope scope(s->ctx->stream);
          EDX_HIP(hipStreamWaitEvent(s->ctx->stream, done_ev, 0));
          if (!s->lz->step(true)) next.push_back(s);
        }
        stepping.swap(next);
      }
      // the checks (stream wait, Ritz problem, restart) of the solves run on
      // up to kBatchHostThreads threads; each solve is touched by one thread
      const int n_act = (int)active.size();
      std::vector<char> cont(n_act, 0);
      std::vector<std::exception_ptr> err(n_act);
      auto check = [&](int t, int n_thr) {
        for (int i = t; i < n_act; i += n_thr) {
          BatchSolve* s = active[i];
          try {
            edx::DeviceGuard dg(s->ctx->device);
            edx::StreamScope scope(s->ctx->stream);
            cont[i] = s->lz->sync_point() ? 0 : 1;
          } catch (...) {
            err[i] = std::current_exception();
          }
        }
      };
      const int n_thr = std::min(n_act, kBatchHostThreads);
      std::vector<std::thread> pool;
      for (int t = 1; t < n_thr; ++t) pool.emplace_back(check, t, n_thr);
      check(0, n_thr);
      for (auto& th : pool) th.join();
      for (const auto& e : err)
        if (e) std::rethrow_exception(e);
      std::vector<BatchSolve*> still;
      for (int i = 0; i < n_act; ++i)
        if (cont[i]) still.push_back(active[i]);
      active.swap(still);
    }
  } catch (...) {
    (void)hipStreamSynchronize(bctx.stream);
    (void)hipEventDestroy(done_ev);
    throw;
  }
  EDX_HIP(hipStreamSynchronize(bctx.stream));
  (void)hipEventDestroy(done_ev);
  if (graph_counts) {
    graph_counts[0] += graphs.captures();
    graph_counts[1] += graphs.replays();
  }
}

void eigs_batch_impl(edx_ctx* bctx, int n, edx_ctx* const* ctxs, const edx_hamiltonian* const* hams,
                     const edx_eigs_options& o, double* vals, double* const* vecs, double* info) {
  const int nev = o.nev;
  EDX_REQUIRE(nev >= 1, EDX_ERR_INVALID, "nev must be >= 1");
  EDX_REQUIRE(o.precision == EDX_PREC_FP64 || o.precision == EDX_PREC_MIXED, EDX_ERR_INVALID, "bad precision");
  EDX_REQUIRE(o.tol > 0.0, EDX_ERR_INVALID, "tol must be positive");
  EDX_REQUIRE(o.max_restarts >= 1, EDX_ERR_INVALID, "max_restarts must be >= 1");
  edx::DeviceGuard dg(bctx->device);
  if (info) std::fill(info, info + 4 * (size_t)n, 0.0);
  int graph_counts[2] = {0, 0};  // round graphs captured, replayed
  struct CountGuard {
    edx_ctx& c;
    const int* g;
    ~CountGuard() {
      c.graph_captures += g[0];
      c.graph_replays += g[1];
    }
  } count_guard{*bctx, graph_counts};
  // sectors of the dense path (and empty ones) are solved one by one
  std::vector<std::unique_ptr<BatchSolve>> own;
  std::vector<BatchSolve*> lanczos;
  for (int i = 0; i < n; ++i) {
    const int64_t D = hams[i]->dim;
    if (D == 0) continue;
    if (D <= 500) {
      edx::StreamScope scope(ctxs[i]->stream);
      dense_solve(*ctxs[i], *hams[i], nev, vals + (size_t)i * nev, vecs ? vecs[i] : nullptr);
      continue;
    }
    EDX_REQUIRE(nev < D - 1, EDX_ERR_INVALID, "nev must be < sector dimension - 1");
    own.emplace_back(new BatchSolve{i, ctxs[i], hams[i], nullptr});
    lanczos.push_back(own.back().get());
  }
  for (BatchSolve* s : lanczos) EDX_HIP(hipEventCreateWithFlags(&s->ev, hipEventDisableTiming));
  struct EventGuard {
    std::vector<BatchSolve*>& v;
    ~EventGuard() {
      for (BatchSolve* s : v)
        if (s->ev) (void)hipEventDestroy(s->ev);
    }
  } eg{lanczos};
  if (lanczos.empty()) return;
  int b = o.block > 0 ? o.block : std::min(nev, 8);
  b = std::max(1, std::min(b, kMaxW));
  Options opt;
  opt.nev = nev;
  opt.block = b;
  opt.basis = o.basis;
  opt.tol = o.tol;
  opt.eig_tol = o.eig_tol;
  opt.max_restarts = o.max_restarts;
  opt.seed = o.seed;
  std::vector<std::pair<const void*, int>> no_start(lanczos.size(), {nullptr, 0});
  if (o.precision == EDX_PREC_MIXED) {
    // complex64 sweeps of every solve together, then the FP64 phase of every
    // solve together, each started as in a single solve (eigs_impl)
    for (BatchSolve* s : lanczos) {
      const int n32 = (int)std::min<int64_t>(std::min(nev, kMaxW), s->h->dim - 2);
      Options o32 = opt;
      o32.nev = n32;
      o32.tol = std::max(o.tol, o.mixed_tol > 0 ? o.mixed_tol : 3e-7);
      o32.eig_tol = 0.0;
      o32.out_cluster = true;
      o32.stall_checks = 8;
      o32.stall_ceiling = std::max(30.0 * o32.tol, 1e-4);
      edx::StreamScope scope(s->ctx->stream);
      s->lz.reset(new BlockLanczos(*s->ctx, *s->h, EDX_PREC_MIXED, o32));
    }
    run_batch_phase(*bctx, lanczos, EDX_PREC_MIXED, no_start, graph_counts);
    std::vector<DevBuf<double2>> starts(lanczos.size());
    std::vector<std::pair<const void*, int>> st(lanczos.size());
    for (size_t i = 0; i < lanczos.size(); ++i) {
      BatchSolve* s = lanczos[i];
      edx::StreamScope scope(s->ctx->stream);
      const Result& r32 = s->lz->result();
      const int n32 = (int)std::min<int64_t>(std::min(nev, kMaxW), s->h->dim - 2);
      const int n_start = std::max(n32, std::min(r32.nvec, kMaxW));
      Options o64 = opt;
      o64.block = std::min(kMaxW, std::max(b, n_start));
      o64.basis = 0;
      o64.check_every = 1;
      o64.gap_ref = r32.ritz;
      o64.gap_ref_res = r32.ritz_res;
      s->matvecs = r32.matvecs;
      s->restarts = r32.restarts;
      const int64_t D = s->h->dim, ld = s->lz->ld();
      starts[i].alloc((size_t)n_start * ld);
      hipLaunchKernelGGL((convert_vec<float2, double2>), dim3(edx::grid_for(D, 256)), dim3(256), 0, s->ctx->stream, D,
                         n_start, (const float2*)s->lz->vec(s->lz->current(), 0), ld, starts[i].p, ld);
      EDX_HIP(hipGetLastError());
      st[i] = {starts[i].p, n_start};
      s->lz.reset(new BlockLanczos(*s->ctx, *s->h, EDX_PREC_FP64, o64));
    }
    run_batch_phase(*bctx, lanczos, EDX_PREC_FP64, st, graph_counts);
  } else {
    opt.check_every = 2;
    for (BatchSolve* s : lanczos) {
      edx::StreamScope scope(s->ctx->stream);
      s->lz.reset(new BlockLanczos(*s->ctx, *s->h, EDX_PREC_FP64, opt));
    }
    run_batch_phase(*bctx, lanczos, EDX_PREC_FP64, no_start, graph_counts);
  }
  for (size_t i = 0; i < lanczos.size(); ++i) {
    BatchSolve* s = lanczos[i];
    edx::StreamScope scope(s->ctx->stream);
    const Result& res = s->lz->result();
    double* v = vals + (size_t)s->idx * nev;
    for (int c = 0; c < nev; ++c) v[c] = res.vals[c];
    const int64_t D = s->h->dim;
    if (vecs && vecs[s->idx])
      edx::copy_to_host(*s->ctx, vecs[s->idx], D * 16, s->lz->vec(s->lz->current(), 0), s->lz->ld() * 16, D * 16,
                        nev);
    EDX_HIP(hipStreamSynchronize(s->ctx->stream));
    if (info) {
      double* f = info + 4 * (size_t)s->idx;
      f[0] = (double)(s->matvecs + res.matvecs);
      f[1] = (double)(s->restarts + res.restarts);
      f[2] = res.max_rel_res;
      f[3] = (double)b;
    }
  }
}

}  // namespace

