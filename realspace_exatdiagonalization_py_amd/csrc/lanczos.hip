// Sector eigensolver: thick-restarted block Lanczos with full (two-pass
// classical Gram-Schmidt) reorthogonalisation.
//
// Reference: diagonalize_block_matrixfree (SRE:2248-2299) forms the dense
// block for n <= 500 and calls LAPACK eigh on its upper triangle, and for
// larger blocks runs ARPACK eigs(which="SR", tol=1e-10) on the CSR operator,
// returning the real parts in ascending order.  The blocks are Hermitian up to
// the complex64 rounding of the stored amplitudes, so the real parts of their
// eigenvalues equal the eigenvalues of the Hermitian part to second order in
// that rounding; this solver works with the Hermitian part (the projected
// matrix is symmetrised) and reproduces ARPACK's real parts.
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
#include <cmath>
#include <cstring>
#include <memory>

#include "dense.h"
#include "objects.h"

using edx::DevBuf;
using edx::dense::cplx;
using edx::dense::Mat;

namespace edx {
void spmm(const edx_hamiltonian& h, hipStream_t st, int precision, int nb, const void* X, void* Y, int64_t ldy);
void interleave(hipStream_t st, int precision, int64_t dim, int nb, const void* src, int64_t ld, void* dst);
void dense_matrix(const edx_hamiltonian& h, hipStream_t st, double2* out);
}  // namespace edx

namespace {

constexpr int kRS = 32;          // rows staged per Gram sub-chunk
constexpr int kMaxVec = 64;      // basis vectors the Gram/update kernels accept (64*16 outputs)
constexpr int kMaxW = 16;        // block width

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <typename V> __device__ __forceinline__ V from_d2(double re, double im);
template <> __device__ __forceinline__ double2 from_d2<double2>(double re, double im) { return make_double2(re, im); }
template <> __device__ __forceinline__ float2 from_d2<float2>(double re, double im) {
  return make_float2((float)re, (float)im);
}

// partial[blk][i*nw + j] = sum over the block's rows of conj(V_i[r]) W_j[r]
template <typename V>
__global__ void __launch_bounds__(256)
gram_partial(int64_t D, int nv, const V* __restrict__ Vb, int64_t ldv, int nw, const V* __restrict__ Wb,
             int64_t ldw, int64_t rows_per_block, double2* __restrict__ partial) {
  extern __shared__ double2 sh[];
  const int RSP = kRS + 1;
  double2* sv = sh;
  double2* sw = sh + (size_t)nv * RSP;
  const int nout = nv * nw;
  double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
  const int64_t rb = (int64_t)blockIdx.x * rows_per_block;
  const int64_t re = min(D, rb + rows_per_block);
  const int nvec = nv + nw;
  for (int64_t r0 = rb; r0 < re; r0 += kRS) {
    for (int idx = threadIdx.x; idx < nvec * kRS; idx += 256) {
      const int vec = idx / kRS, rr = idx % kRS;
      const int64_t r = r0 + rr;
      double2 x = make_double2(0.0, 0.0);
      if (r < re) x = vec < nv ? to_d2(Vb[(int64_t)vec * ldv + r]) : to_d2(Wb[(int64_t)(vec - nv) * ldw + r]);
      if (vec < nv) sv[vec * RSP + rr] = x;
      else sw[(vec - nv) * RSP + rr] = x;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int o = threadIdx.x + 256 * s;
      if (o < nout) {
        const double2* a = sv + (o / nw) * RSP;
        const double2* b = sw + (o % nw) * RSP;
        double xr = ar[s], xi = ai[s];
        for (int rr = 0; rr < kRS; ++rr) {
          const double2 u = a[rr], w = b[rr];
          xr = fma(u.x, w.x, xr);
          xr = fma(u.y, w.y, xr);
          xi = fma(u.x, w.y, xi);
          xi = fma(-u.y, w.x, xi);
        }
        ar[s] = xr;
        ai[s] = xi;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int o = threadIdx.x + 256 * s;
    if (o < nout) partial[(int64_t)blockIdx.x * nout + o] = make_double2(ar[s], ai[s]);
  }
}

__global__ void gram_reduce(int nblk, int nout, const double2* partial, double2* out) {
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= nout) return;
  double xr = 0.0, xi = 0.0;
  for (int b = 0; b < nblk; ++b) {
    const double2 p = partial[(int64_t)b * nout + o];
    xr += p.x;
    xi += p.y;
  }
  out[o] = make_double2(xr, xi);
}

// Out_k[r] = (accumulate ? Out_k[r] : 0) + sign * sum_i V_i[r] S[i][k], k < nk <= 16.
// Every input of row r is read before any output of row r is written, so Out
// may alias V when nk <= 16.
template <typename V>
__global__ void __launch_bounds__(256)
small_gemm(int64_t D, int nv, const V* Vb, int64_t ldv, int nk, V* Ob, int64_t ldo, const double2* S, int ldS,
           double sign, int accumulate) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= D) return;
  double xr[kMaxW], xi[kMaxW];
#pragma unroll
  for (int k = 0; k < kMaxW; ++k) xr[k] = xi[k] = 0.0;
  for (int i = 0; i < nv; ++i) {
    const double2 v = to_d2(Vb[(int64_t)i * ldv + r]);
    const double2* srow = S + (int64_t)i * ldS;
#pragma unroll
    for (int k = 0; k < kMaxW; ++k) {
      if (k < nk) {
        const double2 s = srow[k];
        xr[k] = fma(v.x, s.x, xr[k]);
        xr[k] = fma(-v.y, s.y, xr[k]);
        xi[k] = fma(v.x, s.y, xi[k]);
        xi[k] = fma(v.y, s.x, xi[k]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxW; ++k) {
    if (k < nk) {
      double orr = sign * xr[k], oi = sign * xi[k];
      if (accumulate) {
        const double2 o = to_d2(Ob[(int64_t)k * ldo + r]);
        orr += o.x;
        oi += o.y;
      }
      Ob[(int64_t)k * ldo + r] = from_d2<V>(orr, oi);
    }
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
  double tol = 1e-10;
  int max_restarts = 400;
  uint64_t seed = 1;
};

struct Result {
  std::vector<double> vals;
  int64_t matvecs = 0;
  int restarts = 0;
  double max_rel_res = 0.0;
};

// Block Lanczos on one precision.  Vectors live in two banks (ping-pong for
// the thick restart), each holding `cap` vectors of leading dimension ld.
class BlockLanczos {
 public:
  BlockLanczos(edx_ctx& ctx, const edx_hamiltonian& h, int prec, const Options& opt)
      : ctx_(ctx), h_(h), prec_(prec), opt_(opt) {
    D_ = h.dim;
    b_ = opt.block;
    esz_ = prec == EDX_PREC_FP64 ? 16 : 8;
    ld_ = (D_ + 63) / 64 * 64;
    int base = opt.basis > 0 ? opt.basis : std::max({2 * (opt.nev + b_), 4 * b_, 24});
    base = (base + b_ - 1) / b_ * b_;
    nq_max_ = std::min(base, kMaxVec - b_);
    nq_max_ = std::max(nq_max_ / b_ * b_, opt.nev + b_);
    cap_ = nq_max_ + b_;
    EDX_REQUIRE(cap_ <= kMaxVec, EDX_ERR_INVALID, "Krylov basis too large");
    bank_[0].alloc((size_t)cap_ * ld_ * esz_);
    bank_[1].alloc((size_t)cap_ * ld_ * esz_);
    W_.alloc((size_t)b_ * ld_ * esz_);
    Wsave_.alloc((size_t)b_ * ld_ * esz_);
    X_.alloc((size_t)b_ * D_ * esz_);
    n_blocks_ = (int)std::min<int64_t>(std::max<int64_t>((D_ + 255) / 256, 1), 512);
    rows_per_block_ = ((D_ + n_blocks_ - 1) / n_blocks_ + kRS - 1) / kRS * kRS;
    partial_.alloc((size_t)n_blocks_ * kMaxVec * kMaxW);
    small_.alloc((size_t)kMaxVec * kMaxVec);
    small_out_.alloc((size_t)kMaxVec * kMaxW);
  }

  // Runs to convergence.  `start` (optional, device, same precision,
  // column-major with leading dimension start_ld) seeds the first block with
  // n_start vectors (rest random).  Ritz vectors are left in vec(0..nev).
  Result run(const void* start, int n_start, int64_t start_ld) {
    Result res;
    cur_ = 0;
    int nq = 0;
    // initial block
    random_block(vec(cur_, 0), b_, opt_.seed);
    if (start && n_start > 0) {
      const int ns = std::min(n_start, b_);
      EDX_HIP(hipMemcpy2DAsync(vec(cur_, 0), ld_ * esz_, start, start_ld * esz_, D_ * esz_, ns,
                               hipMemcpyDeviceToDevice, ctx_.stream));
    }
    Mat Rdummy;
    orthonormalize(vec(cur_, 0), 0, nullptr, Rdummy, true);
    Mat M(cap_, cap_), B(b_, cap_);
    std::vector<double> theta;
    Mat Y;
    for (int iter = 0;; ++iter) {
      // W = H Vr
      void* Vr = vec(cur_, nq);
      edx::interleave(ctx_.stream, prec_, D_, b_, Vr, ld_, X_.p);
      edx::spmm(h_, ctx_.stream, prec_, b_, X_.p, W_.p, ld_);
      res.matvecs += b_;
      const int nbasis = nq + b_;
      Mat C1 = gram(vec(cur_, 0), nbasis, W_.p, b_);
      gemm(vec(cur_, 0), nbasis, W_.p, b_, C1, -1.0, true);
      Mat C2 = gram(vec(cur_, 0), nbasis, W_.p, b_);
      gemm(vec(cur_, 0), nbasis, W_.p, b_, C2, -1.0, true);
      for (int i = 0; i < nbasis; ++i)
        for (int j = 0; j < b_; ++j) {
          const cplx c = C1(i, j) + C2(i, j);
          if (i < nq) {
            M(i, nq + j) = c;
            M(nq + j, i) = std::conj(c);
          }
        }
      for (int i = 0; i < b_; ++i)
        for (int j = 0; j < b_; ++j) {
          const cplx cij = C1(nq + i, j) + C2(nq + i, j);
          const cplx cji = C1(nq + j, i) + C2(nq + j, i);
          M(nq + i, nq + j) = 0.5 * (cij + std::conj(cji));
        }
      nq += b_;
      // next residual block
      Mat R;
      orthonormalize(W_.p, nq, vec(cur_, nq), R, false);
      B = Mat(b_, cap_);
      for (int i = 0; i < b_; ++i)
        for (int j = 0; j < b_; ++j) B(i, nq - b_ + j) = R(i, j);
      // Rayleigh-Ritz
      if (nq < opt_.nev) continue;
      ritz(M, nq, theta, Y);
      const int nev = opt_.nev;
      double worst = 0.0;
      bool conv = true;
      for (int c = 0; c < nev; ++c) {
        double rn2 = 0.0;
        for (int i = 0; i < b_; ++i) {
          cplx s(0.0, 0.0);
          for (int j = 0; j < nq; ++j) s += B(i, j) * Y(j, c);
          rn2 += std::norm(s);
        }
        const double rel = std::sqrt(rn2) / std::max(1.0, std::fabs(theta[c]));
        worst = std::max(worst, rel);
        if (!(rel <= opt_.tol)) conv = false;
      }
      res.max_rel_res = worst;
      const bool full = nq + b_ > nq_max_;
      if (conv) {
        // Ritz vectors of the wanted pairs into the other bank, then back to slot 0
        Mat Yk(nq, nev);
        for (int i = 0; i < nq; ++i)
          for (int c = 0; c < nev; ++c) Yk(i, c) = Y(i, c);
        combine(vec(cur_, 0), nq, vec(1 - cur_, 0), nev, Yk);
        cur_ = 1 - cur_;
        res.vals.assign(theta.begin(), theta.begin() + nev);
        res.restarts = restarts_;
        EDX_HIP(hipStreamSynchronize(ctx_.stream));
        return res;
      }
      if (full) {
        EDX_REQUIRE(restarts_ < opt_.max_restarts, EDX_ERR_NOCONV,
                    "block Lanczos did not converge within " + std::to_string(opt_.max_restarts) +
                        " restarts (max relative residual " + std::to_string(worst) + ")");
        int k = nev + (nq - nev) / 2;
        k = std::max(k, nev + std::min(b_, nq - nev));
        k = std::min(k, nq - b_);
        k = std::max(k, nev);
        Mat Yk(nq, k);
        for (int i = 0; i < nq; ++i)
          for (int c = 0; c < k; ++c) Yk(i, c) = Y(i, c);
        combine(vec(cur_, 0), nq, vec(1 - cur_, 0), k, Yk);
        EDX_HIP(hipMemcpy2DAsync(vec(1 - cur_, k), ld_ * esz_, vec(cur_, nq), ld_ * esz_, D_ * esz_, b_,
                                 hipMemcpyDeviceToDevice, ctx_.stream));
        Mat Bn(b_, cap_);
        for (int i = 0; i < b_; ++i)
          for (int c = 0; c < k; ++c) {
            cplx s(0.0, 0.0);
            for (int j = 0; j < nq; ++j) s += B(i, j) * Y(j, c);
            Bn(i, c) = s;
          }
        M = Mat(cap_, cap_);
        for (int c = 0; c < k; ++c) M(c, c) = theta[c];
        B = Bn;
        nq = k;
        cur_ = 1 - cur_;
        ++restarts_;
      }
    }
  }

  // Device address of basis vector i of the current bank.
  void* vec(int bank, int i) { return static_cast<char*>((void*)bank_[bank].p) + (size_t)i * ld_ * esz_; }
  int current() const { return cur_; }
  int64_t ld() const { return ld_; }

 private:
  Mat gram(const void* V, int nv, const void* W, int nw) {
    EDX_REQUIRE(nv <= kMaxVec && nw <= kMaxW, EDX_ERR_INTERNAL, "gram size");
    const size_t smem = (size_t)(nv + nw) * (kRS + 1) * sizeof(double2);
    if (prec_ == EDX_PREC_FP64)
      hipLaunchKernelGGL(gram_partial<double2>, dim3(n_blocks_), dim3(256), smem, ctx_.stream, D_, nv,
                         (const double2*)V, ld_, nw, (const double2*)W, ld_, rows_per_block_, partial_.p);
    else
      hipLaunchKernelGGL(gram_partial<float2>, dim3(n_blocks_), dim3(256), smem, ctx_.stream, D_, nv,
                         (const float2*)V, ld_, nw, (const float2*)W, ld_, rows_per_block_, partial_.p);
    EDX_HIP(hipGetLastError());
    const int nout = nv * nw;
    hipLaunchKernelGGL(gram_reduce, dim3(edx::grid_for(nout, 256)), dim3(256), 0, ctx_.stream, n_blocks_, nout,
                       (const double2*)partial_.p, small_out_.p);
    EDX_HIP(hipGetLastError());
    std::vector<double2> h(nout);
    EDX_HIP(hipMemcpyAsync(h.data(), small_out_.p, nout * sizeof(double2), hipMemcpyDeviceToHost, ctx_.stream));
    EDX_HIP(hipStreamSynchronize(ctx_.stream));
    Mat C(nv, nw);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nw; ++j) C(i, j) = cplx(h[i * nw + j].x, h[i * nw + j].y);
    return C;
  }

  // Out[:, k] (+)= sign * V S  (S is nv x nk, nk <= 16 per launch)
  void gemm(const void* V, int nv, void* Out, int nk, const Mat& S, double sign, bool accumulate) {
    std::vector<double2> hs((size_t)nv * nk);
    for (int i = 0; i < nv; ++i)
      for (int k = 0; k < nk; ++k) hs[(size_t)i * nk + k] = make_double2(S(i, k).real(), S(i, k).imag());
    EDX_HIP(hipMemcpyAsync(small_.p, hs.data(), hs.size() * sizeof(double2), hipMemcpyHostToDevice, ctx_.stream));
    launch_small_gemm(V, nv, Out, nk, small_.p, nk, sign, accumulate);
    EDX_HIP(hipStreamSynchronize(ctx_.stream));
  }

  void launch_small_gemm(const void* V, int nv, void* Out, int nk, const double2* S, int ldS, double sign,
                         bool accumulate) {
    if (prec_ == EDX_PREC_FP64)
      hipLaunchKernelGGL(small_gemm<double2>, dim3(edx::grid_for(D_, 256)), dim3(256), 0, ctx_.stream, D_, nv,
                         (const double2*)V, ld_, nk, (double2*)Out, ld_, S, ldS, sign, accumulate ? 1 : 0);
    else
      hipLaunchKernelGGL(small_gemm<float2>, dim3(edx::grid_for(D_, 256)), dim3(256), 0, ctx_.stream, D_, nv,
                         (const float2*)V, ld_, nk, (float2*)Out, ld_, S, ldS, sign, accumulate ? 1 : 0);
    EDX_HIP(hipGetLastError());
  }

  // Out[:, 0..nk) = V[:, 0..nv) Y (Out must not alias V), tiled by 16 columns.
  void combine(const void* V, int nv, void* Out, int nk, const Mat& Y) {
    std::vector<double2> hs((size_t)nv * nk);
    for (int i = 0; i < nv; ++i)
      for (int k = 0; k < nk; ++k) hs[(size_t)i * nk + k] = make_double2(Y(i, k).real(), Y(i, k).imag());
    DevBuf<double2> S(hs.size());
    EDX_HIP(hipMemcpyAsync(S.p, hs.data(), hs.size() * sizeof(double2), hipMemcpyHostToDevice, ctx_.stream));
    for (int k0 = 0; k0 < nk; k0 += kMaxW) {
      const int nkt = std::min(kMaxW, nk - k0);
      launch_small_gemm(V, nv, static_cast<char*>(Out) + (size_t)k0 * ld_ * esz_, nkt, S.p + k0, nk, 1.0, false);
    }
    EDX_HIP(hipStreamSynchronize(ctx_.stream));
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

  // Orthonormalises the b vectors at W (already orthogonal to the first nq
  // basis vectors of the current bank) into `dst` (may equal W), with
  // W_in = dst R.  Cholesky QR twice; a rank-deficient block is replaced by
  // random directions orthogonalised against the basis, and R is then the
  // projection of the original block onto the new one.
  void orthonormalize(void* W, int nq, void* dst, Mat& R, bool first) {
    if (!dst) dst = W;
    const double tiny = prec_ == EDX_PREC_FP64 ? 1e-24 : 1e-12;
    Mat G1 = gram(W, b_, W, b_);
    Mat R1, R2;
    bool ok = edx::dense::cholesky_upper(G1, R1, tiny);
    if (ok) {
      if (dst != W)
        EDX_HIP(hipMemcpy2DAsync(dst, ld_ * esz_, W, ld_ * esz_, D_ * esz_, b_, hipMemcpyDeviceToDevice,
                                 ctx_.stream));
      gemm(dst, b_, dst, b_, edx::dense::upper_inverse(R1), 1.0, false);
      Mat G2 = gram(dst, b_, dst, b_);
      ok = edx::dense::cholesky_upper(G2, R2, 1e-8);
      if (ok) {
        gemm(dst, b_, dst, b_, edx::dense::upper_inverse(R2), 1.0, false);
        R = Mat(b_, b_);
        for (int i = 0; i < b_; ++i)
          for (int j = 0; j < b_; ++j) {
            cplx s(0.0, 0.0);
            for (int k = 0; k < b_; ++k) s += R2(i, k) * R1(k, j);
            R(i, j) = s;
          }
        return;
      }
    }
    EDX_REQUIRE(!first || nq == 0, EDX_ERR_INTERNAL, "start block orthonormalisation failed");
    // deflation: fresh random block orthogonal to the basis
    EDX_HIP(hipMemcpy2DAsync(Wsave_.p, ld_ * esz_, W, ld_ * esz_, D_ * esz_, b_, hipMemcpyDeviceToDevice,
                             ctx_.stream));
    for (int attempt = 0; attempt < 3; ++attempt) {
      random_block(dst, b_, opt_.seed * 7919 + 104729 * (uint64_t)(++deflations_));
      for (int pass = 0; pass < 2 && nq > 0; ++pass) {
        Mat C = gram(vec(cur_, 0), nq, dst, b_);
        gemm(vec(cur_, 0), nq, dst, b_, C, -1.0, true);
      }
      Mat G = gram(dst, b_, dst, b_);
      Mat Ra;
      if (!edx::dense::cholesky_upper(G, Ra, tiny)) continue;
      gemm(dst, b_, dst, b_, edx::dense::upper_inverse(Ra), 1.0, false);
      Mat G2 = gram(dst, b_, dst, b_);
      Mat Rb;
      if (!edx::dense::cholesky_upper(G2, Rb, 1e-8)) continue;
      gemm(dst, b_, dst, b_, edx::dense::upper_inverse(Rb), 1.0, false);
      R = gram(dst, b_, Wsave_.p, b_);
      return;
    }
    throw edx::Error{EDX_ERR_NOCONV, "could not extend the Krylov basis (sector dimension exhausted)"};
  }

  void ritz(const Mat& M, int nq, std::vector<double>& theta, Mat& Y) {
    Mat A(nq, nq);
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j) A(i, j) = M(i, j);
    edx::dense::herm_eig(A, theta, Y, true);
  }

  edx_ctx& ctx_;
  const edx_hamiltonian& h_;
  int prec_;
  Options opt_;
  int64_t D_ = 0, ld_ = 0;
  int b_ = 1, esz_ = 16, nq_max_ = 0, cap_ = 0, cur_ = 0, restarts_ = 0;
  uint64_t deflations_ = 0;
  int n_blocks_ = 1;
  int64_t rows_per_block_ = 0;
  DevBuf<char> bank_[2];
  DevBuf<char> W_, Wsave_, X_;
  DevBuf<double2> partial_, small_, small_out_;
};

// Dense path for small sectors: explicit block, eigen-decomposition of the
// Hermitian matrix formed from its upper triangle (the reference's
// eigh(UPLO="U") for n <= 500).
void dense_solve(edx_ctx& ctx, const edx_hamiltonian& h, int nev, double* vals, double* vecs) {
  const int64_t D = h.dim;
  DevBuf<double2> dmat((size_t)D * D);
  edx::dense_matrix(h, ctx.stream, dmat.p);
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

}  // namespace

extern "C" {

int edx_eigs(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t nev, int32_t block, double tol, uint64_t seed,
             int32_t precision, double* vals, double* vecs, double* info) {
  return edx::guarded(ctx, [&] {
    EDX_REQUIRE(ctx && ham && vals, EDX_ERR_INVALID, "null argument");
    EDX_REQUIRE(nev >= 1, EDX_ERR_INVALID, "nev must be >= 1");
    EDX_REQUIRE(precision == EDX_PREC_FP64 || precision == EDX_PREC_MIXED, EDX_ERR_INVALID, "bad precision");
    EDX_REQUIRE(tol > 0.0, EDX_ERR_INVALID, "tol must be positive");
    edx::DeviceGuard dg(ctx->device);
    const int64_t D = ham->dim;
    if (info) std::fill(info, info + 4, 0.0);
    if (D == 0) return;
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
    opt.tol = tol;
    opt.seed = seed;
    Result res;
    int64_t mv_total = 0;
    std::unique_ptr<BlockLanczos> fin;
    if (precision == EDX_PREC_MIXED) {
      Options o32 = opt;
      o32.tol = std::max(tol, 2e-5);
      BlockLanczos lz32(*ctx, *ham, EDX_PREC_MIXED, o32);
      Result r32 = lz32.run(nullptr, 0, 0);
      mv_total += r32.matvecs;
      // FP64 continuation from the complex64 Ritz vectors
      Options o64 = opt;
      o64.block = std::min(kMaxW, std::max(b, nev));
      o64.basis = 3 * o64.block;
      fin.reset(new BlockLanczos(*ctx, *ham, EDX_PREC_FP64, o64));
      DevBuf<double2> start((size_t)nev * lz32.ld());
      hipLaunchKernelGGL((convert_vec<float2, double2>), dim3(edx::grid_for(D, 256)), dim3(256), 0, ctx->stream, D,
                         nev, (const float2*)lz32.vec(lz32.current(), 0), lz32.ld(), start.p, lz32.ld());
      EDX_HIP(hipGetLastError());
      res = fin->run(start.p, nev, lz32.ld());
      res.restarts += r32.restarts;
    } else {
      fin.reset(new BlockLanczos(*ctx, *ham, EDX_PREC_FP64, opt));
      res = fin->run(nullptr, 0, 0);
    }
    mv_total += res.matvecs;
    for (int c = 0; c < nev; ++c) vals[c] = res.vals[c];
    if (vecs)
      EDX_HIP(hipMemcpy2DAsync(vecs, D * 16, fin->vec(fin->current(), 0), fin->ld() * 16, D * 16, nev,
                               hipMemcpyDeviceToHost, ctx->stream));
    EDX_HIP(hipStreamSynchronize(ctx->stream));
    if (info) {
      info[0] = (double)mv_total;
      info[1] = (double)res.restarts;
      info[2] = res.max_rel_res;
      info[3] = (double)b;
    }
  });
}

}  // extern "C"
