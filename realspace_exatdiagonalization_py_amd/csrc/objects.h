// Definitions of the opaque C-ABI objects.
#pragma once

#include <exception>
#include <mutex>
#include <new>

#include "edx_internal.h"

struct edx_ctx : public edx::Ctx {};

struct edx_catalog {
  int device = 0;
  int n_site = 0, n_filled = 0, n_g = 0, fermionic = 0, product_phase = 0;
  // 1 when every site phase is +1 or -1: alpha is then +-1, exact in the
  // complex64 rep_amp, and the projection table is Hermitian to FP64 rounding
  int exact_alpha = 1;
  int64_t n_total = 0, n_orbits = 0, n_stab = 0;
  std::vector<int32_t> h_perm;        // host copy (1-based) to validate phase updates
  edx::ObjBuf<uint8_t> perm;          // [n_g][n_site] 0-based targets
  edx::ObjBuf<double> angle;          // [n_g][n_site]
  edx::ObjBuf<double2> phase;         // [n_g][n_site]
  edx::ObjBuf<int32_t> inv;           // [n_g]
  edx::ObjBuf<uint32_t> binom;        // [(n_site+1)][(n_filled+1)]
  edx::ObjBuf<int32_t> rep_rank;      // [C] orbit index of the mask of that rank
  edx::ObjBuf<int32_t> rep_sym;       // [C] largest g with g.rep == mask
  edx::ObjBuf<float2> rep_amp;        // [C] complex64 alpha_g(rep)
  edx::ObjBuf<uint64_t> reps;         // [n_orbits] ascending
  edx::ObjBuf<int32_t> stab_order;    // [n_orbits]
  edx::ObjBuf<uint32_t> stab_off;     // [n_orbits + 1]
  edx::ObjBuf<uint8_t> stab_g;        // [n_stab]
  edx::ObjBuf<double2> stab_phase;    // [n_stab]
  // [C] rep_rank, rep_sym and rep_amp packed per rank, built on first use by
  // an on-the-fly Hamiltonian and rebuilt by every phase refresh
  edx::ObjBuf<edx::RankRec> rank_rec;
  std::mutex rank_rec_mu;

  edx::GroupView view() const {
    return edx::GroupView{n_site, n_filled, n_g, fermionic, product_phase, perm.p, angle.p,
                          phase.p, inv.p, binom.p};
  }
};

struct edx_sector {
  int device = 0;
  const edx_catalog* cat = nullptr;
  int64_t dim = 0;
  edx::ObjBuf<uint64_t> reps;         // [dim]
  edx::ObjBuf<int32_t> stab_order;    // [dim]
  edx::ObjBuf<int32_t> orbit_to_basis;  // [n_orbits]
};

struct edx_hamiltonian {
  int device = 0;
  const edx_sector* sec = nullptr;
  int64_t dim = 0;
  int64_t n_entries = 0;
  int32_t n_bilinear = 0;
  edx::ObjBuf<int64_t> row_ptr;       // [dim + 1]
  edx::ObjBuf<int32_t> col;           // [n_entries]
  edx::ObjBuf<uint16_t> term;         // [n_entries] original term index
  edx::ObjBuf<double2> val;           // [n_entries] complex128
  edx::ObjBuf<double2> hdiag;         // [dim]
  edx::ObjBuf<float2> hdiag32;        // [dim]
  // 1: the sliced copy holds the Hermitian part (H + H^H) / 2 of the table
  // (tables whose alpha were rounded to complex64 are Hermitian only to that
  // rounding); edx_matvec then applies the table itself from the CSR arrays
  int herm_part = 0;
  // sliced-ELL copy used by the SpMM: slices of 64 / sell_g rows, sell_g lanes per row
  int sell_g = 1;
  int64_t sell_slices = 0;
  int64_t sell_entries = 0;           // stored entries including padding
  edx::ObjBuf<int64_t> sell_off;      // [sell_slices + 1]
  edx::ObjBuf<int32_t> sell_col;      // [sell_entries], -1 = padding (packed: words, below)
  edx::ObjBuf<double2> sell_val;      // [sell_entries] (empty when packed)
  edx::ObjBuf<float2> sell_val32;     // [sell_entries] complex64 copy (empty when packed)
  // palette-packed copy (sellpack.hip); the elements are pal[index]
  // (complex128) and pal32[index] (complex64).  1: sell_col holds the 32-bit
  // words column | (palette index << pack_bits), 0xFFFFFFFF = padding;
  // 2 (wide): sell_col keeps the columns and sell_idx holds the indices
  int sell_packed = 0;
  int pack_bits = 0;
  int n_pal = 0;
  edx::ObjBuf<double2> pal;           // [n_pal]
  edx::ObjBuf<float2> pal32;          // [n_pal]
  edx::ObjBuf<uint16_t> sell_idx;     // [sell_entries] (wide layout only)
  // 1: no stored table; every product recomputes the elements from the terms
  // and the catalog (the arrays below), CSR and sliced arrays are empty
  int on_the_fly = 0;
  edx::ObjBuf<int32_t> fly_toff;      // [n_site + 1] terms grouped by source site
  edx::ObjBuf<uint8_t> fly_to;        // [n_bilinear] target bit, grouped order
  edx::ObjBuf<double2> fly_amp;       // [n_bilinear] amplitude, grouped order
  edx::ObjBuf<uint64_t> fly_between;  // [n_bilinear] Jordan-Wigner string, grouped order
  edx::ObjBuf<double2> fly_chi;       // [EDX_MAX_GROUP] characters of the sector's irrep
};

// A captured HIP graph of grouped sector products and its executable form.
struct edx_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace edx {
// Builds cat.rank_rec if it is not there (thread-safe).
void ensure_rank_records(edx_catalog& cat, Ctx& ctx);
// (Re)builds the sliced copy of the table with G lanes per row (1, 2 or 4).
void build_sell(edx_hamiltonian& h, Ctx& ctx, int G);
// Lanes per row for a table of this shape.
int default_sell_lanes(int64_t dim, int64_t n_entries);
// Builds the sliced copy with the context's fixed lane count, the autotuned
// one (timing the layouts on a miss of the tuning table) or the heuristic.
void build_sell_for(edx_hamiltonian& h, Ctx& ctx);
// The same in steps, for builds that write the sliced copy themselves:
// plan_sell picks the lanes (and on a miss of the tuning table marks the key
// as being timed, tune = true); once the table and its sliced copy are
// complete, finish_sell times the layouts if asked to; abandon_sell releases
// the key of a build that failed before that.
struct SellPlan {
  int lanes = 1;
  bool tune = false;
  int prec = 0, nb = 0;
};
SellPlan plan_sell(const edx_hamiltonian& h, Ctx& ctx);
// One product of a grouped SpMM: Y = H X for nb interleaved vectors X, Y
// column-major with leading dimension ldy.
struct SpmmJob {
  const edx_hamiltonian* h;
  int nb;
  const void* X;
  void* Y;
  int64_t ldy;
};
// The jobs (one element precision) on ctx's stream; sliced tables of equal
// block width and lanes per row share launches, with spmm()'s arithmetic.
void spmm_group(const SpmmJob* jobs, int n, Ctx& ctx, int precision);
// Algorithmic bytes of those products (as the kernel timer counts them).
double spmm_group_bytes(const SpmmJob* jobs, int n, int precision);
void finish_sell(edx_hamiltonian& h, Ctx& ctx, SellPlan& p);
void abandon_sell(const edx_hamiltonian& h, SellPlan& p);
// Slice offsets and storage of a sliced copy with G lanes per row (contents
// not written); returns the stored entry count.
int64_t prepare_sell(edx_hamiltonian& h, Ctx& ctx, int G);
// Bit length of dim: the column field of a packed table word.
int pack_bits_for(int64_t dim);
// Replaces h's complete unpacked sliced copy by the palette-packed one when
// ctx allows packing and the values fit; returns true when packed.
bool pack_sell(edx_hamiltonian& h, Ctx& ctx);
// Looks up the device address of edx_matvec_noalloc's static scratch array
// on `device` (current) once per device (noalloc.hip).
void noalloc_init(int device);
}  // namespace edx

namespace edx {

inline void set_error(edx_ctx* ctx, const std::string& msg);
// Thread-local fallback when no ctx is available.
std::string& tls_error();

// Token of the calling thread (address of a thread-local object; never 0).
inline uintptr_t thread_token() {
  thread_local char tag;
  return reinterpret_cast<uintptr_t>(&tag);
}

// Claims ctx for the calling thread for the duration of an API call.  A
// call from a second thread while another thread's call is running on the
// same ctx is refused (ok = false) instead of racing on its stream, staging
// buffers and kernel timer; nested calls on the owning thread are allowed.
struct CtxClaim {
  Ctx* ctx;
  bool ok = true;
  explicit CtxClaim(Ctx* c) : ctx(c) {
    if (!ctx) return;
    const uintptr_t me = thread_token();
    if (ctx->owner.load(std::memory_order_acquire) == me) {
      ++ctx->depth;
      return;
    }
    uintptr_t expected = 0;
    if (ctx->owner.compare_exchange_strong(expected, me, std::memory_order_acq_rel)) {
      ctx->depth = 1;
      return;
    }
    ok = false;
  }
  ~CtxClaim() {
    if (!ctx || !ok) return;
    if (--ctx->depth == 0) ctx->owner.store(0, std::memory_order_release);
  }
  CtxClaim(const CtxClaim&) = delete;
  CtxClaim& operator=(const CtxClaim&) = delete;
};

template <class F>
int guarded(edx_ctx* ctx, F&& f) {
  CtxClaim claim(ctx);
  if (!claim.ok) {
    // the owner thread may be writing ctx->last_error: report through this
    // thread's own error text (edx_last_error(NULL))
    tls_error() = "edx context is in use by another thread: a context serves one thread at a time "
                  "(create one context per thread)";
    return EDX_ERR_BUSY;
  }
  try {
    StreamScope scope(ctx ? ctx->stream : nullptr);
    f();
    return EDX_OK;
  } catch (const Error& e) {
    set_error(ctx, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(ctx, "host allocation failed");
    return EDX_ERR_OOM;
  } catch (const std::exception& e) {
    set_error(ctx, e.what());
    return EDX_ERR_INTERNAL;
  } catch (...) {
    set_error(ctx, "unknown error");
    return EDX_ERR_INTERNAL;
  }
}

inline void set_error(edx_ctx* ctx, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  tls_error() = msg;
}

}  // namespace edx
