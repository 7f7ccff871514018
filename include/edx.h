/*
 * edx — C ABI of the MI355X symmetry-resolved exact-diagonalization path.
 *
 * Plain C types only.  Every entry point returns an int status (EDX_OK or a
 * negative EDX_ERR_* code) and records a message retrievable with
 * edx_last_error(ctx).  Device memory belongs to the objects created here;
 * host buffers passed in belong to the caller.  Work areas inside a call are
 * allocated stream-ordered on the ctx's stream and freed before the call
 * returns; arrays of the objects use ordinary device allocations.  Calls on one ctx are
 * serialised on that ctx's HIP stream; different ctxs may be used from
 * different threads.  A ctx serves one thread at a time: a call made while
 * another thread's call runs on the same ctx returns EDX_ERR_BUSY at once.  "ref" citations are lines of the reference file
 * src/realspace_exactdiagonalization_py/symmetry_resolved_ed.py (SRE) unless
 * another file is named.
 *
 * Conventions shared with the reference:
 *   - mask bit i-1 is site i (1-based sites in all term/permutation inputs);
 *   - masks of fixed popcount are indexed by their colex rank;
 *   - a bilinear term (i, j, t) moves a particle from site i to site j;
 *   - group element g maps the representative to the table entry's mask.
 */
#ifndef EDX_H
#define EDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDX_OK 0
#define EDX_ERR_INVALID (-1)   /* invalid argument          -> ValueError      */
#define EDX_ERR_ASSERT (-2)    /* violated precondition     -> AssertionError  */
#define EDX_ERR_HIP (-3)       /* HIP runtime failure       -> RuntimeError    */
#define EDX_ERR_OOM (-4)       /* device allocation failure -> MemoryError     */
#define EDX_ERR_NOCONV (-5)    /* eigensolver not converged -> RuntimeError    */
#define EDX_ERR_INTERNAL (-6)  /* anything else             -> RuntimeError    */
#define EDX_ERR_BUSY (-7)      /* ctx in use by another thread -> RuntimeError;
                                  its text: edx_last_error(NULL) on the caller */

#define EDX_PREC_FP64 0        /* complex128 vectors and matrix elements        */
#define EDX_PREC_MIXED 1       /* complex64 sweeps, FP64 recurrence/refinement  */

typedef struct edx_ctx edx_ctx;
typedef struct edx_catalog edx_catalog;
typedef struct edx_sector edx_sector;
typedef struct edx_hamiltonian edx_hamiltonian;
typedef struct edx_graph edx_graph;

/* ---- context ------------------------------------------------------------
 * Device context owning one HIP stream.  No reference counterpart (the
 * reference is single-process numpy); every object below is created on a ctx. */
int edx_ctx_create(int32_t device, edx_ctx** out);
int edx_ctx_destroy(edx_ctx* ctx);
const char* edx_last_error(const edx_ctx* ctx);
/* Number of HIP devices visible (0 when none); never fails. */
int32_t edx_device_count(void);
int edx_synchronize(edx_ctx* ctx);

/* ---- orbit catalog --------------------------------------------------------
 * Replaces build_symmetry_orbit_catalog / _build_orbit_catalog_kernel
 * (ref SRE:1298-1404, :433-575).  perm: [n_g*n_site] 1-based targets,
 * perm_phases: [n_g*n_site] complex128 as (re, im) pairs, perm_angles:
 * [n_g*n_site] arg(perm_phases) as numpy.angle returns them.  The operations
 * must form a group (each has its inverse in the list); n_site <= 63 and
 * C(n_site, n_filled) < 2^31. */
int edx_catalog_build(edx_ctx* ctx, int32_t n_site, int32_t n_filled, int32_t n_g,
                      const int32_t* perm, const double* perm_phases, const double* perm_angles,
                      int32_t fermionic, edx_catalog** out);
/* A catalog from host copies of its tables (the layout edx_catalog_fetch
 * writes), for catalogs restored from a checkpoint: the reference pickles the
 * Symmetry_Orbit_Catalog dataclass with its tables (ref SRE:1151-1190,
 * :2530-2563) and its consumers rebuild sector bases from the loaded catalog
 * (ref observables/charge_pump.py:431-462).  The group arguments are those of
 * edx_catalog_build; sizes, offsets, group indices, orbit indices and
 * representatives are validated (EDX_ERR_INVALID). */
int edx_catalog_upload(edx_ctx* ctx, int32_t n_site, int32_t n_filled, int32_t n_g, const int32_t* perm,
                       const double* perm_phases, const double* perm_angles, int32_t fermionic, int64_t n_orbits,
                       int64_t n_stab, const int32_t* rep_rank, const int32_t* rep_sym, const float* rep_amp,
                       const uint64_t* reps, const int32_t* stab_order, const int64_t* stab_offsets,
                       const int32_t* stab_g, const double* stab_phase, edx_catalog** out);
/* Device copy of a catalog (all tables and the current phases).  The flux
 * scan of ed_scan (ref SRE:2658-2703) refreshes the phases of a copy per twist
 * in place, where the reference builds a fresh catalog per twist, so that the
 * caller's catalog keeps its own phases. */
int edx_catalog_clone(edx_ctx* ctx, const edx_catalog* src, edx_catalog** out);
/* Sizes: C(n_site, n_filled), number of orbits, total stabilizer entries. */
int edx_catalog_info(const edx_catalog* cat, int64_t* n_total, int64_t* n_orbits, int64_t* n_stab);
/* Copies the catalog to host buffers (ref Symmetry_Orbit_Catalog fields,
 * SRE:1179-1190).  Any pointer may be NULL to skip that array.
 * rep_rank/rep_sym: int32[n_total]; rep_amp: complex64[n_total] as float
 * pairs; reps: uint64[n_orbits] ascending; stab_order: int32[n_orbits];
 * stab_offsets: int64[n_orbits+1]; stab_g: int32[n_stab] ascending per
 * orbit; stab_phase: complex128[n_stab] as double pairs. */
int edx_catalog_fetch(edx_ctx* ctx, const edx_catalog* cat, int32_t* rep_rank, int32_t* rep_sym,
                      float* rep_amp, uint64_t* reps, int32_t* stab_order, int64_t* stab_offsets,
                      int32_t* stab_g, double* stab_phase);
/* Replaces update_orbit_stabilizer_phases (ref SRE:1220-1269): new per-site
 * phases for the same permutations; stabilizer phases and table amplitudes
 * recomputed as running products.  Different permutations -> EDX_ERR_INVALID. */
int edx_catalog_update_phases(edx_ctx* ctx, edx_catalog* cat, const int32_t* perm,
                              const double* perm_phases, const double* perm_angles);
/* The same refresh with the formulas of a fresh build (alpha in exponent
 * form, ref SRE:538-551): the tables then equal those edx_catalog_build gives
 * at the new phases, bit for bit, without repeating the orbit scan.  Used by
 * the flux scan of ed_scan, where the reference builds a fresh catalog per
 * twist (ref SRE:2686-2691). */
int edx_catalog_rephase(edx_ctx* ctx, edx_catalog* cat, const int32_t* perm,
                        const double* perm_phases, const double* perm_angles);
/* Batched O(1) canonical lookup (ref get_canonical / CanonicalMap,
 * SRE:1595-1632): for n masks of popcount n_filled, the representative mask,
 * the table's g (rep -> mask) and alpha (complex128 widened from the stored
 * complex64).  Masks of another popcount -> EDX_ERR_INVALID. */
int edx_catalog_lookup(edx_ctx* ctx, const edx_catalog* cat, int64_t n, const uint64_t* masks, uint64_t* rep,
                       int32_t* g, double* alpha);
int edx_catalog_destroy(edx_catalog* cat);

/* ---- symmetry sector (irrep filter) ----------------------------------------
 * Replaces build_symmetry_sector_basis / _is_orbit_compatible
 * (ref SRE:1484-1546): keep orbit o iff |chi(h) - alpha_h| <= atol for all h in
 * its stabilizer, in catalog order.  chi: [n_g] complex128 (re, im). */
int edx_sector_build(edx_ctx* ctx, const edx_catalog* cat, const double* chi, double atol,
                     edx_sector** out, int64_t* dim);
/* reps: uint64[dim]; stab_order: int32[dim]; orbit_to_basis: int32[n_orbits]
 * (-1 outside the sector).  NULL skips. */
int edx_sector_fetch(edx_ctx* ctx, const edx_sector* sec, uint64_t* reps, int32_t* stab_order,
                     int32_t* orbit_to_basis);
int edx_sector_destroy(edx_sector* sec);

/* ---- sector Hamiltonian (projection table) ---------------------------------
 * Replaces build_matrixfree_projection_table / _vectorized_block_triplets and
 * the CSR of MatrixFreeHamiltonian (ref SRE:595-693, :1719-1769, :1913-1963).
 * Terms: 1-based sites; amplitudes complex128 as (re, im).  Off-diagonal
 * entries t*s_JW*alpha*conj(chi(g))*sqrt(stab_row/stab_col) are stored per row
 * sorted by (column, term); the density diagonal separately.  The sector and
 * catalog must outlive the Hamiltonian. */
int edx_hamiltonian_build(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec,
                          const double* chi, int32_t n_bilinear, const int32_t* i_from,
                          const int32_t* i_to, const double* t, int32_t n_density,
                          const int32_t* v_i, const int32_t* v_j, const double* v,
                          edx_hamiltonian** out, int64_t* n_entries);
/* The same operator with no stored table (replaces the reference's
 * matrix-free product _apply_hamiltonian_kernel, ref SRE:826-923, which
 * recomputes the elements at every call).  Keeps the terms grouped by source
 * site, the irrep characters and the density diagonal; every SpMM walks the
 * valid hops of each row's representative, gathers the target's orbit, group
 * element and alpha from the catalog and forms the element as the stored
 * build does, in FP64.  The first such build on a catalog adds to it a packed
 * copy of its rank tables (16 bytes per mask, kept and refreshed with the
 * catalog's phases); a product after a refresh to a twist is refused.  n_entries receives the entry count the stored table
 * would hold.  Requires a catalog whose site phases are all +1 or -1
 * (EDX_ERR_INVALID otherwise): the product uses H[r][x] = conj(H[x][r]).
 * edx_hamiltonian_fetch_table / _fetch_csr / _set_spmv_group / _autotune
 * return EDX_ERR_INVALID for such a Hamiltonian; edx_matvec, edx_eigs* and
 * edx_matvec_noalloc accept it. */
int edx_hamiltonian_build_on_the_fly(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec,
                                     const double* chi, int32_t n_bilinear, const int32_t* i_from,
                                     const int32_t* i_to, const double* t, int32_t n_density,
                                     const int32_t* v_i, const int32_t* v_j, const double* v,
                                     edx_hamiltonian** out, int64_t* n_entries);
/* Projection table in the reference's triplet form: row/col int64[n_entries],
 * vals complex128[n_entries] (re, im), h_diag complex128[dim].  Triplets come
 * grouped by row, sorted by (column, term) within a row.  NULL skips. */
int edx_hamiltonian_fetch_table(edx_ctx* ctx, const edx_hamiltonian* ham, int64_t* row, int64_t* col,
                                int32_t* term, double* vals, double* h_diag);
/* Rows [row_begin, row_end) of the same table in CSR form (the off-diagonal
 * part of the CSR that MatrixFreeHamiltonian builds, ref SRE:1958-1963):
 * row_ptr int64[n+1] starting at 0, col int32 and vals complex128 of the
 * rows' entries (row_ptr[n] of them, sorted by (column, term) within a row),
 * h_diag complex128[n].  col/vals/h_diag may be NULL. */
int edx_hamiltonian_fetch_csr(edx_ctx* ctx, const edx_hamiltonian* ham, int64_t row_begin, int64_t row_end,
                              int64_t* row_ptr, int32_t* col, double* vals, double* h_diag);
/* Sector block given as a host CSR (off-diagonal part, any order within a
 * row) plus a diagonal: the matrix-mode block of build_ed_Hamiltonian_
 * symmetry_block (ref SRE:2102-2174) placed on the device for edx_eigs and
 * edx_matvec.  row_ptr int64[dim+1], col int32[nnz], vals/h_diag complex128. */
int edx_hamiltonian_from_csr(edx_ctx* ctx, int64_t dim, int64_t nnz, const int64_t* row_ptr, const int32_t* col,
                             const double* vals, const double* h_diag, edx_hamiltonian** out);
/* y = H x for n_vec host vectors of length dim (complex128, column-major
 * [n_vec][dim]); replaces MatrixFreeHamiltonian.__call__ and
 * apply_hamiltonian (ref SRE:1965-1976, :1979-2069). */
int edx_matvec(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t n_vec, const double* x, double* y);
/* Same on device-resident buffers (row-major [dim][n_vec] interleaved input,
 * column-major [n_vec][dim] output; complex128, or complex64 with
 * EDX_PREC_MIXED).  Pointers are device addresses.  Both apply the table
 * itself; for a table built on a twisted catalog (site phases other than +-1,
 * whose alpha the catalog rounds to complex64) the eigensolver's sliced copy
 * holds the Hermitian part (H + H^H) / 2 instead, and these two read the CSR
 * arrays. */
int edx_matvec_device(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t precision, int32_t n_vec,
                      const void* x_dev, void* y_dev);
/* Allocation-free edx_matvec.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, hosts that forbid hipMalloc).  Same arguments, results and
 * errors as edx_matvec, but its three work areas (3 x 16 x dim bytes per
 * vector, vectors processed in batches of up to 16) come from one static
 * 96 MiB device array per GPU, held from the first context on that GPU until
 * the process ends; dim > 2 097 152 -> EDX_ERR_INVALID.  Calls on one GPU
 * share that array: the library serialises them within a process and each
 * call completes on the ctx stream before it returns; separate processes
 * have separate arrays. */
int edx_matvec_noalloc(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t n_vec, const double* x, double* y);
/* Grouped products on device buffers: y_dev[i] = H_i x_dev[i] for n sector
 * Hamiltonians, n_vec <= 16 vectors each (layouts as edx_matvec_device,
 * leading dimension dim_i), in one launch per group of sliced tables of equal
 * lanes per row (on-the-fly tables one by one).  Applies the eigensolver's
 * operator (the Hermitian part for twisted tables, see edx_matvec_device).
 * edx_matvec_group runs it and waits; edx_matvec_group_graph captures the
 * same launches into a HIP graph (no memory operations in it), which
 * edx_graph_launch enqueues on a ctx stream without waiting: every launch
 * reads the buffers' current contents and overwrites y (same results, bit for
 * bit, as edx_matvec_group).  The Hamiltonians and buffers must outlive the
 * graph.  No reference counterpart (the reference applies one sector at a
 * time, SRE:1884-1976). */
int edx_matvec_group(edx_ctx* ctx, int32_t n, const edx_hamiltonian* const* hams, int32_t precision,
                     int32_t n_vec, const void* const* x_dev, void* const* y_dev);
int edx_matvec_group_graph(edx_ctx* ctx, int32_t n, const edx_hamiltonian* const* hams, int32_t precision,
                           int32_t n_vec, const void* const* x_dev, void* const* y_dev, edx_graph** out);
int edx_graph_launch(edx_ctx* ctx, const edx_graph* graph);
int edx_graph_destroy(edx_graph* graph);
int edx_hamiltonian_destroy(edx_hamiltonian* ham);

/* ---- eigensolver ----------------------------------------------------------
 * Replaces diagonalize_block_matrixfree (ref SRE:2248-2299): the nev lowest
 * eigenvalues (ascending, real) of the sector block, by thick-restarted block
 * Lanczos with full reorthogonalisation.  The operator is the Hermitian part
 * of the block (the block itself unless its catalog is twisted; the
 * reference's eigs returns the real parts of the block's eigenvalues, the same
 * numbers up to the complex64 rounding of alpha).  block = 0 picks min(nev, 8).
 * precision: EDX_PREC_FP64 or EDX_PREC_MIXED (complex64 sweeps then FP64
 * refinement).  vecs (optional, may be NULL): complex128 [nev][dim].
 * info (optional): [0] matvecs (vector count), [1] restarts,
 * [2] max residual / max(1,|E|), [3] block size. */
int edx_eigs(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t nev, int32_t block, double tol,
             uint64_t seed, int32_t precision, double* vals, double* vecs, double* info);

/* Options of edx_eigs_ex (edx_eigs fills them from its arguments).
 *   tol      relative residual target ||Hx - Ex|| / max(1,|E|) (ARPACK's tol)
 *   eig_tol  > 0: a pair also counts as converged once the residual-squared-
 *            over-gap bound of its eigenvalue cluster is below eig_tol (the
 *            eigenvalue is then accurate to eig_tol; the vector to ~sqrt)
 *   mixed_tol residual at which EDX_PREC_MIXED hands over to FP64 (default 3e-7;
 *            earlier when the complex64 sweeps stall below max(30 x mixed_tol,
 *            1e-4), as they do under a target below their rounding floor)
 *   basis    Ritz basis size (0: automatic), block as in edx_eigs */
typedef struct edx_eigs_options {
  int32_t nev;
  int32_t block;
  int32_t basis;
  int32_t precision;
  int32_t max_restarts;
  double tol;
  double eig_tol;
  double mixed_tol;
  uint64_t seed;
} edx_eigs_options;
void edx_eigs_default_options(edx_eigs_options* opts);
int edx_eigs_ex(edx_ctx* ctx, const edx_hamiltonian* ham, const edx_eigs_options* opts, double* vals,
                double* vecs, double* info);
/* Batched sector solves: n Hamiltonians (any dimensions; shared options)
 * solved together from the calling thread, as edx_eigs_ex solves each.  Solve
 * i runs on its own context ctxs[i] (distinct, not ctx, on ctx's device; each
 * is claimed for the call, EDX_ERR_BUSY if another thread holds one).  Every
 * round, the block-step SpMMs of the solves that are stepping go to grouped
 * launches on ctx's stream (sliced tables of equal block width and lanes per
 * row in one kernel), replayed from a captured HIP graph of that
 * set of solves (edx_batch_graph_stats), and the rest of each block step runs
 * on the solve's own stream.  vals: [n][nev]; vecs (optional): n pointers,
 * each NULL or complex128 [nev][dim_i]; info (optional): [n][4] as edx_eigs_ex.  Results
 * are the same, bit for bit, as edx_eigs_ex on the same Hamiltonian with the
 * same schedule.  No reference counterpart (the reference solves sectors one
 * after another with ARPACK). */
int edx_eigs_batch(edx_ctx* ctx, int32_t n, edx_ctx* const* ctxs, const edx_hamiltonian* const* hams,
                   const edx_eigs_options* opts, double* vals, double* const* vecs, double* info);
/* Round graphs of the batched solves driven from ctx so far: the grouped
 * SpMM of a round (one launch over the sliced tables of every stepping solve)
 * is captured into a HIP graph once per set of stepping
 * solves in a phase and replayed by every round with that set.  captures /
 * replays (optional) receive the totals.  No reference counterpart. */
int edx_batch_graph_stats(edx_ctx* ctx, int64_t* captures, int64_t* replays);

/* ---- eigenvector consumers -----------------------------------------------------
 * Replaces _expand_sector_state_to_fock_amplitudes (ref
 * observables/entanglement_spectrum.py:82-144): coeff (dim complex) expanded
 * to Fock amplitudes, pruned (|amp| <= atol) and normalised.  convention 0 is
 * the reference's formula, sum_g c/sqrt(|G||Stab|) conj(chi(g)) alpha_g;
 * convention 1 uses chi(g) conj(alpha_g), under which eigenvectors of the
 * sector block expand to eigenvectors of the Fock-space Hamiltonian (the two
 * agree at k = 0 without twist).  Outputs the n_out nonzero amplitudes in
 * ascending mask order; with masks or amps NULL only n_out is returned.
 * A near-zero norm gives EDX_ERR_INVALID, as the reference's ValueError. */
int edx_fock_expand(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec, const double* chi,
                    const double* coeff, double atol, int32_t convention, int64_t capacity, uint64_t* masks,
                    double* amps, int64_t* n_out);
/* Replaces _position_operator_matrix (ref observables/charge_pump.py:102-162):
 * for each column of sec_from, rows[col] = row of the same representative in
 * sec_to with vals[col] = (1/|G|) sum_g chi_to(g) conj(chi_from(g)) U(g.s),
 * U(m) = prod over occupied sites of site_phases (n_site complex); rows[col]
 * = -1 when the orbit is not in sec_to or |elem| <= 1e-13.  Both sectors come
 * from cat. */
int edx_position_operator(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec_to,
                          const edx_sector* sec_from, const double* chi_to, const double* chi_from,
                          const double* site_phases, int64_t* rows, double* vals);
/* Replaces the matrix loop of off_diagonal_long_range_order (ref
 * observables/off_diagonal_long_range_order.py:80-140): rho (n_site x n_site
 * complex, row-major) with rho[i][j] = <psi| a_(i+1)^dag a_(j+1) |psi> for the
 * sector state coeff, using the table projection of project_to_sector
 * (SRE:1660-1690) and the Jordan-Wigner sign of hopping_phase_for_stats. */
int edx_one_body_density(edx_ctx* ctx, const edx_catalog* cat, const edx_sector* sec, const double* chi,
                         const double* coeff, double* rho);

/* ---- SpMV launch shape -------------------------------------------------------
 * Lanes of a wavefront that share one matrix row in the sliced-ELL copy of the
 * table (1, 2 or 4; slices hold 64 / lanes rows).  Setting it on a built
 * Hamiltonian rebuilds that copy.  No reference counterpart (launch-shape
 * tuning of the device product). */
int edx_set_spmv_group(edx_ctx* ctx, int32_t lanes);                 /* default for new builds; 0 = auto */
int edx_hamiltonian_set_spmv_group(edx_ctx* ctx, edx_hamiltonian* ham, int32_t lanes);
int edx_hamiltonian_info(const edx_hamiltonian* ham, int64_t* dim, int64_t* n_entries, int32_t* spmv_group);

/* ---- Palette-packed tables ---------------------------------------------------
 * A sector table holds few distinct elements (hopping amplitude x sign x
 * character x stabilizer-norm ratio, SRE:595-693).  With packing on (the
 * default) a table whose distinct complex128 values number at most
 * min(1024, 2^(32 - bit length of dim)) stores its sliced copy as one 32-bit
 * word per slot, column | palette index << bit length of dim (layout 1), and
 * its values once; the SpMM then streams 4 bytes per entry instead of 12
 * (complex64) or 20 (complex128).  A table of up to 3072 values that do not
 * fit those bits (e.g. the Hermitian parts of twisted blocks) keeps its
 * int32 columns and stores a 16-bit palette index per slot (layout 2, 6 bytes
 * per entry).  Both give results bit-identical to the unpacked copy, which
 * tables of more values keep.  No reference counterpart (storage layout of
 * the device product). */
/* for tables built on ctx from now on; a new context starts with packing on,
 * or off when the environment variable EDX_TABLE_PACKING is "0" */
int edx_set_table_packing(edx_ctx* ctx, int32_t enable);
/* packed: 0 unpacked, 1 32-bit words, 2 columns and 16-bit indices;
 * n_palette: its distinct values;
 * stored_slots: slots of the sliced copy including padding; table_bytes:
 * device bytes of the sliced copy (offsets, words or columns and elements,
 * palettes).  Any output may be NULL. */
int edx_hamiltonian_layout(const edx_hamiltonian* ham, int32_t* packed, int32_t* n_palette, int64_t* stored_slots,
                           int64_t* table_bytes);

/* ---- SpMM autotuner --------------------------------------------------------
 * With tuning on, every Hamiltonian build (edx_hamiltonian_build,
 * edx_hamiltonian_from_csr) picks the lanes per row of its sliced copy from a
 * process-wide table keyed by (precision, block width nb, floor(log2 dim),
 * mean row length in quarter octaves); on a miss it times the 1-, 2- and
 * 4-lane layouts (warm-up plus seven individually timed launches of the
 * SpMM at that precision and width, on the build's stream; the fastest
 * launch counts), records the three times and keeps the fastest.  A fixed edx_set_spmv_group choice overrides tuning; sectors with
 * dim <= 500 (dense path) are not tuned.  edx_autotune_table returns, per
 * entry, keys[5*i..] = (precision, nb, dim_log2, row_quarter_octaves, lanes)
 * and ms[3*i..] = average launch times of the 1/2/4-lane layouts; n is the
 * table size (entries beyond capacity are not written).
 * edx_hamiltonian_autotune times and rebuilds one Hamiltonian without the
 * table (the Hamiltonian must not be in use by another call meanwhile).
 * No reference counterpart (launch-shape tuning of the device product). */
int edx_set_autotune(int32_t enable, int32_t precision, int32_t nb);
int edx_autotune_table(int32_t capacity, int32_t* keys, float* ms, int32_t* n);
int edx_autotune_clear(void);
int edx_hamiltonian_autotune(edx_ctx* ctx, edx_hamiltonian* ham, int32_t precision, int32_t nb, int32_t* lanes,
                             float* ms);
/* Eigensolver schedules.  With tuning on, each eigensolve of dim >= 65536
 * and block width <= 8 looks up (precision, block, floor(log2 dim), basis
 * size) in a second process-wide table; on a miss it times, on its own
 * Krylov banks at the shape of a block step three quarters into the basis,
 * the Gram of [Q, W] against W on vector FMAs and (complex64) on the FP32
 * matrix cores, and the three block-update variants (vectors per prefetch
 * stage, rows per lane: 2/2, 2/1, 4/1), and keeps the fastest of each when
 * it beats the default (matrix cores for complex64 from 2^20 rows; update
 * 2/2 for complex64, 2/1 for complex128) by more than 3 %.  The
 * update variants sum every output in the same order, so they give
 * identical results.  edx_schedule_table returns keys[6*i..] = (precision,
 * block, dim_log2, basis, gram, update) with gram 1 = vector FMAs, 2 =
 * matrix cores and update 1..3, and ms[5*i..] = fastest launch times of the
 * two Gram schedules (0 when not timed) and the three updates.
 * edx_set_solver_schedule fixes a context's choices (0 = tuned or default),
 * overriding the table.  edx_autotune_clear clears both tables.
 * No reference counterpart. */
int edx_schedule_table(int32_t capacity, int32_t* keys, float* ms, int32_t* n);
int edx_set_solver_schedule(edx_ctx* ctx, int32_t gram, int32_t update);
/* Combination of device vectors, the restart step of the eigensolver that
 * replaces diagonalize_block_matrixfree (ref SRE:2248-2299): out[:, k] =
 * sum_j v[:, j] y[j][k] for j < nv, k < nk (1..240 each), v and out
 * column-major with leading dimension ld > dim (complex128 for
 * EDX_PREC_FP64, complex64 for EDX_PREC_MIXED; device addresses; out must
 * not overlap v), y host row-major complex128 (re, im pairs, nv x nk).
 * schedule 1 sums in FP64 on vector FMAs (the result is the exact sum
 * rounded once); 2 runs on the FP32 matrix cores (complex64 only; needs
 * ld >= dim rounded up to 32, ld even, 16-byte aligned buffers, and
 * nv <= 112 for nk > 32, nv <= 204 otherwise): coefficients rounded to
 * FP32, FP32 sums of 8
 * terms added in FP64, so each output errs by a few FP32 roundings of its
 * terms; 0 picks 2 for complex64 combinations of more than 16 outputs from
 * 65536 rows when it applies, else 1.  Synchronous; timed by
 * edx_kernel_stats as class 4 (vector FMAs) or 5 (matrix cores).  edx_set_combine_schedule fixes the schedule of the
 * eigensolver's combinations of more than 16 outputs on a context (0 =
 * automatic).  No reference counterpart as entry points. */
int edx_combine_vectors(edx_ctx* ctx, int32_t precision, const void* v_dev, int64_t ld, int64_t dim, int32_t nv,
                        const double* y, int32_t nk, void* out_dev, int32_t schedule);
int edx_set_combine_schedule(edx_ctx* ctx, int32_t schedule);

/* ---- host-only helper (no device needed) ------------------------------------
 * Eigen-decomposition of a Hermitian n x n matrix (row-major (re, im) pairs)
 * from its upper (upper != 0, as LAPACK eigh(UPLO="U") in ref SRE:2281) or
 * lower triangle; w ascending; z (optional) eigenvectors as columns. */
int edx_host_herm_eig(int32_t n, const double* a, int32_t upper, double* w, double* z);

/* ---- live kernel timing -------------------------------------------------------
 * No reference counterpart.  With timing enabled every complex128 SpMM (class
 * 0), complex64 SpMM (1), Gram (2), block update (3) and combination of more
 * than 16 outputs (restarts, Ritz vectors, edx_combine_vectors) on FP64 vector
 * FMAs (4) or on the FP32 matrix cores (5) launch on the ctx stream is
 * bracketed by HIP events; ms/launches/bytes (arrays of
 * EDX_KERNEL_CLASSES, NULL skips) receive the totals (bytes = algorithmic
 * bytes of each launch).  enable: 1 on, 0 off, -1 keep; reset != 0 clears the
 * totals first. */
#define EDX_KERNEL_CLASSES 6
int edx_kernel_stats(edx_ctx* ctx, int32_t enable, int32_t reset, double* ms, int64_t* launches, double* bytes);

/* Arithmetic of the same timed launches (flops[EDX_KERNEL_CLASSES], totals
 * since the last reset): 8 flops per complex multiply-add of the Gram,
 * update and combination classes (FP64 FMAs, or FP32 matrix-core products
 * for the complex64 matrix-core Gram and class 5), 0 for the SpMM classes.
 * No reference counterpart. */
int edx_kernel_flops(edx_ctx* ctx, double* flops);

/* ---- device helpers for benchmarks and tests ------------------------------ */
int edx_device_alloc(edx_ctx* ctx, uint64_t bytes, void** out);
int edx_device_free(edx_ctx* ctx, void* p);
int edx_memcpy_h2d(edx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int edx_memcpy_d2h(edx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Best time of n_rep launches (after one warm-up) of a plain streaming read of
 * `bytes` (a positive multiple of 16) at src: the practical read ceiling the
 * bandwidth-bound kernels are compared with.  No reference counterpart. */
int edx_time_stream_read(edx_ctx* ctx, const void* src, uint64_t bytes, int32_t n_rep, double* best_ms);
/* Free and total device memory of the ctx's device (hipMemGetInfo); used by
 * the serving loop to admit work by HBM footprint. */
int edx_mem_info(edx_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* Device-to-device copy on the ctx stream (runtime blit kernel); a plain copy
 * of known size, used to calibrate memory counters. */
int edx_device_copy(edx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Average duration (ms) of the last n_rep launches of the eigensolver's SpMM
 * (y_dev = its operator x_dev, layouts as edx_matvec_device) timed with HIP
 * events on the ctx stream, inputs resident in device memory. */
int edx_time_matvec(edx_ctx* ctx, const edx_hamiltonian* ham, int32_t precision, int32_t n_vec,
                    int32_t n_rep, const void* x_dev, void* y_dev, double* avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* EDX_H */
