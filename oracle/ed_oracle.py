"""CPU oracle for the symmetry-resolved ED path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker (or as the timed
CPU baseline).  The product path never calls it: the package raises when its
HIP library is missing instead of falling back to anything here.

This is a numpy/scipy restatement of the reference algorithm
(``src/realspace_exactdiagonalization_py/symmetry_resolved_ed.py``, written
``SRE`` below), array based: a group is ``perm`` (``(nG, N)`` int, 1-based
targets) plus ``phases`` (``(nG, N)`` complex128).  Floating-point steps keep
the reference's operation order so that values agree bit for bit where the
reference's own numpy calls are deterministic.

Deliberate differences from the reference, each named where it applies:

* ranks for ``N > 42`` use a correct O(N) colex rank; the reference passes
  ``n_site=0`` to its fallback there and every rank becomes 0 (``SRE:665``,
  ``:896``) and its catalog scatter hands a 2-D array to a 1-D routine
  (``SRE:554-555``).  Below ``N = 42`` both give the LinTable rank, which
  equals the colex rank.

Pinning (each function is checked in ``tests/test_oracle_golden.py`` against
the reference's notebook/CSV/test data or against an independent property):
see the docstring of each function.  Nothing here is "parity unpinned" except
what is stated at ``update_orbit_stabilizer_phases`` for nonzero flux.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

LIN_TABLE_MAX_N_SITE = 42


# ---------------------------------------------------------------------------
# Combinadics  (SRE:110-211, 377-426)
# ---------------------------------------------------------------------------

def binom_table(n_site: int, n_filled: int) -> np.ndarray:
    """``C(b, j)`` for ``b <= n_site``, ``j <= n_filled`` (SRE:143-158).
    Pinned: closed form (math.comb)."""
    t = np.zeros((n_site + 1, n_filled + 1), dtype=np.int64)
    for b in range(n_site + 1):
        for j in range(min(b, n_filled) + 1):
            t[b, j] = math.comb(b, j)
    return t


def rank_colex(m: int, bt: np.ndarray) -> int:
    """``sum_j C(b_j, j+1)`` over set bits ``b_0 < b_1 < ...`` (SRE:1193-1217).
    Pinned: design.ipynb cell 12 (N=4,k=2 table) and bijection onto [0, C)."""
    r, j = 0, 1
    while m:
        low = m & -m
        r += int(bt[low.bit_length() - 1, j])
        j += 1
        m ^= low
    return r


def lin_table(n_site: int, n_filled: int):
    """Split tables ``L[hi] + R[lo]`` (SRE:166-211).
    Pinned: design.ipynb cell 14 (N=12,k=3 sizes 64+64, n_right 6, 0 mismatches)."""
    n_right = n_site // 2
    n_left = n_site - n_right
    right = np.zeros(1 << n_right, dtype=np.int64)
    for v in range(1 << n_right):
        r, j, tmp = 0, 1, v
        while tmp:
            low = tmp & -tmp
            r += math.comb(low.bit_length() - 1, j)
            j += 1
            tmp ^= low
        right[v] = r
    left = np.zeros(1 << n_left, dtype=np.int64)
    acc = 0
    for u in range(1 << n_left):
        left[u] = acc
        rem = n_filled - u.bit_count()
        if 0 <= rem <= n_right:
            acc += math.comb(n_right, rem)
    return left, right, n_right


def rank_vec(masks: np.ndarray, n_site: int, bt: np.ndarray, lin=None) -> np.ndarray:
    """Vectorised colex rank of any-shaped uint64 array.  Uses the LinTable when
    given (``SRE:397-401``), else an O(N) bit walk valid for all ``N <= 63``."""
    masks = np.asarray(masks, dtype=np.uint64)
    if lin is not None:
        left, right, nr = lin
        lo = (np.uint64(1) << np.uint64(nr)) - np.uint64(1)
        return left[masks >> np.uint64(nr)] + right[masks & lo]
    flat = masks.reshape(-1)
    rank = np.zeros(flat.shape[0], dtype=np.int64)
    j = np.ones(flat.shape[0], dtype=np.int64)
    for i in range(n_site):
        bit = ((flat >> np.uint64(i)) & np.uint64(1)).astype(bool)
        if bit.any():
            rank[bit] += bt[i, j[bit]]
            j[bit] += 1
    return rank.reshape(masks.shape)


def unrank_chunk(ranks: np.ndarray, n_site: int, n_filled: int, bt: np.ndarray) -> np.ndarray:
    """Masks of colex ranks (SRE:377-394).  Pinned: unrank(rank(m)) == m."""
    masks = np.zeros(len(ranks), dtype=np.uint64)
    r = np.asarray(ranks, dtype=np.int64).copy()
    for j in range(n_filled, 0, -1):
        b = np.searchsorted(bt[:, j], r, side="right") - 1
        b = np.minimum(b, n_site - 1)
        masks |= np.uint64(1) << b.astype(np.uint64)
        r -= bt[b, j]
    return masks


def gosper_next(x: int) -> int:
    """Next mask of equal popcount (SRE:1117-1129).  Pinned: design.ipynb cell 5."""
    c = x & -x
    r = x + c
    return (((r ^ x) >> 2) // c) | r


# ---------------------------------------------------------------------------
# Group action  (SRE:977-1008, 1087-1114, 346-374)
# ---------------------------------------------------------------------------

def apply_op(m: int, perm_row, phase_row, fermionic: bool):
    """``(g.m, alpha_g(m))``, phase as a running complex128 product over occupied
    sites in ascending order times the inversion parity (SRE:977-1008).
    Pinned: test_core.py:54-74, design.ipynb cell 18."""
    tmp, new, phase, parity = m, 0, 1.0 + 0.0j, 1
    while tmp:
        low = tmp & -tmp
        i = low.bit_length() - 1
        p = int(perm_row[i])
        if fermionic and (new >> p).bit_count() % 2 == 1:
            parity = -parity
        new |= 1 << (p - 1)
        phase *= complex(phase_row[i])
        tmp ^= low
    return new, phase * parity


def canonical_rep_scalar(m: int, perm, phases, fermionic: bool, identity_idx: int = 0):
    """Min over the group with strict ``<`` so the first g wins (SRE:1087-1114).
    Pinned: test_core.py:76-87."""
    best, best_g, best_a = m, identity_idx, 1.0 + 0.0j
    for g in range(perm.shape[0]):
        s, a = apply_op(m, perm[g], phases[g], fermionic)
        if s < best:
            best, best_g, best_a = s, g, a
    return best, best_g, best_a


def byte_tables(n_site: int, perm: np.ndarray, phases: np.ndarray):
    """Byte-lookup image/angle tables (SRE:346-374): angles summed bit-ascending
    within the byte, starting at 0.0."""
    n_bytes = (n_site + 7) // 8
    n_g = perm.shape[0]
    ptbl = np.zeros((n_g, n_bytes, 256), dtype=np.uint64)
    atbl = np.zeros((n_g, n_bytes, 256), dtype=np.float64)
    ang = np.angle(phases)
    for g in range(n_g):
        for b in range(n_bytes):
            for v in range(256):
                out, a = 0, 0.0
                for i in range(8):
                    si = 8 * b + i
                    if si >= n_site:
                        break
                    if (v >> i) & 1:
                        out |= 1 << (int(perm[g, si]) - 1)
                        a += ang[g, si]
                ptbl[g, b, v] = out
                atbl[g, b, v] = a
    return ptbl, atbl, n_bytes


# ---------------------------------------------------------------------------
# Orbit catalog  (SRE:433-575, 1298-1404)
# ---------------------------------------------------------------------------

@dataclass
class Catalog:
    n_site: int
    n_filled: int
    perm: np.ndarray
    phases: np.ndarray
    fermionic: bool
    reps: np.ndarray                       # uint64 ascending
    stab_order: np.ndarray                 # int64
    stab_g: list = field(default_factory=list)      # list of int64 arrays
    stab_phase: list = field(default_factory=list)  # list of complex128 arrays
    rep_rank: np.ndarray | None = None     # int32 [C]
    rep_sym: np.ndarray | None = None      # int32 [C]
    rep_amp: np.ndarray | None = None      # complex64 [C]
    bt: np.ndarray | None = None
    lin: tuple | None = None

    def rank(self, masks):
        return rank_vec(masks, self.n_site, self.bt, self.lin)


def build_catalog(n_site: int, n_filled: int, perm, phases, fermionic: bool,
                  chunk: int = 1 << 16) -> Catalog:
    """Orbit-stabilizer decomposition and dense ``(rep, g, alpha)`` tables.

    Representative = min image; table entry of ``g.rep`` written for g ascending
    so the largest g wins; ``alpha = complex64(exp(i sum angle) * sign)`` with
    the byte-ordered angle sum; stabilizer = g with ``g.rep == rep`` ascending
    (SRE:483-571).
    Pinned: design.ipynb cell 16 (Haldane 2x3: 38 orbits, rep[3] = 21), the
    reference CSV orbit counts, closed-form character counts, and the
    table invariant ``apply_op(rep, g) == (mask, alpha)`` checked per state.
    """
    perm = np.asarray(perm, dtype=np.int64)
    phases = np.asarray(phases, dtype=np.complex128)
    n_g = perm.shape[0]
    n_total = math.comb(n_site, n_filled)
    bt = binom_table(n_site, n_filled)
    lin = lin_table(n_site, n_filled) if n_site <= LIN_TABLE_MAX_N_SITE else None
    rep_rank = np.full(n_total, -1, dtype=np.int32)
    rep_sym = np.zeros(n_total, dtype=np.int32)
    rep_amp = np.zeros(n_total, dtype=np.complex64)
    cat = Catalog(n_site, n_filled, perm, phases, fermionic, None, None, [], [],
                  rep_rank, rep_sym, rep_amp, bt, lin)
    perm_bits = (np.uint64(1) << (perm - 1).astype(np.uint64))
    reps: list[int] = []
    stab_order: list[int] = []

    if n_filled == 0 or n_filled == n_site:
        m = (1 << n_filled) - 1
        gs, ps = [], []
        for g in range(n_g):
            s, a = apply_op(m, perm[g], phases[g], fermionic)
            if s == m:
                gs.append(g)
                ps.append(a)
        reps.append(m)
        stab_order.append(len(gs))
        cat.stab_g.append(np.asarray(gs, dtype=np.int64))
        cat.stab_phase.append(np.asarray(ps, dtype=np.complex128))
        for g in range(n_g):
            s, a = apply_op(m, perm[g], phases[g], fermionic)
            idx = rank_colex(s, bt)
            rep_rank[idx] = 0
            rep_sym[idx] = g
            rep_amp[idx] = np.complex64(a)
    else:
        ptbl, atbl, n_bytes = byte_tables(n_site, perm, phases)
        inv = np.zeros((n_g, n_site, n_site), dtype=np.int64)
        if fermionic:
            for g in range(n_g):
                for i in range(n_site):
                    for j in range(i + 1, n_site):
                        if perm[g, i] > perm[g, j]:
                            inv[g, i, j] = 1
        for start in range(0, n_total, chunk):
            stop = min(start + chunk, n_total)
            masks = unrank_chunk(np.arange(start, stop, dtype=np.int64), n_site, n_filled, bt)
            mn = np.full(stop - start, np.iinfo(np.uint64).max, dtype=np.uint64)
            for g in range(n_g):
                img = np.zeros(stop - start, dtype=np.uint64)
                for b in range(n_bytes):
                    img |= np.take(ptbl[g, b], ((masks >> np.uint64(8 * b)) & np.uint64(0xFF)).astype(np.int64))
                mn = np.minimum(mn, img)
            pos = np.nonzero(mn == masks)[0]
            if pos.size == 0:
                continue
            first = len(reps)
            rm = masks[pos]
            imgs = np.zeros((n_g, pos.size), dtype=np.uint64)
            angs = np.zeros((n_g, pos.size), dtype=np.float64)
            for g in range(n_g):
                for b in range(n_bytes):
                    by = ((rm >> np.uint64(8 * b)) & np.uint64(0xFF)).astype(np.int64)
                    imgs[g] |= np.take(ptbl[g, b], by)
                    angs[g] += np.take(atbl[g, b], by)
            ph = np.exp(1j * angs)
            if fermionic:
                bits = ((rm[:, None] >> np.arange(n_site, dtype=np.uint64)) & np.uint64(1)).astype(np.int64)
                for g in range(n_g):
                    n_inv = (bits * (bits @ inv[g])).sum(axis=1)
                    ph[g, (n_inv & 1) == 1] *= -1.0 + 0.0j
            rk = cat.rank(imgs)
            orbit_ids = np.arange(first, first + pos.size, dtype=np.int64)
            for g in range(n_g):
                rep_rank[rk[g]] = orbit_ids
                rep_sym[rk[g]] = g
                rep_amp[rk[g]] = ph[g].astype(np.complex64)
            fix = imgs == rm[None, :]
            for j in range(pos.size):
                gs = np.nonzero(fix[:, j])[0]
                reps.append(int(rm[j]))
                stab_order.append(len(gs))
                cat.stab_g.append(gs.astype(np.int64))
                cat.stab_phase.append(ph[gs, j].astype(np.complex128))
        del perm_bits
    cat.reps = np.asarray(reps, dtype=np.uint64)
    cat.stab_order = np.asarray(stab_order, dtype=np.int64)
    return cat


def table_entry(m: int, n_site: int, perm, phases, fermionic: bool, tables=None):
    """``(rep, g, alpha)`` of one mask as the catalog stores it, without the
    catalog: rep = min image over the group, g = the largest element with
    ``g.rep == m``, alpha = complex64 of the exp-form phase of g on rep
    (byte-ordered angle sum, times -1 for an odd number of inversions).  For
    spot checks of catalogs too large for ``build_catalog``.  Pinned: equals
    ``build_catalog``'s tables on every mask of the small cases (tests)."""
    perm = np.asarray(perm, dtype=np.int64)
    n_g = perm.shape[0]
    ptbl, atbl, n_bytes = tables if tables is not None else byte_tables(n_site, perm, np.asarray(phases))

    def image(x, g):
        out = 0
        for b in range(n_bytes):
            out |= int(ptbl[g, b, (x >> (8 * b)) & 0xFF])
        return out

    rep = min(image(m, g) for g in range(n_g))
    g_tab = max(g for g in range(n_g) if image(rep, g) == m)
    ang = 0.0
    for b in range(n_bytes):
        ang += atbl[g_tab, b, (rep >> (8 * b)) & 0xFF]
    a = np.exp(1j * np.float64(ang))
    if fermionic:
        occ = [i for i in range(n_site) if (rep >> i) & 1]
        n_inv = sum(1 for x in range(len(occ)) for y in range(x + 1, len(occ))
                    if perm[g_tab, occ[x]] > perm[g_tab, occ[y]])
        if n_inv & 1:
            a = a * (-1.0 + 0.0j)
    return rep, g_tab, np.complex64(a)


def update_orbit_stabilizer_phases(cat: Catalog, new_phases, fermionic: bool) -> Catalog:
    """In-place flux refresh (SRE:1220-1269): stabilizer phases and table
    amplitudes recomputed as running products (``apply_op``) for the group with
    the same permutations and new per-site phases; table g unchanged
    (largest-g-wins again).  Pinned at zero flux by equality with a fresh
    catalog; at nonzero flux by ``|alpha_prod - alpha_fresh| <= 1 ulp(c64)`` and
    the per-state invariant.  Exact complex64 values at nonzero flux: parity
    unpinned against the reference itself (no stored values exist)."""
    new_phases = np.asarray(new_phases, dtype=np.complex128)
    cat.phases = new_phases
    for o in range(len(cat.reps)):
        r = int(cat.reps[o])
        for j, g in enumerate(cat.stab_g[o]):
            _, a = apply_op(r, cat.perm[g], new_phases[g], fermionic)
            cat.stab_phase[o][j] = a
    n_g = cat.perm.shape[0]
    for o in range(len(cat.reps)):
        r = int(cat.reps[o])
        for g in range(n_g):
            s, a = apply_op(r, cat.perm[g], new_phases[g], fermionic)
            idx = rank_colex(s, cat.bt)
            if cat.rep_rank[idx] == o:
                cat.rep_sym[idx] = g
                cat.rep_amp[idx] = np.complex64(a)
    return cat


# ---------------------------------------------------------------------------
# Irrep filter  (SRE:1484-1546)
# ---------------------------------------------------------------------------

def sector_basis(cat: Catalog, chi, atol: float = 1e-12):
    """Orbits with ``|chi(h) - alpha_h| <= atol`` for all stabilizer h, in
    catalog order.  Returns ``(reps, stab_order, orbit_to_basis)``.
    Pinned: reference CSV ``sector_dim`` and closed-form dims."""
    chi = np.asarray(chi, dtype=np.complex128)
    keep = np.zeros(len(cat.reps), dtype=bool)
    for o in range(len(cat.reps)):
        ok = True
        for g, a in zip(cat.stab_g[o], cat.stab_phase[o]):
            if not abs(chi[g] - a) <= atol:
                ok = False
                break
        keep[o] = ok
    o2b = np.full(len(cat.reps), -1, dtype=np.int64)
    o2b[keep] = np.arange(int(keep.sum()))
    return cat.reps[keep].copy(), cat.stab_order[keep].copy(), o2b


# ---------------------------------------------------------------------------
# Projection table, matvec, eigensolve  (SRE:595-693, 1884-1976, 2248-2299)
# ---------------------------------------------------------------------------

def term_arrays(terms):
    if not terms:
        return (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.complex128))
    return (np.asarray([t[0] for t in terms], dtype=np.int64),
            np.asarray([t[1] for t in terms], dtype=np.int64),
            np.asarray([complex(t[2]) for t in terms], dtype=np.complex128))


def projection_table(cat: Catalog, basis_reps, basis_stab, o2b, bilinear, density, chi,
                     fermionic: bool, with_terms: bool = False):
    """``(row, col, vals, h_diag)`` with
    ``vals = t * s_JW * alpha * conj(chi(g)) * sqrt(stab_row / stab_col)``, the
    term moving a particle from its first site to its second and the JW string
    taken on the column representative (SRE:631-693).
    Pinned: design.ipynb cells 22/24 (Haldane 2x3 k=0: 1120 entries, CSR nnz
    756 with the diagonal, energies) and the golden energies."""
    chi = np.asarray(chi, dtype=np.complex128)
    reps = np.asarray(basis_reps, dtype=np.uint64)
    stab = np.asarray(basis_stab, dtype=np.int64)
    D = reps.shape[0]
    fi, ti, tt = term_arrays(bilinear)
    vi, vj, vv = term_arrays(density)
    hop = (np.uint64(1) << (fi - 1).astype(np.uint64)) | (np.uint64(1) << (ti - 1).astype(np.uint64))
    frm = np.uint64(1) << (fi - 1).astype(np.uint64)
    h_diag = np.zeros(D, dtype=np.complex128)
    for k in range(vi.shape[0]):
        both = ((reps & (np.uint64(1) << np.uint64(vi[k] - 1))) != 0) & \
               ((reps & (np.uint64(1) << np.uint64(vj[k] - 1))) != 0)
        h_diag[both] += vv[k]
    in_basis = o2b >= 0
    rows, cols, vals, terms = [], [], [], []
    for k in range(fi.shape[0]):
        cid = np.nonzero((reps & hop[k]) == frm[k])[0]
        if cid.size == 0:
            continue
        idx = cat.rank(reps[cid] ^ hop[k])
        orb = cat.rep_rank[idx]
        keep = np.nonzero(in_basis[orb])[0]
        if keep.size == 0:
            continue
        r = o2b[orb[keep]]
        c = cid[keep]
        g = cat.rep_sym[idx[keep]]
        a = cat.rep_amp[idx[keep]]
        s = np.ones(keep.size, dtype=np.complex128)
        if fermionic:
            lo, hi = min(int(fi[k]), int(ti[k])), max(int(fi[k]), int(ti[k]))
            between = ((1 << (hi - lo - 1)) - 1) << lo if hi - lo > 1 else 0
            flip = (np.bitwise_count(reps[c] & np.uint64(between)) & 1) == 1
            s[flip] = -1.0 + 0.0j
        v = tt[k] * s * a * np.conj(chi[g]) * np.sqrt(stab[r] / stab[c])
        rows.append(r)
        cols.append(c)
        vals.append(v)
        terms.append(np.full(keep.size, k, dtype=np.int64))
    if rows:
        out = (np.concatenate(rows).astype(np.int64), np.concatenate(cols).astype(np.int64),
               np.concatenate(vals), h_diag)
        return out + (np.concatenate(terms),) if with_terms else out
    out = (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.complex128), h_diag)
    return out + (np.zeros(0, np.int64),) if with_terms else out


def csr_hamiltonian(row, col, vals, h_diag, D: int) -> sp.csr_matrix:
    """CSR with duplicates summed, diagonal added, zeros eliminated (SRE:1958-1963)."""
    H = sp.csr_matrix((vals, (row, col)), shape=(D, D))
    H.setdiag(H.diagonal() + h_diag)
    H.eliminate_zeros()
    return H


def reference_start_vector(n: int) -> np.ndarray:
    """Seeded complex Gaussian start vector (SRE:2291-2293)."""
    rng = np.random.default_rng(12345)
    x0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return x0 / np.linalg.norm(x0)


def diagonalize(H: sp.csr_matrix, nev: int):
    """Dense ``eigh(UPLO='U')`` for ``n <= 500`` (matvec against unit vectors),
    else ``eigs(which='SR', tol=1e-10, v0=seeded)``; real parts ascending
    (SRE:2248-2299).  Pinned: test_energy_alignment.py values, CSV energies."""
    n = H.shape[0]
    if n == 0:
        return np.zeros(0), np.zeros((0, 0), dtype=np.complex128)
    if n <= 500:
        Hd = np.empty((n, n), dtype=np.complex128)
        x = np.zeros(n, dtype=np.complex128)
        for j in range(n):
            x[j] = 1.0
            Hd[:, j] = H @ x
            x[j] = 0.0
        w, v = np.linalg.eigh(Hd, UPLO="U")
        k = min(nev, n)
        return w[:k].real, v[:, :k]
    op = spla.LinearOperator((n, n), matvec=lambda x: H @ np.ascontiguousarray(x, dtype=np.complex128),
                             dtype=np.complex128)
    w, v = spla.eigs(op, k=nev, which="SR", v0=reference_start_vector(n), tol=1e-10)
    o = np.argsort(w.real)
    return w.real[o], v[:, o]


# ---------------------------------------------------------------------------
# Independent checks (not restatements): identity-group Fock matrix
# ---------------------------------------------------------------------------

def fock_hamiltonian(n_site: int, n_filled: int, bilinear, density, fermionic: bool) -> sp.csr_matrix:
    """Plain fixed-filling Fock-basis matrix of the engine's operator (term
    ``(i, j, t)`` moves a particle from i to j with amplitude t, JW string
    between them).  Used to check that the union of sector spectra equals the
    full spectrum; independent of the orbit machinery."""
    bt = binom_table(n_site, n_filled)
    C = math.comb(n_site, n_filled)
    masks = unrank_chunk(np.arange(C, dtype=np.int64), n_site, n_filled, bt)
    fi, ti, tt = term_arrays(bilinear)
    vi, vj, vv = term_arrays(density)
    diag = np.zeros(C, dtype=np.complex128)
    for k in range(vi.shape[0]):
        both = ((masks >> np.uint64(vi[k] - 1)) & np.uint64(1)).astype(bool) & \
               ((masks >> np.uint64(vj[k] - 1)) & np.uint64(1)).astype(bool)
        diag[both] += vv[k]
    rows, cols, vals = [np.arange(C)], [np.arange(C)], [diag]
    for k in range(fi.shape[0]):
        b_i = np.uint64(1) << np.uint64(fi[k] - 1)
        b_j = np.uint64(1) << np.uint64(ti[k] - 1)
        ok = ((masks & b_i) != 0) & ((masks & b_j) == 0)
        c = np.nonzero(ok)[0]
        new = masks[c] ^ (b_i | b_j)
        s = np.ones(c.size)
        if fermionic:
            lo, hi = min(int(fi[k]), int(ti[k])), max(int(fi[k]), int(ti[k]))
            between = ((1 << (hi - lo - 1)) - 1) << lo if hi - lo > 1 else 0
            s[(np.bitwise_count(masks[c] & np.uint64(between)) & 1) == 1] = -1.0
        rows.append(rank_vec(new, n_site, bt))
        cols.append(c)
        vals.append(tt[k] * s)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(C, C))


def sector_dims_closed_form(n_site: int, n_filled: int, perm, chis, fermionic: bool):
    """``dim_k = (1/|G|) sum_g conj(chi_k(g)) Tr(U_g)`` with the trace over the
    fixed-filling space computed from the cycle structure of each permutation
    (zero flux only).  Independent of the catalog code."""
    perm = np.asarray(perm, dtype=np.int64)
    n_g = perm.shape[0]
    traces = []
    for g in range(n_g):
        seen = [False] * n_site
        cycles = []
        for i in range(n_site):
            if not seen[i]:
                L, j = 0, i
                while not seen[j]:
                    seen[j] = True
                    j = int(perm[g, j]) - 1
                    L += 1
                cycles.append(L)
        # count fixed-filling configurations invariant under g: choose whole cycles
        # summing to n_filled; fermionic sign of a filled L-cycle is (-1)^(L-1)
        poly = {0: 1}
        for L in cycles:
            new = dict(poly)
            w = (-1) ** (L - 1) if fermionic else 1
            for n, c in poly.items():
                new[n + L] = new.get(n + L, 0) + c * w
            poly = new
        traces.append(poly.get(n_filled, 0))
    traces = np.asarray(traces, dtype=np.complex128)
    dims = []
    for chi in chis:
        d = np.sum(np.conj(np.asarray(chi)) * traces) / n_g
        dims.append(int(round(d.real)))
    return dims


# ---------------------------------------------------------------------------
# Eigenvector consumers  (observables/entanglement_spectrum.py:82-144,
# observables/charge_pump.py:41-162, observables/off_diagonal_long_range_order.py:43-140)
# ---------------------------------------------------------------------------

def fock_expand(cat: Catalog, reps, stab, chi, c, atol: float = 1e-12, convention: str = "reference") -> dict:
    """Sector eigenvector -> normalised Fock amplitudes ``{mask: amp}``
    (entanglement_spectrum.py:82-144): for every basis column with
    ``|c| > atol`` and every g ascending, add
    ``(c / sqrt(|G| |Stab|)) * conj(chi(g)) * alpha_g`` at ``g.rep``; prune
    ``|amp| <= atol``; renormalise.  ``convention="hamiltonian"`` uses
    ``chi(g) * conj(alpha_g)`` instead.  Pinned: at k = 0 without twist both
    give the Fock-space eigenvector of ``fock_hamiltonian`` with the sector
    eigenvalue; for any sector and twist the "hamiltonian" expansion does; the
    reference expansion of c is the conjugate of the "hamiltonian" expansion
    of conj(c) (tests)."""
    n_g = cat.perm.shape[0]
    chi = np.asarray(chi, dtype=np.complex128)
    amps: dict = {}
    for col, s in enumerate(reps):
        coeff = c[col]
        if abs(coeff) <= atol:
            continue
        pref = coeff / np.sqrt(n_g * int(stab[col]))
        for g in range(n_g):
            m, ph = apply_op(int(s), cat.perm[g], cat.phases[g], cat.fermionic)
            if convention == "reference":
                amp = pref * np.conj(chi[g]) * ph
            else:
                amp = pref * chi[g] * np.conj(ph)
            amps[m] = amps.get(m, 0.0 + 0.0j) + amp
    for m in list(amps):
        if abs(amps[m]) <= atol:
            del amps[m]
    norm2 = sum(abs(a) ** 2 for a in amps.values())
    if norm2 <= atol:
        raise ValueError("Expanded state has near-zero norm.")
    inv = 1.0 / np.sqrt(norm2)
    for m in amps:
        amps[m] *= inv
    return amps


def position_phase(m: int, site_phases) -> complex:
    """``prod over occupied sites (ascending) of site_phases`` (charge_pump.py:81-99)."""
    z = 1.0 + 0.0j
    while m:
        low = m & -m
        z *= site_phases[low.bit_length() - 1]
        m ^= low
    return z


def position_operator(cat: Catalog, reps_to, reps_from, chi_to, chi_from, site_phases):
    """``<chi_to| U |chi_from>`` as COO ``(rows, cols, vals)`` (charge_pump.py:102-162):
    column ``col`` (rep s of the source basis) pairs with the row of s in the
    target basis; ``elem = (1/|G|) sum_g chi_to(g) conj(chi_from(g)) U(g.s)``,
    kept when ``|elem| > 1e-13``.  Pinned: at zero twist and ``U = 1`` it is the
    identity on shared orbits of equal irreps and zero between different
    irreps (character orthogonality), checked in the tests."""
    n_g = cat.perm.shape[0]
    where = {int(s): i for i, s in enumerate(reps_to)}
    inv_ng = 1.0 / n_g
    rows, cols, vals = [], [], []
    for col, s in enumerate(reps_from):
        row = where.get(int(s))
        if row is None:
            continue
        elem = 0.0 + 0.0j
        for g in range(n_g):
            m, _ = apply_op(int(s), cat.perm[g], cat.phases[g], cat.fermionic)
            elem += chi_to[g] * np.conj(chi_from[g]) * position_phase(m, site_phases)
        elem *= inv_ng
        if abs(elem) > 1e-13:
            rows.append(row)
            cols.append(col)
            vals.append(elem)
    return np.asarray(rows, np.int64), np.asarray(cols, np.int64), np.asarray(vals, np.complex128)


def one_body_density_matrix(cat: Catalog, reps, stab, o2b, chi, c):
    """``rho[i, j] = <psi| a_{i+1}^dag a_{j+1} |psi>`` of a sector state
    (off_diagonal_long_range_order.py:80-140), restated with the table lookup
    of project_to_sector (SRE:1660-1690) and the JW sign of
    hopping_phase_for_stats (SRE:1777-1810).  Pinned: equals the one-body
    density matrix of the Fock expansion of the same state (tests)."""
    n = cat.n_site
    n_g = cat.perm.shape[0]
    chi = np.asarray(chi, dtype=np.complex128)
    rho = np.zeros((n, n), dtype=np.complex128)
    inv_ng = 1.0 / n_g
    inv_sqrt_ng = 1.0 / np.sqrt(n_g)
    for col, s in enumerate(reps):
        c_col = c[col]
        if c_col == 0:
            continue
        diag_w = abs(c_col) ** 2 * inv_ng
        ket_norm = inv_sqrt_ng / np.sqrt(int(stab[col]))
        for g in range(n_g):
            km, kph = apply_op(int(s), cat.perm[g], cat.phases[g], cat.fermionic)
            ket_amp = c_col * (ket_norm * np.conj(chi[g]) * kph)
            tj = km
            while tj:
                lj = tj & -tj
                j = lj.bit_length()           # 1-based
                tj ^= lj
                rho[j - 1, j - 1] += diag_w
                for i in range(1, n + 1):
                    if i == j or (km >> (i - 1)) & 1:
                        continue
                    nm = (km & ~(1 << (j - 1))) | (1 << (i - 1))
                    idx = rank_colex(nm, cat.bt)
                    o = int(cat.rep_rank[idx])
                    row = int(o2b[o]) if o >= 0 else -1
                    if row < 0:
                        continue
                    coeff = complex(cat.rep_amp[idx]) * np.conj(chi[int(cat.rep_sym[idx])])
                    bra_amp = c[row] * (np.sqrt(int(stab[row])) * inv_sqrt_ng * np.conj(coeff))
                    sign = 1.0
                    if cat.fermionic:
                        lo, hi = min(i, j), max(i, j)
                        between = (((1 << (hi - lo - 1)) - 1) << lo) if hi - lo > 1 else 0
                        if bin(km & between).count("1") % 2:
                            sign = -1.0
                    rho[i - 1, j - 1] += np.conj(bra_amp) * ket_amp * sign
    return rho
