"""GPU: the path at the sizes of the larger workloads, checked through
properties that do not need an oracle run at that size.

* Haldane honeycomb 3x6 (36 sites), 9 hard-core bosons, translation group of
  order 18: C(36,9) = 94 143 280 masks.
* Checkerboard 4x6x2 (48 sites), 8 fermions, order 24: C(48,8) = 377 348 994
  masks (beyond the reference's N <= 42 rank tables).

Checks: orbit counts and every sector dimension against the closed-form
character count; table entries of sampled masks against the one-mask
restatement of the catalog rules; the flux refresh against a fresh catalog
at the same twist; one sector solved in mixed precision against FP64 with
true residuals (the 36-site one also at a twist); one full-size sector of
the Hubbard 3x4 workload against the oracle's ARPACK solve.
"""

from __future__ import annotations

import math
import time

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from oracle import ed_oracle as O
from realspace_exatdiagonalization_py_amd import models as M
from realspace_exatdiagonalization_py_amd import second_quantized_model as SQ
from realspace_exatdiagonalization_py_amd.groups import (build_identity_group, build_identity_irrep_list,
                                                         build_irrep_list, build_translation_group, group_arrays)

pytestmark = pytest.mark.gpu


def _haldane36():
    return M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[3, 6]), 9


def _checkerboard48():
    p = {"t": 1.0, "t'_1": 0.3, "t'_2": -0.3, "t''": -0.2, "phi_over_2pi": 1 / 8, "V1": 1.0, "V2": 0.0, "V3": 0.0}
    return M.build_zero_flux_fermionic_fci_second_quantized_model(sample_size=[4, 6], params=p), 8


def _check_catalog(model, ne, n_orbits_expected, n_sample, seed):
    t0 = time.perf_counter()
    G = build_translation_group(model.lattice)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    N = model.lattice.n_site
    print(f"[scale] N={N} Ne={ne}: catalog {gc.n_orbits} orbits in {time.perf_counter() - t0:.2f}s", flush=True)
    assert gc.n_total == math.comb(N, ne)
    assert gc.n_orbits == n_orbits_expected
    perm, ph = group_arrays(G)
    ferm = model.particle_statistics.name == "FERMIONIC"
    irreps = build_irrep_list(G, model.lattice)
    dims = O.sector_dims_closed_form(N, ne, perm, [ir.values for ir in irreps], ferm)
    got = [ed.build_symmetry_sector_basis(gc, ir).dim for ir in irreps]
    assert got == dims
    print(f"[scale] sector dims {min(dims)}..{max(dims)} match the character count", flush=True)
    rng = np.random.default_rng(seed)
    ranks = rng.choice(gc.n_total, size=n_sample, replace=False)
    bt = O.binom_table(N, ne)
    masks = O.unrank_chunk(np.sort(ranks).astype(np.int64), N, ne, bt)
    rep, g, a = ed.CanonicalMap(G, model.particle_statistics, gc).lookup(masks)
    tables = O.byte_tables(N, np.asarray(perm, dtype=np.int64), np.asarray(ph))
    for i, m in enumerate(masks):
        r_ref, g_ref, a_ref = O.table_entry(int(m), N, perm, ph, ferm, tables)
        assert (int(rep[i]), int(g[i])) == (r_ref, g_ref)
        assert np.complex64(a[i]) == a_ref
    return gc, G


def test_haldane36_catalog_flux_refresh_and_solve():
    model, ne = _haldane36()
    gc, G = _check_catalog(model, ne, 5_230_280, 400, seed=1)
    # flux refresh in place against a fresh catalog at the same twist
    theta = [0.25, 0.0]
    Gt = build_translation_group(model.lattice, theta)
    fresh = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=Gt,
                                            particle_statistics=model.particle_statistics)
    t0 = time.perf_counter()
    ed.update_orbit_stabilizer_phases(gc, Gt, model.particle_statistics)
    print(f"[scale] flux refresh {time.perf_counter() - t0:.2f}s", flush=True)
    np.testing.assert_array_equal(gc.rep_rank_table, fresh.rep_rank_table)
    np.testing.assert_array_equal(gc.rep_sym_table, fresh.rep_sym_table)
    a, b = gc.rep_amp_table, fresh.rep_amp_table
    ulp = np.spacing(np.abs(b.real)) + np.spacing(np.abs(b.imag))
    assert np.all(np.abs(a - b) <= 2 * ulp)
    del fresh
    # one sector at that twist (twisted terms, refreshed catalog): complex64
    # sweeps + FP64 refinement against FP64, on the block's Hermitian part
    tmodel, _ = _haldane36()
    SQ.update_second_quantized_model_with_twisted_phases(tmodel, twisted_phases_over_2π=theta)
    tbasis = ed.build_symmetry_sector_basis(gc, build_irrep_list(Gt, tmodel.lattice)[2])
    Ht = ed.MatrixFreeHamiltonian(tbasis, tmodel.bilinear_terms, tmodel.density_density_terms, None)
    t0 = time.perf_counter()
    w32, v32 = ed.diagonalize_block_matrixfree(Ht, Ht.n, nev=5, precision=ed.PREC_MIXED, eig_tol=1e-11)
    t1 = time.perf_counter()
    w64, _ = ed.diagonalize_block_matrixfree(Ht, Ht.n, nev=5, precision=ed.PREC_FP64, want_vectors=False)
    print(f"[scale] twisted sector D={tbasis.dim}: mixed {t1 - t0:.2f}s fp64 {time.perf_counter() - t1:.2f}s "
          f"|dE| {np.max(np.abs(w32 - w64)):.1e} E={w32.tolist()}", flush=True)
    assert np.max(np.abs(w32 - w64)) <= 1e-10
    R = Ht(v32) - v32 * w32
    assert np.max(np.linalg.norm(R, axis=0)) <= 1e-5
    del Ht, tbasis, v32, R
    # one sector at zero twist: complex64 sweeps + FP64 refinement against FP64
    ed.update_orbit_stabilizer_phases(gc, G, model.particle_statistics)
    irrep = build_irrep_list(G, model.lattice)[1]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    t0 = time.perf_counter()
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    print(f"[scale] sector D={basis.dim} table {H._dev.n_entries} entries in {time.perf_counter() - t0:.2f}s",
          flush=True)
    t0 = time.perf_counter()
    w32, v32 = ed.diagonalize_block_matrixfree(H, H.n, nev=5, precision=ed.PREC_MIXED, eig_tol=1e-12)
    t1 = time.perf_counter()
    w64, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=5, precision=ed.PREC_FP64, eig_tol=1e-12,
                                             want_vectors=False)
    t2 = time.perf_counter()
    print(f"[scale] mixed {t1 - t0:.2f}s fp64 {t2 - t1:.2f}s E={w32.tolist()}", flush=True)
    assert np.max(np.abs(w32 - w64)) <= 1e-10
    R = H(v32) - v32 * w32
    assert np.max(np.linalg.norm(R, axis=0)) <= 1e-5


def test_checkerboard48_catalog_and_sectors():
    model, ne = _checkerboard48()
    _check_catalog(model, ne, 15_724_214, 200, seed=2)


def test_checkerboard48_sector_solve():
    """Momentum sectors of the 48-site checkerboard model (D ~ 15.7 M, ~1.3 G
    stored entries): the 10 lowest levels in mixed precision against the FP64
    solve, with true residuals of the returned vectors; k = 0 at the library
    defaults, a k != 0 sector at the benchmark's settings."""
    model, ne = _checkerboard48()
    G = build_translation_group(model.lattice)
    t0 = time.perf_counter()
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    irrep = build_irrep_list(G, model.lattice)[0]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    print(f"[scale] checkerboard48 k=0: D={basis.dim} table {H._dev.n_entries} entries, catalog+basis+table "
          f"{time.perf_counter() - t0:.2f}s", flush=True)
    t0 = time.perf_counter()
    w32, v32 = ed.diagonalize_block_matrixfree(H, H.n, nev=10, precision=ed.PREC_MIXED, eig_tol=1e-12)
    t1 = time.perf_counter()
    w64, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=10, precision=ed.PREC_FP64, eig_tol=1e-12,
                                             want_vectors=False)
    t2 = time.perf_counter()
    print(f"[scale] checkerboard48 mixed {t1 - t0:.2f}s ({H.last_info[0]:.0f} matvecs) fp64 {t2 - t1:.2f}s "
          f"|dE| {np.max(np.abs(w32 - w64)):.1e} E={w32.tolist()}", flush=True)
    assert np.max(np.abs(w32 - w64)) <= 1e-10
    R = H(v32) - v32 * w32
    assert np.max(np.linalg.norm(R, axis=0)) <= 1e-5
    del H, basis, v32, R
    # a sector of nonzero momentum with the benchmark's solver settings (block
    # width 4, complex64-to-FP64 handover at 1.5e-8) against the library's
    # FP64 solve (block width 8, ARPACK's residual rule)
    irrep = build_irrep_list(G, model.lattice)[13]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    t0 = time.perf_counter()
    w4, v4 = ed.diagonalize_block_matrixfree(H, H.n, nev=10, precision=ed.PREC_MIXED, eig_tol=1e-11, block=4,
                                             mixed_tol=1.5e-8)
    t1 = time.perf_counter()
    mv4 = int(H.last_info[0])
    w64, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=10, precision=ed.PREC_FP64, want_vectors=False)
    print(f"[scale] checkerboard48 sector 13 D={basis.dim}: block 4 mixed {t1 - t0:.2f}s ({mv4} matvecs) fp64 "
          f"{time.perf_counter() - t1:.2f}s ({H.last_info[0]:.0f} matvecs) |dE| {np.max(np.abs(w4 - w64)):.1e}",
          flush=True)
    assert np.max(np.abs(w4 - w64)) <= 1e-10
    R = H(v4) - v4 * w4
    assert np.max(np.linalg.norm(R, axis=0) / np.maximum(1.0, np.abs(w4))) <= 1e-6


@pytest.mark.parametrize("nev", [6, 4])
def test_heisenberg28_blocks_above_a_million_rows(nev):
    """A sector of the 28-site Heisenberg chain (D ~ 1.43 M, above the 2^20
    rows from which the Gram uses its larger tiles at block widths 4 and 6):
    mixed precision against FP64 within 1e-10, and the true residuals of the
    returned FP64 vectors."""
    model = M.build_heisenberg_second_quantized_model(28)
    G = build_translation_group(model.lattice)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=14, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    irrep = build_irrep_list(G, model.lattice)[3]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    assert basis.dim >= 1 << 20
    print(f"[scale] heisenberg28 k=3 D={basis.dim} nev={nev}", flush=True)
    w32, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=nev, precision=ed.PREC_MIXED, eig_tol=1e-12,
                                             want_vectors=False)
    w64, v64 = ed.diagonalize_block_matrixfree(H, H.n, nev=nev)
    assert np.max(np.abs(w32 - w64)) <= 1e-10, (w32, w64)
    R = H._dev.matvec(np.ascontiguousarray(v64)) - v64 * w64
    assert np.max(np.linalg.norm(R, axis=0)) <= 1e-9 * max(1.0, np.abs(w64).max())


def _hubbard34_full_space(eps=None):
    model = M.build_spinful_hubbard_second_quantized_model([3, 4], 1.0, 4.0, onsite_disorder=eps)
    G = build_identity_group(model.lattice.n_site)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=12, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(gc, build_identity_irrep_list()[0])
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    return model, basis, H


def test_hubbard3x4_full_space_clean_and_disordered():
    """Hubbard 3x4, U = 4, 12 fermions on the identity group (D = C(24, 12) =
    2 704 156), clean and with seeded on-site disorder eps_s ~ U[-0.1, 0.1]
    on both spin orbitals of each site.

    * Clean: the 6 lowest levels equal the 6 lowest of the union of the 12
      momentum sectors, each solved for its own 6 lowest.
    * Disordered: mixed precision against FP64 within 1e-10; true residuals;
      every level within Weyl's bounds E_i + min V <= E_i(eps) <= E_i + max V
      (V = sum eps n is diagonal, its extremes are the sums of the 12 smallest
      and the 12 largest orbital energies); the clean ground state's Rayleigh
      quotient under the disordered table equals E_0 + <V>, with <V> summed
      on the host over the basis masks (the disorder sits on the right
      orbitals), and lies above the disordered ground level."""
    model, basis, Hc = _hubbard34_full_space()
    assert basis.dim == math.comb(24, 12)
    print(f"[scale] hubbard3x4 identity group D={basis.dim} table {Hc._dev.n_entries} entries", flush=True)
    t0 = time.perf_counter()
    wc, vc = ed.diagonalize_block_matrixfree(Hc, Hc.n, nev=6, precision=ed.PREC_MIXED, eig_tol=1e-12)
    print(f"[scale] clean full space {time.perf_counter() - t0:.2f}s E={wc.tolist()}", flush=True)
    masks = basis.representative_masks.copy()
    del Hc

    Gt = build_translation_group(model.lattice)
    gt = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=12, symmetry_group=Gt,
                                         particle_statistics=model.particle_statistics)
    parts = []
    for irrep in build_irrep_list(Gt, model.lattice):
        b = ed.build_symmetry_sector_basis(gt, irrep)
        H = ed.MatrixFreeHamiltonian(b, model.bilinear_terms, model.density_density_terms, None)
        w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=6, precision=ed.PREC_MIXED, eig_tol=1e-12,
                                               want_vectors=False)
        parts.append(w)
        del H
    union = np.sort(np.concatenate(parts))[:6]
    print(f"[scale] union of 12 sectors {union.tolist()}", flush=True)
    assert np.max(np.abs(union - wc)) <= 1e-9, (union, wc)
    del gt

    eps = np.random.default_rng(2024).uniform(-0.1, 0.1, 12)
    model_d, basis_d, Hd = _hubbard34_full_space(eps)
    np.testing.assert_array_equal(basis_d.representative_masks, masks)
    t0 = time.perf_counter()
    wd32, _ = ed.diagonalize_block_matrixfree(Hd, Hd.n, nev=6, precision=ed.PREC_MIXED, eig_tol=1e-12,
                                              want_vectors=False)
    t1 = time.perf_counter()
    wd, vd = ed.diagonalize_block_matrixfree(Hd, Hd.n, nev=6)
    print(f"[scale] disordered mixed {t1 - t0:.2f}s fp64 {time.perf_counter() - t1:.2f}s E={wd.tolist()}",
          flush=True)
    assert np.max(np.abs(wd32 - wd)) <= 1e-10, (wd32, wd)
    R = Hd._dev.matvec(np.ascontiguousarray(vd)) - vd * wd
    assert np.max(np.linalg.norm(R, axis=0)) <= 1e-9 * max(1.0, np.abs(wd).max())

    # orbital energies from the disorder terms (1-based orbital i is mask bit i - 1)
    e_orb = np.zeros(24)
    for i, j, v in model_d.density_density_terms:
        if i == j:
            e_orb[i - 1] += v.real
    assert np.count_nonzero(e_orb) == 24
    lo, hi = np.sort(e_orb)[:12].sum(), np.sort(e_orb)[12:].sum()
    assert np.all(wd >= wc + lo - 1e-9) and np.all(wd <= wc + hi + 1e-9), (wd - wc, lo, hi)

    V = np.zeros(masks.shape[0])
    for b in range(24):
        V += e_orb[b] * ((masks >> np.uint64(b)) & np.uint64(1)).astype(np.float64)
    psi = np.ascontiguousarray(vc[:, 0])
    rq = float(np.vdot(psi, Hd(psi)).real) / float(np.vdot(psi, psi).real)
    want = float(wc[0]) + float(np.sum(np.abs(psi) ** 2 * V)) / float(np.vdot(psi, psi).real)
    print(f"[scale] Rayleigh quotient of the clean ground state {rq:.12f}, E0 + <V> {want:.12f}", flush=True)
    assert abs(rq - want) <= 1e-9
    assert rq >= wd[0] - 1e-10


def _levels(w, tol=1e-8):
    """Distinct levels of an ascending spectrum (values within tol merged) and
    their multiplicities."""
    out, mult = [], []
    for x in w:
        if out and abs(x - out[-1]) <= tol:
            mult[-1] += 1
        else:
            out.append(float(x))
            mult.append(1)
    return np.array(out), mult


def test_hubbard3x4_full_sector_against_oracle():
    """One full-size momentum sector of the Hubbard 3x4 workload (U = 4, 12
    fermions, D ~ 2.25e5, 5.6 M table entries): the device's 6 lowest levels,
    FP64 and mixed precision, against the oracle's own catalog, table and
    ARPACK call (eigs, which='SR', tol=1e-10, the reference's seeded start).
    The sector holds spin triplets, of which ARPACK may return fewer copies
    than there are, so levels are compared as sets: every ARPACK level lies
    within 1e-10 of a device level, the device holds at least as many copies
    of each (below the highest level, which the cut at 6 may split), and
    every device level below ARPACK's highest is an ARPACK level."""
    model = M.build_spinful_hubbard_second_quantized_model([3, 4], 1.0, 4.0)
    G = build_translation_group(model.lattice)
    irrep = build_irrep_list(G, model.lattice)[5]
    perm, ph = group_arrays(G)
    t0 = time.perf_counter()
    oc = O.build_catalog(24, 12, perm, ph, True)
    reps, stab, o2b = O.sector_basis(oc, irrep.values)
    r, c, v, hd = O.projection_table(oc, reps, stab, o2b, model.bilinear_terms, model.density_density_terms,
                                     irrep.values, True)
    Href = O.csr_hamiltonian(r, c, v, hd, len(reps))
    w_ref, _ = O.diagonalize(Href, 6)
    print(f"[scale] hubbard3x4 sector 5 oracle D={len(reps)} {time.perf_counter() - t0:.1f}s E={w_ref.tolist()}",
          flush=True)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=12, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    np.testing.assert_array_equal(basis.representative_masks, reps)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    lo, mo = _levels(w_ref)
    for prec in (ed.PREC_FP64, ed.PREC_MIXED):
        mixed = prec == ed.PREC_MIXED
        w, vecs = ed.diagonalize_block_matrixfree(H, H.n, nev=6, precision=prec, eig_tol=1e-11 if mixed else 0.0)
        ld, md = _levels(w)
        print(f"[scale] device precision {prec}: E={w.tolist()}", flush=True)
        for x, m in zip(lo, mo):
            j = int(np.argmin(np.abs(ld - x)))
            assert abs(ld[j] - x) <= 1e-10, (x, w, w_ref)
            if x < lo[-1]:
                assert md[j] >= m, (x, w, w_ref)
        for x in ld[ld < lo[-1] - 1e-8]:
            assert np.min(np.abs(lo - x)) <= 1e-10, (x, w, w_ref)
        R = Href @ vecs - vecs * w
        assert np.max(np.linalg.norm(R, axis=0)) <= (1e-5 if mixed else 1e-6)
