"""Device path vs the CPU oracle on the same seeded inputs.

Integer outputs (ranks, orbit indices, g, representatives, stabilizers,
sector dimensions, projection-table structure) must be bit-exact.  Floating
point: complex64 amplitudes exact at zero flux and within one float32 ulp at
nonzero flux (device sincos vs libm); complex128 matrix elements within 4e-16
relative; H x within 1e-13 relative; eigenvalues within 1e-10 absolute of the
exact levels and of the oracle's ARPACK/LAPACK values (the accuracy the
mixed-precision path is held to against FP64).
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import is_fermionic, make_case, random_vector
from oracle import ed_oracle as O
import realspace_exatdiagonalization_py_amd as ed
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list, group_arrays
from realspace_exatdiagonalization_py_amd.symmetry_resolved_ed import DENSE_CUTOFF

pytestmark = pytest.mark.gpu

CATALOG_CASES = ["heis16", "hal23", "hub23", "hal23_twist", "cb23_twist", "heis10_identity", "hal23_ne0",
                 "cb23_full_twist", "hal33_twist", "hub34"]


def _both(name):
    model, ne, G = make_case(name)
    perm, ph = group_arrays(G)
    oc = O.build_catalog(model.lattice.n_site, ne, perm, ph, is_fermionic(model))
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    return model, ne, G, oc, gc


def _amp_close(a_dev, a_ref, exact):
    if exact:
        return np.array_equal(a_dev, a_ref)
    ulp = np.spacing(np.abs(a_ref.real).astype(np.float32)) + np.spacing(np.abs(a_ref.imag).astype(np.float32))
    return bool(np.all(np.abs(a_dev - a_ref) <= 2 * ulp + 1e-45))


@pytest.mark.parametrize("name", CATALOG_CASES)
def test_catalog_bit_exact(name):
    print(f"[catalog] {name}", flush=True)
    model, ne, G, oc, gc = _both(name)
    assert gc.n_total == len(oc.rep_rank)
    assert gc.n_orbits == len(oc.reps)
    np.testing.assert_array_equal(gc.representative_masks, oc.reps)
    np.testing.assert_array_equal(gc.rep_rank_table, oc.rep_rank)
    np.testing.assert_array_equal(gc.rep_sym_table, oc.rep_sym)
    assert gc.stabilizer_order_list == [int(x) for x in oc.stab_order]
    assert gc.stabilizer_g_indices_list == [[int(g) for g in s] for s in oc.stab_g]
    # alpha is a sign, exact in complex64, when every site phase of every
    # operation is +1 or -1; otherwise within one float32 ulp per component
    ph = np.asarray(group_arrays(G)[1])
    exact_alpha = bool(np.all(ph.imag == 0.0) and np.all(np.abs(ph.real) == 1.0))
    assert _amp_close(gc.rep_amp_table, oc.rep_amp, exact=exact_alpha)
    ph_dev = np.concatenate([np.asarray(p, complex) for p in gc.stabilizer_phases_list])
    ph_ref = np.concatenate([np.asarray(p, complex) for p in oc.stab_phase])
    assert np.abs(ph_dev - ph_ref).max(initial=0.0) <= 4e-16


@pytest.mark.parametrize("name", ["heis16", "hub23", "hal23_twist", "cb23_twist", "hub34"])
def test_sector_filter_bit_exact(name):
    print(f"[sector] {name}", flush=True)
    model, ne, G, oc, gc = _both(name)
    for irrep in build_irrep_list(G, model.lattice):
        reps, stab, o2b = O.sector_basis(oc, irrep.values)
        b = ed.build_symmetry_sector_basis(gc, irrep)
        assert b.dim == len(reps), irrep.label
        np.testing.assert_array_equal(b.representative_masks, reps)
        np.testing.assert_array_equal(np.asarray(b.stabilizer_order_list), stab)
        np.testing.assert_array_equal(b.orbit_to_basis, o2b)


def _oracle_block(model, oc, irrep, fermionic):
    reps, stab, o2b = O.sector_basis(oc, irrep.values)
    r, c, v, hd, k = O.projection_table(oc, reps, stab, o2b, model.bilinear_terms, model.density_density_terms,
                                        irrep.values, fermionic, with_terms=True)
    return reps, r, c, v, hd, k


@pytest.mark.parametrize("name", ["heis16", "hub23", "hal23", "hal23_twist", "cb23_twist", "hub34",
                                  "hub23_disorder"])
def test_projection_table(name):
    print(f"[table] {name}", flush=True)
    model, ne, G, oc, gc = _both(name)
    ferm = is_fermionic(model)
    irreps = build_irrep_list(G, model.lattice)
    for irrep in irreps[:3]:
        reps, r, c, v, hd, k = _oracle_block(model, oc, irrep, ferm)
        basis = ed.build_symmetry_sector_basis(gc, irrep)
        cmap = ed.CanonicalMap(G, model.particle_statistics, gc)
        H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, cmap)
        row, col, term, vals, hdg = H._dev.fetch_table()
        assert len(row) == len(r)
        o_ref = np.lexsort((k, c, r))
        o_dev = np.lexsort((term, col, row))
        np.testing.assert_array_equal(row[o_dev], r[o_ref])
        np.testing.assert_array_equal(col[o_dev], c[o_ref])
        np.testing.assert_array_equal(term[o_dev], k[o_ref])
        scale = np.maximum(np.abs(v[o_ref]), 1e-300)
        assert np.max(np.abs(vals[o_dev] - v[o_ref]) / scale, initial=0.0) <= 4e-16
        np.testing.assert_array_equal(hdg, hd)
        # the reference-order triplets of build_matrixfree_projection_table
        rr, cc, vv, hh = ed.build_matrixfree_projection_table(basis, model.bilinear_terms,
                                                              model.density_density_terms, cmap)
        np.testing.assert_array_equal(rr, r)
        np.testing.assert_array_equal(cc, c)
        assert np.max(np.abs(vv - v), initial=0.0) <= 4e-16 * max(1.0, np.abs(v).max(initial=0.0))
        # H x against the oracle CSR
        Href = O.csr_hamiltonian(r, c, v, hd, len(reps))
        for seed in (1, 2):
            x = random_vector(len(reps), seed)
            y_ref = Href @ x
            y = H(x)
            assert np.linalg.norm(y - y_ref) <= 1e-13 * max(1.0, np.linalg.norm(y_ref))
        X = np.stack([random_vector(len(reps), s) for s in range(5)], axis=1)
        Y = H._dev.matvec(X)
        assert np.linalg.norm(Y - Href @ X) <= 1e-13 * max(1.0, np.linalg.norm(Href @ X))


def test_canonical_lookup():
    model, ne, G, oc, gc = _both("hub23")
    cmap = ed.CanonicalMap(G, model.particle_statistics, gc)
    masks = O.unrank_chunk(np.arange(len(oc.rep_rank)), model.lattice.n_site, ne, oc.bt)
    rep, g, a = cmap.lookup(masks)
    np.testing.assert_array_equal(rep, oc.reps[oc.rep_rank])
    np.testing.assert_array_equal(g, oc.rep_sym)
    np.testing.assert_array_equal(a, oc.rep_amp.astype(np.complex128))
    m = int(masks[17])
    assert ed.get_canonical(cmap, m) == (int(oc.reps[oc.rep_rank[17]]), int(oc.rep_sym[17]),
                                         complex(oc.rep_amp[17]))
    # outside the filling -> the O(|G|) scan, as in the reference
    assert ed.get_canonical(cmap, 0b1) == ed.get_canonical_representative(0b1, G, model.particle_statistics)


@pytest.mark.parametrize("name", ["hal23", "cb23_twist", "hub23"])
def test_update_orbit_stabilizer_phases(name):
    """Catalog built at zero flux, refreshed to the twisted group, vs the oracle's refresh."""
    print(f"[update] {name}", flush=True)
    model, ne, G = make_case(name)
    from realspace_exatdiagonalization_py_amd.groups import build_translation_group

    G0 = build_translation_group(model.lattice, [0.0] * model.lattice.dim)
    Gt = build_translation_group(model.lattice, [0.23, 0.61])
    ferm = is_fermionic(model)
    perm, ph0 = group_arrays(G0)
    _, pht = group_arrays(Gt)
    oc = O.build_catalog(model.lattice.n_site, ne, perm, ph0, ferm)
    O.update_orbit_stabilizer_phases(oc, pht, ferm)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G0,
                                         particle_statistics=model.particle_statistics)
    assert ed.update_orbit_stabilizer_phases(gc, G0, model.particle_statistics) is gc  # same object: no-op
    ed.update_orbit_stabilizer_phases(gc, Gt, model.particle_statistics)
    np.testing.assert_array_equal(gc.rep_sym_table, oc.rep_sym)
    np.testing.assert_array_equal(gc.rep_amp_table, oc.rep_amp)  # product form: same roundings
    ph_dev = np.concatenate([np.asarray(p, complex) for p in gc.stabilizer_phases_list])
    ph_ref = np.concatenate([np.asarray(p, complex) for p in oc.stab_phase])
    np.testing.assert_array_equal(ph_dev, ph_ref)
    # a different permutation part is rejected
    from realspace_exatdiagonalization_py_amd.groups import Finite_Symmetry_Group
    bad = Finite_Symmetry_Group("translations", [Gt.operations[0]] + list(reversed(Gt.operations[1:])))
    with pytest.raises(ValueError):
        ed.update_orbit_stabilizer_phases(gc, bad, model.particle_statistics)


def _solve_both(name, irrep_index, nev, precision=ed.PREC_FP64):
    model, ne, G, oc, gc = _both(name)
    ferm = is_fermionic(model)
    irrep = build_irrep_list(G, model.lattice)[irrep_index]
    reps, r, c, v, hd, k = _oracle_block(model, oc, irrep, ferm)
    Href = O.csr_hamiltonian(r, c, v, hd, len(reps))
    w_ref, _ = O.diagonalize(Href, nev)
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    cmap = ed.CanonicalMap(G, model.particle_statistics, gc)
    H_op, n = ed.hamiltonian_linear_operator(basis, model.bilinear_terms, model.density_density_terms, cmap)
    w, vecs = ed.diagonalize_block_matrixfree(H_op, n, nev=nev, precision=precision)
    return w, vecs, w_ref, Href


@pytest.mark.parametrize("name,irrep,nev", [("heis16", 0, 4), ("hal23", 0, 3), ("hal23_twist", 1, 3),
                                            ("hub24", 1, 6), ("cb23_twist", 0, 4), ("hub23_disorder", 0, 6)])
def test_eigs_vs_oracle(name, irrep, nev):
    print(f"[eigs] {name} irrep={irrep} nev={nev}", flush=True)
    w, vecs, w_ref, Href = _solve_both(name, irrep, nev)
    # exact levels.  Up to the dense cutoff the reference's own call is dense
    # (eigh of the upper triangle) and exact.  Above it, the dense spectrum of
    # the block's Hermitian part (the block itself unless twisted; a twisted
    # block's eigenvalues have the same real parts to second order in its
    # complex64 rounding): the reference's ARPACK call can return too few
    # copies of a degenerate level (a spin triplet of the disordered case
    # comes back twice).  Where it returns all of them the device values
    # match it too.
    if Href.shape[0] <= DENSE_CUTOFF:
        w_exact = w_ref
    else:
        w_exact = np.linalg.eigvalsh(((Href + Href.getH()) * 0.5).toarray())[:nev]
    assert np.max(np.abs(w - w_exact)) <= 1e-10, (w, w_exact)
    if np.max(np.abs(w_ref - w_exact)) <= 1e-10:
        assert np.max(np.abs(w - w_ref)) <= 1e-10, (w, w_ref)
    else:
        assert name == "hub23_disorder", (w_ref, w_exact)
    # eigenvector residuals
    R = Href @ vecs - vecs * w
    assert np.max(np.linalg.norm(R, axis=0)) <= 1e-6


def _sliced_spmm(H, X):
    """The eigensolver's SpMM (one timed launch) on the columns of X."""
    from realspace_exatdiagonalization_py_amd import _edx
    import ctypes as C

    dh = H._dev
    n, nb = X.shape
    Xd = _edx.DeviceArray(dh.ctx, n * nb * 16)
    Yd = _edx.DeviceArray(dh.ctx, n * nb * 16)
    Xd.upload(np.ascontiguousarray(X, dtype=np.complex128))  # row-major [dim][nb]: interleaved
    ms = C.c_double()
    dh.ctx.check(_edx.lib.edx_time_matvec(dh.ctx.handle, dh.handle, ed.PREC_FP64, nb, 1, Xd.ptr, Yd.ptr,
                                          C.byref(ms)))
    return Yd.download(np.empty((nb, n), np.complex128)).T


@pytest.mark.parametrize("name,irrep", [("heis16", 0), ("hub24", 1), ("hal23", 0), ("hub34", 2)])
def test_sliced_copy_written_with_the_table(name, irrep):
    """The sliced copy that ptab_values writes while building the table (with
    its padding) is the one rebuilt from the CSR arrays: the same layout gives
    bitwise the same SpMM."""
    from realspace_exatdiagonalization_py_amd import _edx

    model, ne, G, oc, gc = _both(name)
    ir = build_irrep_list(G, model.lattice)[irrep]
    basis = ed.build_symmetry_sector_basis(gc, ir)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms,
                                 ed.CanonicalMap(G, model.particle_statistics, gc))
    dh = H._dev
    n = basis.dim
    print(f"[sell] {name} irrep={irrep} dim={n}", flush=True)
    X = np.stack([random_vector(n, s) for s in range(6)], axis=1)
    g0 = dh.spmv_group
    y_built = _sliced_spmm(H, X)
    for g in [x for x in (1, 2, 4) if x != g0] + [g0]:  # rebuilt from the CSR arrays, ending on g0
        dh.ctx.check(_edx.lib.edx_hamiltonian_set_spmv_group(dh.ctx.handle, dh.handle, g))
        y_g = _sliced_spmm(H, X)
        assert np.linalg.norm(y_g - y_built) <= 1e-13 * np.linalg.norm(y_built), g
    np.testing.assert_array_equal(y_g, y_built)


def test_twisted_block_solves_its_hermitian_part():
    """A twisted catalog rounds alpha to complex64 (as the reference stores
    it), so the block is Hermitian only to that rounding.  H x applies the
    block itself; the eigensolver's SpMM applies (H + H^H) / 2, whose lowest
    eigenvalues both precisions return within 1e-10.  The reference's eigs
    returns the real parts of the block's eigenvalues: the same numbers to
    second order in the rounding (the multiplets the rounding splits here are
    split by ~1e-10; tests/test_gpu_twisted.py lists them), so within 1e-10
    too."""
    import scipy.sparse as sp

    model, ne, G, oc, gc = _both("hub24_twist")
    irrep = build_irrep_list(G, model.lattice)[1]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    cmap = ed.CanonicalMap(G, model.particle_statistics, gc)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, cmap)
    n = basis.dim
    print(f"[herm] hub24_twist irrep=1 dim={n}", flush=True)
    assert n > 500  # past the dense cutoff: the block Lanczos path
    row, col, term, vals, hdg = H._dev.fetch_table()
    Hraw = sp.coo_matrix((vals, (row, col)), shape=(n, n)).toarray() + np.diag(hdg)
    Hh = (Hraw + Hraw.conj().T) / 2
    assert np.abs(Hraw - Hraw.conj().T).max() > 1e-12  # the case does carry the rounding
    X = np.stack([random_vector(n, s) for s in range(3)], axis=1)
    assert np.linalg.norm(H._dev.matvec(X) - Hraw @ X) <= 1e-13 * np.linalg.norm(Hraw @ X)
    assert np.linalg.norm(_sliced_spmm(H, X) - Hh @ X) <= 1e-13 * np.linalg.norm(Hh @ X)
    exact = np.linalg.eigvalsh(Hh)[:6]
    w64, vecs = ed.diagonalize_block_matrixfree(H, n, nev=6)
    w32, _ = ed.diagonalize_block_matrixfree(H, n, nev=6, precision=ed.PREC_MIXED)
    assert np.max(np.abs(w64 - exact)) <= 1e-10, (w64, exact)
    assert np.max(np.abs(w32 - exact)) <= 1e-10, (w32, exact)
    assert np.max(np.linalg.norm(Hh @ vecs - vecs * w64, axis=0)) <= 1e-6
    # the reference's convention on the oracle's own block
    reps, r, c, v, hd, k = _oracle_block(model, oc, irrep, is_fermionic(model))
    w_ref, _ = O.diagonalize(O.csr_hamiltonian(r, c, v, hd, len(reps)), 6)
    assert np.max(np.abs(w64 - w_ref)) <= 1e-10, (w64, w_ref)


def test_twisted_matrix_mode_matches_matrixfree():
    """Matrix mode (the block's triplets as a device CSR) solves the Hermitian
    part of a twisted block too, so both modes agree within 1e-10; the scipy
    block of build_ed_Hamiltonian_symmetry_block solved by
    diagonalize_block_arpack gives the same levels."""
    from fractions import Fraction

    model, ne, G = make_case("hub24_twist")
    data = ed.build_ed_data(model, filling_fraction=Fraction(ne, model.lattice.n_site), symmetry_group=G)
    label = data.irrep_list[1].label
    w_mat, v_mat = ed.ed_scan_at_irrep_matrix(label, data, nev=5)
    w_mf, _ = ed.ed_scan_at_irrep_matrixfree(label, data, nev=5)
    assert data.sector_dims[1] > 500
    assert np.max(np.abs(w_mat - w_mf)) <= 1e-10, (w_mat, w_mf)
    basis = ed.build_symmetry_sector_basis(data.orbit_catalog, data.irrep_list[1])
    Hs = ed.build_ed_Hamiltonian_symmetry_block(basis, model.bilinear_terms, model.density_density_terms,
                                                ed.CanonicalMap(G, model.particle_statistics, data.orbit_catalog))
    w_sp, _ = ed.diagonalize_block_arpack(Hs, nev=5)
    assert np.max(np.abs(w_sp - w_mat)) <= 1e-10, (w_sp, w_mat)


def test_mixed_handover_below_the_complex64_floor():
    """mixed_tol far below what complex64 sweeps can reach: they stall, hand
    over to FP64 anyway, and the result still matches FP64 within 1e-10."""
    import ctypes as C
    from realspace_exatdiagonalization_py_amd import _edx

    model, ne, G, oc, gc = _both("hub24")
    irrep = build_irrep_list(G, model.lattice)[1]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms,
                                 ed.CanonicalMap(G, model.particle_statistics, gc))
    dh = H._dev
    assert dh.dim > 500
    out = {}
    for prec, mixed_tol in ((ed.PREC_FP64, 1e-6), (ed.PREC_MIXED, 1e-9)):
        vals, info = np.zeros(6), np.zeros(4)
        o = _edx.eigs_options(nev=6, precision=prec, tol=1e-10, mixed_tol=mixed_tol, seed=12345, max_restarts=2000)
        dh.ctx.check(_edx.lib.edx_eigs_ex(dh.ctx.handle, dh.handle, C.byref(o), _edx.ptr(vals), None,
                                          _edx.ptr(info)))
        out[prec] = (vals, info)
    print(f"[handover] matvecs fp64 {out[ed.PREC_FP64][1][0]:.0f} mixed {out[ed.PREC_MIXED][1][0]:.0f}", flush=True)
    assert np.max(np.abs(out[ed.PREC_MIXED][0] - out[ed.PREC_FP64][0])) <= 1e-10
    assert out[ed.PREC_MIXED][1][1] < 200  # restarts: the stall guard ended the sweeps


@pytest.mark.parametrize("nev", [12, 20])
def test_many_levels_wide_blocks(nev):
    """nev above the block width and above 16 (the complex64 sweeps resolve 16
    pairs, the FP64 phase starts from them in a block of up to 16): both
    precisions against the dense spectrum of the oracle's block."""
    model, ne, G, oc, gc = _both("hub24")
    irrep = build_irrep_list(G, model.lattice)[1]
    reps, r, c, v, hd, k = _oracle_block(model, oc, irrep, True)
    dense = O.csr_hamiltonian(r, c, v, hd, len(reps)).toarray()
    exact = np.linalg.eigvalsh((dense + dense.conj().T) / 2)[:nev]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms,
                                 ed.CanonicalMap(G, model.particle_statistics, gc))
    print(f"[wide] hub24 irrep=1 dim={basis.dim} nev={nev}", flush=True)
    for prec in (ed.PREC_FP64, ed.PREC_MIXED):
        w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=nev, precision=prec)
        assert np.max(np.abs(w - exact)) <= 1e-10, (prec, w, exact)


def test_eigs_mixed_precision_cfg1():
    w64, _, w_ref, _ = _solve_both("heis16", 0, 4)
    w32, _, _, _ = _solve_both("heis16", 0, 4, precision=ed.PREC_MIXED)
    assert np.max(np.abs(w32 - w_ref)) <= 1e-10
    assert np.max(np.abs(w64 - w_ref)) <= 1e-10


def test_degenerate_multiplets_hubbard():
    """Spin multiplets appear with full multiplicity (S_z is not fixed)."""
    model, ne, G, oc, gc = _both("hub24")
    irrep = build_irrep_list(G, model.lattice)[0]
    reps, r, c, v, hd, k = _oracle_block(model, oc, irrep, True)
    dense = O.csr_hamiltonian(r, c, v, hd, len(reps)).toarray()
    exact = np.linalg.eigvalsh((dense + dense.conj().T) / 2)[:8]
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    cmap = ed.CanonicalMap(G, model.particle_statistics, gc)
    H_op, n = ed.hamiltonian_linear_operator(basis, model.bilinear_terms, model.density_density_terms, cmap)
    w, _ = ed.diagonalize_block_matrixfree(H_op, n, nev=8)
    assert np.max(np.abs(w - exact)) <= 1e-10, (w, exact)


def test_block_width_and_handover_options():
    """``block`` and ``mixed_tol`` of diagonalize_block_matrixfree on a 2x4
    Hubbard sector whose 8 lowest levels hold spin multiplets: every block at
    least as wide as the largest multiplicity gives the exact levels (1e-10),
    in FP64 and in mixed precision with handovers 1e-8 and 1e-6; a narrower
    block may return fewer copies of a multiplet, but every value it returns
    is an exact level (the block's limit, documented, never a wrong value)."""
    model, ne, G, oc, gc = _both("hub24")
    irrep = build_irrep_list(G, model.lattice)[0]
    reps, r, c, v, hd, k = _oracle_block(model, oc, irrep, True)
    dense = O.csr_hamiltonian(r, c, v, hd, len(reps)).toarray()
    spectrum = np.linalg.eigvalsh((dense + dense.conj().T) / 2)
    nev = 8
    exact = spectrum[:nev]
    runs = np.split(exact, np.nonzero(np.diff(exact) > 1e-8)[0] + 1)
    mult = max(len(x) for x in runs)  # copies wanted of the most degenerate level
    print(f"[block] hub24 irrep=0 dim={len(reps)} multiplicities {[len(x) for x in runs]}", flush=True)
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    for b in sorted({mult, mult + 1, 8}):
        w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=nev, block=b)
        assert np.max(np.abs(w - exact)) <= 1e-10, (b, w, exact)
        assert int(H.last_info[3]) == b
        for mtol in (1e-8, 1e-6):
            w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=nev, block=b, precision=ed.PREC_MIXED,
                                                   mixed_tol=mtol, eig_tol=1e-11)
            assert np.max(np.abs(w - exact)) <= 1e-10, (b, mtol, w, exact)
    if mult > 2:
        w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=nev, block=mult - 1)
        dist = np.min(np.abs(w[:, None] - spectrum[None, :]), axis=1)
        print(f"[block] block {mult - 1}: {w.tolist()}", flush=True)
        assert np.max(dist) <= 1e-9, (w, exact)


def test_golden_energy_alignment():
    """Julia golden energies of the reference's test_energy_alignment.py (ATOL 1e-9), on the GPU."""
    import json
    from pathlib import Path

    gold = json.loads((Path(__file__).parent / "golden" / "reference_values.json").read_text())["energy_alignment"]
    for name, key in (("hal34", "JULIA_REF_BOSON_34"), ("cb34", "JULIA_REF_FERMION_34")):
        model, ne, G = make_case(name)
        data = ed.build_ed_data(model, filling_fraction=__import__("fractions").Fraction(ne, model.lattice.n_site),
                                symmetry_group=G)
        labels = [tuple(int(x) for x in k.split(",")) for k in gold[key]]
        ed.ed_scan(data, nev=5, mode="matrixfree", scanned_sectors=labels)
        for lab in labels:
            idx = [i for i, ir in enumerate(data.irrep_list) if ir.label == lab][0]
            vals = data.ed_scan_res[idx][0]
            np.testing.assert_allclose(vals, gold[key][f"{lab[0]},{lab[1]}"], atol=gold["ATOL"], rtol=0)


def test_dense_path_reference_values():
    import json
    from pathlib import Path

    gold = json.loads((Path(__file__).parent / "golden" / "reference_values.json").read_text())
    model, ne, G = make_case("hal23")
    data = ed.build_ed_data(model, filling_fraction=__import__("fractions").Fraction(ne, 12), symmetry_group=G)
    vals, _ = ed.ed_scan_at_irrep_matrixfree((0, 0), data, nev=3)
    assert abs(vals[0] - gold["test_core_haldane_2x3_E0"]) <= 1e-9
    vals_m, _ = ed.ed_scan_at_irrep_matrix((0, 0), data, nev=3)
    np.testing.assert_allclose(vals_m, vals, atol=1e-8)
    assert data.sector_dims[0] == gold["design_notebook"]["haldane_2x3_k0_block"]["dim"]


def test_empty_sector_and_errors():
    model, ne, G = make_case("hub23")
    data = ed.build_ed_data(model, filling_fraction=__import__("fractions").Fraction(ne, model.lattice.n_site),
                            symmetry_group=G)
    # a character table that no orbit matches gives an empty sector
    irrep = ed.OneDim_Irrep("none", np.full(len(G.operations), 0.3 + 0.1j))
    basis = ed.build_symmetry_sector_basis(data.orbit_catalog, irrep)
    assert basis.dim == 0
    with pytest.raises(ValueError):
        ed.ed_scan(data, mode="bogus")
    with pytest.raises(ValueError):
        ed.ed_scan_at_irrep_matrixfree((9, 9), data)
    cmap = ed.CanonicalMap(G, model.particle_statistics, data.orbit_catalog)
    b0 = ed.build_symmetry_sector_basis(data.orbit_catalog, data.irrep_list[0])
    with pytest.raises(ValueError):  # i == j bilinear term
        ed.MatrixFreeHamiltonian(b0, [(1, 1, 1.0)], [], cmap)
    with pytest.raises(ValueError):  # site out of range
        ed.MatrixFreeHamiltonian(b0, [(1, 99, 1.0)], [], cmap)


def test_concurrent_sectors_match_sequential():
    """Sectors solved four at a time on separate streams give the results of
    the one-at-a-time scan (same start vectors, same arithmetic per stream)."""
    from fractions import Fraction

    model, ne, G = make_case("hub24")
    seq = ed.build_ed_data(model, filling_fraction=Fraction(ne, model.lattice.n_site), symmetry_group=G)
    ed.ed_scan(seq, nev=4, mode="matrixfree")
    con = ed.build_ed_data(model, filling_fraction=Fraction(ne, model.lattice.n_site), symmetry_group=G)
    ed.ed_scan(con, nev=4, mode="matrixfree", concurrent_sectors=4)
    assert sorted(con.ed_scan_res) == sorted(seq.ed_scan_res) == list(range(len(seq.irrep_list)))
    assert con.sector_dims == seq.sector_dims
    for idx in seq.ed_scan_res:
        np.testing.assert_allclose(con.ed_scan_res[idx][0], seq.ed_scan_res[idx][0], atol=1e-10, rtol=0)
    mix = ed.build_ed_data(model, filling_fraction=Fraction(ne, model.lattice.n_site), symmetry_group=G)
    ed.ed_scan(mix, nev=4, mode="matrixfree_mixed", concurrent_sectors=3)
    for idx in seq.ed_scan_res:
        np.testing.assert_allclose(mix.ed_scan_res[idx][0], seq.ed_scan_res[idx][0], atol=1e-10, rtol=0)
