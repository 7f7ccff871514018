"""GPU: the reference API entry points that the sector solve does not
exercise by itself, against the oracle on the same seeded inputs.

* ``project_to_sector`` / ``project_to_unnormalized_sector`` (SRE:1662-1716):
  every fixed-filling mask, against the oracle's table entry and scalar scan;
* ``apply_hamiltonian`` / ``apply_hamiltonian_with_threads`` (SRE:1979-2069):
  ``y`` overwritten in place with the oracle CSR product (1e-13 relative);
* ``edx_matvec_device`` through the C ABI (complex128 and complex64);
* ``full_ed`` (SRE:2966-2992) against the lowest levels of the plain Fock
  matrix (1e-10);
* ``ed_scan(mode="matrix")`` against the Julia goldens of the reference's
  ``test/test_energy_alignment.py:89, :119`` (ATOL 1e-9, all sectors);
* checkpoints (SRE:2530-2563): a saved scan loads with its catalog, and a
  sector basis built from the loaded catalog (the reference consumers'
  pattern, observables/charge_pump.py:431-462) equals the original;
* the flux scan (SRE:2608-2703): per-twist files named as the reference
  names them, each twist's catalog equal bit for bit to a fresh catalog at
  that twist, eigenvalues equal to per-twist fresh solves (1e-10), the
  caller's catalog unchanged, finished twists skipped on a rerun;
* a scan whose k-th sector fails, rerun: finished sectors are not solved
  again (SRE:2485-2501).
"""

from __future__ import annotations

import json
import os
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import is_fermionic, make_case, random_vector
from oracle import ed_oracle as O
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd import models as M
from realspace_exatdiagonalization_py_amd import symmetry_resolved_ed as sre
from realspace_exatdiagonalization_py_amd._edx import lib, ptr
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list, build_translation_group, group_arrays

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_values.json").read_text())


def _data(name):
    model, ne, G = make_case(name)
    return ed.build_ed_data(model, filling_fraction=Fraction(ne, model.lattice.n_site), symmetry_group=G)


@pytest.mark.parametrize("name", ["hub23", "hal23_twist"])
def test_project_to_sector_against_oracle(name):
    print(f"[project] {name}", flush=True)
    model, ne, G = make_case(name)
    perm, ph = group_arrays(G)
    ferm = is_fermionic(model)
    oc = O.build_catalog(model.lattice.n_site, ne, perm, ph, ferm)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    cmap = ed.CanonicalMap(G, model.particle_statistics, gc)
    masks = O.unrank_chunk(np.arange(len(oc.rep_rank)), model.lattice.n_site, ne, oc.bt)
    for irrep in build_irrep_list(G, model.lattice)[:4]:
        reps, _, o2b = O.sector_basis(oc, irrep.values)
        basis = ed.build_symmetry_sector_basis(gc, irrep)
        chi = np.asarray(irrep.values)
        for r, m in enumerate(masks):
            got = ed.project_to_sector(int(m), basis, cmap)
            row = o2b[oc.rep_rank[r]]
            if row < 0:
                assert got is None
                continue
            # table path: (row, alpha * conj(chi(g))) with the complex64 alpha widened
            want = complex(np.complex128(oc.rep_amp[r]) * np.conj(chi[oc.rep_sym[r]]))
            assert got[0] == row and abs(got[1] - want) <= 1e-15, (irrep.label, int(m), got, want)
            # scalar path (SRE:1693-1716): the O(|G|) scan, g maps m -> rep
            rep_s, g_s, a_s = O.canonical_rep_scalar(int(m), perm, ph, ferm)
            got_u = ed.project_to_unnormalized_sector(int(m), basis, model.particle_statistics)
            idx = int(np.searchsorted(reps, np.uint64(rep_s)))
            assert idx < len(reps) and int(reps[idx]) == rep_s
            assert got_u[0] == idx and abs(got_u[1] - a_s * np.conj(chi[g_s])) <= 1e-15
        # a mask of another particle number is outside the filling of the table
        assert ed.project_to_unnormalized_sector(0, basis, model.particle_statistics) is None or ne == 0


def _oracle_sector(model, ne, G, irrep_idx):
    perm, ph = group_arrays(G)
    ferm = is_fermionic(model)
    oc = O.build_catalog(model.lattice.n_site, ne, perm, ph, ferm)
    irrep = build_irrep_list(G, model.lattice)[irrep_idx]
    reps, stab, o2b = O.sector_basis(oc, irrep.values)
    r, c, v, hd = O.projection_table(oc, reps, stab, o2b, model.bilinear_terms, model.density_density_terms,
                                     irrep.values, ferm)
    return irrep, O.csr_hamiltonian(r, c, v, hd, len(reps))


@pytest.mark.parametrize("name", ["heis16", "hub24_twist"])
def test_apply_hamiltonian_in_place(name):
    model, ne, G = make_case(name)
    irrep, Href = _oracle_sector(model, ne, G, 1)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    cmap = ed.CanonicalMap(G, model.particle_statistics, gc)
    x = random_vector(basis.dim, 3)
    y = np.full(basis.dim, np.nan + 0j)
    out = ed.apply_hamiltonian(y, x, basis, model.bilinear_terms, model.density_density_terms, cmap)
    assert out is y
    y_ref = Href @ x
    assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)
    y2 = np.zeros(basis.dim, dtype=np.complex128)
    ed.apply_hamiltonian_with_threads(y2, x, basis, model.bilinear_terms, model.density_density_terms, cmap,
                                      np.zeros((4, basis.dim), dtype=np.complex128))
    np.testing.assert_array_equal(y2, y)


def test_matvec_device_c_abi():
    """edx_matvec_device: interleaved device input, column-major device output."""
    model, ne, G = make_case("hub24")
    irrep, Href = _oracle_sector(model, ne, G, 2)
    gc = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                         particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(gc, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    dh = H._dev
    D, nb = dh.dim, 3
    X = np.stack([random_vector(D, s) for s in range(nb)], axis=1)  # (D, nb)
    Y_ref = Href @ X
    for prec, dt, tol in ((_edx.PREC_FP64, np.complex128, 1e-13), (_edx.PREC_MIXED, np.complex64, 1e-6)):
        xs = np.ascontiguousarray(X.astype(dt))  # row-major [D][nb]: the interleaved layout
        dx = _edx.DeviceArray(dh.ctx, xs.nbytes)
        dy = _edx.DeviceArray(dh.ctx, xs.nbytes)
        dx.upload(xs)
        dh.ctx.check(lib.edx_matvec_device(dh.ctx.handle, dh.handle, prec, nb, dx.ptr, dy.ptr))
        ys = dy.download(np.empty((nb, D), dtype=dt))  # column-major [nb][D]
        err = np.linalg.norm(ys.T.astype(np.complex128) - Y_ref) / np.linalg.norm(Y_ref)
        assert err <= tol, (prec, err)


def test_full_ed_against_fock_spectrum():
    model = M.build_heisenberg_second_quantized_model(12)
    vals, vecs = ed.full_ed(model, 6, nev=6)
    F = O.fock_hamiltonian(12, 6, model.bilinear_terms, model.density_density_terms, False).toarray()
    w = np.linalg.eigvalsh((F + F.conj().T) / 2)[:6]
    np.testing.assert_allclose(vals, w, atol=1e-10, rtol=0)
    assert vecs.shape == (924, 6)
    R = F @ vecs - vecs * vals
    assert np.max(np.linalg.norm(R, axis=0)) <= 1e-8


def test_matrix_mode_julia_goldens():
    """ed_scan(mode="matrix") over every sector, as test_energy_alignment.py does."""
    gold = GOLD["energy_alignment"]
    for name, key in (("hal34", "JULIA_REF_BOSON_34"), ("cb34", "JULIA_REF_FERMION_34")):
        print(f"[matrix mode] {name}", flush=True)
        data = _data(name)
        ed.ed_scan(data, nev=5, mode="matrix")
        assert sorted(data.ed_scan_res) == list(range(len(data.irrep_list)))
        for lab_s, ref in gold[key].items():
            lab = tuple(int(x) for x in lab_s.split(","))
            idx = sre._find_irrep_idx(data, lab)
            vals = data.ed_scan_res[idx][0]
            np.testing.assert_allclose(vals[:len(ref)], ref, atol=gold["ATOL"], rtol=0)


def test_checkpoint_roundtrip_with_catalog(tmp_path):
    data = _data("hub24")
    path = str(tmp_path / "scan.pkl")
    labels = [ir.label for ir in data.irrep_list[:3]]
    ed.ed_scan(data, nev=3, mode="matrixfree", scanned_sectors=labels, checkpoint_path=path)
    loaded = ed.load_checkpoint(path)
    assert sorted(loaded.ed_scan_res) == [0, 1, 2]
    for i in range(3):
        np.testing.assert_array_equal(loaded.ed_scan_res[i][0], data.ed_scan_res[i][0])
    cat0, cat1 = data.orbit_catalog, loaded.orbit_catalog
    assert cat1._dev_obj is None  # restored from host tables
    for a in ("rep_rank_table", "rep_sym_table", "rep_amp_table", "representative_masks"):
        np.testing.assert_array_equal(getattr(cat1, a), getattr(cat0, a))
    assert cat1.stabilizer_g_indices_list == cat0.stabilizer_g_indices_list
    # the reference consumers' pattern: a sector basis from the loaded catalog
    irrep = loaded.irrep_list[1]
    b1 = ed.build_symmetry_sector_basis(loaded.orbit_catalog, irrep)
    b0 = ed.build_symmetry_sector_basis(cat0, data.irrep_list[1])
    np.testing.assert_array_equal(b1.representative_masks, b0.representative_masks)
    model = loaded.second_quantized_model
    H = ed.MatrixFreeHamiltonian(b1, model.bilinear_terms, model.density_density_terms, None)
    w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=3)
    np.testing.assert_allclose(w, data.ed_scan_res[1][0], atol=1e-10, rtol=0)
    # rerun on the same path resumes from the file (no sector solved again)
    fresh = _data("hub24")
    calls = []
    orig = sre.ed_scan_at_irrep_matrixfree
    try:
        sre.ed_scan_at_irrep_matrixfree = lambda *a, **k: calls.append(a) or orig(*a, **k)
        ed.ed_scan(fresh, nev=3, mode="matrixfree", scanned_sectors=labels, checkpoint_path=path)
    finally:
        sre.ed_scan_at_irrep_matrixfree = orig
    assert calls == [] and sorted(fresh.ed_scan_res) == [0, 1, 2]


def test_failed_sector_then_rerun_skips_finished(tmp_path, monkeypatch):
    data = _data("hub23")
    path = str(tmp_path / "partial.pkl")
    orig = sre.ed_scan_at_irrep_matrixfree
    solved = []

    def flaky(label, ed_data, **kw):
        if len(solved) == 2:
            raise RuntimeError("sector failed")
        solved.append(label)
        return orig(label, ed_data, **kw)

    monkeypatch.setattr(sre, "ed_scan_at_irrep_matrixfree", flaky)
    with pytest.raises(RuntimeError, match="sector failed"):
        ed.ed_scan(data, nev=2, mode="matrixfree", checkpoint_path=path)
    assert sorted(ed.load_checkpoint(path).ed_scan_res) == [0, 1]
    solved_before = list(solved)
    monkeypatch.setattr(sre, "ed_scan_at_irrep_matrixfree",
                        lambda label, ed_data, **kw: solved.append(label) or orig(label, ed_data, **kw))
    ed.ed_scan(data, nev=2, mode="matrixfree", checkpoint_path=path, overwrite=True)
    again = solved[len(solved_before):]
    assert set(again).isdisjoint(solved_before)
    assert sorted(data.ed_scan_res) == list(range(len(data.irrep_list)))
    assert sorted(ed.load_checkpoint(path).ed_scan_res) == list(range(len(data.irrep_list)))


def test_flux_scan_matches_fresh_per_twist(tmp_path):
    model = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[3, 4])
    ne = 6
    G0 = build_translation_group(model.lattice)
    data = ed.build_ed_data(model, filling_fraction=Fraction(ne, 24), symmetry_group=G0)
    before = {a: np.array(getattr(data.orbit_catalog, a), copy=True)
              for a in ("rep_rank_table", "rep_sym_table", "rep_amp_table")}
    thetas = [0.0, 0.25, 0.5]
    labels = [ir.label for ir in data.irrep_list[:3]]
    d = str(tmp_path / "flux")
    paths = ed.ed_scan(data, nev=3, mode="matrixfree", scanned_sectors=labels, checkpoint_dir=d,
                       twisted_phases_over_2π_list=thetas)
    # the caller's catalog keeps its phases
    for a, v in before.items():
        np.testing.assert_array_equal(getattr(data.orbit_catalog, a), v)
    for θ, p in zip(thetas, paths):
        flux = [θ, 0.0]
        assert os.path.basename(p) == ed.ed_scan_checkpoint_filename(model, flux, Fraction(ne, 24))
        loaded = ed.load_checkpoint(p)
        # a fresh model, group and catalog at this twist, as the reference builds them
        m2 = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[3, 4])
        ed.update_second_quantized_model_with_twisted_phases(m2, twisted_phases_over_2π=flux)
        fresh = ed.build_ed_data(m2, filling_fraction=Fraction(ne, 24),
                                 symmetry_group=build_translation_group(m2.lattice, flux))
        for a in ("rep_rank_table", "rep_sym_table", "rep_amp_table"):
            np.testing.assert_array_equal(getattr(loaded.orbit_catalog, a), getattr(fresh.orbit_catalog, a))
        assert loaded.orbit_catalog.stabilizer_phases_list == fresh.orbit_catalog.stabilizer_phases_list
        ed.ed_scan(fresh, nev=3, mode="matrixfree", scanned_sectors=labels)
        for idx in range(3):
            np.testing.assert_allclose(loaded.ed_scan_res[idx][0], fresh.ed_scan_res[idx][0], atol=1e-10, rtol=0)
        # the consumers' pattern on the loaded twist
        b = ed.build_symmetry_sector_basis(loaded.orbit_catalog, loaded.irrep_list[0])
        assert b.dim == loaded.sector_dims[0]
    # finished twists are skipped on a rerun
    calls = []
    orig = sre._ed_scan_sectors
    try:
        sre._ed_scan_sectors = lambda *a, **k: calls.append(a) or orig(*a, **k)
        again = ed.ed_scan(data, nev=3, mode="matrixfree", scanned_sectors=labels, checkpoint_dir=d,
                           twisted_phases_over_2π_list=thetas + [0.75])
    finally:
        sre._ed_scan_sectors = orig
    assert again[:3] == paths and len(calls) == 1
