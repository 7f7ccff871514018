"""GPU: the on-the-fly sector Hamiltonian (no stored table) against the
stored table and the oracle.

The reference's matrix-free product (SRE:826-923) recomputes every element
at each call; ``edx_hamiltonian_build_on_the_fly`` does the same on the
device.  Per case (fermions and bosons, spin chain, Hubbard, checkerboard,
Haldane-type bosons, one sector past the dense cutoff and one below it):

* the entry count equals the stored table's;
* ``H X`` for 1, 5 and 16 vectors through ``edx_matvec`` equals the stored
  table's product within 1e-14 relative (elements formed by the same
  formula; the on-the-fly product sums conj(H[x][r]) over the hops of row r,
  so the two differ by FP64 rounding only) and the oracle's CSR product
  within 1e-13;
* ``edx_matvec_device`` in complex64 within 1e-6 of the FP64 product;
* the lowest levels in FP64 and mixed precision equal the stored table's
  within 1e-10;
* ``edx_matvec_noalloc`` on the on-the-fly Hamiltonian equals ``edx_matvec``;
* the table accessors refuse it, a twisted catalog is refused, and
  ``table="auto"`` keeps the stored table when it fits;
* after its catalog's phases are refreshed to a twist the product is refused,
  and after a refresh back to zero twist it equals the first product bit for
  bit (the packed rank records are rebuilt with the phases).
"""

from __future__ import annotations

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import is_fermionic, make_case, random_vector
from oracle import ed_oracle as O
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd import symmetry_resolved_ed as sre
from realspace_exatdiagonalization_py_amd._edx import lib, ptr
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list, group_arrays

pytestmark = pytest.mark.gpu


def _pair(name, sector):
    model, ne, G = make_case(name)
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    irrep = build_irrep_list(G, model.lattice)[sector]
    basis = ed.build_symmetry_sector_basis(cat, irrep)
    Hs = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    Hf = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None,
                                  table="on_the_fly")
    return model, ne, G, irrep, Hs, Hf


def _oracle_csr(model, ne, G, irrep):
    perm, ph = group_arrays(G)
    ferm = is_fermionic(model)
    oc = O.build_catalog(model.lattice.n_site, ne, perm, ph, ferm)
    reps, stab, o2b = O.sector_basis(oc, irrep.values)
    r, c, v, hd = O.projection_table(oc, reps, stab, o2b, model.bilinear_terms, model.density_density_terms,
                                     irrep.values, ferm)
    return O.csr_hamiltonian(r, c, v, hd, len(reps))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name,sector,oracle", [
    ("heis16", 1, True),   # spin chain, D = 810: past the dense cutoff
    ("hub24", 2, True),    # spinful Hubbard fermions (Jordan-Wigner strings)
    ("cb34", 0, True),     # checkerboard fermions, longer-range hops
    ("hal34", 1, True),    # Haldane-type bosons
    ("hub34", 3, False),   # 3x4 Hubbard, D ~ 2.3e5
])
def test_on_the_fly_product_equals_stored_table(name, sector, oracle):
    model, ne, G, irrep, Hs, Hf = _pair(name, sector)
    print(f"[fly] {name} k={irrep.label} D={Hs.n} entries={Hs._dev.n_entries}", flush=True)
    assert Hf.on_the_fly and not Hs.on_the_fly
    assert Hf.n == Hs.n and Hf._dev.n_entries == Hs._dev.n_entries
    D = Hs.n
    Href = _oracle_csr(model, ne, G, irrep) if oracle else None
    for nb in (1, 5, 16):
        X = np.stack([random_vector(D, 100 + s) for s in range(nb)], axis=1)
        ys, yf = Hs._dev.matvec(X), Hf._dev.matvec(X)
        assert _rel(yf, ys) <= 1e-14, (nb, _rel(yf, ys))
        if Href is not None:
            assert _rel(yf, Href @ X) <= 1e-13
    # complex64 product on device-resident vectors
    nb = 4
    X = np.stack([random_vector(D, 200 + s) for s in range(nb)], axis=1)
    y64 = Hf._dev.matvec(X)
    xs = np.ascontiguousarray(X.astype(np.complex64))
    dh = Hf._dev
    dx, dy = _edx.DeviceArray(dh.ctx, xs.nbytes), _edx.DeviceArray(dh.ctx, xs.nbytes)
    dx.upload(xs)
    dh.ctx.check(lib.edx_matvec_device(dh.ctx.handle, dh.handle, _edx.PREC_MIXED, nb, dx.ptr, dy.ptr))
    ys32 = dy.download(np.empty((nb, D), dtype=np.complex64))
    assert _rel(ys32.T.astype(np.complex128), y64) <= 1e-6


@pytest.mark.parametrize("name,sector,nev", [("heis16", 1, 5), ("hub24", 2, 6), ("hal34", 1, 5)])
def test_on_the_fly_levels_equal_stored_table(name, sector, nev):
    _model, _ne, _G, irrep, Hs, Hf = _pair(name, sector)
    for prec in (ed.PREC_FP64, ed.PREC_MIXED):
        ws, _ = ed.diagonalize_block_matrixfree(Hs, Hs.n, nev=nev, precision=prec, want_vectors=False)
        wf, _ = ed.diagonalize_block_matrixfree(Hf, Hf.n, nev=nev, precision=prec, want_vectors=False)
        print(f"[fly] {name} k={irrep.label} prec={prec} max|dE| {np.max(np.abs(wf - ws)):.2e}", flush=True)
        assert np.max(np.abs(wf - ws)) <= 1e-10, (prec, wf, ws)


def test_on_the_fly_dense_path_and_vectors():
    """A sector below the dense cutoff: the dense block is formed from the
    on-the-fly product (16 unit vectors at a time) and solved as the stored
    table's is."""
    _model, _ne, _G, irrep, Hs, Hf = _pair("heis12", 1)
    assert Hs.n <= sre.DENSE_CUTOFF
    ws, vs = ed.diagonalize_block_matrixfree(Hs, Hs.n, nev=4)
    wf, vf = ed.diagonalize_block_matrixfree(Hf, Hf.n, nev=4)
    assert np.max(np.abs(wf - ws)) <= 1e-12
    for j in range(4):
        r = Hf(vf[:, j]) - wf[j] * vf[:, j]
        assert np.linalg.norm(r) <= 1e-10 * max(1.0, abs(wf[j]))


def test_on_the_fly_noalloc_equals_matvec():
    _model, _ne, _G, _irrep, _Hs, Hf = _pair("heis16", 1)
    dh = Hf._dev
    rng = np.random.default_rng(5)
    for n_vec in (3, 20):
        x = np.ascontiguousarray(rng.standard_normal((n_vec, dh.dim)) + 1j * rng.standard_normal((n_vec, dh.dim)))
        y_ref, y = np.empty_like(x), np.full_like(x, np.nan)
        dh.ctx.check(lib.edx_matvec(dh.ctx.handle, dh.handle, n_vec, ptr(x), ptr(y_ref)))
        dh.ctx.check(lib.edx_matvec_noalloc(dh.ctx.handle, dh.handle, n_vec, ptr(x), ptr(y)))
        np.testing.assert_array_equal(y, y_ref)


def test_on_the_fly_refusals_and_auto(monkeypatch):
    _model, _ne, _G, _irrep, Hs, Hf = _pair("hub24", 1)
    with pytest.raises(ValueError, match="stores no table"):
        Hf._dev.fetch_table()
    with pytest.raises(ValueError, match="stores no table"):
        Hf._dev.fetch_csr(0, 4)
    with pytest.raises(ValueError, match="no sliced table"):
        Hf._dev.autotune()
    with pytest.raises(ValueError, match="no sliced table"):
        Hf._dev.ctx.check(lib.edx_hamiltonian_set_spmv_group(Hf._dev.ctx.handle, Hf._dev.handle, 2))
    # plenty of free memory: "auto" stores the table
    Ha = ed.MatrixFreeHamiltonian(Hs.basis, Hs.bilinear_terms, Hs.density_terms, None, table="auto")
    assert not Ha.on_the_fly
    # no room for a table: "auto" goes on the fly, with the same product
    monkeypatch.setattr(sre, "_AUTO_TABLE_SHARE", 0.0)
    Hz = ed.MatrixFreeHamiltonian(Hs.basis, Hs.bilinear_terms, Hs.density_terms, None, table="auto")
    assert Hz.on_the_fly
    x = random_vector(Hs.n, 9)
    assert _rel(Hz(x), Hs(x)) <= 1e-14
    monkeypatch.undo()
    # twisted catalog: alpha rounded to complex64, the block is not exactly Hermitian
    model, ne, G = make_case("hub23_twist")
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[1])
    with pytest.raises(ValueError, match="site phases are all"):
        ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None, table="on_the_fly")
    Ht = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None, table="auto")
    assert not Ht.on_the_fly


def test_on_the_fly_follows_catalog_phase_refresh():
    from realspace_exatdiagonalization_py_amd.groups import build_translation_group
    from realspace_exatdiagonalization_py_amd import second_quantized_model as SQ

    model, ne, G = make_case("hal34")
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[1])
    Hf = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None,
                                  table="on_the_fly")
    x = random_vector(Hf.n, 21)
    y0 = Hf(x)
    m2, _ne, _G = make_case("hal34")
    SQ.update_second_quantized_model_with_twisted_phases(m2, twisted_phases_over_2π=[0.3, 0.0])
    ed.update_orbit_stabilizer_phases(cat, build_translation_group(m2.lattice, [0.3, 0.0]),
                                      model.particle_statistics)
    with pytest.raises(ValueError, match="refreshed to a twist"):
        Hf(x)
    ed.update_orbit_stabilizer_phases(cat, G, model.particle_statistics)
    np.testing.assert_array_equal(Hf(x), y0)


def test_on_the_fly_diagonal_only_and_one_state():
    """No bilinear terms (a diagonal block), and a one-state sector."""
    model, ne, G = make_case("hub24")
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[0])
    Hs = ed.MatrixFreeHamiltonian(basis, [], model.density_density_terms, None)
    Hf = ed.MatrixFreeHamiltonian(basis, [], model.density_density_terms, None, table="on_the_fly")
    assert Hf._dev.n_entries == Hs._dev.n_entries == 0
    x = random_vector(Hs.n, 5)
    np.testing.assert_array_equal(Hf(x), Hs(x))
    m0, _ne, G0 = make_case("hal23_ne0")  # no particles: one state, H = 0
    cat0 = ed.build_symmetry_orbit_catalog(second_quantized_model=m0, n_filled=0, symmetry_group=G0,
                                           particle_statistics=m0.particle_statistics)
    b0 = ed.build_symmetry_sector_basis(cat0, build_irrep_list(G0, m0.lattice)[0])
    assert b0.dim == 1
    H1 = ed.MatrixFreeHamiltonian(b0, m0.bilinear_terms, m0.density_density_terms, None, table="on_the_fly")
    w, _ = ed.diagonalize_block_matrixfree(H1, H1.n, nev=1)
    assert w.shape == (1,) and abs(w[0]) <= 1e-15
