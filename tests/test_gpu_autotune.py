"""GPU: the SpMM autotuner (lanes per row of the sliced table picked by
timing).  The tuned layout must be the fastest of the three it timed, be
reused for builds of the same shape, and leave the eigenvalues unchanged:
every layout stores the same entries, so eigenvalues agree to 1e-10."""

from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from realspace_exatdiagonalization_py_amd import _edx

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuning():
    _edx.autotune_clear()
    _edx.set_autotune(True, _edx.PREC_MIXED, 4)
    yield
    _edx.set_autotune(False)
    _edx.autotune_clear()


def _data():
    model = ed.build_spinful_hubbard_second_quantized_model([3, 3], 1.0, 4.0)
    G = ed.build_translation_group(model.lattice)
    return model, ed.build_ed_data(model, filling_fraction=Fraction(9, 18), symmetry_group=G)


def _hamiltonian(model, data, idx):
    basis = ed.build_symmetry_sector_basis(data.orbit_catalog, data.irrep_list[idx])
    return ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)


def test_tuned_layout_is_fastest_reused_and_exact(tuning):
    model, data = _data()
    H0 = _hamiltonian(model, data, 0)
    table = _edx.autotune_table()
    assert len(table) == 1
    e = table[0]
    print("[autotune]", e, flush=True)
    assert e["precision"] == _edx.PREC_MIXED and e["block"] == 4
    assert all(t > 0 for t in e["ms"].values())
    assert e["lanes"] == min(e["ms"], key=e["ms"].get)
    assert H0._dev.spmv_group == e["lanes"]
    # a second build of the same sector hits the table (no new entry)
    H0b = _hamiltonian(model, data, 0)
    assert len(_edx.autotune_table()) == 1
    assert H0b._dev.spmv_group == e["lanes"]
    # eigenvalues with every fixed layout agree with the tuned one
    w_t, _ = ed.diagonalize_block_matrixfree(H0, H0.n, nev=4, precision=ed.PREC_FP64, want_vectors=False)
    _edx.set_autotune(False)
    for lanes in (1, 2, 4):
        H0._dev.ctx.check(_edx.lib.edx_hamiltonian_set_spmv_group(H0._dev.ctx.handle, H0._dev.handle, lanes))
        w, _ = ed.diagonalize_block_matrixfree(H0, H0.n, nev=4, precision=ed.PREC_FP64, want_vectors=False)
        assert np.max(np.abs(w - w_t)) <= 1e-10, lanes


def test_direct_tuning_of_one_hamiltonian(tuning):
    model, data = _data()
    _edx.set_autotune(False)
    H = _hamiltonian(model, data, 1)
    x = np.random.default_rng(3).normal(size=(H.n, 3)) + 0j
    y0 = H._dev.matvec(x)
    r = H._dev.autotune(_edx.PREC_FP64, 3)
    assert r["lanes"] in (1, 2, 4) and r["lanes"] == min(r["ms"], key=r["ms"].get)
    d, n_e, g = _edx.C.c_int64(), _edx.C.c_int64(), _edx.C.c_int32()
    _edx.lib.edx_hamiltonian_info(H._dev.handle, _edx.C.byref(d), _edx.C.byref(n_e), _edx.C.byref(g))
    assert g.value == r["lanes"]
    np.testing.assert_allclose(H._dev.matvec(x), y0, rtol=0, atol=1e-12)
    assert len(_edx.autotune_table()) == 0  # direct tuning does not touch the table
