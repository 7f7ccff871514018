"""GPU: batched sector solves (``edx_eigs_batch``).

A batch of uneven sectors — Hubbard 2x4 (D ~ 1.6e3), Heisenberg 16 (810),
Haldane bosons 3x4 (1.1e4, once with a stored table and once on the fly),
Hubbard 3x4 (2.3e5), a twisted Hubbard 2x4 sector (the Hermitian-part
table) and a sector of the dense path (Heisenberg 12, D <= 500) — is solved
together and one by one, in FP64 and in mixed precision.  Every solve runs
the same kernels on the same inputs in both cases (the grouped SpMM,
replayed from captured round graphs, applies each table with the
single-table kernel's arithmetic), so eigenvalues, eigenvectors and
matvec counts must be identical bit for bit; the round graphs are counted.  The sector
loop ``solve_sectors_batched`` must equal ``ed_scan``'s results the same
way, and invalid batches are refused.
"""

from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import make_case
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd._edx import lib
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list

pytestmark = pytest.mark.gpu


def _operator(name, sector, table="stored"):
    model, ne, G = make_case(name)
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[sector])
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None, table=table)
    H._keep = cat
    return H


@pytest.fixture(scope="module")
def operators():
    ops = [_operator("hub24", s) for s in range(4)]
    ops += [_operator("heis16", 1), _operator("hal34", 1), _operator("hal34", 2, table="on_the_fly"),
            _operator("hub34", 2), _operator("hub24_twist", 1), _operator("heis12", 1)]
    assert ops[6].on_the_fly and ops[-1].n <= 500 and ops[7].n > 2e5
    return ops


@pytest.mark.parametrize("precision", [ed.PREC_FP64, ed.PREC_MIXED])
def test_batched_solves_equal_single_solves(operators, precision):
    nev = 5
    eig_tol = 1e-12 if precision == ed.PREC_MIXED else 0.0
    cap0, rep0 = _edx.batch_graph_stats(_edx.current_context())
    got = ed.diagonalize_blocks_batched(operators, nev=nev, precision=precision, eig_tol=eig_tol)
    infos = [H.last_info.copy() for H in operators]
    # every round's grouped SpMM was a replay of a captured graph, and rounds
    # with an unchanged set of stepping solves reused their set's graph
    cap, rep = _edx.batch_graph_stats(_edx.current_context())
    rounds = sum(int(i[0] / i[3]) for i in infos if i[3] > 0)
    print(f"[batch] round graphs: {cap - cap0} captured, {rep - rep0} replayed", flush=True)
    assert 0 < cap - cap0 < rep - rep0 <= rounds
    for H, (w, v), inf in zip(operators, got, infos):
        w1, v1 = ed.diagonalize_block_matrixfree(H, H.n, nev=nev, precision=precision, eig_tol=eig_tol)
        print(f"[batch] D={H.n} prec={precision} matvecs={inf[0]:.0f} E0={w[0]:.12f}", flush=True)
        np.testing.assert_array_equal(w, w1)
        np.testing.assert_array_equal(v, v1)
        if H.n > 500:
            np.testing.assert_array_equal(inf, H.last_info)
        # ARPACK's residual rule in FP64; the eigenvalue bound 1e-12 leaves the
        # vectors accurate to about sqrt(1e-12 x gap)
        r = H(v[:, 0]) - w[0] * v[:, 0]
        assert np.linalg.norm(r) <= (1e-8 if eig_tol == 0.0 else 1e-5) * max(1.0, abs(w[0]))


def test_solve_sectors_batched_equals_scan():
    model, ne, G = make_case("hub24")
    frac = Fraction(ne, model.lattice.n_site)
    a = ed.build_ed_data(model, filling_fraction=frac, symmetry_group=G)
    b = ed.build_ed_data(model, filling_fraction=frac, symmetry_group=G)
    ed.solve_sectors_batched(a, nev=4, precision=ed.PREC_MIXED, table="stored")
    ed.ed_scan(b, nev=4, mode="matrixfree_mixed")
    assert sorted(a.ed_scan_res) == sorted(b.ed_scan_res) == list(range(len(a.irrep_list)))
    assert a.sector_dims == b.sector_dims
    for i in a.ed_scan_res:
        np.testing.assert_array_equal(a.ed_scan_res[i][0], b.ed_scan_res[i][0])
        np.testing.assert_array_equal(a.ed_scan_res[i][1], b.ed_scan_res[i][1])


def test_invalid_batches_are_refused(operators):
    H = operators[0]
    bctx = _edx.current_context()
    opts = _edx.eigs_options(nev=2)
    vals = np.zeros((2, 2))
    ham_arr = (_edx.C.c_void_p * 2)(H._dev.handle, H._dev.handle)
    c = ed.symmetry_resolved_ed._batch_contexts(1)[0]
    same = (_edx.C.c_void_p * 2)(c.handle, c.handle)
    with pytest.raises(ValueError, match="distinct"):
        bctx.check(lib.edx_eigs_batch(bctx.handle, 2, same, ham_arr, _edx.C.byref(opts),
                                      vals.ctypes.data_as(_edx.C.c_void_p), None, None))
    own = (_edx.C.c_void_p * 1)(bctx.handle)
    with pytest.raises(ValueError, match="own context"):
        bctx.check(lib.edx_eigs_batch(bctx.handle, 1, own, ham_arr, _edx.C.byref(opts),
                                      vals.ctypes.data_as(_edx.C.c_void_p), None, None))


def test_batch_edge_cases():
    """A batch of one, a batch holding empty and one-state sectors (every
    site filled: one sector of dimension 1, the others empty), and a batch
    whose operators are on the fly only."""
    H = _operator("hub24", 0)
    [(w, v)] = ed.diagonalize_blocks_batched([H], nev=3)
    w1, v1 = ed.diagonalize_block_matrixfree(H, H.n, nev=3)
    np.testing.assert_array_equal(w, w1)
    np.testing.assert_array_equal(v, v1)
    model, ne, G = make_case("cb23_full_twist")
    data = ed.build_ed_data(model, filling_fraction=Fraction(ne, model.lattice.n_site), symmetry_group=G)
    res = ed.solve_sectors_batched(data, nev=2, precision=ed.PREC_FP64, table="stored")
    assert sorted(data.sector_dims) == [0] * (len(data.sector_dims) - 1) + [1]
    assert sum(len(r[0]) for r in res) == 1
    fly = [_operator("hal34", s, table="on_the_fly") for s in (0, 1)]
    got = ed.diagonalize_blocks_batched(fly, nev=3, precision=ed.PREC_MIXED, eig_tol=1e-12)
    for Hf, (w, v) in zip(fly, got):
        w1, v1 = ed.diagonalize_block_matrixfree(Hf, Hf.n, nev=3, precision=ed.PREC_MIXED, eig_tol=1e-12)
        np.testing.assert_array_equal(w, w1)
