"""GPU: ``edx_matvec_noalloc`` (work areas in a static device array, no
device allocation) against ``edx_matvec`` through the C ABI.

Both run the same kernels on the same inputs, so y = H x must agree bit for
bit.  Checked on two calls in a row with the second smaller (a 16-site
Heisenberg sector, then a 12-site one), on a twisted table (whose product is
applied from the CSR arrays), on more than 16 vectors (batched) and on a
block placed on the device from a host CSR; a dimension beyond the scratch
limit is rejected with ValueError.
"""

from __future__ import annotations

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import make_case
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd._edx import lib, ptr
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list

pytestmark = pytest.mark.gpu


def _sector_hamiltonian(name, sector):
    model, ne, G = make_case(name)
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[sector])
    return ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)


def _both(dh, n_vec, seed):
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(rng.standard_normal((n_vec, dh.dim)) + 1j * rng.standard_normal((n_vec, dh.dim)))
    y_ref = np.empty_like(x)
    y = np.full_like(x, np.nan)
    dh.ctx.check(lib.edx_matvec(dh.ctx.handle, dh.handle, n_vec, ptr(x), ptr(y_ref)))
    dh.ctx.check(lib.edx_matvec_noalloc(dh.ctx.handle, dh.handle, n_vec, ptr(x), ptr(y)))
    return y_ref, y


def test_noalloc_matches_matvec_twice_second_smaller():
    big = _sector_hamiltonian("heis16", 0)
    small = _sector_hamiltonian("heis12", 1)
    assert big.n > small.n > 0
    y_ref, y = _both(big._dev, 5, seed=1)
    np.testing.assert_array_equal(y, y_ref)
    y_ref, y = _both(small._dev, 2, seed=2)
    np.testing.assert_array_equal(y, y_ref)


def test_noalloc_twisted_table_and_batches():
    H = _sector_hamiltonian("heis12_twist", 1)
    y_ref, y = _both(H._dev, 18, seed=3)  # two batches (16 + 2)
    np.testing.assert_array_equal(y, y_ref)
    y_ref, y = _both(H._dev, 1, seed=4)
    np.testing.assert_array_equal(y, y_ref)


def _csr_block(n, nnz_per_row, seed):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n), nnz_per_row)
    cols = rng.integers(0, n, size=rows.shape[0])
    keep = rows != cols
    import scipy.sparse as sp

    A = sp.csr_matrix((rng.standard_normal(keep.sum()) + 1j * rng.standard_normal(keep.sum()),
                       (rows[keep], cols[keep])), shape=(n, n))
    A = (A + A.getH()) * 0.5 + sp.diags(rng.standard_normal(n))
    return ed.symmetry_resolved_ed._csr_to_device(A)


def test_noalloc_host_csr_block():
    dh = _csr_block(3000, 7, seed=5)
    y_ref, y = _both(dh, 4, seed=6)
    np.testing.assert_array_equal(y, y_ref)


def test_noalloc_rejects_dimension_beyond_scratch():
    ctx = _edx.Context.default()
    n = 2_097_153  # one vector needs 48 * n bytes > 96 MiB
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    diag = np.zeros(n, dtype=np.complex128)
    h = _edx.C.c_void_p()
    ctx.check(lib.edx_hamiltonian_from_csr(ctx.handle, n, 0, ptr(row_ptr), None, None, ptr(diag),
                                           _edx.C.byref(h)))
    dh = ed.symmetry_resolved_ed._DeviceHamiltonian(ctx, h, ())
    x = np.zeros((1, n), dtype=np.complex128)
    y = np.zeros_like(x)
    with pytest.raises(ValueError, match="scratch limit"):
        ctx.check(lib.edx_matvec_noalloc(ctx.handle, dh.handle, 1, ptr(x), ptr(y)))
    # the regular entry point still serves it
    ctx.check(lib.edx_matvec(ctx.handle, dh.handle, 1, ptr(x), ptr(y)))
