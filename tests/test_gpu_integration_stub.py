"""GPU: the reference-side ctypes binding shown in INTEGRATION.md, run as
written.

The code block that starts with ``# _edx_gpu.py`` is taken from
INTEGRATION.md, the library path is filled in, and the module is executed.
Its ``eigs_gpu`` then solves sector blocks that the oracle builds in the
reference's way (CSR from the projection table): an untwisted Heisenberg
sector past the dense cutoff in FP64 and mixed precision, and a twisted
Hubbard sector (the stub's Hermitian-part rule).  The eigenvalues must equal
the package's own solve of the same sector within 1e-10, the eigenvectors
must have relative residual <= 1e-8 under the oracle's CSR, and the stub's
error mapping must raise ValueError for an invalid table.
"""

from __future__ import annotations

import ctypes as C
import re
import types
from pathlib import Path

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import is_fermionic, make_case
from oracle import ed_oracle as O
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list, group_arrays

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def _stub_module():
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```python\n(.*?)```", text, flags=re.S)
    src = next(b for b in blocks if b.lstrip().startswith("# _edx_gpu.py"))
    assert 'C.CDLL("libedx.so")' in src
    src = src.replace('C.CDLL("libedx.so")', f"C.CDLL({str(_edx.LIB_PATH)!r})")
    mod = types.ModuleType("_edx_gpu_stub")
    exec(compile(src, "INTEGRATION.md:_edx_gpu.py", "exec"), mod.__dict__)
    return mod


def _oracle_block(name, sector):
    model, ne, G = make_case(name)
    perm, ph = group_arrays(G)
    ferm = is_fermionic(model)
    oc = O.build_catalog(model.lattice.n_site, ne, perm, ph, ferm)
    irrep = build_irrep_list(G, model.lattice)[sector]
    reps, stab, o2b = O.sector_basis(oc, irrep.values)
    r, c, v, hd = O.projection_table(oc, reps, stab, o2b, model.bilinear_terms, model.density_density_terms,
                                     irrep.values, ferm)
    return model, ne, G, irrep, O.csr_hamiltonian(r, c, v, hd, len(reps))


def _package_levels(model, ne, G, irrep, nev):
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, irrep)
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=nev, want_vectors=False)
    return w


@pytest.mark.parametrize("name,sector,nev", [("heis16", 1, 4), ("hub24_twist", 1, 5)])
def test_integration_stub_solves_reference_blocks(name, sector, nev):
    stub = _stub_module()
    model, ne, G, irrep, Href = _oracle_block(name, sector)
    assert Href.shape[0] > 500
    w_pkg = _package_levels(model, ne, G, irrep, nev)
    Hh = Href if name == "heis16" else (Href + Href.getH()) * 0.5
    for prec in (0, 1):
        vals, vecs = stub.eigs_gpu(Href, nev, precision=prec)
        print(f"[stub] {name} prec={prec} max|dE| vs package {np.max(np.abs(vals - w_pkg)):.2e}", flush=True)
        assert np.max(np.abs(vals - w_pkg)) <= 1e-10, (prec, vals, w_pkg)
        R = Hh @ vecs - vecs * vals
        assert np.max(np.linalg.norm(R, axis=0) / np.maximum(1.0, np.abs(vals))) <= 1e-8


def test_integration_stub_error_mapping():
    stub = _stub_module()
    if not stub._ctx:
        stub._check(stub._lib.edx_ctx_create(0, C.byref(stub._ctx)))
    rp = np.array([1, 1, 1], dtype=np.int64)  # row_ptr must start at 0
    diag = np.zeros(2, dtype=np.complex128)
    h = C.c_void_p()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    with pytest.raises(ValueError, match="row_ptr"):
        stub._check(stub._lib.edx_hamiltonian_from_csr(stub._ctx, 2, 0, ptr(rp), None, None, ptr(diag), C.byref(h)))
    assert stub._EXC[-2] is AssertionError and stub._EXC[-4] is MemoryError
