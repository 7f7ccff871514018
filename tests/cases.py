"""Seeded model cases shared by the CPU and GPU tests.

Each case returns ``(model, n_filled, group)`` built with the package's own
lattice/term builders (which follow the reference conventions).
"""

from __future__ import annotations

import numpy as np

from realspace_exatdiagonalization_py_amd import models as M
from realspace_exatdiagonalization_py_amd import second_quantized_model as SQ
from realspace_exatdiagonalization_py_amd.groups import build_identity_group, build_translation_group, group_arrays


def _twisted(model, theta):
    SQ.update_second_quantized_model_with_twisted_phases(model, twisted_phases_over_2π=list(theta))
    return model, build_translation_group(model.lattice, list(theta))


def make_case(name: str):
    if name == "heis16":  # spin-1/2 chain N=16, Sz=0, C16
        m = M.build_heisenberg_second_quantized_model(16)
        return m, 8, build_translation_group(m.lattice)
    if name == "heis12":
        m = M.build_heisenberg_second_quantized_model(12)
        return m, 6, build_translation_group(m.lattice)
    if name == "heis10_identity":
        m = M.build_heisenberg_second_quantized_model(10)
        return m, 5, build_identity_group(m.lattice.n_site)
    if name == "hub34":  # spinful Hubbard 3x4, U=4, half filling
        m = M.build_spinful_hubbard_second_quantized_model([3, 4], 1.0, 4.0)
        return m, 12, build_translation_group(m.lattice)
    if name == "hub23":
        m = M.build_spinful_hubbard_second_quantized_model([2, 3], 1.0, 8.0)
        return m, 6, build_translation_group(m.lattice)
    if name == "hub24":
        m = M.build_spinful_hubbard_second_quantized_model([2, 4], 1.0, 4.0)
        return m, 8, build_translation_group(m.lattice)
    if name == "hal23":
        m = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[2, 3])
        return m, 3, build_translation_group(m.lattice)
    if name == "hub23_twist":
        m = M.build_spinful_hubbard_second_quantized_model([2, 3], 1.0, 4.0)
        m, G = _twisted(m, (0.3, 0.21))
        return m, 6, G
    if name == "hub24_twist":  # sectors of ~1.6k states: past the dense cutoff
        m = M.build_spinful_hubbard_second_quantized_model([2, 4], 1.0, 4.0)
        m, G = _twisted(m, (0.31, 0.17))
        return m, 8, G
    if name == "heis12_twist":
        m = M.build_heisenberg_second_quantized_model(12)
        m, G = _twisted(m, (0.27, 0.0))
        return m, 6, G
    if name == "heis18_twist":  # sectors of ~2.7k states
        m = M.build_heisenberg_second_quantized_model(18)
        m, G = _twisted(m, (0.27, 0.0))
        return m, 9, G
    if name == "hal34_twist":  # sectors of ~11.2k states
        m = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[3, 4])
        m, G = _twisted(m, (0.31, 0.0))
        return m, 6, G
    if name == "cb34_twist":
        p = dict(M.my_optimal_param)
        p["t''"] = -0.2
        m = M.build_checkerboard_second_quantized_model([3, 4], p)
        m, G = _twisted(m, (0.37, 0.0))
        return m, 4, G
    if name == "hal34":
        m = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[3, 4])
        return m, 6, build_translation_group(m.lattice)
    if name == "hal23_twist":
        m = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[2, 3])
        m, G = _twisted(m, (0.3, 0.0))
        return m, 3, G
    if name == "hal33_twist":
        m = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[3, 3])
        m, G = _twisted(m, (0.41, 0.0))
        return m, 4, G
    if name == "cb23_twist":
        m = M.build_checkerboard_second_quantized_model([2, 3], dict(M.my_optimal_param))
        m, G = _twisted(m, (0.37, 0.0))
        return m, 3, G
    if name == "cb34":
        p = dict(M.my_optimal_param)
        p["t''"] = -0.2
        m = M.build_checkerboard_second_quantized_model([3, 4], p)
        return m, 4, build_translation_group(m.lattice)
    if name == "hal23_ne0":
        m = M.build_zero_flux_bosonic_fci_second_quantized_model(sample_size=[2, 3])
        return m, 0, build_translation_group(m.lattice)
    if name == "cb23_full_twist":  # every site filled, fermions, twisted: product-form phases
        m = M.build_checkerboard_second_quantized_model([2, 3], dict(M.my_optimal_param))
        m, G = _twisted(m, (0.29, 0.0))
        return m, 12, G
    if name == "hub23_disorder":  # seeded on-site disorder, identity group: D = C(12, 6) = 924
        eps = np.random.default_rng(11).uniform(-0.1, 0.1, 6)
        m = M.build_spinful_hubbard_second_quantized_model([2, 3], 1.0, 4.0, onsite_disorder=eps)
        return m, 6, build_identity_group(m.lattice.n_site)
    raise KeyError(name)


def is_fermionic(model) -> bool:
    return model.particle_statistics.name == "FERMIONIC"


def oracle_catalog(name: str):
    from oracle import ed_oracle as O

    model, ne, G = make_case(name)
    perm, ph = group_arrays(G)
    return O.build_catalog(model.lattice.n_site, ne, perm, ph, is_fermionic(model))


def random_vector(n: int, seed: int = 7) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def oracle_sector_levels(model, n_filled, group, irrep, nev, extra=4):
    """The oracle's solve of one sector in the reference's way (CSR table,
    ``eigs(SR)`` with the seeded start vector, real parts ascending; dense
    ``eigh`` for dims <= 500) and the exact lowest ``nev + extra`` levels of
    the block's Hermitian part (dense for dims <= 3000, else ``eigsh`` at
    tolerance 1e-13).  Returns ``(dim, w_reference_method, w_hermitian_part)``."""
    import scipy.sparse.linalg as spla

    from oracle import ed_oracle as O

    perm, ph = group_arrays(group)
    ferm = is_fermionic(model)
    oc = O.build_catalog(model.lattice.n_site, n_filled, perm, ph, ferm)
    reps, stab, o2b = O.sector_basis(oc, irrep.values)
    r, c, v, hd = O.projection_table(oc, reps, stab, o2b, model.bilinear_terms, model.density_density_terms,
                                     irrep.values, ferm)
    D = len(reps)
    H = O.csr_hamiltonian(r, c, v, hd, D)
    w_ref, _ = O.diagonalize(H, nev)
    Hh = ((H + H.getH()) * 0.5).tocsr()
    k = min(nev + extra, D - 2)
    if D <= 3000:
        herm = np.linalg.eigvalsh(Hh.toarray())[:k]
    else:
        herm = np.sort(spla.eigsh(Hh, k=k, which="SA", tol=1e-13, v0=O.reference_start_vector(D))[0])
    return D, np.asarray(w_ref), np.asarray(herm)


def isolated_levels(herm, nev, gap=1e-6):
    """Indices i < nev whose level is farther than ``gap`` from its neighbours
    in the exact spectrum ``herm`` (ascending; the last one is not judged
    above if it is the end of the list)."""
    out = []
    for i in range(min(nev, len(herm))):
        lo = herm[i] - herm[i - 1] if i > 0 else np.inf
        hi = herm[i + 1] - herm[i] if i + 1 < len(herm) else 0.0
        if lo > gap and hi > gap:
            out.append(i)
    return out
