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
