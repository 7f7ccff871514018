"""Consumers of sector eigenvectors, on the device tables.

Mirrors the hot parts of the reference's observables
(``src/realspace_exactdiagonalization_py/observables/``):

* ``_expand_sector_state_to_fock_amplitudes`` (entanglement_spectrum.py:82-144),
  used by the entanglement spectrum and the many-body Chern number;
* ``many_body_position_phases``, ``_position_phase_for_mask`` and
  ``_position_operator_matrix`` (charge_pump.py:41-162);
* ``off_diagonal_long_range_order`` and its bootstrap ``_bootstrap_ed_sector``
  (off_diagonal_long_range_order.py:43-140, static_structure_factor.py:380-453).

The per-state loops run as HIP kernels over the catalog's representative
tables (``edx_fock_expand``, ``edx_position_operator``,
``edx_one_body_density``); the host keeps the reference's argument checks,
return types and exceptions.
"""

from __future__ import annotations

import cmath
import math
from fractions import Fraction

import numpy as np
import scipy.sparse as sp

from . import _edx
from ._edx import lib, ptr
from .groups import build_identity_group, build_translation_group, group_order
from .second_quantized_model import (
    Particle_Statistics,
    Real_Space_Second_Quantized_Model,
    update_second_quantized_model_with_twisted_phases,
)
from .symmetry_resolved_ed import (
    Symmetry_Resolved_ED_Data,
    Symmetry_Sector_Basis,
    build_ed_data,
    build_symmetry_sector_basis,
    ed_scan,
)

#: ``convention`` values of ``expand_sector_state_to_fock``.
REFERENCE, HAMILTONIAN = "reference", "hamiltonian"


def _check_basis(basis: Symmetry_Sector_Basis, particle_statistics: Particle_Statistics | None = None):
    cat = basis.catalog
    if basis.symmetry_group is not cat.symmetry_group:
        raise ValueError("the catalog was refreshed to another group after this basis was built; "
                         "rebuild the basis from the current catalog")
    if particle_statistics is not None and (particle_statistics is Particle_Statistics.FERMIONIC) != cat.fermionic:
        raise ValueError("particle statistics differ from those the catalog was built with")
    return cat


def expand_sector_state_to_fock(c, basis: Symmetry_Sector_Basis, *, atol: float = 1e-12,
                                convention: str = REFERENCE) -> tuple[np.ndarray, np.ndarray]:
    """Fock amplitudes of a sector state as arrays ``(masks, amps)``, masks
    ascending, pruned at ``atol`` and normalised.

    ``convention="reference"`` is the reference's formula
    ``sum_g c/sqrt(|G||Stab|) conj(chi(g)) alpha_g U_g|s>``.  With the table
    convention of the sector blocks (g maps the representative to the state,
    coefficient ``alpha conj(chi(g))``) the eigenvectors of a k != 0 block, or
    of any twisted block, expand to eigenvectors of the Fock-space Hamiltonian
    only under ``convention="hamiltonian"`` (``chi(g) conj(alpha_g)``); the two
    agree at k = 0 without twist."""
    if convention not in (REFERENCE, HAMILTONIAN):
        raise ValueError(f"convention must be {REFERENCE!r} or {HAMILTONIAN!r}")
    cat = _check_basis(basis)
    c = np.ascontiguousarray(np.asarray(c, dtype=np.complex128).reshape(-1))
    if c.shape[0] != basis.dim:
        raise ValueError(f"state has {c.shape[0]} amplitudes for a sector of dimension {basis.dim}")
    chi = np.ascontiguousarray(np.asarray(basis.irrep.values, dtype=np.complex128))
    ctx = _edx.current_context(cat._dev.ctx)
    conv = 0 if convention == REFERENCE else 1
    n = _edx.C.c_int64()
    ctx.check(lib.edx_fock_expand(ctx.handle, cat._dev.handle, basis._dev.handle, ptr(chi), ptr(c), float(atol), conv,
                                  0, None, None, _edx.C.byref(n)))
    masks = np.empty(n.value, dtype=np.uint64)
    amps = np.empty(n.value, dtype=np.complex128)
    ctx.check(lib.edx_fock_expand(ctx.handle, cat._dev.handle, basis._dev.handle, ptr(chi), ptr(c), float(atol), conv,
                                  n.value, ptr(masks), ptr(amps), _edx.C.byref(n)))
    return masks, amps


def _expand_sector_state_to_fock_amplitudes(c, basis: Symmetry_Sector_Basis,
                                            particle_statistics: Particle_Statistics, *,
                                            atol: float = 1e-12) -> dict[int, complex]:
    """``{mask: amplitude}`` of a sector eigenvector (entanglement_spectrum.py:82-144)."""
    _check_basis(basis, particle_statistics)
    masks, amps = expand_sector_state_to_fock(c, basis, atol=atol, convention=REFERENCE)
    return {int(m): complex(a) for m, a in zip(masks, amps)}


def many_body_position_phases(lattice, direction: int, *, include_sublattice: bool = True) -> np.ndarray:
    """Per-site phases ``exp(2 pi i x_j / L)`` of Resta's position operator
    (charge_pump.py:41-78)."""
    if not 1 <= direction <= lattice.dim:
        raise ValueError(f"direction must be in 1..{lattice.dim}.")
    L = float(lattice.sample_size[direction - 1])
    phases = np.zeros(lattice.n_site, dtype=np.complex128)
    for i, (cell, isub) in enumerate(lattice.site_list):
        x = float(cell[direction - 1])
        if include_sublattice:
            x += float(lattice.sub_crys_list[isub - 1][direction - 1])
        phases[i] = cmath.exp(2j * math.pi * x / L)
    return phases


def _position_phase_for_mask(m: int, site_phases: np.ndarray) -> complex:
    """Product of the phases of the occupied sites (charge_pump.py:81-99)."""
    z = 1.0 + 0.0j
    tmp = int(m)
    while tmp:
        lsb = tmp & -tmp
        z *= site_phases[lsb.bit_length() - 1]
        tmp ^= lsb
    return z


def _position_operator_matrix(basis_to: Symmetry_Sector_Basis, basis_from: Symmetry_Sector_Basis,
                              site_phases: np.ndarray, particle_statistics: Particle_Statistics) -> sp.csr_matrix:
    """Projected position operator ``<chi_to| U |chi_from>`` (charge_pump.py:102-162)."""
    G = basis_from.symmetry_group
    assert basis_to.symmetry_group.name == G.name
    assert group_order(basis_to.symmetry_group) == group_order(G)
    assert basis_to.symmetry_group.n_site == G.n_site
    cat = _check_basis(basis_from, particle_statistics)
    _check_basis(basis_to, particle_statistics)
    if basis_to.catalog is not cat:
        raise ValueError("both bases must come from the same orbit catalog")
    n_to, n_from = basis_to.dim, basis_from.dim
    if n_from == 0:
        return sp.csr_matrix((n_to, n_from), dtype=np.complex128)
    ph = np.ascontiguousarray(np.asarray(site_phases, dtype=np.complex128))
    if ph.shape[0] != cat.n_site:
        raise ValueError("site_phases must have one entry per site")
    chi_to = np.ascontiguousarray(np.asarray(basis_to.irrep.values, dtype=np.complex128))
    chi_from = np.ascontiguousarray(np.asarray(basis_from.irrep.values, dtype=np.complex128))
    rows = np.empty(n_from, dtype=np.int64)
    vals = np.empty(n_from, dtype=np.complex128)
    ctx = _edx.current_context(cat._dev.ctx)
    ctx.check(lib.edx_position_operator(ctx.handle, cat._dev.handle, basis_to._dev.handle, basis_from._dev.handle,
                                        ptr(chi_to), ptr(chi_from), ptr(ph), ptr(rows), ptr(vals)))
    keep = rows >= 0
    cols = np.nonzero(keep)[0]
    return sp.csr_matrix((vals[keep], (rows[keep], cols)), shape=(n_to, n_from))


def one_body_density_matrix(basis: Symmetry_Sector_Basis, c, particle_statistics: Particle_Statistics) -> np.ndarray:
    """``rho[i, j] = <psi| a_(i+1)^dag a_(j+1) |psi>`` of the sector state
    ``c``, 0-based ``(n_site, n_site)`` (off_diagonal_long_range_order.py:80-140)."""
    cat = _check_basis(basis, particle_statistics)
    c = np.ascontiguousarray(np.asarray(c, dtype=np.complex128).reshape(-1))
    if c.shape[0] != basis.dim:
        raise ValueError(f"state has {c.shape[0]} amplitudes for a sector of dimension {basis.dim}")
    chi = np.ascontiguousarray(np.asarray(basis.irrep.values, dtype=np.complex128))
    rho = np.zeros((cat.n_site, cat.n_site), dtype=np.complex128)
    ctx = _edx.current_context(cat._dev.ctx)
    ctx.check(lib.edx_one_body_density(ctx.handle, cat._dev.handle, basis._dev.handle, ptr(chi), ptr(c), ptr(rho)))
    return rho


def _bootstrap_ed_sector(model: Real_Space_Second_Quantized_Model, sector_label, filling_fraction: Fraction,
                         target_eigval_idx: int, ed_mode: str, ed_data: Symmetry_Resolved_ED_Data | None):
    """``(ed_data, basis, c)`` for a sector eigenvector (static_structure_factor.py:380-453)."""
    lattice = model.lattice
    n_site = lattice.n_site
    if ed_data is None:
        flux0 = [0.0] * lattice.dim
        update_second_quantized_model_with_twisted_phases(model, twisted_phases_over_2π=flux0)
        G = build_identity_group(n_site) if sector_label == "identity" else build_translation_group(lattice, flux0)
        ed_data = build_ed_data(model, filling_fraction=filling_fraction, symmetry_group=G)
        scanned = None if sector_label == "identity" else [sector_label]
        ed_scan(ed_data, nev=max(target_eigval_idx, 2), mode=ed_mode, scanned_sectors=scanned)
    irrep_idx = None
    for idx, irrep in enumerate(ed_data.irrep_list):
        if irrep.label == sector_label:
            irrep_idx = idx
            break
    if irrep_idx is None:
        raise ValueError(f"Sector {sector_label!r} not found in ed_data.irrep_list.")
    if irrep_idx not in ed_data.ed_scan_res:
        raise ValueError(f"Sector {sector_label!r} was not scanned.")
    _, vecs = ed_data.ed_scan_res[irrep_idx]
    if target_eigval_idx > vecs.shape[1]:
        raise ValueError(f"target_eigval_idx={target_eigval_idx} exceeds {vecs.shape[1]} eigenvectors.")
    basis = build_symmetry_sector_basis(ed_data.orbit_catalog, ed_data.irrep_list[irrep_idx])
    c = vecs[:, target_eigval_idx - 1]
    return ed_data, basis, c


def off_diagonal_long_range_order(model: Real_Space_Second_Quantized_Model, sector_label, *,
                                  target_eigval_idx: int = 1, filling_fraction: Fraction, ed_mode: str = "matrix",
                                  ed_data: Symmetry_Resolved_ED_Data | None = None) -> np.ndarray:
    """One-body density matrix of a sector eigenstate
    (off_diagonal_long_range_order.py:43-140)."""
    label = "identity" if sector_label == "identity" else tuple(int(x) for x in sector_label)
    ed_data, basis, c = _bootstrap_ed_sector(model, label, filling_fraction, target_eigval_idx, ed_mode, ed_data)
    return one_body_density_matrix(basis, c, model.particle_statistics)
