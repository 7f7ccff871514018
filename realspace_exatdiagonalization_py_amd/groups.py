"""Symmetry operations, finite groups and 1-D irreps (host side).

Mirrors the reference's section 1/2/7/18/19 API (``symmetry_resolved_ed.py``):
``Symmetry_Operation`` (:930-971), ``apply_operation_to_mask`` (:977-1008),
``Finite_Symmetry_Group`` (:1016-1061), ``group_order`` (:1070),
``get_canonical_representative`` (:1087-1114), ``OneDim_Irrep``
(:1412-1433), ``build_identity_group`` (:2711), ``build_translation_group``
(:2724-2790), the irrep builders (:2815-2898).  These are O(|G| N) host
computations that define the group order, the per-site phases and the
characters the device kernels consume.
"""

from __future__ import annotations

import cmath
import itertools
import math
from dataclasses import dataclass
from typing import Any

import numpy as np

from .bitwise_operations import COMPLEX_ONE, Mask
from .second_quantized_model import Particle_Statistics


@dataclass(init=False)
class Symmetry_Operation:
    """Permutation ``perm`` (1-based targets) with per-site phases ``eta_g(i)``."""

    label: Any
    perm: list[int]
    perm_phases: np.ndarray

    def __init__(self, label: Any, perm: list[int], *, perm_phases=None):
        n = len(perm)
        ph = (np.full(n, COMPLEX_ONE, dtype=np.complex128) if perm_phases is None
              else np.asarray(perm_phases, dtype=np.complex128))
        assert ph.shape[0] == n, "Symmetry_Operation: perm_phases must have length n_site."
        self.label = label
        self.perm = [int(p) for p in perm]
        self.perm_phases = ph

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return f"Symmetry_Operation(label={self.label!r}, perm={self.perm})"


def apply_operation_to_mask(m: Mask, op: Symmetry_Operation,
                            particle_statistics: Particle_Statistics) -> tuple[Mask, complex]:
    """Image of ``m`` under ``op`` and the accumulated phase: running complex
    product of ``eta`` over occupied sites (ascending), times the parity of the
    permutation restricted to the occupied sites for fermions."""
    fermionic = particle_statistics is Particle_Statistics.FERMIONIC
    tmp, new, phase, parity = m, 0, COMPLEX_ONE, 1
    while tmp:
        low = tmp & -tmp
        i = low.bit_length() - 1
        p = op.perm[i]
        if fermionic and (new >> p).bit_count() & 1:
            parity = -parity
        new |= 1 << (p - 1)
        phase *= op.perm_phases[i]
        tmp ^= low
    return new, phase * parity


@dataclass(init=False)
class Finite_Symmetry_Group:
    """Ordered list of operations; ``identity_idx`` is 0-based."""

    name: str
    n_site: int
    operations: list[Symmetry_Operation]
    identity_idx: int

    def __init__(self, name: str, ops: list[Symmetry_Operation], *, identity_idx: int = 0):
        if len(ops) == 0:
            raise ValueError("Finite_Symmetry_Group: ops must not be empty.")
        n = len(ops[0].perm)
        for op in ops:
            if len(op.perm) != n or op.perm_phases.shape[0] != n:
                raise ValueError("Finite_Symmetry_Group: all operations must have "
                                 "perm/perm_phases of length n_site.")
        if ops[identity_idx].perm != list(range(1, n + 1)):
            raise ValueError("Finite_Symmetry_Group: the identity operation must have "
                             "the trivial permutation.")
        self.name = name
        self.n_site = n
        self.operations = ops
        self.identity_idx = identity_idx

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return f"Finite_Symmetry_Group(name={self.name!r}, n_site={self.n_site}, |G|={len(self.operations)})"


def group_order(G: Finite_Symmetry_Group) -> int:
    return len(G.operations)


def group_arrays(G: Finite_Symmetry_Group) -> tuple[np.ndarray, np.ndarray]:
    """``(perm, phases)``: ``(nG, N)`` int32 1-based targets and complex128 phases."""
    perm = np.asarray([op.perm for op in G.operations], dtype=np.int32)
    phases = np.asarray([op.perm_phases for op in G.operations], dtype=np.complex128)
    return perm, phases


def get_canonical_representative(m: Mask, G: Finite_Symmetry_Group,
                                 stats: Particle_Statistics) -> tuple[Mask, int, complex]:
    """Smallest mask of the orbit, the first g reaching it (g maps m to it), its phase."""
    best, best_g, best_a = m, G.identity_idx, COMPLEX_ONE
    for g, op in enumerate(G.operations):
        s, a = apply_operation_to_mask(m, op, stats)
        if s < best:
            best, best_g, best_a = s, g, a
    return best, best_g, best_a


@dataclass(init=False)
class OneDim_Irrep:
    label: Any
    values: np.ndarray

    def __init__(self, label: Any, values):
        self.label = label
        self.values = np.asarray(values, dtype=np.complex128)

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return f"OneDim_Irrep(label={self.label!r}, |values|={len(self.values)})"


def build_identity_group(n_site: int) -> Finite_Symmetry_Group:
    return Finite_Symmetry_Group("identity", [Symmetry_Operation("identity", list(range(1, n_site + 1)))],
                                 identity_idx=0)


def build_translation_group(lattice, θ=None) -> Finite_Symmetry_Group:
    """Translations in ``itertools.product`` order (first direction outermost),
    with gauge-covariant phases ``eta_g(i) = exp(2 pi i sum_d theta_d (c_d - c'_d) / L_d)``
    where ``c'`` is the translated (wrapped) cell of site i."""
    L = list(lattice.sample_size)
    if θ is None or len(θ) == 0:
        tw = getattr(lattice, "twisted_phases_over_2π", [])
        th = [float(x) for x in tw] if tw else [0.0] * lattice.dim
    else:
        if len(θ) != lattice.dim:
            raise ValueError(f"θ must have length {lattice.dim}.")
        th = [float(x) for x in θ]
    Lf = [float(x) for x in L]
    ops = []
    for shift in itertools.product(*[range(n) for n in L]):
        perm = [0] * lattice.n_site
        ph = np.zeros(lattice.n_site, dtype=np.complex128)
        for i, (cell, isub) in enumerate(lattice.site_list):
            sc = tuple((c + s) % Ld for c, s, Ld in zip(cell, shift, L))
            perm[i] = lattice.site_to_index_map[(sc, isub)]
            ph[i] = cmath.exp(2j * math.pi * sum(t * (c - s2) / Ld
                                                 for t, c, s2, Ld in zip(th, cell, sc, Lf)))
        ops.append(Symmetry_Operation(tuple(shift), perm, perm_phases=ph))
    return Finite_Symmetry_Group("translations", ops, identity_idx=0)


def build_identity_irrep_list() -> list[OneDim_Irrep]:
    return [OneDim_Irrep("identity", [1.0 + 0.0j])]


def build_translation_irrep_list(G: Finite_Symmetry_Group, lattice) -> list[OneDim_Irrep]:
    """Momentum irreps ``chi_(k1,k2)(d1,d2) = exp(2 pi i (k1 d1/L1 + k2 d2/L2))``, k1-major."""
    assert G.name == "translations" and group_order(G) == math.prod(lattice.sample_size)
    L1, L2 = lattice.sample_size
    out = []
    for k1 in range(L1):
        for k2 in range(L2):
            vals = [cmath.exp(2j * math.pi * (k1 * op.label[0] / L1 + k2 * op.label[1] / L2))
                    for op in G.operations]
            out.append(OneDim_Irrep((k1, k2), vals))
    return out


def build_irrep_list(G: Finite_Symmetry_Group, lattice) -> list[OneDim_Irrep]:
    if G.name == "identity":
        return build_identity_irrep_list()
    if G.name == "translations":
        return build_translation_irrep_list(G, lattice)
    raise ValueError(f"Irrep construction for group {G.name!r} is not yet implemented.")


__all__ = [
    "Symmetry_Operation", "apply_operation_to_mask", "Finite_Symmetry_Group", "group_order",
    "group_arrays", "get_canonical_representative", "OneDim_Irrep", "build_identity_group",
    "build_translation_group", "build_identity_irrep_list", "build_translation_irrep_list",
    "build_irrep_list",
]
