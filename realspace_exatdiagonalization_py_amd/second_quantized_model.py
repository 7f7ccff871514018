"""Second-quantized model container (API input of the ED path).

Mirrors the reference's ``second_quantized_model.py``:
``Particle_Statistics`` (:29-45), ``Real_Space_Second_Quantized_Model``
(:48-85) and ``update_second_quantized_model_with_twisted_phases``
(:88-130).  Terms are ``(i, j, amplitude)`` tuples with 1-based sites; a
bilinear term moves a particle from site ``i`` to site ``j`` in the ED
engine, a density term ``(i, j, V)`` adds ``V`` when both sites are occupied
(``i == j`` gives ``V n_i``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from enum import Enum
from typing import Any


class Particle_Statistics(Enum):
    """Hard-core bosons (no signs) or spinless fermions (JW and permutation signs)."""

    BOSONIC = "Bosonic"
    FERMIONIC = "Fermionic"

    def __repr__(self) -> str:  # pragma: no cover - cosmetic
        return f"Particle_Statistics.{self.name}"


@dataclass
class Real_Space_Second_Quantized_Model:
    params: dict[str, Any]
    lattice: Any
    tb_model: Any
    particle_statistics: Particle_Statistics
    bilinear_terms: list[tuple[int, int, complex]] = field(default_factory=list)
    density_density_terms: list[tuple[int, int, complex]] = field(default_factory=list)


def update_second_quantized_model_with_twisted_phases(
    second_quantized_model: Real_Space_Second_Quantized_Model,
    *,
    twisted_phases_over_2π: list[float],
) -> Real_Space_Second_Quantized_Model:
    """In-place: regenerate the bilinear terms with twisted boundary phases."""
    from .lattice import generate_bilinear_terms

    lattice = second_quantized_model.lattice
    if len(twisted_phases_over_2π) != lattice.dim:
        raise ValueError(f"twisted_phases_over_2π must have length {lattice.dim}.")
    second_quantized_model.bilinear_terms = generate_bilinear_terms(
        second_quantized_model.tb_model,
        twisted_phases_over_2π=list(twisted_phases_over_2π),
    )
    lattice.twisted_phases_over_2π = list(twisted_phases_over_2π)
    return second_quantized_model


__all__ = [
    "Particle_Statistics",
    "Real_Space_Second_Quantized_Model",
    "update_second_quantized_model_with_twisted_phases",
]
