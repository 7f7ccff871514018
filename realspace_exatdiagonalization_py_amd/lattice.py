"""Real-space lattices and tight-binding term generation.

The reference takes its lattice and its hopping-term generator from the
``tightbinding-py`` package (pinned 0.2.0 as an editable path dependency,
``pyproject.toml:26-27``), which is not distributed with it.  This module is
this package's own implementation of the small part of that interface the ED
path consumes:

* ``initialize_real_space_lattice`` -> an object with ``sample_size``,
  ``n_sub``, ``dim``, ``n_site``, ``site_list`` (``((cell), isub)`` with a
  1-based ``isub``), ``site_to_index_map`` (1-based) and
  ``twisted_phases_over_2π``; used by ``build_translation_group``
  (reference ``symmetry_resolved_ed.py:2754-2790``) and the model builders.
* ``initialize_real_space_tightbinding_model`` / ``add_hopping_term`` /
  ``generate_bilinear_terms`` / ``full_hopping_map`` -> the bond list and its
  expansion into directed ``(i, j, t)`` terms (call sites
  ``models/heisenberg.py:72``, ``models/hubbard.py:77``,
  ``models/bosonic_fci.py:117``, ``second_quantized_model.py:120-125``).

Conventions (each pinned by a test against data the reference holds):

* Sites are ordered cell-major with the sublattice index innermost, and cells
  with the FIRST coordinate fastest: ``index = 1 + n_sub*(x + L1*y) + (isub-1)``
  (pinned by the flux-aware translation phases printed in the reference's
  ``doc/design.ipynb`` cell 27 and by the Haldane 2x3 representative list of
  cell 16).
* A bond ``((c_from, s_from), (c_to, s_to), t)`` is translated over every cell
  and emitted as the directed pair ``(i_from, i_to, t)`` followed by
  ``(i_to, i_from, conj(t))`` (term count 144 for Haldane 2x3,
  ``test/test_core.py:106-111``).
* With twisted boundaries, the term ``(i_from, i_to, t)`` of a bond that winds
  ``w_d`` times around periodic direction ``d`` carries
  ``exp(-2 pi i sum_d theta_d w_d)``.  The sign is the one for which the
  reference's flux-aware translation group (``symmetry_resolved_ed.py:2776-2788``)
  commutes with the Hamiltonian the reference's engine builds from the terms
  (hop from the first index to the second, ``symmetry_resolved_ed.py:635-638``);
  the reference's own sign lives in the absent package, so twisted-term parity
  with it is unpinned.
"""

from __future__ import annotations

import cmath
import math
from dataclasses import dataclass, field
from typing import Any

Cell = tuple[int, ...]
Site = tuple[Cell, int]


@dataclass
class Real_Space_Lattice:
    sample_size: list[int]
    brav_vec_list: list[list[float]]
    sub_crys_list: list[list[float]]
    lattice_name: str
    pbc_indicator: list[bool]
    site_list: list[Site] = field(default_factory=list)
    site_to_index_map: dict[Site, int] = field(default_factory=dict)
    twisted_phases_over_2π: list[float] = field(default_factory=list)

    @property
    def dim(self) -> int:
        return len(self.sample_size)

    @property
    def n_sub(self) -> int:
        return len(self.sub_crys_list)

    @property
    def n_site(self) -> int:
        return len(self.site_list)

    @property
    def n_cell(self) -> int:
        return math.prod(self.sample_size)


def _cells_first_fastest(sample_size: list[int]) -> list[Cell]:
    """All cells with the first coordinate varying fastest."""
    cells: list[Cell] = [()]
    for L in sample_size:
        cells = [c + (x,) for x in range(L) for c in cells]
    return cells


def initialize_real_space_lattice(
    *,
    sample_size: list[int],
    brav_vec_list: list[list[float]] | None = None,
    sub_crys_list: list[list[float]] | None = None,
    lattice_name: str = "lattice",
    pbc_indicator: list[bool] | None = None,
) -> Real_Space_Lattice:
    """Build a finite lattice of ``prod(sample_size)`` cells x ``n_sub`` sites."""
    dim = len(sample_size)
    if brav_vec_list is None:
        brav_vec_list = [[1.0 if a == b else 0.0 for b in range(dim)] for a in range(dim)]
    if sub_crys_list is None:
        sub_crys_list = [[0.0] * dim]
    if pbc_indicator is None:
        pbc_indicator = [True] * dim
    lat = Real_Space_Lattice(
        sample_size=[int(x) for x in sample_size],
        brav_vec_list=[list(v) for v in brav_vec_list],
        sub_crys_list=[list(v) for v in sub_crys_list],
        lattice_name=lattice_name,
        pbc_indicator=list(pbc_indicator),
    )
    for cell in _cells_first_fastest(lat.sample_size):
        for isub in range(1, lat.n_sub + 1):
            lat.site_list.append((cell, isub))
            lat.site_to_index_map[(cell, isub)] = len(lat.site_list)
    lat.twisted_phases_over_2π = [0.0] * dim
    return lat


@dataclass
class Real_Space_TightBinding_Model:
    lattice: Real_Space_Lattice
    model_name: str
    hopping_list: list[tuple[tuple[Site, Site], complex]] = field(default_factory=list)

    @property
    def full_hopping_map(self) -> dict[tuple[Site, Site], complex]:
        """Directed bonds over all cells (both directions), zero twist.

        Amplitudes of bonds that join the same ordered pair of sites (the direct
        and the wrapped bond of a length-2 periodic direction) add up; the
        reference's Hubbard 2x2 benchmark energy is reproduced only that way."""
        out: dict[tuple[Site, Site], complex] = {}
        for (s_from, s_to, t, _w) in _expanded_bonds(self):
            out[(s_from, s_to)] = out.get((s_from, s_to), 0j) + t
            out[(s_to, s_from)] = out.get((s_to, s_from), 0j) + t.conjugate()
        return out


def initialize_real_space_tightbinding_model(
    lattice: Real_Space_Lattice, *, model_name: str = "tb_model"
) -> Real_Space_TightBinding_Model:
    return Real_Space_TightBinding_Model(lattice=lattice, model_name=model_name)


def add_hopping_term(tb_model: Real_Space_TightBinding_Model, term) -> None:
    """Append one bond ``(((c_from, s_from), (c_to, s_to)), t)`` (cells relative)."""
    (site_from, site_to), t = term
    c_from, s_from = site_from
    c_to, s_to = site_to
    tb_model.hopping_list.append(
        (((tuple(int(x) for x in c_from), int(s_from)),
          (tuple(int(x) for x in c_to), int(s_to))), complex(t))
    )


def _expanded_bonds(tb_model: Real_Space_TightBinding_Model):
    """Yield ``(site_from, site_to, t, winding)`` for every cell and bond.

    ``winding[d]`` counts how often the bond wraps around direction ``d``
    going from ``site_from`` to ``site_to``.  Bonds leaving the sample along
    an open (non-periodic) direction are dropped.
    """
    lat = tb_model.lattice
    L = lat.sample_size
    for cell in _cells_first_fastest(L):
        for ((c_from, s_from), (c_to, s_to)), t in tb_model.hopping_list:
            raw_from = [a + b for a, b in zip(cell, c_from)]
            raw_to = [a + b for a, b in zip(cell, c_to)]
            keep = True
            for d in range(lat.dim):
                if not lat.pbc_indicator[d] and not (
                    0 <= raw_from[d] < L[d] and 0 <= raw_to[d] < L[d]
                ):
                    keep = False
            if not keep:
                continue
            winding = tuple(
                math.floor(raw_to[d] / L[d]) - math.floor(raw_from[d] / L[d])
                for d in range(lat.dim)
            )
            site_from = (tuple(x % Ld for x, Ld in zip(raw_from, L)), s_from)
            site_to = (tuple(x % Ld for x, Ld in zip(raw_to, L)), s_to)
            yield site_from, site_to, t, winding


def generate_bilinear_terms(
    tb_model: Real_Space_TightBinding_Model,
    *,
    twisted_phases_over_2π: list[float] | None = None,
) -> list[tuple[int, int, complex]]:
    """Directed ``(i, j, t)`` terms (1-based) of every bond and its conjugate."""
    lat = tb_model.lattice
    theta = [0.0] * lat.dim if twisted_phases_over_2π is None else [
        float(x) for x in twisted_phases_over_2π
    ]
    if len(theta) != lat.dim:
        raise ValueError(f"twisted_phases_over_2π must have length {lat.dim}.")
    idx = lat.site_to_index_map
    terms: list[tuple[int, int, complex]] = []
    for site_from, site_to, t, winding in _expanded_bonds(tb_model):
        arg = sum(th * w for th, w in zip(theta, winding))
        tw = t * cmath.exp(-2j * math.pi * arg) if arg != 0.0 else t
        terms.append((idx[site_from], idx[site_to], tw))
        terms.append((idx[site_to], idx[site_from], tw.conjugate()))
    return terms


__all__ = [
    "Real_Space_Lattice",
    "Real_Space_TightBinding_Model",
    "initialize_real_space_lattice",
    "initialize_real_space_tightbinding_model",
    "add_hopping_term",
    "generate_bilinear_terms",
]
