"""Model builders (term lists) following the reference's conventions.

Each builder reproduces the term list of the reference builder it names, on
this package's own lattice module:

* ``build_heisenberg_second_quantized_model`` - ``models/heisenberg.py:40-95``
  (bilinear ``J/2`` on NN bonds, density ``J`` on bonds and ``-J/2`` on sites).
* ``build_spinful_hubbard_second_quantized_model`` - ``models/hubbard.py:29-101``
  (interleaved spin vertices, hopping ``-t``, on-site ``U`` as a density term).
* ``build_zero_flux_bosonic_fci_second_quantized_model`` -
  ``models/bosonic_fci.py:44-176`` (Haldane honeycomb hard-core bosons).
* ``build_zero_flux_fermionic_fci_second_quantized_model`` and
  ``build_phase_explore_fermionic_checkerboard_model`` -
  ``models/fermionic_fci.py:56-219`` and ``:599-758`` (checkerboard fermions).

The ``build_*_ed_data`` helpers wrap them with the translation group and
``build_ed_data``, as the reference's helpers do.
"""

from __future__ import annotations

import cmath
import math
from fractions import Fraction
from typing import Any

import numpy as np

from .lattice import (
    add_hopping_term,
    generate_bilinear_terms,
    initialize_real_space_lattice,
    initialize_real_space_tightbinding_model,
)
from .second_quantized_model import Particle_Statistics, Real_Space_Second_Quantized_Model

# Parameter sets (values as in the reference's model modules).
params_DNSheng: dict[str, float] = {
    "t": 1.0, "t'": 0.60, "t''": -0.58, "phi_over_2pi": 0.2, "V1": 0.0, "V2": 0.0,
}
params_Sun_Gu_Katsura_Sarma: dict[str, float] = {
    "t": 1.0,
    "t'_1": 1 / (2 + math.sqrt(2)),
    "t'_2": -1 / (2 + math.sqrt(2)),
    "t''": 1 / (2 + 2 * math.sqrt(2)),
    "phi_over_2pi": 1 / 8,
    "V1": 2.0, "V2": 1.0, "V3": 0.0,
}
my_optimal_param: dict[str, float] = {
    "t": -1.0,
    "t'_1": -1.4 / (2 + math.sqrt(2)),
    "t'_2": 1.4 / (2 + math.sqrt(2)),
    "t''": -1.0 / (2 + 2 * math.sqrt(2)),
    "phi_over_2pi": 1 / 8,
    "V1": 2.0, "V2": 0.45, "V3": 0.2, "λ": 2.8,
}
CHECKERBOARD_TPP_DENOMINATOR = 2 + 2 * math.sqrt(2)


# ---------------------------------------------------------------------------
# Heisenberg chain (hard-core boson mapping)
# ---------------------------------------------------------------------------

def build_heisenberg_second_quantized_model(N: int, J: float = 1.0) -> Real_Space_Second_Quantized_Model:
    lat = initialize_real_space_lattice(
        sample_size=[N, 1],
        brav_vec_list=[[1.0, 0.0], [0.0, 1.0]],
        sub_crys_list=[[0.0, 0.0]],
        lattice_name="1D_Chain",
        pbc_indicator=[True, False],
    )
    tb = initialize_real_space_tightbinding_model(lat, model_name="Heisenberg_Chain")
    add_hopping_term(tb, ((((0, 0), 1), ((1, 0), 1)), J / 2))
    bilinear = generate_bilinear_terms(tb, twisted_phases_over_2π=[0.0] * lat.dim)
    idx = lat.site_to_index_map
    density: list[tuple[int, int, complex]] = []
    for x in range(N):
        density.append((idx[((x, 0), 1)], idx[(((x + 1) % N, 0), 1)], complex(J)))
    for x in range(N):
        i = idx[((x, 0), 1)]
        density.append((i, i, complex(-J / 2)))
    return Real_Space_Second_Quantized_Model(
        params={"J": J, "N_site": N}, lattice=lat, tb_model=tb,
        particle_statistics=Particle_Statistics.BOSONIC,
        bilinear_terms=bilinear, density_density_terms=density,
    )


# ---------------------------------------------------------------------------
# Spinful Fermi-Hubbard on the square lattice
# ---------------------------------------------------------------------------

def build_spinful_hubbard_second_quantized_model(
    sample_size: list[int], t: float = 1.0, U: float = 8.0,
    *, onsite_disorder: np.ndarray | None = None,
) -> Real_Space_Second_Quantized_Model:
    """Spinful Hubbard model; ``onsite_disorder[s]`` (optional) adds ``eps_s n_s``
    on both spin vertices of spatial site ``s`` as density terms ``(i, i, eps)``."""
    Lx, Ly = sample_size
    r_data = initialize_real_space_lattice(
        sample_size=sample_size, brav_vec_list=[[1.0, 0.0], [0.0, 1.0]],
        sub_crys_list=[[0.0, 0.0]], lattice_name="Square_Hubbard", pbc_indicator=[True, True],
    )
    tb_spinless = initialize_real_space_tightbinding_model(r_data, model_name="hubbard_spinless")
    add_hopping_term(tb_spinless, ((((0, 0), 1), ((1, 0), 1)), complex(-t)))
    add_hopping_term(tb_spinless, ((((0, 0), 1), ((0, 1), 1)), complex(-t)))
    lat = initialize_real_space_lattice(
        sample_size=sample_size, brav_vec_list=[[1.0, 0.0], [0.0, 1.0]],
        sub_crys_list=[[0.0, 0.0], [0.0, 0.0]], lattice_name="Square_Hubbard_Spinful",
        pbc_indicator=[True, True],
    )

    def sv(x: int, y: int, s: int) -> int:
        return lat.site_to_index_map[((x, y), s)]

    bilinear: list[tuple[int, int, complex]] = []
    for (site_from, site_to), tamp in tb_spinless.full_hopping_map.items():
        x1, y1 = site_from[0]
        x2, y2 = site_to[0]
        bilinear.append((sv(x1, y1, 1), sv(x2, y2, 1), complex(tamp)))
        bilinear.append((sv(x1, y1, 2), sv(x2, y2, 2), complex(tamp)))
    density = [(sv(x, y, 1), sv(x, y, 2), complex(U)) for y in range(Ly) for x in range(Lx)]
    if onsite_disorder is not None:
        eps = np.asarray(onsite_disorder, dtype=np.float64).reshape(-1)
        if eps.shape[0] != Lx * Ly:
            raise ValueError("onsite_disorder must have one value per spatial site.")
        for y in range(Ly):
            for x in range(Lx):
                e = complex(eps[x + Lx * y])
                for s in (1, 2):
                    i = sv(x, y, s)
                    density.append((i, i, e))
    # the spinful tight-binding description (each spin hops as the spinless
    # model does), from which twisted-boundary terms are generated
    tb = initialize_real_space_tightbinding_model(lat, model_name="Hubbard")
    for s in (1, 2):
        add_hopping_term(tb, ((((0, 0), s), ((1, 0), s)), complex(-t)))
        add_hopping_term(tb, ((((0, 0), s), ((0, 1), s)), complex(-t)))
    return Real_Space_Second_Quantized_Model(
        params={"t": t, "U": U}, lattice=lat, tb_model=tb,
        particle_statistics=Particle_Statistics.FERMIONIC,
        bilinear_terms=bilinear, density_density_terms=density,
    )


# ---------------------------------------------------------------------------
# Bosonic FCI on the Haldane honeycomb
# ---------------------------------------------------------------------------

def build_haldane_tb_model(sample_size: list[int], params: dict[str, Any]):
    t, t1, t2 = params["t"], params["t'"], params["t''"]
    phi = 2 * math.pi * params["phi_over_2pi"]
    lat = initialize_real_space_lattice(
        sample_size=sample_size,
        brav_vec_list=[[1.0, 0.0], [1 / 2, math.sqrt(3) / 2]],
        sub_crys_list=[[0.0, 0.0], [1 / 3, 1 / 3]],
        lattice_name="Haldane_Honeycomb", pbc_indicator=[True, True],
    )
    tb = initialize_real_space_tightbinding_model(lat, model_name="haldane_boson_FCI")
    e = cmath.exp
    for c_to in ((0, 0), (0, -1), (-1, 0)):
        add_hopping_term(tb, ((((0, 0), 1), (c_to, 2)), -t))
    add_hopping_term(tb, ((((0, 0), 1), ((1, 0), 1)), -t1 * e(1j * phi)))
    add_hopping_term(tb, ((((0, 0), 1), ((0, 1), 1)), -t1 * e(-1j * phi)))
    add_hopping_term(tb, ((((0, 0), 1), ((-1, 1), 1)), -t1 * e(1j * phi)))
    add_hopping_term(tb, ((((0, 0), 2), ((1, 0), 2)), -t1 * e(-1j * phi)))
    add_hopping_term(tb, ((((0, 0), 2), ((0, 1), 2)), -t1 * e(1j * phi)))
    add_hopping_term(tb, ((((0, 0), 2), ((-1, 1), 2)), -t1 * e(-1j * phi)))
    add_hopping_term(tb, ((((0, 0), 2), ((1, 1), 1)), -t2))
    add_hopping_term(tb, ((((1, 0), 1), ((0, 1), 2)), -t2))
    add_hopping_term(tb, ((((1, 0), 2), ((0, 1), 1)), -t2))
    return tb


def build_zero_flux_bosonic_fci_second_quantized_model(
    *, sample_size: list[int] | None = None, params: dict[str, Any] | None = None,
) -> Real_Space_Second_Quantized_Model:
    sample_size = [3, 4] if sample_size is None else list(sample_size)
    params = dict(params_DNSheng) if params is None else dict(params)
    tb = build_haldane_tb_model(sample_size, params)
    lat = tb.lattice
    bilinear = generate_bilinear_terms(tb, twisted_phases_over_2π=[0.0] * lat.dim)
    V1, V2 = params["V1"], params["V2"]
    density: list[tuple[int, int, complex]] = []
    seen: set[tuple[int, int]] = set()

    def add_pair(i: int, j: int, V: float) -> None:
        if i == j:
            return
        pair = (min(i, j), max(i, j))
        if pair in seen:
            return
        seen.add(pair)
        density.append((pair[0], pair[1], complex(V)))

    L = lat.sample_size
    if V1 != 0.0:
        for (cell, isub) in lat.site_list:
            if isub != 1:
                continue
            i = lat.site_to_index_map[(cell, 1)]
            for shift in ((0, 0), (0, -1), (-1, 0)):
                cto = tuple((c + s) % Ld for c, s, Ld in zip(cell, shift, L))
                add_pair(i, lat.site_to_index_map[(cto, 2)], V1)
    if V2 != 0.0:
        for (cell, isub) in lat.site_list:
            for shift in ((1, 0), (0, 1), (-1, 1)):
                cto = tuple((c + s) % Ld for c, s, Ld in zip(cell, shift, L))
                add_pair(lat.site_to_index_map[(cell, isub)], lat.site_to_index_map[(cto, isub)], V2)
    return Real_Space_Second_Quantized_Model(
        params=dict(params), lattice=lat, tb_model=tb,
        particle_statistics=Particle_Statistics.BOSONIC,
        bilinear_terms=bilinear, density_density_terms=density,
    )


# ---------------------------------------------------------------------------
# Fermionic FCI on the checkerboard lattice
# ---------------------------------------------------------------------------

def _checkerboard_tb(sample_size: list[int], p: dict[str, Any]):
    lat = initialize_real_space_lattice(
        lattice_name="checkerboard", sample_size=sample_size,
        brav_vec_list=[[1.0, 0.0], [0.0, 1.0]], sub_crys_list=[[0.5, 0], [0.0, 0.5]],
        pbc_indicator=[True, True],
    )
    tb = initialize_real_space_tightbinding_model(lat, model_name="checkerboard")
    t, t1_1, t1_2, t2 = p["t"], p["t'_1"], p["t'_2"], p["t''"]
    phi = 2 * math.pi * p["phi_over_2pi"]
    e = cmath.exp
    add_hopping_term(tb, ((((0, 0), 1), ((0, 0), 2)), -t * e(-1j * phi)))
    add_hopping_term(tb, ((((0, 0), 1), ((1, 0), 2)), -t * e(1j * phi)))
    add_hopping_term(tb, ((((0, 0), 2), ((0, 1), 1)), -t * e(-1j * phi)))
    add_hopping_term(tb, ((((1, 0), 2), ((0, 1), 1)), -t * e(1j * phi)))
    add_hopping_term(tb, ((((0, 0), 1), ((1, 0), 1)), -t1_1))
    add_hopping_term(tb, ((((0, 0), 1), ((0, 1), 1)), -t1_2))
    add_hopping_term(tb, ((((0, 0), 2), ((1, 0), 2)), -t1_2))
    add_hopping_term(tb, ((((0, 0), 2), ((0, 1), 2)), -t1_1))
    add_hopping_term(tb, ((((0, 0), 1), ((1, 1), 1)), -t2))
    add_hopping_term(tb, ((((0, 0), 2), ((1, 1), 2)), -t2))
    add_hopping_term(tb, ((((1, 0), 2), ((0, 1), 2)), -t2))
    add_hopping_term(tb, ((((0, 1), 1), ((1, 0), 1)), -t2))
    return tb


def build_fermionic_checkerboard_tb_model(
    *, sample_size: list[int] | None = None, params: dict[str, Any] | None = None,
    flip_bands: bool = True,
):
    sample_size = [3, 4] if sample_size is None else list(sample_size)
    p = dict(params_Sun_Gu_Katsura_Sarma if params is None else params)
    if flip_bands:
        for key in ("t", "t'_1", "t'_2", "t''"):
            p[key] = -p[key]
    return _checkerboard_tb(sample_size, p)


def _checkerboard_density(lat, V1: float, V2: float, V3: float):
    L1, L2 = lat.sample_size
    out: list[tuple[int, int, complex]] = []
    for i_from, (cf, sf) in enumerate(lat.site_list, start=1):
        for i_to, (ct, st) in enumerate(lat.site_list, start=1):
            if i_from == i_to:
                continue
            sh = tuple((b - a) % L for a, b, L in zip(cf, ct, lat.sample_size))
            if V1 != 0.0 and sf == 1 and st == 2 and sh in ((0, 0), (1, 0), (0, L2 - 1), (1, L2 - 1)):
                out.append((i_from, i_to, complex(V1)))
            if V2 != 0.0 and sf == st and sh in ((1, 0), (0, 1)):
                out.append((i_from, i_to, complex(V2)))
            if V3 != 0.0 and sf == st and sh in ((1, 1), (1, L2 - 1)):
                out.append((i_from, i_to, complex(V3)))
    return out


def build_zero_flux_fermionic_fci_second_quantized_model(
    *, sample_size: list[int] | None = None, params: dict[str, Any] | None = None,
) -> Real_Space_Second_Quantized_Model:
    sample_size = [3, 4] if sample_size is None else list(sample_size)
    p = dict(params_Sun_Gu_Katsura_Sarma if params is None else params)
    tb = build_fermionic_checkerboard_tb_model(sample_size=sample_size, params=p)
    lat = tb.lattice
    bilinear = generate_bilinear_terms(tb, twisted_phases_over_2π=[0.0] * lat.dim)
    density = _checkerboard_density(lat, p["V1"], p["V2"], p["V3"])
    return Real_Space_Second_Quantized_Model(
        params=p, lattice=lat, tb_model=tb, particle_statistics=Particle_Statistics.FERMIONIC,
        bilinear_terms=bilinear, density_density_terms=density,
    )


def build_checkerboard_second_quantized_model(
    sample_size: list[int], params: dict[str, Any],
) -> Real_Space_Second_Quantized_Model:
    """Checkerboard fermions with pre-signed hoppings and ``λ``-scaled
    interactions (the model half of ``build_phase_explore_fermionic_checkerboard_model``)."""
    p = dict(params)
    tb = _checkerboard_tb(list(sample_size), p)
    lat = tb.lattice
    bilinear = generate_bilinear_terms(tb, twisted_phases_over_2π=[0.0] * lat.dim)
    lam = p.get("λ", 1.0)
    density = _checkerboard_density(lat, p["V1"] * lam, p["V2"] * lam, p["V3"] * lam)
    return Real_Space_Second_Quantized_Model(
        params=p, lattice=lat, tb_model=tb, particle_statistics=Particle_Statistics.FERMIONIC,
        bilinear_terms=bilinear, density_density_terms=density,
    )


# ---------------------------------------------------------------------------
# ED-data helpers (translation group + build_ed_data)
# ---------------------------------------------------------------------------

def build_heisenberg_ed_data(N: int, J: float = 1.0):
    from .symmetry_resolved_ed import build_ed_data, build_translation_group

    model = build_heisenberg_second_quantized_model(N, J)
    G = build_translation_group(model.lattice)
    return build_ed_data(model, filling_fraction=Fraction(N // 2, model.lattice.n_site), symmetry_group=G)


def build_spinful_hubbard_ed_data(sample_size: list[int], t: float = 1.0, U: float = 8.0):
    from .symmetry_resolved_ed import build_ed_data, build_translation_group

    model = build_spinful_hubbard_second_quantized_model(sample_size, t, U)
    G = build_translation_group(model.lattice)
    n_filled = sample_size[0] * sample_size[1]
    return build_ed_data(model, filling_fraction=Fraction(n_filled, model.lattice.n_site), symmetry_group=G)


def build_bosonic_fci_ed_data(
    *, sample_size: list[int] | None = None, params: dict[str, Any] | None = None,
    filling_fraction_per_band=Fraction(1, 2),
):
    from .symmetry_resolved_ed import build_ed_data, build_translation_group

    model = build_zero_flux_bosonic_fci_second_quantized_model(sample_size=sample_size, params=params)
    ff = Fraction(filling_fraction_per_band) / model.lattice.n_sub
    return build_ed_data(model, filling_fraction=ff, symmetry_group=build_translation_group(model.lattice))


def build_phase_explore_fermionic_checkerboard_model(
    params: dict[str, Any], sample_size: list[int], filling_fraction_per_band,
):
    from .symmetry_resolved_ed import build_ed_data, build_translation_group

    model = build_checkerboard_second_quantized_model(sample_size, params)
    lat = model.lattice
    ff = Fraction(filling_fraction_per_band) / lat.n_sub
    n_filled = int(lat.n_site * ff)
    ed_data = build_ed_data(model, filling_fraction=ff, symmetry_group=build_translation_group(lat))
    return model, ed_data, n_filled, ff


__all__ = [
    "params_DNSheng", "params_Sun_Gu_Katsura_Sarma", "my_optimal_param",
    "CHECKERBOARD_TPP_DENOMINATOR",
    "build_heisenberg_second_quantized_model", "build_spinful_hubbard_second_quantized_model",
    "build_haldane_tb_model", "build_zero_flux_bosonic_fci_second_quantized_model",
    "build_fermionic_checkerboard_tb_model", "build_zero_flux_fermionic_fci_second_quantized_model",
    "build_checkerboard_second_quantized_model",
    "build_heisenberg_ed_data", "build_spinful_hubbard_ed_data", "build_bosonic_fci_ed_data",
    "build_phase_explore_fermionic_checkerboard_model",
]
