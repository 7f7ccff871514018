"""MI355X-native symmetry-resolved exact diagonalization.

Drop-in for the ED engine API of ``realspace_exactdiagonalization_py``
(bit primitives, model container, groups/irreps, orbit catalog, irrep
filter, CanonicalMap, projection-table Hamiltonian, sector eigensolver and
scans), computed by the HIP library ``libedx.so`` on the GPU.  Importing the
package requires the library (``python -m
realspace_exatdiagonalization_py_amd.csrc.build``); device calls require a
visible GPU and raise otherwise.
"""

from .bitwise_operations import (
    COMPLEX_ONE, MAX_N_SITE_UINT64, Mask, bitmask_of_site, decode_bit_mask_to_configuration,
    decode_bit_mask_to_configuration_inplace, empty_site_for_mask, empty_site_iter_for_mask,
    encode_configuration_to_bit_mask, filled_site_iter_for_mask, is_site_empty, is_site_occupied,
    n_occupied_for_mask, occupy_site_for_mask,
)
from .second_quantized_model import (
    Particle_Statistics, Real_Space_Second_Quantized_Model, update_second_quantized_model_with_twisted_phases,
)
from .lattice import (
    add_hopping_term, generate_bilinear_terms, initialize_real_space_lattice,
    initialize_real_space_tightbinding_model,
)
from .symmetry_resolved_ed import (
    CanonicalMap, Finite_Symmetry_Group, MatrixFreeHamiltonian, OneDim_Irrep, Symmetry_Operation,
    Symmetry_Orbit_Catalog, Symmetry_Resolved_ED_Data, Symmetry_Sector_Basis, apply_hamiltonian,
    apply_hamiltonian_with_threads, apply_operation_to_mask, build_ed_Hamiltonian_symmetry_block, build_ed_data,
    build_identity_group, build_identity_irrep_list, build_irrep_list, build_matrixfree_projection_table,
    build_symmetry_orbit_catalog, build_symmetry_sector_basis, build_translation_group,
    build_translation_irrep_list, diagonalize_block_arpack, diagonalize_block_dense, diagonalize_block_matrixfree,
    diagonalize_blocks_batched, solve_sectors_batched,
    ed_scan, ed_scan_at_irrep_matrix, ed_scan_at_irrep_matrixfree, ed_scan_checkpoint_filename, full_ed,
    get_canonical, get_canonical_representative, group_order, hamiltonian_linear_operator, hopping_phase_for_stats,
    load_checkpoint, populate_canonical_map, print_spectrum, project_to_sector, project_to_unnormalized_sector,
    rank_combination, run_concurrently, save_checkpoint, update_orbit_stabilizer_phases,
)
from . import models, observables
from .observables import (
    expand_sector_state_to_fock, many_body_position_phases, off_diagonal_long_range_order, one_body_density_matrix,
)
from .models import (
    CHECKERBOARD_TPP_DENOMINATOR, build_bosonic_fci_ed_data, build_fermionic_checkerboard_tb_model,
    build_heisenberg_ed_data, build_heisenberg_second_quantized_model, build_phase_explore_fermionic_checkerboard_model,
    build_spinful_hubbard_ed_data, build_spinful_hubbard_second_quantized_model,
    build_zero_flux_bosonic_fci_second_quantized_model, build_zero_flux_fermionic_fci_second_quantized_model,
    params_DNSheng, params_Sun_Gu_Katsura_Sarma,
)
from ._edx import PREC_FP64, PREC_MIXED

__version__ = "0.1.0"
