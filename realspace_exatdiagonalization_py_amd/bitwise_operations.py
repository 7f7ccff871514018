"""Occupation-number bitmask primitives.

Site ``i`` (1-based) is bit ``i - 1`` of an unsigned 64-bit mask, so a
configuration ``|n_1 .. n_N>`` is ``m = sum_i n_i 2^(i-1)``.  The public API
takes and returns Python ints; device kernels use ``uint64``.

Mirrors ``utils/bitwise_operations.py`` of the reference
(``bitmask_of_site`` :37, ``occupy_site_for_mask`` :50, ``empty_site_for_mask``
:63, ``is_site_occupied`` :76, ``is_site_empty`` :89, ``n_occupied_for_mask``
:102, the iterators :114-151, encode/decode :154-219, ``MAX_N_SITE_UINT64``
:34).
"""

from __future__ import annotations

from collections.abc import Iterable, Iterator

import numpy as np

Mask = int
COMPLEX_ONE = 1.0 + 0.0j
#: Largest vertex count the 64-bit mask kernels accept.
MAX_N_SITE_UINT64 = 63


def bitmask_of_site(i: int) -> Mask:
    """Mask with only (1-based) site ``i`` occupied."""
    return 1 << (i - 1)


def occupy_site_for_mask(m: Mask, i: int) -> Mask:
    return m | bitmask_of_site(i)


def empty_site_for_mask(m: Mask, i: int) -> Mask:
    return m & ~bitmask_of_site(i)


def is_site_occupied(m: Mask, i: int) -> bool:
    return (m >> (i - 1)) & 1 == 1


def is_site_empty(m: Mask, i: int) -> bool:
    return (m >> (i - 1)) & 1 == 0


def n_occupied_for_mask(m: Mask) -> int:
    return int(m).bit_count()


def filled_site_iter_for_mask(m: Mask, n_site: int) -> Iterator[int]:
    """1-based occupied sites of ``m`` in ascending order (lazy)."""
    return (i for i in range(1, n_site + 1) if (m >> (i - 1)) & 1)


def empty_site_iter_for_mask(m: Mask, n_site: int) -> Iterator[int]:
    """1-based empty sites of ``m`` in ascending order (lazy)."""
    return (i for i in range(1, n_site + 1) if not (m >> (i - 1)) & 1)


def encode_configuration_to_bit_mask(occ: Iterable[int]) -> Mask:
    m = 0
    for i in occ:
        m |= 1 << (int(i) - 1)
    return m


def decode_bit_mask_to_configuration(m: Mask, n_site: int) -> list[int]:
    return list(filled_site_iter_for_mask(m, n_site))


def decode_bit_mask_to_configuration_inplace(occ: list[int], m: Mask, n_site: int) -> list[int]:
    """Overwrite ``occ`` with the occupied sites of ``m``; ``len(occ)`` must equal popcount."""
    assert len(occ) == n_occupied_for_mask(m), (
        "decode_bit_mask_to_configuration_inplace: length(occ) must equal "
        "the number of occupied sites in the mask."
    )
    for n, i in enumerate(filled_site_iter_for_mask(m, n_site)):
        occ[n] = i
    return occ


def as_uint64_array(mask_list) -> np.ndarray:
    return np.asarray(mask_list, dtype=np.uint64)


__all__ = [
    "Mask", "COMPLEX_ONE", "MAX_N_SITE_UINT64", "bitmask_of_site",
    "occupy_site_for_mask", "empty_site_for_mask", "is_site_occupied",
    "is_site_empty", "n_occupied_for_mask", "filled_site_iter_for_mask",
    "empty_site_iter_for_mask", "encode_configuration_to_bit_mask",
    "decode_bit_mask_to_configuration",
    "decode_bit_mask_to_configuration_inplace", "as_uint64_array",
]
