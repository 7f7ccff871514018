"""GPU: the measurement entry points the benchmark relies on.

* ``edx_kernel_stats`` / ``edx_kernel_flops``: after a timed solve every
  eigensolver class has launches, time, algorithmic bytes and (Gram, block
  update, combinations) flops, with flops >= 8 D per launch; a
  reset clears them; with the timer off nothing is counted.
* ``edx_time_stream_read``: the streaming-read reference returns a rate
  between 1 TB/s and the 8 TB/s peak and refuses sizes that are not whole
  16-byte words.
"""

from __future__ import annotations

import ctypes as C

import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import make_case
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list

pytestmark = pytest.mark.gpu


def _hamiltonian():
    model, ne, G = make_case("hub24")
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[1])
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    H._keep = cat
    return H


def test_kernel_timer_classes_and_flops():
    ctx = _edx.current_context()
    H = _hamiltonian()
    _edx.kernel_stats(ctx, enable=0, reset=True)  # cleared, off
    ed.diagonalize_block_matrixfree(H, H.n, nev=6, precision=ed.PREC_MIXED, want_vectors=False)
    off = _edx.kernel_stats(ctx)
    assert all(v == (0.0, 0, 0.0, 0.0) for v in off.values()), off  # nothing counted while off
    _edx.kernel_stats(ctx, enable=1, reset=True)
    ed.diagonalize_block_matrixfree(H, H.n, nev=6, precision=ed.PREC_MIXED, want_vectors=False)
    ks = _edx.kernel_stats(ctx, enable=0)
    print(f"[timer] D={H.n} {ks}", flush=True)
    for cls in ("spmm_c64", "spmm_c128", "gram", "gemm"):
        ms, n, nbytes, flops = ks[cls]
        assert n > 0 and ms > 0 and nbytes > 0, (cls, ks[cls])
        if cls.startswith("spmm"):
            assert flops == 0
        else:
            # 8 flops per complex multiply-add over at least one vector pair
            # per row: flops >= 8 D per launch, and per byte moved at most
            # 8 x (vectors per launch) ~ a few hundred
            assert flops >= 8 * H.n * n and flops <= 1000 * nbytes, (cls, ks[cls])
    for cls in ("combine", "combine_mfma"):
        ms, n, nbytes, flops = ks[cls]
        if n:  # restarts of this solve went through this combination class
            assert flops > 0 and nbytes > 0
    assert _edx.kernel_stats(ctx) == ks  # off: unchanged by reading
    cleared = _edx.kernel_stats(ctx, reset=True)  # a reset clears before reading
    assert all(v == (0.0, 0, 0.0, 0.0) for v in cleared.values()), cleared


def test_stream_read_reference():
    ctx = _edx.current_context()
    n = 1 << 30
    a = _edx.DeviceArray(ctx, n)  # contents do not matter: only read
    ms = C.c_double()
    ctx.check(_edx.lib.edx_time_stream_read(ctx.handle, a.ptr, n, 3, C.byref(ms)))
    gbps = n / ms.value / 1e6
    print(f"[stream] {gbps:.0f} GB/s", flush=True)
    assert 1000.0 < gbps < 8500.0
    with pytest.raises(ValueError):
        ctx.check(_edx.lib.edx_time_stream_read(ctx.handle, a.ptr, n - 8, 3, C.byref(ms)))
