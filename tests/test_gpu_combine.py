"""GPU: combinations of basis vectors (the eigensolver's restart step).

``edx_combine_vectors`` on device buffers against a float64 numpy product
of the same inputs:

* vector schedule (FP64 sums): every component is the exact sum rounded
  once to the vector precision;
* matrix-core schedule (complex64 on the FP32 matrix cores): every
  component within 24 FP32 roundings of the sum of its terms' magnitudes
  (coefficients rounded to FP32, FP32 sums of 8 products per step, steps
  added in FP64, one final rounding), and the whole result
  within 5e-7 relative in the Frobenius norm;
* the automatic schedule equals the matrix-core result bit for bit on
  complex64 shapes of more than 16 outputs from 65536 rows, the vector
  result otherwise;
* rows past ``dim`` of the output are never written; ragged last strips,
  several column groups (more than 64 outputs) and every tile count are
  covered; bad arguments raise.

And one solve per schedule (mixed precision, restarts of 33 outputs) gives
the same eigenvalues, its combinations timed in that schedule's class.
"""

from __future__ import annotations

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import make_case
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list

pytestmark = pytest.mark.gpu

U32 = 2.0**-24
SENTINEL = np.complex64(3.0 - 7.0j)


def _run(ctx, prec, V, ld, dim, Y, schedule):
    """Combination of V's columns (host array [nv][ld]) by Y on the device;
    returns the output [nk][ld] (rows past dim keep the sentinel)."""
    nv, nk = Y.shape
    dt = np.complex128 if prec == ed.PREC_FP64 else np.complex64
    vin = _edx.DeviceArray(ctx, V.nbytes)
    vin.upload(V)
    out0 = np.full((nk, ld), SENTINEL, dtype=dt)
    vout = _edx.DeviceArray(ctx, out0.nbytes)
    vout.upload(out0)
    _edx.combine_vectors(ctx, prec, vin, ld, dim, Y, vout, schedule)
    return vout.download(np.empty((nk, ld), dtype=dt))


def _inputs(rng, dim, ld, nv, nk, dt):
    V = np.empty((nv, ld), dtype=dt)
    V[:, :dim] = (rng.standard_normal((nv, dim)) + 1j * rng.standard_normal((nv, dim))) / np.sqrt(2 * dim)
    V[:, dim:] = 5.0 + 2.0j  # padding rows: read by whole strips, never combined into rows < dim
    Y = (rng.standard_normal((nv, nk)) + 1j * rng.standard_normal((nv, nk))) / np.sqrt(nv)
    return V, Y


SHAPES = [  # dim, nv, nk
    (40, 5, 17),          # two strips, the second ragged; 3 tiles
    (1000, 13, 21),       # small: the automatic schedule stays on vector FMAs
    (70001, 104, 57),     # restart shape of a 104-vector basis; ragged last strip
    (65573, 33, 70),      # two column groups
    (131072, 60, 33),     # 5 tiles
    (65600, 7, 120),      # nv not a multiple of 4; two full groups
    (80000, 204, 20),     # largest basis the staging takes at <= 32 outputs
]


@pytest.mark.parametrize("dim,nv,nk", SHAPES)
def test_combine_complex64_schedules(dim, nv, nk):
    ctx = _edx.current_context()
    rng = np.random.default_rng(dim + 7 * nv + nk)
    ld = (dim + 64) // 64 * 64  # > dim, as the eigensolver lays out its vectors
    V, Y = _inputs(rng, dim, ld, nv, nk, np.complex64)
    Vd = V[:, :dim].astype(np.complex128)
    ref = Y.T @ Vd  # [nk][dim], float64
    terms = np.abs(Y.T) @ np.abs(Vd)  # sum over j of |y_jk| |v_j[r]|
    outs = {}
    for sched in (1, 2, 0):
        out = _run(ctx, ed.PREC_MIXED, V, ld, dim, Y, sched)
        assert np.all(out[:, dim:] == SENTINEL), f"schedule {sched} wrote rows past dim"
        outs[sched] = out[:, :dim]
    vec, mat = outs[1].astype(np.complex128), outs[2].astype(np.complex128)
    for part in (np.real, np.imag):
        # one rounding of the FP64 sum
        err = np.abs(part(vec) - part(ref))
        assert np.all(err <= U32 * np.abs(part(ref)) + 1e-13 * terms), float(err.max())
        err = np.abs(part(mat) - part(ref))
        assert np.all(err <= 24 * U32 * terms), float((err / terms).max())
    rel = np.linalg.norm(mat - ref) / np.linalg.norm(ref)
    print(f"[combine] dim={dim} nv={nv} nk={nk} matrix-core rel err {rel:.2e}", flush=True)
    assert rel <= 5e-7
    auto_matrix = dim >= 65536 and nk > 16
    assert np.array_equal(outs[0], outs[2] if auto_matrix else outs[1])


@pytest.mark.parametrize("dim,nv,nk", [(1000, 13, 21), (70001, 40, 57)])
def test_combine_complex128(dim, nv, nk):
    ctx = _edx.current_context()
    rng = np.random.default_rng(dim + nv)
    ld = (dim + 64) // 64 * 64  # > dim, as the eigensolver lays out its vectors
    V, Y = _inputs(rng, dim, ld, nv, nk, np.complex128)
    ref = Y.T @ V[:, :dim]
    terms = np.abs(Y.T) @ np.abs(V[:, :dim])
    for sched in (0, 1):
        out = _run(ctx, ed.PREC_FP64, V, ld, dim, Y, sched)
        assert np.all(out[:, dim:] == SENTINEL)
        assert np.all(np.abs(out[:, :dim] - ref) <= 1e-14 * terms)


def test_combine_arguments_checked():
    ctx = _edx.current_context()
    dim, ld, nv, nk = 1000, 1024, 8, 20
    rng = np.random.default_rng(3)
    V, Y = _inputs(rng, dim, ld, nv, nk, np.complex64)
    vin = _edx.DeviceArray(ctx, V.nbytes)
    vin.upload(V)
    out = _edx.DeviceArray(ctx, nk * ld * 16)
    v64 = _edx.DeviceArray(ctx, nv * ld * 16)
    with pytest.raises(ValueError):  # matrix cores take complex64 only
        _edx.combine_vectors(ctx, ed.PREC_FP64, v64, ld, dim, Y[:, :4], out, 2)
    with pytest.raises(ValueError):  # unknown schedule
        _edx.combine_vectors(ctx, ed.PREC_MIXED, vin, ld, dim, Y, out, 3)
    with pytest.raises(ValueError):  # ld must exceed dim
        _edx.combine_vectors(ctx, ed.PREC_MIXED, vin, dim, dim, Y, out, 1)
    with pytest.raises(ValueError):  # ld below dim rounded up to 32
        _edx.combine_vectors(ctx, ed.PREC_MIXED, vin, 1001, 1000, Y, out, 2)
    with pytest.raises(ValueError):  # basis too large for the coefficient staging
        big = np.zeros((120, 40), dtype=np.complex128)
        vbig = _edx.DeviceArray(ctx, 120 * ld * 8)
        _edx.combine_vectors(ctx, ed.PREC_MIXED, vbig, ld, dim, big, out, 2)
    with pytest.raises(ValueError):  # output overlapping the input
        _edx.combine_vectors(ctx, ed.PREC_MIXED, vin, ld, dim, Y[:, :4], vin, 1)
    with pytest.raises(ValueError):
        _edx.set_combine_schedule(ctx, 3)


def test_solve_same_levels_under_both_combine_schedules():
    model, ne, G = make_case("hub24")
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[1])
    H = ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None)
    ctx = _edx.current_context()
    levels = {}
    try:
        for sched in (1, 2):
            _edx.set_combine_schedule(ctx, sched)
            _edx.kernel_stats(ctx, enable=1, reset=True)
            w, _ = ed.diagonalize_block_matrixfree(H, H.n, nev=6, precision=ed.PREC_MIXED, want_vectors=False)
            ks = _edx.kernel_stats(ctx, enable=0)
            ms, n, _, _ = ks["combine" if sched == 1 else "combine_mfma"]
            print(f"[combine solve] D={H.n} schedule {sched}: {n} combinations, {ms:.2f} ms, matvecs "
                  f"{H.last_info[0]:.0f}", flush=True)
            assert n > 0  # the complex64 restarts went through the combination of that schedule
            if sched == 1:
                assert ks["combine_mfma"][1] == 0  # (complex128 ones always run on vector FMAs)
            levels[sched] = np.asarray(w)
    finally:
        _edx.set_combine_schedule(ctx, 0)
    assert np.max(np.abs(levels[1] - levels[2])) <= 1e-10
