"""GPU: grouped sector products and their captured HIP graph.

Six sector Hamiltonians of uneven dimension (Hubbard 2x4 sectors, Heisenberg
16, Haldane bosons 3x4 with a stored table and one on the fly; D = 810 …
11 240) are applied to device-resident blocks of 4 vectors in complex128 and
complex64:

* ``edx_matvec_group`` (one grouped launch for the stored tables) equals, bit
  for bit, each table's own SpMM (``edx_time_matvec`` with one repetition,
  the eigensolver's operator);
* the graph captured by ``edx_matvec_group_graph`` is replayed four times,
  with other device work between the replays and new inputs before two of
  them: every replay equals the eager grouped product of the current inputs
  bit for bit, and the outputs are overwritten each time (cleared in
  between).
"""

from __future__ import annotations

import numpy as np
import pytest

import realspace_exatdiagonalization_py_amd as ed
from cases import make_case
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd._edx import lib
from realspace_exatdiagonalization_py_amd.groups import build_irrep_list

pytestmark = pytest.mark.gpu

NB = 4


def _operator(name, sector, table="stored"):
    model, ne, G = make_case(name)
    cat = ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=ne, symmetry_group=G,
                                          particle_statistics=model.particle_statistics)
    basis = ed.build_symmetry_sector_basis(cat, build_irrep_list(G, model.lattice)[sector])
    return ed.MatrixFreeHamiltonian(basis, model.bilinear_terms, model.density_density_terms, None, table=table)


@pytest.fixture(scope="module")
def operators():
    ops = [_operator("hub24", s) for s in range(3)]
    ops += [_operator("heis16", 1), _operator("hal34", 1), _operator("hal34", 2, table="on_the_fly")]
    assert ops[-1].on_the_fly
    return ops


def _inputs(ops, dt, seed):
    rng = np.random.default_rng(seed)
    return [np.ascontiguousarray((rng.standard_normal((H.n, NB)) + 1j * rng.standard_normal((H.n, NB))).astype(dt))
            for H in ops]


def _single(ctx, H, prec, x_host, dt):
    dx = _edx.DeviceArray(ctx, x_host.nbytes)
    dy = _edx.DeviceArray(ctx, x_host.nbytes)
    dx.upload(x_host)
    ms = _edx.C.c_double()
    ctx.check(lib.edx_time_matvec(ctx.handle, H._dev.handle, prec, NB, 1, dx.ptr, dy.ptr, _edx.C.byref(ms)))
    return dy.download(np.empty((NB, H.n), dtype=dt))


@pytest.mark.parametrize("prec,dt", [(_edx.PREC_FP64, np.complex128), (_edx.PREC_MIXED, np.complex64)])
def test_grouped_products_and_graph_replays(operators, prec, dt):
    ctx = _edx.current_context()
    hams = [H._dev for H in operators]
    xs_host = _inputs(operators, dt, 11)
    dxs = [_edx.DeviceArray(ctx, x.nbytes) for x in xs_host]
    dys = [_edx.DeviceArray(ctx, x.nbytes) for x in xs_host]
    for d, x in zip(dxs, xs_host):
        d.upload(x)
    _edx.matvec_group(ctx, hams, prec, NB, dxs, dys)
    eager = [d.download(np.empty((NB, H.n), dtype=dt)) for d, H in zip(dys, operators)]
    for H, x, y in zip(operators, xs_host, eager):
        np.testing.assert_array_equal(y, _single(ctx, H, prec, x, dt))
    g = _edx.MatvecGroupGraph(ctx, hams, prec, NB, dxs, dys)
    zeros = [np.zeros((NB, H.n), dtype=dt) for H in operators]
    for replay in range(4):
        if replay in (1, 3):  # new inputs: the graph reads the buffers' current contents
            xs_host = _inputs(operators, dt, 100 + replay)
            for d, x in zip(dxs, xs_host):
                d.upload(x)
            _edx.matvec_group(ctx, hams, prec, NB, dxs, dys)
            eager = [d.download(np.empty((NB, H.n), dtype=dt)) for d, H in zip(dys, operators)]
        for d, z in zip(dys, zeros):
            d.upload(z)
        _single(ctx, operators[0], prec, xs_host[0], dt)  # other device work between replays
        g.launch()
        g.synchronize()
        got = [d.download(np.empty((NB, H.n), dtype=dt)) for d, H in zip(dys, operators)]
        for y, e in zip(got, eager):
            np.testing.assert_array_equal(y, e)
        print(f"[graph] prec={prec} replay {replay}: {len(got)} sector products equal the eager ones", flush=True)


def test_group_arguments_checked(operators):
    ctx = _edx.current_context()
    d = _edx.DeviceArray(ctx, 64)
    with pytest.raises(ValueError, match="vectors per product"):
        _edx.matvec_group(ctx, [operators[0]._dev], _edx.PREC_FP64, 17, [d], [d])
    with pytest.raises(ValueError, match="precision"):
        _edx.matvec_group(ctx, [operators[0]._dev], 7, 1, [d], [d])


def test_group_with_an_empty_sector():
    """A grouped product over an empty sector and a one-state sector."""
    ctx = _edx.current_context()
    m0, _ne, G0 = make_case("hal23_ne0")
    cat0 = ed.build_symmetry_orbit_catalog(second_quantized_model=m0, n_filled=0, symmetry_group=G0,
                                           particle_statistics=m0.particle_statistics)
    irreps = build_irrep_list(G0, m0.lattice)
    ops = [ed.MatrixFreeHamiltonian(ed.build_symmetry_sector_basis(cat0, ir), m0.bilinear_terms,
                                    m0.density_density_terms, None) for ir in irreps[:2]]
    assert sorted(H.n for H in ops) == [0, 1]
    xs = [_edx.DeviceArray(ctx, max(H.n, 1) * 16 * NB) for H in ops]
    ys = [_edx.DeviceArray(ctx, max(H.n, 1) * 16 * NB) for H in ops]
    for x in xs:
        x.upload(np.ones(NB, dtype=np.complex128))
    _edx.matvec_group(ctx, [H._dev for H in ops], _edx.PREC_FP64, NB, xs, ys)
    g = _edx.MatvecGroupGraph(ctx, [H._dev for H in ops], _edx.PREC_FP64, NB, xs, ys)
    g.launch()
    g.synchronize()
    for H, y in zip(ops, ys):
        if H.n:
            got = y.download(np.empty((NB, 1), dtype=np.complex128))
            np.testing.assert_array_equal(got, np.zeros((NB, 1)))  # N = 0: the empty state, H = 0
