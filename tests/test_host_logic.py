"""Host-side logic that needs no device: context routing, the concurrent
sector runner and the ctypes binding's declarations against ``include/edx.h``."""

from __future__ import annotations

import re
import threading
import time
from pathlib import Path

import pytest

import realspace_exatdiagonalization_py_amd as ed
from realspace_exatdiagonalization_py_amd import _edx
from realspace_exatdiagonalization_py_amd import symmetry_resolved_ed as sre

ROOT = Path(__file__).resolve().parents[1]


class _FakeLib:
    """Stands in for the library's context calls (no device in this process)."""

    def __init__(self):
        self.created = 0
        self.synced = []
        self.handles = set()

    def edx_device_count(self):
        return 1

    def edx_ctx_create(self, device, out):
        self.created += 1
        out._obj.value = 1000 + self.created
        self.handles.add(out._obj.value)
        return 0

    def edx_ctx_destroy(self, h):
        return 0

    def edx_synchronize(self, h):
        self.synced.append(h)
        return 0

    def edx_last_error(self, h):
        return b"fake"


@pytest.fixture
def fake_lib(monkeypatch):
    fake = _FakeLib()
    monkeypatch.setattr(_edx, "lib", fake)
    monkeypatch.setattr(sre, "lib", fake)
    monkeypatch.setattr(_edx.Context, "_default", None)
    monkeypatch.setattr(sre, "_WORKER_CONTEXTS", {})
    yield fake
    # contexts made here hold fake handles: keep the real library from destroying them
    for c in list(_edx.Context._live):
        if c.handle is not None and c.handle.value in fake.handles:
            c.handle = None


def test_default_context_and_routing(fake_lib):
    c0 = _edx.Context.default()
    assert _edx.Context.default() is c0
    assert _edx.current_context() is c0
    c1 = _edx.Context(0)
    with _edx.use_context(c1):
        assert _edx.current_context() is c1
        assert _edx.current_context(c0) is c1
        seen = []
        t = threading.Thread(target=lambda: seen.append(_edx.current_context()))
        t.start()
        t.join()
        assert seen == [c0]  # routing is per thread
    assert _edx.current_context() is c0
    assert _edx.current_context(c1) is c1


def test_run_concurrently_uses_one_context_per_worker(fake_lib):
    seen = {}
    lock = threading.Lock()

    def job(i):
        time.sleep(0.01)
        with lock:
            seen[i] = (threading.get_ident(), _edx.current_context().handle.value)

    ed.run_concurrently(range(12), job, 4)
    assert sorted(seen) == list(range(12))
    by_thread = {}
    for tid, h in seen.values():
        by_thread.setdefault(tid, set()).add(h)
    assert 1 < len(by_thread) <= 4
    assert all(len(hs) == 1 for hs in by_thread.values())  # a thread keeps its stream
    handles = [h for hs in by_thread.values() for h in hs]
    assert len(set(handles)) == len(handles)  # streams are not shared between threads
    assert len(fake_lib.synced) >= 2  # every worker stream is drained at the end
    # contexts are reused by the next scan
    created = fake_lib.created
    ed.run_concurrently(range(8), job, 4)
    assert fake_lib.created == created


def test_run_concurrently_reraises_and_stops(fake_lib):
    done = []

    def job(i):
        if i == 3:
            raise ValueError("sector 3")
        time.sleep(0.005)
        done.append(i)

    with pytest.raises(ValueError, match="sector 3"):
        ed.run_concurrently(range(40), job, 3)
    assert len(done) < 39
    with pytest.raises(KeyError):
        ed.run_concurrently([1], lambda i: {}[i], 1)


def _header_functions():
    text = (ROOT / "include" / "edx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return set(re.findall(r"\b(edx_[a-z0-9_]+)\s*\(", text))


def test_binding_declares_every_header_function():
    declared = _header_functions()
    assert declared == set(_edx.SIGNATURES), (declared ^ set(_edx.SIGNATURES))


def test_library_exports_every_header_function():
    import ctypes

    lib = ctypes.CDLL(str(_edx.LIB_PATH))
    missing = [n for n in sorted(_header_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_hermitian_part_of_a_block():
    """Matrix mode hands the device Lanczos (H + H^H) / 2 when a block is not
    Hermitian beyond FP64 rounding, and the block itself otherwise."""
    import numpy as np
    import scipy.sparse as sp

    rng = np.random.default_rng(3)
    A = sp.random(40, 40, density=0.2, random_state=4, dtype=np.float64) * (1 + 0.5j)
    H = (A + A.getH()).tocsr()
    np.testing.assert_array_equal(sre._hermitian_part(H).toarray(), H.toarray())  # Hermitian: unchanged
    E = H.copy()
    E.data = E.data * (1 + 1e-8 * rng.standard_normal(E.data.shape))  # rounding-sized asymmetry
    Hh = sre._hermitian_part(E)
    np.testing.assert_allclose(Hh.toarray(), (E.toarray() + E.toarray().conj().T) / 2, rtol=0, atol=1e-15)
    assert abs(Hh - Hh.getH()).max() == 0.0


def test_more_than_63_sites_is_refused_before_any_device_work():
    """The 64-bit mask limit (the reference's N <= 63 check, SRE:1328-1332)
    raises ValueError before the library is asked for anything."""
    model = ed.build_heisenberg_second_quantized_model(64)
    G = ed.build_translation_group(model.lattice)
    with pytest.raises(ValueError, match="64-bit"):
        ed.build_symmetry_orbit_catalog(second_quantized_model=model, n_filled=32, symmetry_group=G,
                                        particle_statistics=model.particle_statistics)


def test_table_mode_host_rules():
    """The on-the-fly product's catalog condition (every site phase exactly +1
    or -1, as the library checks) and the table-mode argument, on the host."""
    from cases import make_case

    for name, exact in (("hub24", True), ("heis16", True), ("hal34", True), ("hub24_twist", False),
                        ("cb34_twist", False)):
        _m, _ne, G = make_case(name)
        assert sre._phases_exact(G) is exact, name
    with pytest.raises(ValueError, match="table must be one of"):
        sre._build_device_hamiltonian(None, [], [], None, table="dense")
