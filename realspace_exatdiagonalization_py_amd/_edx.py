"""ctypes binding of ``libedx.so`` (declared in ``include/edx.h``).

The library is the only compute path of this package: importing this module
raises ``ImportError`` if the library was not built, and every call that
needs a device raises ``RuntimeError`` when none is visible.  Status codes map
to the Python exceptions the reference raises for the same conditions
(``ValueError`` / ``AssertionError``), plus ``MemoryError`` and
``RuntimeError`` for device failures.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
import weakref
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("EDX_LIBRARY", Path(__file__).resolve().parent / "libedx.so"))

if not LIB_PATH.exists():
    raise ImportError(
        f"edx HIP library not found at {LIB_PATH}; build it with "
        "`python -m realspace_exatdiagonalization_py_amd.csrc.build` (hipcc, gfx950)"
    )

lib = C.CDLL(str(LIB_PATH))

OK = 0
ERR_INVALID, ERR_ASSERT, ERR_HIP, ERR_OOM, ERR_NOCONV, ERR_INTERNAL, ERR_BUSY = -1, -2, -3, -4, -5, -6, -7
PREC_FP64, PREC_MIXED = 0, 1

P = C.c_void_p
i32, i64, u64, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
PP = C.POINTER(C.c_void_p)
pi32 = C.POINTER(C.c_int32)
pi64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "edx_ctx_create": (C.c_int, [i32, PP]),
    "edx_ctx_destroy": (C.c_int, [P]),
    "edx_last_error": (C.c_char_p, [P]),
    "edx_device_count": (i32, []),
    "edx_synchronize": (C.c_int, [P]),
    "edx_catalog_build": (C.c_int, [P, i32, i32, i32, P, P, P, i32, PP]),
    "edx_catalog_upload": (C.c_int, [P, i32, i32, i32, P, P, P, i32, i64, i64, P, P, P, P, P, P, P, P, PP]),
    "edx_catalog_clone": (C.c_int, [P, P, PP]),
    "edx_catalog_info": (C.c_int, [P, pi64, pi64, pi64]),
    "edx_catalog_fetch": (C.c_int, [P, P, P, P, P, P, P, P, P, P]),
    "edx_catalog_update_phases": (C.c_int, [P, P, P, P, P]),
    "edx_catalog_rephase": (C.c_int, [P, P, P, P, P]),
    "edx_catalog_lookup": (C.c_int, [P, P, i64, P, P, P, P]),
    "edx_catalog_destroy": (C.c_int, [P]),
    "edx_sector_build": (C.c_int, [P, P, P, f64, PP, pi64]),
    "edx_sector_fetch": (C.c_int, [P, P, P, P, P]),
    "edx_sector_destroy": (C.c_int, [P]),
    "edx_hamiltonian_build": (C.c_int, [P, P, P, P, i32, P, P, P, i32, P, P, P, PP, pi64]),
    "edx_hamiltonian_build_on_the_fly": (C.c_int, [P, P, P, P, i32, P, P, P, i32, P, P, P, PP, pi64]),
    "edx_hamiltonian_fetch_table": (C.c_int, [P, P, P, P, P, P, P]),
    "edx_hamiltonian_fetch_csr": (C.c_int, [P, P, i64, i64, P, P, P, P]),
    "edx_hamiltonian_from_csr": (C.c_int, [P, i64, i64, P, P, P, P, PP]),
    "edx_matvec": (C.c_int, [P, P, i32, P, P]),
    "edx_matvec_device": (C.c_int, [P, P, i32, i32, P, P]),
    "edx_matvec_noalloc": (C.c_int, [P, P, i32, P, P]),
    "edx_hamiltonian_destroy": (C.c_int, [P]),
    "edx_eigs": (C.c_int, [P, P, i32, i32, f64, u64, i32, P, P, P]),
    "edx_set_spmv_group": (C.c_int, [P, i32]),
    "edx_hamiltonian_set_spmv_group": (C.c_int, [P, P, i32]),
    "edx_hamiltonian_info": (C.c_int, [P, pi64, pi64, pi32]),
    "edx_set_table_packing": (C.c_int, [P, i32]),
    "edx_hamiltonian_layout": (C.c_int, [P, pi32, pi32, pi64, pi64]),
    "edx_set_autotune": (C.c_int, [i32, i32, i32]),
    "edx_autotune_table": (C.c_int, [i32, P, P, pi32]),
    "edx_autotune_clear": (C.c_int, []),
    "edx_schedule_table": (C.c_int, [i32, P, P, pi32]),
    "edx_eigs_batch": (C.c_int, [P, i32, P, P, P, P, P, P]),
    "edx_batch_graph_stats": (C.c_int, [P, pi64, pi64]),
    "edx_matvec_group": (C.c_int, [P, i32, P, i32, i32, P, P]),
    "edx_matvec_group_graph": (C.c_int, [P, i32, P, i32, i32, P, P, PP]),
    "edx_graph_launch": (C.c_int, [P, P]),
    "edx_graph_destroy": (C.c_int, [P]),
    "edx_set_solver_schedule": (C.c_int, [P, i32, i32]),
    "edx_combine_vectors": (C.c_int, [P, i32, P, C.c_int64, C.c_int64, i32, P, i32, P, i32]),
    "edx_set_combine_schedule": (C.c_int, [P, i32]),
    "edx_hamiltonian_autotune": (C.c_int, [P, P, i32, i32, pi32, P]),
    "edx_host_herm_eig": (C.c_int, [i32, P, i32, P, P]),
    "edx_device_alloc": (C.c_int, [P, u64, PP]),
    "edx_device_free": (C.c_int, [P, P]),
    "edx_memcpy_h2d": (C.c_int, [P, P, P, u64]),
    "edx_memcpy_d2h": (C.c_int, [P, P, P, u64]),
    "edx_device_copy": (C.c_int, [P, P, P, u64]),
    "edx_mem_info": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "edx_time_matvec": (C.c_int, [P, P, i32, i32, i32, P, P, C.POINTER(C.c_double)]),
    "edx_kernel_stats": (C.c_int, [P, i32, i32, P, P, P]),
    "edx_kernel_flops": (C.c_int, [P, P]),
    "edx_time_stream_read": (C.c_int, [P, P, u64, i32, P]),
    "edx_eigs_ex": (C.c_int, [P, P, P, P, P, P]),
    "edx_eigs_default_options": (None, [P]),
    "edx_fock_expand": (C.c_int, [P, P, P, P, P, f64, i32, i64, P, P, pi64]),
    "edx_position_operator": (C.c_int, [P, P, P, P, P, P, P, P, P]),
    "edx_one_body_density": (C.c_int, [P, P, P, P, P, P]),
}


class EigsOptions(C.Structure):
    """Mirror of ``edx_eigs_options`` (include/edx.h)."""

    _fields_ = [("nev", C.c_int32), ("block", C.c_int32), ("basis", C.c_int32), ("precision", C.c_int32),
                ("max_restarts", C.c_int32), ("tol", C.c_double), ("eig_tol", C.c_double),
                ("mixed_tol", C.c_double), ("seed", C.c_uint64)]

for _name, (_res, _args) in SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def eigs_options(**kw) -> EigsOptions:
    o = EigsOptions()
    lib.edx_eigs_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def ptr(a: np.ndarray | None):
    """Data pointer of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def raise_for(rc: int, ctx=None) -> None:
    if rc == OK:
        return
    # a refused concurrent call reports through the calling thread's own text
    msg = lib.edx_last_error(None if rc == ERR_BUSY else ctx)
    text = msg.decode() if msg else f"edx error {rc}"
    if rc == ERR_INVALID:
        raise ValueError(text)
    if rc == ERR_ASSERT:
        raise AssertionError(text)
    if rc == ERR_OOM:
        raise MemoryError(text)
    raise RuntimeError(text)


class Context:
    """One device context (HIP stream).  ``Context.default()`` is shared per process."""

    _default: "Context | None" = None
    _lock = threading.Lock()
    _live: "weakref.WeakSet[Context]" = weakref.WeakSet()
    _live_lock = threading.Lock()

    def __init__(self, device: int = 0):
        n = lib.edx_device_count()
        if n <= 0:
            raise RuntimeError("edx: no HIP device is visible; the ED path runs only on the GPU")
        h = C.c_void_p()
        raise_for(lib.edx_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device
        with Context._live_lock:
            Context._live.add(self)

    @classmethod
    def default(cls) -> "Context":
        with cls._lock:
            if cls._default is None:
                cls._default = Context(int(os.environ.get("EDX_DEVICE", "0")))
            return cls._default

    def check(self, rc: int) -> None:
        raise_for(rc, self.handle)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            lib.edx_ctx_destroy(h)
            self.handle = None


_local = threading.local()


class use_context:
    """Routes the device calls made by this thread to ``ctx`` (its own HIP
    stream), e.g. to solve several sectors concurrently from worker threads::

        with use_context(Context(0)):
            ed.ed_scan_at_irrep_matrixfree(label, data)
    """

    def __init__(self, ctx: "Context"):
        self.ctx = ctx

    def __enter__(self):
        self.prev = getattr(_local, "ctx", None)
        _local.ctx = self.ctx
        return self.ctx

    def __exit__(self, *exc):
        _local.ctx = self.prev
        return False


def current_context(default: "Context | None" = None) -> "Context":
    """The thread's context set by ``use_context``, else ``default``, else the process default."""
    ctx = getattr(_local, "ctx", None)
    if ctx is not None:
        return ctx
    return default if default is not None else Context.default()


KERNEL_CLASSES = ("spmm_c128", "spmm_c64", "gram", "gemm", "combine", "combine_mfma")


def kernel_stats(ctx: "Context", enable: int = -1, reset: bool = False) -> dict:
    """Totals of the live kernel timer: {class: (ms, launches, bytes, flops)}
    (algorithmic bytes and FP64 / matrix-core flops of the timed launches)."""
    k = len(KERNEL_CLASSES)
    ms = np.zeros(k)
    n = np.zeros(k, dtype=np.int64)
    b = np.zeros(k)
    f = np.zeros(k)
    ctx.check(lib.edx_kernel_stats(ctx.handle, enable, 1 if reset else 0, ptr(ms), ptr(n), ptr(b)))
    ctx.check(lib.edx_kernel_flops(ctx.handle, ptr(f)))  # after a reset: the cleared totals too
    return {k: (float(ms[i]), int(n[i]), float(b[i]), float(f[i])) for i, k in enumerate(KERNEL_CLASSES)}


def batch_graph_stats(ctx: "Context") -> tuple[int, int]:
    """Round graphs of the batched solves driven from ``ctx`` so far:
    (captured, replayed)."""
    cap, rep = C.c_int64(), C.c_int64()
    ctx.check(lib.edx_batch_graph_stats(ctx.handle, C.byref(cap), C.byref(rep)))
    return int(cap.value), int(rep.value)


def set_autotune(enable: bool = True, precision: int = PREC_MIXED, block: int = 6) -> None:
    """Turns the SpMM autotuner on or off for Hamiltonian builds of this
    process: each new build's lanes per row come from the tuning table for
    (precision, block width, size, row length), timed on a miss."""
    raise_for(lib.edx_set_autotune(1 if enable else 0, int(precision), int(block)))


def autotune_table() -> list[dict]:
    """Entries of the tuning table: key, chosen lanes and the average SpMM
    launch time (ms) of the 1-, 2- and 4-lane layouts."""
    n = C.c_int32()
    raise_for(lib.edx_autotune_table(0, None, None, C.byref(n)))
    cap = n.value
    keys = np.zeros(5 * max(cap, 1), dtype=np.int32)
    ms = np.zeros(3 * max(cap, 1), dtype=np.float32)
    raise_for(lib.edx_autotune_table(cap, ptr(keys), ptr(ms), C.byref(n)))
    out = []
    for i in range(min(cap, n.value)):
        k = keys[5 * i:5 * i + 5]
        out.append({"precision": int(k[0]), "block": int(k[1]), "dim_log2": int(k[2]),
                    "row_quarter_octaves": int(k[3]), "lanes": int(k[4]),
                    "ms": {1: float(ms[3 * i]), 2: float(ms[3 * i + 1]), 4: float(ms[3 * i + 2])}})
    return out


def autotune_clear() -> None:
    raise_for(lib.edx_autotune_clear())


GRAM_SCHEDULES = {1: "vector", 2: "matrix"}
UPDATE_VARIANTS = {1: (2, 2), 2: (2, 1), 3: (4, 1)}  # (vectors per prefetch stage, rows per lane)


def schedule_table() -> list[dict]:
    """Entries of the eigensolver schedule table: key, kept Gram schedule
    (1 vector FMAs, 2 FP32 matrix cores) and update variant (1..3), and the
    fastest timed launch (ms) of each candidate (0.0: not timed)."""
    n = C.c_int32()
    raise_for(lib.edx_schedule_table(0, None, None, C.byref(n)))
    cap = n.value
    keys = np.zeros(6 * max(cap, 1), dtype=np.int32)
    ms = np.zeros(5 * max(cap, 1), dtype=np.float32)
    raise_for(lib.edx_schedule_table(cap, ptr(keys), ptr(ms), C.byref(n)))
    out = []
    for i in range(min(cap, n.value)):
        k, m = keys[6 * i:6 * i + 6], ms[5 * i:5 * i + 5]
        out.append({"precision": int(k[0]), "block": int(k[1]), "dim_log2": int(k[2]), "basis": int(k[3]),
                    "gram": int(k[4]), "update": int(k[5]),
                    "gram_ms": {1: float(m[0]), 2: float(m[1])},
                    "update_ms": {1: float(m[2]), 2: float(m[3]), 3: float(m[4])}})
    return out


def set_solver_schedule(ctx: "Context", gram: int = 0, update: int = 0) -> None:
    """Fixes the eigensolver schedule of solves on ``ctx`` (0: tuned or default)."""
    ctx.check(lib.edx_set_solver_schedule(ctx.handle, int(gram), int(update)))


def set_combine_schedule(ctx: "Context", schedule: int = 0) -> None:
    """Fixes how solves on ``ctx`` form combinations of more than 16 vectors
    (restarts, Ritz vectors): 1 FP64 vector FMAs, 2 FP32 matrix cores
    (complex64), 0 automatic."""
    ctx.check(lib.edx_set_combine_schedule(ctx.handle, int(schedule)))


def combine_vectors(ctx: "Context", precision: int, v: "DeviceArray", ld: int, dim: int, y, out: "DeviceArray",
                    schedule: int = 0) -> None:
    """out[:, k] = sum_j v[:, j] y[j, k] on device buffers (column-major,
    leading dimension ``ld``; complex128 for PREC_FP64, complex64 for
    PREC_MIXED); ``y`` is an (nv, nk) complex array on the host."""
    y = np.ascontiguousarray(y, dtype=np.complex128)
    if y.ndim != 2:
        raise ValueError("y must have shape (nv, nk)")
    nv, nk = y.shape
    esz = 16 if precision == PREC_FP64 else 8
    if nv * ld * esz > v.nbytes or nk * ld * esz > out.nbytes:
        raise ValueError("buffer smaller than its vectors")
    ctx.check(lib.edx_combine_vectors(ctx.handle, int(precision), v.ptr, int(ld), int(dim), int(nv),
                                      y.ctypes.data_as(C.c_void_p), int(nk), out.ptr, int(schedule)))


def set_table_packing(ctx: "Context", enable: bool = True) -> None:
    """Palette-packed sliced tables for Hamiltonians built on ``ctx`` from now on."""
    ctx.check(lib.edx_set_table_packing(ctx.handle, 1 if enable else 0))


def hamiltonian_layout(handle) -> dict:
    """Layout of a built Hamiltonian's sliced copy: ``packed`` (0 unpacked,
    1 32-bit words, 2 columns and 16-bit palette indices), palette size,
    stored slots (with padding) and device bytes."""
    packed, n_pal = C.c_int32(), C.c_int32()
    slots, nbytes = C.c_int64(), C.c_int64()
    raise_for(lib.edx_hamiltonian_layout(handle, C.byref(packed), C.byref(n_pal), C.byref(slots), C.byref(nbytes)))
    return {"packed": packed.value, "n_palette": n_pal.value, "stored_slots": slots.value,
            "table_bytes": nbytes.value}


def mem_info(ctx: "Context") -> tuple[int, int]:
    """(free, total) bytes of device memory."""
    f, t = C.c_uint64(), C.c_uint64()
    ctx.check(lib.edx_mem_info(ctx.handle, C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def kernel_stats_all(enable: int = -1, reset: bool = False) -> dict:
    """``kernel_stats`` summed over every live context (all streams of the
    process).  Call it while no solve is running: a context in use by
    another thread raises ``RuntimeError`` (``EDX_ERR_BUSY``) rather than
    being read while its owner writes it."""
    with Context._live_lock:
        ctxs = list(Context._live)
    tot = {k: [0.0, 0, 0.0, 0.0] for k in KERNEL_CLASSES}
    for c in ctxs:
        for k, v in kernel_stats(c, enable, reset).items():
            for m in range(4):
                tot[k][m] += v[m]
    return {k: tuple(v) for k, v in tot.items()}


class DeviceArray:
    """Raw device allocation (used by benchmarks/tests for HBM-resident vectors)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.check(lib.edx_device_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.ctx.check(lib.edx_memcpy_h2d(self.ctx.handle, self.ptr, ptr(a), a.nbytes))

    def download(self, out: np.ndarray) -> np.ndarray:
        assert out.flags["C_CONTIGUOUS"] and out.nbytes <= self.nbytes
        self.ctx.check(lib.edx_memcpy_d2h(self.ctx.handle, ptr(out), self.ptr, out.nbytes))
        return out

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            lib.edx_device_free(self.ctx.handle, p)
            self.ptr = None


def _group_args(hams, xs, ys):
    n = len(hams)
    assert n == len(xs) == len(ys) and n >= 1
    return (n, (C.c_void_p * n)(*[h.handle for h in hams]), (C.c_void_p * n)(*[x.ptr.value for x in xs]),
            (C.c_void_p * n)(*[y.ptr.value for y in ys]))


def matvec_group(ctx: "Context", hams, precision: int, n_vec: int, xs, ys) -> None:
    """``ys[i] = H_i xs[i]`` for device Hamiltonians ``hams`` (``_DeviceHamiltonian``)
    and ``DeviceArray`` buffers, grouped launches; waits."""
    n, h, x, y = _group_args(hams, xs, ys)
    ctx.check(lib.edx_matvec_group(ctx.handle, n, h, int(precision), int(n_vec), x, y))


class MatvecGroupGraph:
    """The grouped products of ``matvec_group`` captured once into a HIP graph;
    ``launch()`` enqueues a replay on ``ctx``'s stream (``synchronize()``
    waits).  Hamiltonians and buffers are kept alive with the graph."""

    def __init__(self, ctx: "Context", hams, precision: int, n_vec: int, xs, ys):
        self.ctx = ctx
        self._keep = (list(hams), list(xs), list(ys))
        n, h, x, y = _group_args(hams, xs, ys)
        g = C.c_void_p()
        ctx.check(lib.edx_matvec_group_graph(ctx.handle, n, h, int(precision), int(n_vec), x, y, C.byref(g)))
        self.handle = g

    def launch(self) -> None:
        self.ctx.check(lib.edx_graph_launch(self.ctx.handle, self.handle))

    def synchronize(self) -> None:
        self.ctx.check(lib.edx_synchronize(self.ctx.handle))

    def __del__(self):
        if getattr(self, "handle", None):
            lib.edx_graph_destroy(self.handle)
            self.handle = None


def host_herm_eig(a: np.ndarray, upper: bool = True):
    """Hermitian eigen-decomposition by the library's host solver (no device)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    n = a.shape[0]
    w = np.zeros(n)
    z = np.zeros((n, n), dtype=np.complex128)
    raise_for(lib.edx_host_herm_eig(n, ptr(a), 1 if upper else 0, ptr(w), ptr(z)))
    return w, z


def exported_symbols() -> list[str]:
    return sorted(SIGNATURES)
