"""Serving loop for sector solves on one GPU (continuous batching).

Requests name a model, a filling, a sector label, a twist and how many
eigenpairs they need.  They are admitted as they arrive, whenever a worker is
free and the request's estimated device-memory footprint fits the budget that
is not yet committed to requests in flight.  Each worker owns a HIP stream, so
admitted solves overlap on the device.  Representative tables are built once
per (lattice, statistics, filling, group, twist) and kept on the device (an
LRU cache with a byte budget), as are the tables and Krylov state of each solve
while it runs.

There is no reference counterpart: the reference solves sectors one at a time
in ``ed_scan`` (``SRE:2462-2522``).  A solve here is the same computation as
``ed_scan_at_irrep_matrixfree`` (``SRE:2402-2459``) on the device, so a result
does not depend on what else is in flight (same start vectors, same arithmetic
per stream).
"""

from __future__ import annotations

import collections
import math
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

from . import _edx
from .groups import build_identity_group, build_irrep_list, build_translation_group, group_order
from .second_quantized_model import Particle_Statistics, Real_Space_Second_Quantized_Model


@dataclass(frozen=True, eq=False)
class SolveRequest:
    """One sector solve.  ``sector`` is an irrep label of the translation group
    (``(k1, k2)``) or ``"identity"``; ``twist`` is ``twisted_phases_over_2π``
    (``None``: no twist).  The model must not be mutated while in flight."""

    model: Real_Space_Second_Quantized_Model
    n_filled: int
    sector: Any
    nev: int = 5
    twist: tuple | None = None
    precision: str = "mixed"
    eig_tol: float = 1e-12
    want_vectors: bool = False
    tag: Any = None


@dataclass
class SolveResult:
    request: SolveRequest
    eigenvalues: np.ndarray
    eigenvectors: np.ndarray | None
    dim: int
    submitted: float
    started: float
    finished: float
    footprint_bytes: int
    info: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.submitted

    @property
    def queue_wait(self) -> float:
        return self.started - self.submitted

    @property
    def service(self) -> float:
        return self.finished - self.started


def estimate_footprint(n_site: int, n_filled: int, n_group: int, n_bilinear: int, nev: int) -> int:
    """Device bytes of one sector solve, estimated before its tables exist:
    sector dimension about C(N, Ne)/|G|, stored entries about
    D * T * p(1-p) N/(N-1) with p = Ne/N (a hop needs an occupied source and
    an empty target), projection table in CSR (22 B/entry) and sliced copy
    (28 B/entry, 20 % padding), 8 B/entry of sort keys during the build, and
    two Krylov banks of 5(nev+b)+2b complex128 vectors."""
    C = math.comb(n_site, n_filled)
    D = max(1, C // max(n_group, 1))
    p = n_filled / n_site if n_site else 0.0
    E = int(D * n_bilinear * p * (1.0 - p) * n_site / max(n_site - 1, 1))
    b = min(nev, 8)
    nvec = 5 * (nev + b) + 2 * b
    return int(E * (22 + 1.2 * 28 + 8) + D * (2 * nvec * 16 + 64))


@dataclass
class _CatalogEntry:
    group: Any
    catalog: Any
    irreps: dict
    nbytes: int
    lock: threading.Lock = field(default_factory=threading.Lock)


class SectorSolveServer:
    """Admits :class:`SolveRequest` s as they arrive and runs them on
    ``workers`` streams.  ``submit`` returns a ``concurrent.futures.Future``
    resolving to a :class:`SolveResult` (or raising the solve's exception).

    hbm_budget: bytes that admitted requests may occupy together (default:
    80 % of the free device memory at start).  A request larger than the
    whole budget runs alone.

    policy: ``"fifo"`` serves in arrival order; requests behind the head may
    start first when they fit and the head does not, at most ``max_skip``
    times for any one head request.  ``"sjf"`` serves the admissible request
    of least estimated footprint (which grows with the sector dimension, so
    roughly the shortest job) first; every older request it passes ages by
    one, and a request passed ``max_skip`` times goes next (waiting for room
    if it must), so none starves.
    """

    def __init__(self, workers: int = 8, *, hbm_budget: int | None = None, catalog_cache_bytes: int | None = None,
                 max_skip: int = 8, policy: str = "fifo", solve_fn: Callable | None = None,
                 context_factory: Callable[[], Any] | None = None,
                 estimate_fn: Callable[[SolveRequest], int] | None = None):
        if workers < 1:
            raise ValueError("workers must be >= 1")
        if policy not in ("fifo", "sjf"):
            raise ValueError("policy must be 'fifo' or 'sjf'")
        self._policy = policy
        self._cv = threading.Condition()
        self._pending: collections.deque = collections.deque()
        self._in_use = 0
        self._running = 0
        self._closed = False
        self._max_skip = max_skip
        self._solve_fn = solve_fn or self._solve_on_device
        self._estimate_fn = estimate_fn
        factory = context_factory if context_factory is not None else self._new_context
        self._contexts = [factory() for _ in range(workers)]
        if hbm_budget is None:
            free, _ = _edx.mem_info(self._contexts[0])
            hbm_budget = int(0.8 * free)
        self._budget = int(hbm_budget)
        self._catalog_budget = int(catalog_cache_bytes if catalog_cache_bytes is not None else self._budget // 4)
        self._catalogs: collections.OrderedDict = collections.OrderedDict()
        self._cat_lock = threading.Lock()
        self._terms: collections.OrderedDict = collections.OrderedDict()  # twisted term lists, LRU
        self._max_term_sets = 256
        self._done: list[SolveResult] = []
        self._peak_in_use = 0
        self._t0 = time.perf_counter()
        # solve_alone runs on the caller's thread: its own context, so that it
        # never shares a worker's stream, staging buffers or kernel timer
        self._alone_context = factory()
        self._alone_lock = threading.Lock()
        self._threads = [threading.Thread(target=self._worker, args=(c,), daemon=True) for c in self._contexts]
        for t in self._threads:
            t.start()

    @staticmethod
    def _new_context():
        return _edx.Context(_edx.Context.default().device)

    # ---- public API ------------------------------------------------------------
    def submit(self, req: SolveRequest) -> Future:
        if req.precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        fut: Future = Future()
        est = self._estimate(req)
        with self._cv:
            if self._closed:
                raise RuntimeError("server is shut down")
            self._pending.append([req, fut, time.perf_counter(), est, 0])
            self._cv.notify_all()
        return fut

    def shutdown(self, wait: bool = True) -> None:
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            for t in self._threads:
                t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.shutdown(wait=True)
        return False

    def stats(self) -> dict:
        with self._cv:
            done = list(self._done)
        if not done:
            return {"completed": 0}
        lat = np.array([r.latency for r in done]) * 1e3
        t_first = min(r.submitted for r in done)
        t_last = max(r.finished for r in done)
        return {
            "completed": len(done),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "mean_queue_ms": float(np.mean([r.queue_wait for r in done]) * 1e3),
            "solves_per_s": len(done) / max(t_last - t_first, 1e-9),
            "peak_admitted_bytes": self._peak_in_use,
            "hbm_budget_bytes": self._budget,
        }

    # ---- scheduling ------------------------------------------------------------
    def _fits(self, ent) -> bool:
        return self._in_use + ent[3] <= self._budget or self._running == 0

    def _take(self):
        """Next admissible entry (called with the lock held), or None."""
        if not self._pending:
            return None
        if self._policy == "sjf":
            return self._take_sjf()
        head = self._pending[0]
        limit = 1 if head[4] >= self._max_skip else len(self._pending)
        for idx in range(min(limit, len(self._pending))):
            ent = self._pending[idx]
            if self._fits(ent):
                del self._pending[idx]
                if idx > 0:
                    head[4] += 1
                return ent
        return None

    def _take_sjf(self):
        for idx, ent in enumerate(self._pending):  # a request passed max_skip times goes next
            if ent[4] >= self._max_skip:
                if not self._fits(ent):
                    return None
                del self._pending[idx]
                return ent
        best = None
        for idx, ent in enumerate(self._pending):
            if self._fits(ent) and (best is None or ent[3] < self._pending[best][3]):
                best = idx
        if best is None:
            return None
        for idx in range(best):  # the older requests passed over
            self._pending[idx][4] += 1
        ent = self._pending[best]
        del self._pending[best]
        return ent

    def _worker(self, ctx) -> None:
        while True:
            with self._cv:
                ent = None
                while ent is None:
                    ent = self._take()
                    if ent is None:
                        if self._closed and not self._pending:
                            return
                        self._cv.wait()
                self._in_use += ent[3]
                self._running += 1
                self._peak_in_use = max(self._peak_in_use, self._in_use)
            req, fut, t_sub, est, _ = ent
            t_start = time.perf_counter()
            try:
                if fut.set_running_or_notify_cancel():
                    vals, vecs, dim, info = self._solve_fn(self, req, ctx)
                    res = SolveResult(req, vals, vecs, dim, t_sub, t_start, time.perf_counter(), est, info)
                    with self._cv:
                        self._done.append(res)
                    fut.set_result(res)
            except BaseException as e:  # noqa: BLE001 - handed to the caller's future
                fut.set_exception(e)
            finally:
                with self._cv:
                    self._in_use -= est
                    self._running -= 1
                    self._cv.notify_all()

    # ---- model preparation -----------------------------------------------------
    @staticmethod
    def _group_key(req: SolveRequest):
        lat = req.model.lattice
        twist = tuple(float(x) for x in req.twist) if req.twist is not None else None
        kind = "identity" if req.sector == "identity" else "translation"
        return (getattr(lat, "lattice_name", ""), tuple(lat.sample_size), lat.n_site, len(lat.sub_crys_list),
                tuple(getattr(lat, "pbc_indicator", ())), req.model.particle_statistics, int(req.n_filled), kind,
                twist)

    def _estimate(self, req: SolveRequest) -> int:
        if self._estimate_fn is not None:
            return int(self._estimate_fn(req))
        lat = req.model.lattice
        n_g = 1 if req.sector == "identity" else int(np.prod(lat.sample_size))
        return estimate_footprint(lat.n_site, int(req.n_filled), n_g, len(req.model.bilinear_terms), int(req.nev))

    def _bilinear_terms(self, req: SolveRequest):
        if req.twist is None or not any(req.twist):
            return req.model.bilinear_terms
        # the entry holds the model itself, so its id cannot be reused by
        # another model while the entry exists
        key = (id(req.model), tuple(float(x) for x in req.twist))
        with self._cat_lock:
            hit = self._terms.get(key)
            if hit is not None and hit[0] is req.model:
                self._terms.move_to_end(key)
                return hit[1]
        from .lattice import generate_bilinear_terms

        if req.model.tb_model is None:
            raise ValueError("a twist needs the model's tight-binding description (tb_model)")
        terms = generate_bilinear_terms(req.model.tb_model, twisted_phases_over_2π=list(req.twist))
        with self._cat_lock:
            self._terms[key] = (req.model, terms)
            self._terms.move_to_end(key)
            while len(self._terms) > self._max_term_sets:
                self._terms.popitem(last=False)
        return terms

    def _catalog(self, req: SolveRequest) -> _CatalogEntry:
        from .symmetry_resolved_ed import build_symmetry_orbit_catalog

        key = self._group_key(req)
        with self._cat_lock:
            ent = self._catalogs.get(key)
            if ent is None:
                ent = _CatalogEntry(None, None, {}, 0)
                self._catalogs[key] = ent
            self._catalogs.move_to_end(key)
        with ent.lock:
            if ent.catalog is None:
                lat = req.model.lattice
                if req.sector == "identity":
                    G = build_identity_group(lat.n_site)
                else:
                    twist = list(req.twist) if req.twist is not None else [0.0] * lat.dim
                    G = build_translation_group(lat, twist)
                cat = build_symmetry_orbit_catalog(second_quantized_model=req.model, n_filled=int(req.n_filled),
                                                   symmetry_group=G,
                                                   particle_statistics=req.model.particle_statistics)
                ent.irreps = {ir.label: ir for ir in build_irrep_list(G, lat)}
                ent.group, ent.catalog = G, cat
                ent.nbytes = 16 * cat.n_total + 24 * cat.n_orbits
                self._evict(keep=key)
        return ent

    def _evict(self, keep) -> None:
        with self._cat_lock:
            total = sum(e.nbytes for e in self._catalogs.values())
            for k in list(self._catalogs):
                if total <= self._catalog_budget:
                    break
                if k == keep:
                    continue
                total -= self._catalogs.pop(k).nbytes  # solves in flight keep their catalog alive

    # ---- the solve ---------------------------------------------------------------
    @staticmethod
    def _solve_on_device(server: "SectorSolveServer", req: SolveRequest, ctx):
        from .symmetry_resolved_ed import (MatrixFreeHamiltonian, _device_eigs, build_symmetry_sector_basis)

        with _edx.use_context(ctx):
            ent = server._catalog(req)
            label = "identity" if req.sector == "identity" else tuple(int(x) for x in req.sector)
            if req.sector == "identity":
                label = next(iter(ent.irreps))
            if label not in ent.irreps:
                raise ValueError(f"Sector {req.sector!r} not found in the irrep list.")
            basis = build_symmetry_sector_basis(ent.catalog, ent.irreps[label])
            if basis.dim == 0:
                return np.zeros(0), (np.zeros((0, 0), np.complex128) if req.want_vectors else None), 0, None
            H = MatrixFreeHamiltonian(basis, server._bilinear_terms(req), req.model.density_density_terms, None)
            prec = _edx.PREC_MIXED if req.precision == "mixed" else _edx.PREC_FP64
            vals, vecs, info = _device_eigs(H._dev, int(req.nev), precision=prec, eig_tol=float(req.eig_tol),
                                            want_vectors=req.want_vectors)
            return vals, vecs, basis.dim, info

    def solve_alone(self, req: SolveRequest) -> SolveResult:
        """The same solve on the calling thread, outside the queue (reference
        point for checking that results do not depend on the mix)."""
        with self._alone_lock:
            t = time.perf_counter()
            vals, vecs, dim, info = self._solve_fn(self, req, self._alone_context)
        return SolveResult(req, vals, vecs, dim, t, t, time.perf_counter(), self._estimate(req), info)


def poisson_arrivals(n: int, rate_per_s: float, seed: int = 0) -> np.ndarray:
    """Arrival times (s) of a seeded Poisson stream."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))


def serve_stream(server: SectorSolveServer, requests, arrivals) -> list[SolveResult]:
    """Submits ``requests`` at the given arrival times (s from now) and waits
    for every result."""
    t0 = time.perf_counter()
    futs = []
    for req, t in zip(requests, arrivals):
        delay = t0 + float(t) - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        futs.append(server.submit(req))
    return [f.result() for f in futs]


__all__ = ["SolveRequest", "SolveResult", "SectorSolveServer", "estimate_footprint", "poisson_arrivals",
           "serve_stream"]
