"""The serving form of the assembly path: assemble-and-apply requests with continuous
batching on one GPU.

A request names a mesh (``n^3`` hexahedra of the unit cube), an order (P1 or P2) and a seed;
the server draws the seeded parameter field ``m`` and direction ``x`` on the device,
assembles the forward Jacobian ``A`` of ``exp(m) grad u . grad p - f p`` (Dirichlet x = 0, 1,
folded) in the precision form (FP32 contraction, FP32 values) and returns ``y = A x`` to the
host.  The reference assembles one problem per call (``benchmarks/bench_assembly.py``); this
is the request loop around that call:

* meshes, patterns, elimination plans and the CSR index arrays of every (n, order) are built
  once and stay resident (``ModelCache``);
* each in-flight request owns a slot: a HIP stream of default priority and a set of device
  buffers (element matrices, values, m, x, y, its seed) taken from its model's pool;
* a request's whole work (seed to the device, m and x drawn from it, element matrices,
  scatter, y = A x, y to pinned host memory) is captured once per buffer set into a HIP
  graph; a request then costs the host a seed write, one replay and an event, instead of a
  launch per kernel (the small requests are bound by those launches);
* admission is size-aware: the waiting request with the earliest deadline (its arrival plus
  a multiple of its problem's device time) goes first, so small requests do not queue behind
  large ones, and the large problems hold at most a few slots at once; a request's buffers
  must fit the HBM budget beside those in flight;
* waiting requests of one small problem are batched into one launch: a batched model holds
  k disjoint copies of the mesh (concatenated element ranges and dofs), so the k requests'
  element matrices, scatter and product are one kernel each, in one graph replay;
* admitted requests run concurrently on their streams, so their kernels overlap on the
  device;
* a finished request's result leaves its pinned buffer for memory of its own, a large one
  on a copy thread (megabytes of host memory to touch, which would otherwise hold up the
  loop's admissions and completions), and its slot and buffers free once it has;
* latency is taken from a request's arrival to its result being the caller's.

Every kernel of a request is deterministic (no float atomics), and a copy of a batched model
computes what the single model does, so a request's result does not depend on what else is in
flight or batched with it: the same bits alone, in the mix and in a batch.
"""

import collections
import concurrent.futures
import gc
import time

import numpy as np
import torch

from . import _lib
from ._lib import ADJOINT, STATE
from .elementbatch import MeshBatches
from .problem import PDEVariationalProblem
from .workloads import Workload

SIZES = (8, 16, 24, 32, 48, 64)
#: results of at most this many values are copied on the loop itself (cheaper than a hand-off)
INLINE_RESULT = 1 << 16
#: the problems batched into one launch (n up to this), and the batch sizes kept
BATCH_MAX_N = 32
BATCHES = (8, 4, 2)


class Request:
    """One assemble-and-apply request; ``arrival`` in seconds from the start of the stream."""

    __slots__ = ("rid", "n", "order", "seed", "arrival", "admitted", "done", "y", "model",
                 "deadline", "batch")

    def __init__(self, rid, n, order, seed, arrival=0.0):
        self.rid, self.n, self.order, self.seed = int(rid), int(n), int(order), int(seed)
        self.arrival = float(arrival)
        self.admitted = self.done = self.deadline = None
        self.y = self.model = None
        self.batch = 1          # requests of the launch that served it

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def poisson_requests(count, rate, seed=0, sizes=SIZES):
    """``count`` requests arriving as a Poisson process of ``rate`` per second: n uniform
    over ``sizes``, order P1 or P2 with probability 0.5, one seed each."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    ns = rng.choice(np.asarray(sizes), count)
    orders = rng.integers(1, 3, count)
    seeds = rng.integers(0, 2 ** 31 - 1, count)
    return [Request(k, ns[k], orders[k], seeds[k], t[k]) for k in range(count)]


def _seed_tensor(seed, device):
    return torch.tensor([int(seed)], dtype=torch.int64, device=device)


def device_field(n, seed, device, passes=3):
    """The seeded parameter field on the ``(n+1)^3`` vertex lattice, drawn on the device
    (``hm_serving_field``): N(0, 1) noise, ``passes`` of the 7-point average over the
    neighbours that exist, then mean 0 and population std 1 (the field of the benchmarks,
    ``workloads.smoothed_noise``)."""
    n1 = int(n) + 1
    work = torch.empty(2 * n1 ** 3 + 2, dtype=torch.float64, device=device)
    out = torch.empty(n1 ** 3, dtype=torch.float64, device=device)
    sd = _seed_tensor(seed, device)
    _lib.check(_lib.lib().hm_serving_field(_lib.ptr(sd), n1, int(passes), _lib.ptr(work), None,
                                           _lib.ptr(out), _lib.stream_handle(None)), "field")
    return out


def device_direction(ndof, seed, device):
    """The seeded direction ``x``: ``ndof`` N(0, 1) values (``hm_serving_normal``, salt 1)."""
    out = torch.empty(int(ndof), dtype=torch.float64, device=device)
    sd = _seed_tensor(seed, device)
    _lib.check(_lib.lib().hm_serving_normal(_lib.ptr(sd), 1, int(ndof), _lib.ptr(out),
                                            _lib.stream_handle(None)), "direction")
    return out


class _TiledMesh:
    """``k`` disjoint copies of a mesh: the vertices repeated, copy ``c``'s elements on its own
    vertices (what ``ElementGroup`` reads)."""

    def __init__(self, mesh, k):
        nv = mesh.verts.shape[0]
        self.verts = np.ascontiguousarray(np.tile(mesh.verts, (k, 1)))
        self.elem_verts = np.ascontiguousarray(np.concatenate(
            [mesh.elem_verts + c * nv for c in range(k)]).astype(np.int32))
        self.ne = k * mesh.ne


class _TiledSpace:
    """A space on ``k`` disjoint copies of its mesh: copy ``c``'s dofs ``c ndof + d``."""

    def __init__(self, space, k):
        self.order, self.local_nodes, self.nd = space.order, space.local_nodes, space.nd
        self.vdim = 1
        self.ndof = k * space.ndof
        self.edofs = np.ascontiguousarray(np.concatenate(
            [space.edofs + c * space.ndof for c in range(k)]).astype(np.int64))


class _TiledBC:
    def __init__(self, bc, ndof, k):
        self.ess_tdof = np.concatenate([bc.ess_tdof + c * ndof for c in range(k)])
        self.ess = self.ess_tdof


class Buffers:
    """The device memory of one launch of a model (its ``copies`` requests), and the HIP
    graph of that work on it (captured at the first use when the model runs graphs)."""

    def __init__(self, model):
        dev = model.device
        k = model.prob.kernel
        e_dtype = torch.float64 if model.precision == "fp64" else torch.float32
        n1 = model.n + 1
        self.E = torch.empty((k.group.ne, k.nd[ADJOINT], k.nd[STATE]), dtype=e_dtype, device=dev)
        self.data = torch.empty(model.pattern.nnz, dtype=e_dtype, device=dev)
        self.m = torch.empty(model.ndof_m, dtype=torch.float64, device=dev)
        self.x = torch.empty(model.ndof, dtype=torch.float64, device=dev)
        self.y = torch.empty(model.ndof, dtype=torch.float64, device=dev)
        self.work = torch.empty(2 * n1 ** 3 + 2, dtype=torch.float64, device=dev)
        self.seed_dev = torch.zeros(model.copies, dtype=torch.int64, device=dev)
        self.seed_host = torch.zeros(model.copies, dtype=torch.int64, pin_memory=True)
        self.y_host = torch.empty(model.ndof, dtype=torch.float64, pin_memory=True)
        self.graph = None


class Model:
    """Everything of one (n, order) that stays resident: the mesh and space tables, the
    pattern and its elimination plan, the CSR index arrays, and a pool of request buffers.
    ``copies`` > 1: the batched model of ``copies`` requests at once, on that many disjoint
    copies of the mesh (one pattern, block-diagonal; one launch of every kernel)."""

    def __init__(self, n, order, precision, device, graphs=False, copies=1):
        self.n, self.order, self.precision, self.device = n, order, precision, device
        self.graphs = bool(graphs)
        self.copies = k = int(copies)
        W = Workload(n, order=order)
        self.W = W
        if k == 1:
            spaces, batches, bc = W.spaces, W.batches, W.bc
        else:
            Vu, Vm = _TiledSpace(W.Vu, k), _TiledSpace(W.Vm, k)
            spaces = [Vu, Vm, Vu]
            batches = MeshBatches(_TiledMesh(W.mesh, k), W.batches.groups[0].degree)
            bc = _TiledBC(W.bc, W.Vu.ndof, k)
        self.prob = PDEVariationalProblem(spaces, W.density, batches, bc, route="two-pass")
        blk = self.prob.block_assembly(ADJOINT, STATE, precision=precision)
        self.plan = blk.plan
        self.dplan = blk.dplan
        self.pattern = blk.pattern
        self.blk = blk
        self.ndof1, self.ndof_m1 = W.Vu.ndof, W.Vm.ndof      # one copy's
        self.ndof = k * self.ndof1
        self.ndof_m = k * self.ndof_m1
        # CSR index arrays on the device (the matrix-vector product reads them)
        self.indptr = self.pattern.dev("indptr").to(device)
        self.indices = self.pattern.dev("indices").to(device)
        # lattice order -> dof order of the parameter space (one copy), and the fields A does
        # not read
        self.lattice_to_dof_m = torch.as_tensor(W.Vm.lattice_to_dof.astype(np.int64)).to(device)
        self.zeros_u = torch.zeros(self.ndof, dtype=torch.float64, device=device)
        self.pool = []
        kk = self.prob.kernel
        esz = 8 if precision == "fp64" else 4
        self.request_bytes = (esz * kk.group.ne * kk.nd[ADJOINT] * kk.nd[STATE] + esz *
                              self.pattern.nnz + 8 * (3 * self.ndof + self.ndof_m) +
                              8 * (2 * (n + 1) ** 3 + 3))
        #: device time of one launch (seconds), measured when the cache is warmed
        self.cost = None

    def take(self):
        return self.pool.pop() if self.pool else Buffers(self)

    def give(self, b):
        self.pool.append(b)

    def work(self, bufs):
        """A launch's work on the current stream, reading its seeds from ``bufs.seed_host``:
        the seeds to the device, every copy's m and x drawn from its seed, element matrices,
        scatter, y = A x, y to the host.  The same calls are captured into the buffers'
        graph."""
        L = _lib.lib()
        st = _lib.stream_handle(None)
        bufs.seed_dev.copy_(bufs.seed_host, non_blocking=True)
        for c in range(self.copies):
            _lib.check(L.hm_serving_field(_lib.ptr(bufs.seed_dev[c:]), self.n + 1, 3,
                                          _lib.ptr(bufs.work), _lib.ptr(self.lattice_to_dof_m),
                                          _lib.ptr(bufs.m[c * self.ndof_m1:]), st),
                       "serving field")
            _lib.check(L.hm_serving_normal(_lib.ptr(bufs.seed_dev[c:]), 1, self.ndof1,
                                           _lib.ptr(bufs.x[c * self.ndof1:]), st),
                       "serving direction")
        loc = [self.zeros_u, bufs.m, self.zeros_u]
        self.prob.kernel.element_matrices(ADJOINT, STATE, loc, precision=self.precision,
                                          out=bufs.E)
        self.pattern.data([bufs.E], plan=self.dplan, out=bufs.data)
        code = _lib.FP64 if bufs.data.dtype == torch.float64 else _lib.FP32
        _lib.check(L.hm_csr_matvec(self.ndof, _lib.ptr(self.indptr), _lib.ptr(self.indices),
                                   _lib.ptr(bufs.data), code, _lib.ptr(bufs.x), _lib.ptr(bufs.y),
                                   st), "serving matvec")
        bufs.y_host.copy_(bufs.y, non_blocking=True)

    def capture(self, bufs):
        """The buffers' graph: one eager run on a side stream, then the capture of the same
        calls (kernels and copies only: no memset nodes)."""
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.work(bufs)
        side.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side, capture_error_mode="thread_local"):
            self.work(bufs)
        torch.cuda.current_stream().wait_stream(side)
        bufs.graph = g

    def launch(self, reqs, bufs, stream):
        """The requests' work (one a copy) on ``stream`` (a graph replay when the model runs
        graphs); returns the completion event.  The buffers' previous launch has completed,
        so their seed slots are free to write."""
        if len(reqs) != self.copies:
            raise ValueError("%d requests for a model of %d copies" % (len(reqs), self.copies))
        for c, r in enumerate(reqs):
            bufs.seed_host[c] = r.seed
        with torch.cuda.stream(stream):
            if self.graphs:
                if bufs.graph is None:
                    self.capture(bufs)
                bufs.graph.replay()
            else:
                self.work(bufs)
            ev = torch.cuda.Event()
            ev.record(stream)
        return ev

    def result(self, bufs, c):
        """Copy ``c``'s product in the pinned buffer (a view)."""
        return bufs.y_host[c * self.ndof1:(c + 1) * self.ndof1]


class ModelCache:
    """The resident models; ``graphs``: requests run as replays of captured HIP graphs (one
    per buffer set), so a request costs the host a few calls instead of one per kernel;
    ``batching``: the small problems also keep batched models of ``BATCHES`` copies."""

    def __init__(self, precision="fp32", device=None, graphs=True, batching=True):
        self.precision = precision
        self.device = device or _lib.require_device()
        self.graphs = bool(graphs)
        self.batching = bool(batching)
        self.models = {}

    def get(self, n, order, copies=1):
        key = (int(n), int(order), int(copies))
        m = self.models.get(key)
        if m is None:
            m = self.models[key] = Model(key[0], key[1], self.precision, self.device,
                                         graphs=self.graphs, copies=key[2])
        return m

    def batch_sizes(self, n):
        """The batched models a problem of ``n^3`` elements keeps (largest first)."""
        return BATCHES if self.batching and int(n) <= BATCH_MAX_N else ()

    def warm(self, sizes=SIZES, orders=(1, 2), buffers=0, batch_buffers=2):
        """Builds the models (with ``buffers`` buffer sets each, ``batch_buffers`` for the
        batched ones, their graphs captured, so that no request waits for a capture) and
        measures each model's device time a launch (``Model.cost``: the median of a few
        replays alone)."""
        for n in sizes:
            for o in orders:
                for k in (1,) + tuple(self.batch_sizes(n)):
                    m = self.get(n, o, k)
                    made = [m.take() for _ in range(int(buffers if k == 1 else batch_buffers))]
                    for b in made:
                        if m.graphs and b.graph is None:
                            m.capture(b)
                    if made:
                        m.cost = _device_time(m, made[0])
                    for b in made:
                        m.give(b)
        torch.cuda.synchronize()


def _device_time(model, bufs, reps=5):
    """Median device time of one launch of ``model`` alone on the current stream."""
    reqs = [Request(0, model.n, model.order, 1 + c) for c in range(model.copies)]
    st = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        a.record(st)
        ev = model.launch(reqs, bufs, st)
        b = torch.cuda.Event(enable_timing=True)
        b.record(st)
        b.synchronize()
        times.append(a.elapsed_time(b) * 1e-3)
    return float(np.median(times))


class _Job:
    """One launch: its requests (one a copy of the model), slot, buffers and event."""

    __slots__ = ("reqs", "model", "slot", "bufs", "event", "copy", "large")

    def __init__(self, reqs, model, slot, bufs, large):
        self.reqs, self.model, self.slot, self.bufs = reqs, model, slot, bufs
        self.large = large
        self.event = self.copy = None


class ServingEngine:
    """Continuous batching over ``slots`` streams within ``hbm_budget`` bytes of request
    buffers.

    Admission: the waiting request with the earliest deadline, ``arrival + slack * cost`` (its
    problem's device time a launch alone, ``Model.cost``), is taken first; with it every
    waiting request of its problem, up to the largest batched model that many fill (the rest
    wait for the next launch).  Problems above ``large`` seconds of device time hold at most
    ``large_slots`` slots at once, so the small ones keep streams of their own.  Without
    measured costs (an unwarmed cache) the order is the arrival order."""

    def __init__(self, cache, slots=8, hbm_budget=48 << 30, slack=4.0, large=0.3e-3,
                 large_slots=3):
        self.cache = cache
        self.slots = [torch.cuda.Stream(device=cache.device) for _ in range(int(slots))]
        self.budget = int(hbm_budget)
        self.slack, self.large, self.large_slots = float(slack), float(large), int(large_slots)
        self.peak_bytes = 0
        self.peak_inflight = 0
        self.launches = 0

    def run(self, requests, keep_results=True, poll_sleep=20e-6):
        """Serve ``requests``; returns them with ``admitted``, ``done``, ``batch`` and (when
        ``keep_results``) ``y`` set.  Times are seconds from the start."""
        for r in requests:
            if self.cache.get(r.n, r.order).request_bytes > self.budget:
                raise ValueError("request %d needs more than the HBM budget" % r.rid)
        # the objects made so far (models, plans, buffers) leave the collector's view, so
        # that a collection during the loop walks only what the loop makes
        gc.collect()
        gc.freeze()
        try:
            return self._run(requests, keep_results, poll_sleep)
        finally:
            gc.unfreeze()

    def _run(self, requests, keep_results, poll_sleep):
        with concurrent.futures.ThreadPoolExecutor(max_workers=2) as copier:
            return self._loop(requests, keep_results, poll_sleep, copier)

    def _pick(self, waiting, large_inflight):
        """The index of the request to admit next, or None: the earliest deadline, passing
        over the large problems while they hold their slots."""
        best = None
        for k, r in enumerate(waiting):
            if r.model.cost is not None and r.model.cost > self.large and \
                    large_inflight >= self.large_slots:
                continue
            if best is None or r.deadline < waiting[best].deadline:
                best = k
        return best

    def _loop(self, requests, keep_results, poll_sleep, copier):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = []
        free = list(range(len(self.slots)))
        inflight = []
        copying = []
        used = 0
        large_inflight = 0
        t0 = time.perf_counter()

        def finish(job):
            nonlocal used, large_inflight
            done = time.perf_counter() - t0
            for r in job.reqs:
                r.done = done
            job.model.give(job.bufs)
            used -= job.model.request_bytes
            large_inflight -= job.large
            free.append(job.slot)
            job.bufs = job.event = job.copy = None

        while pending or waiting or inflight or copying:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                r.model = self.cache.get(r.n, r.order)
                cost = r.model.cost or 0.0
                r.deadline = r.arrival + self.slack * cost
                waiting.append(r)
            while waiting and free:
                k = self._pick(waiting, large_inflight)
                if k is None:
                    break
                r = waiting[k]
                one = r.model.request_bytes
                if used + one > self.budget:
                    break
                group = [r]
                sizes = self.cache.batch_sizes(r.n)
                if sizes:
                    same = sorted((w for w in waiting if w is not r and w.n == r.n
                                   and w.order == r.order), key=lambda w: w.deadline)
                    # the largest batch the waiting requests fill and the budget holds (a
                    # batch of b takes less than b single requests' buffers)
                    fill = next((b for b in sizes
                                 if b <= 1 + len(same) and used + b * one <= self.budget), 1)
                    group += same[:fill - 1]
                model = self.cache.get(r.n, r.order, len(group))
                taken = {id(w) for w in group}
                waiting = [w for w in waiting if id(w) not in taken]
                large = int(r.model.cost is not None and r.model.cost > self.large)
                job = _Job(group, model, free.pop(), model.take(), large)
                used += model.request_bytes
                large_inflight += large
                adm = time.perf_counter() - t0
                for w in group:
                    w.admitted = adm
                    w.batch = len(group)
                job.event = model.launch(group, job.bufs, self.slots[job.slot])
                self.launches += 1
                inflight.append(job)
                self.peak_bytes = max(self.peak_bytes, used)
                self.peak_inflight = max(self.peak_inflight, len(inflight))
            # completions: the results to their own memory (a large one on the copy thread),
            # then the slot
            still = []
            for job in inflight:
                if job.event.query():
                    if keep_results and job.bufs.y_host.numel() > INLINE_RESULT:
                        job.copy = copier.submit(job.bufs.y_host.clone)
                        copying.append(job)
                    else:
                        if keep_results:
                            for c, w in enumerate(job.reqs):
                                w.y = job.model.result(job.bufs, c).clone()
                        finish(job)
                else:
                    still.append(job)
            inflight = still
            still = []
            for job in copying:
                if job.copy.done():
                    y = job.copy.result()
                    n1 = job.model.ndof1
                    for c, w in enumerate(job.reqs):
                        w.y = y if job.model.copies == 1 else y[c * n1:(c + 1) * n1]
                    finish(job)
                else:
                    still.append(job)
            copying = still
            if (not waiting or not free) and (inflight or copying):
                time.sleep(poll_sleep)
            elif not inflight and not waiting and pending:
                time.sleep(min(poll_sleep, max(0.0, pending[0].arrival - now)))
        return requests


def reference_product(cache, req, precision="fp64"):
    """``y = A x`` of a request in FP64 (or ``precision``), alone on the current stream."""
    model = cache.get(req.n, req.order)
    if precision != model.precision:
        refs = cache.__dict__.setdefault("refs", {})
        key = (precision, req.n, req.order)
        if key not in refs:
            refs[key] = Model(req.n, req.order, precision, cache.device)
        model = refs[key]
    bufs = model.take()
    stream = torch.cuda.current_stream()
    ev = model.launch([req], bufs, stream)
    ev.synchronize()
    y = bufs.y_host.clone()
    model.give(bufs)
    return y


def summary(requests, wall=None):
    """Throughput and latency percentiles of a served stream; ``by_model`` splits the
    latencies by problem (``"64^3 P2"`` ...), largest first: the largest products (17 MB of
    result at 64^3 P2) set the tail of a mixed stream."""
    lat = np.array([r.latency for r in requests if r.latency is not None])
    done = [r for r in requests if r.done is not None]
    span = wall if wall is not None else (max(r.done for r in done) - min(r.arrival for r in done))
    nnz = sum(r.model.pattern.nnz for r in done if r.model is not None)
    groups = {}
    for r in done:
        groups.setdefault((r.n, r.order), []).append(r.latency)
    by_model = {"%d^3 P%d" % key: {"requests": len(v), "p50_ms": 1e3 * float(np.percentile(v, 50)),
                                   "p99_ms": 1e3 * float(np.percentile(v, 99))}
                for key, v in sorted(groups.items(), key=lambda kv: (-kv[0][0], -kv[0][1]))}
    batched = sum(1 for r in done if r.batch > 1)
    return {"requests": len(done), "batched": batched, "throughput_req_s": len(done) / span,
            "throughput_nnz_s": nnz / span,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "mean_ms": 1e3 * float(lat.mean()), "max_ms": 1e3 * float(lat.max()), "by_model": by_model}


def error_vs_fp64(cache, requests, count=8):
    """Largest relative error (2-norm) of served products against FP64 ones, over the first
    ``count`` requests of distinct (n, order)."""
    seen = set()
    worst = 0.0
    for r in requests:
        key = (r.n, r.order)
        if key in seen or r.y is None:
            continue
        seen.add(key)
        y64 = reference_product(cache, r, "fp64")
        worst = max(worst, float(torch.linalg.vector_norm(r.y - y64) / torch.linalg.vector_norm(y64)))
        if len(seen) >= count:
            break
    return worst


def capacity(eng, count=96, rounds=3, seed=0):
    """Saturated throughput: closed bursts of ``count`` mixed requests all arriving at once;
    the best of ``rounds`` (the first burst also pays for first touches of the host and
    device memory the loop uses)."""
    best = 0.0
    for k in range(rounds):
        burst = poisson_requests(count, 1e6, seed=seed + k)
        for r in burst:
            r.arrival = 0.0
        t0 = time.perf_counter()
        eng.run(burst, keep_results=False)
        best = max(best, len(burst) / (time.perf_counter() - t0))
    return best


def bench_sample(seconds=20.0, rates=(0.5, 0.8, 0.95), slots=8, seed=0):
    """A bounded serving run for the benchmark line: Poisson arrivals at ``rates`` (fractions
    of the measured capacity) for about ``seconds`` in all; returns the capacity, latency and
    throughput figures of every rate (the first also at the top level) and the error of the
    served products against FP64."""
    cache = ModelCache("fp32")
    cache.warm(buffers=slots)
    eng = ServingEngine(cache, slots=slots)
    cap = capacity(eng, seed=seed + 1)
    per_rate = []
    reqs = []
    for k, f in enumerate(rates):
        rate = f * cap
        count = max(32, min(4000, int(rate * seconds / len(rates))))
        reqs = poisson_requests(count, rate, seed=seed + 5 + k)
        eng.launches = 0
        eng.run(reqs)
        got = summary(reqs)
        got.update({"rate_fraction": f, "rate_req_s": rate, "launches": eng.launches})
        per_rate.append(got)
    out = dict(per_rate[0])
    out.update({"capacity_req_s": cap, "slots": slots,
                "precision": "fp32 contraction and values",
                "admission": "earliest deadline (arrival + %g x device time), large problems "
                             "on at most %d slots, same-problem batches of %s for n <= %d"
                             % (eng.slack, eng.large_slots, "/".join(map(str, BATCHES)),
                                BATCH_MAX_N),
                "p99_ms_by_rate": {"%g" % r["rate_fraction"]: r["p99_ms"] for r in per_rate},
                "p50_ms_by_rate": {"%g" % r["rate_fraction"]: r["p50_ms"] for r in per_rate},
                "rates": per_rate,
                "max_rel_err_vs_fp64": error_vs_fp64(cache, reqs)})
    return out


def main(argv=None):
    """``python -m hippymfem_amd.serving --requests 2000 --rates 0.5,0.8,0.95``: Poisson
    arrivals at fractions of the measured saturated throughput (or absolute rates with
    ``--absolute``), arrival seeds 5, 6, 7."""
    import argparse
    import json

    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rates", default="0.5,0.8,0.95")
    ap.add_argument("--absolute", action="store_true")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--eager", action="store_true", help="launch every kernel (no graphs)")
    ap.add_argument("--no-batching", action="store_true", help="one launch a request")
    a = ap.parse_args(argv)
    cache = ModelCache("fp32", graphs=not a.eager, batching=not a.no_batching)
    cache.warm(buffers=a.slots)
    eng = ServingEngine(cache, slots=a.slots)
    cap = capacity(eng, seed=7)
    print(json.dumps({"capacity_req_s": cap}), flush=True)
    for k, f in enumerate(float(x) for x in a.rates.split(",")):
        rate = f if a.absolute else f * cap
        reqs = poisson_requests(a.requests, rate, seed=5 + k)
        eng.launches = 0
        eng.run(reqs)
        out = summary(reqs)
        out.update({"rate_fraction": None if a.absolute else f, "rate_req_s": rate,
                    "launches": eng.launches,
                    "max_rel_err_vs_fp64": error_vs_fp64(cache, reqs)})
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
