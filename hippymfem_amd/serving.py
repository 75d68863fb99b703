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
* requests are admitted in arrival order as soon as a slot is free and their buffers fit
  the HBM budget beside those in flight; admitted requests run concurrently on their
  streams, so their kernels overlap on the device;
* a finished request's result leaves its pinned buffer for memory of its own, a large one
  on a copy thread (megabytes of host memory to touch, which would otherwise hold up the
  loop's admissions and completions), and its slot and buffers free once it has;
* latency is taken from a request's arrival to its result being the caller's.

Every kernel of a request is deterministic (no float atomics), so a request's result does not
depend on what else is in flight: the same bits alone and in the mix.
"""

import collections
import concurrent.futures
import gc
import time

import numpy as np
import torch

from . import _lib
from ._lib import ADJOINT, STATE
from .problem import PDEVariationalProblem
from .workloads import Workload

SIZES = (8, 16, 24, 32, 48, 64)
#: results of at most this many values are copied on the loop itself (cheaper than a hand-off)
INLINE_RESULT = 1 << 16


class Request:
    """One assemble-and-apply request; ``arrival`` in seconds from the start of the stream."""

    __slots__ = ("rid", "n", "order", "seed", "arrival", "admitted", "done", "y", "slot",
                 "bufs", "event", "model", "copy")

    def __init__(self, rid, n, order, seed, arrival=0.0):
        self.rid, self.n, self.order, self.seed = int(rid), int(n), int(order), int(seed)
        self.arrival = float(arrival)
        self.admitted = self.done = None
        self.y = self.slot = self.bufs = self.event = self.model = self.copy = None

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


class Buffers:
    """The device memory of one in-flight request of a model, and the HIP graph of the
    request's work on it (captured at the first use when the model runs graphs)."""

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
        self.seed_dev = torch.zeros(1, dtype=torch.int64, device=dev)
        self.seed_host = torch.zeros(1, dtype=torch.int64, pin_memory=True)
        self.y_host = torch.empty(model.ndof, dtype=torch.float64, pin_memory=True)
        self.graph = None


class Model:
    """Everything of one (n, order) that stays resident: the mesh and space tables, the
    pattern and its elimination plan, the CSR index arrays, and a pool of request buffers."""

    def __init__(self, n, order, precision, device, graphs=False):
        self.n, self.order, self.precision, self.device = n, order, precision, device
        self.graphs = bool(graphs)
        W = Workload(n, order=order)
        self.W = W
        self.prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
        blk = self.prob.block_assembly(ADJOINT, STATE, precision=precision)
        self.plan = blk.plan
        self.dplan = blk.dplan
        self.pattern = blk.pattern
        self.blk = blk
        self.ndof = W.Vu.ndof
        self.ndof_m = W.Vm.ndof
        # CSR index arrays on the device (the matrix-vector product reads them)
        self.indptr = self.pattern.dev("indptr").to(device)
        self.indices = self.pattern.dev("indices").to(device)
        # lattice order -> dof order of the parameter space, and the fields A does not read
        self.lattice_to_dof_m = torch.as_tensor(W.Vm.lattice_to_dof.astype(np.int64)).to(device)
        self.zeros_u = torch.zeros(self.ndof, dtype=torch.float64, device=device)
        self.pool = []
        k = self.prob.kernel
        esz = 8 if precision == "fp64" else 4
        self.request_bytes = (esz * k.group.ne * k.nd[ADJOINT] * k.nd[STATE] + esz *
                              self.pattern.nnz + 8 * (3 * self.ndof + self.ndof_m) +
                              8 * (2 * (n + 1) ** 3 + 3))

    def take(self):
        return self.pool.pop() if self.pool else Buffers(self)

    def give(self, b):
        self.pool.append(b)

    def work(self, bufs):
        """A request's work on the current stream, reading its seed from ``bufs.seed_host``:
        the seed to the device, m and x drawn from it, element matrices, scatter, y = A x,
        y to the host.  The same calls are captured into the buffers' graph."""
        L = _lib.lib()
        st = _lib.stream_handle(None)
        bufs.seed_dev.copy_(bufs.seed_host, non_blocking=True)
        _lib.check(L.hm_serving_field(_lib.ptr(bufs.seed_dev), self.n + 1, 3, _lib.ptr(bufs.work),
                                      _lib.ptr(self.lattice_to_dof_m), _lib.ptr(bufs.m), st),
                   "serving field")
        _lib.check(L.hm_serving_normal(_lib.ptr(bufs.seed_dev), 1, self.ndof, _lib.ptr(bufs.x), st),
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

    def launch(self, req, bufs, stream):
        """The request's work on its stream (a graph replay when the model runs graphs);
        returns the completion event.  The buffers' previous request has completed, so its
        seed slot is free to write."""
        bufs.seed_host[0] = req.seed
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


class ModelCache:
    """The resident models; ``graphs``: requests run as replays of captured HIP graphs (one
    per buffer set), so a request costs the host a few calls instead of one per kernel."""

    def __init__(self, precision="fp32", device=None, graphs=True):
        self.precision = precision
        self.device = device or _lib.require_device()
        self.graphs = bool(graphs)
        self.models = {}

    def get(self, n, order):
        key = (int(n), int(order))
        m = self.models.get(key)
        if m is None:
            m = self.models[key] = Model(key[0], key[1], self.precision, self.device,
                                         graphs=self.graphs)
        return m

    def warm(self, sizes=SIZES, orders=(1, 2), buffers=0):
        """Builds the models and, with ``buffers``, that many buffer sets each (their graphs
        captured), so that no request waits for a capture."""
        for n in sizes:
            for o in orders:
                m = self.get(n, o)
                made = [m.take() for _ in range(int(buffers))]
                for b in made:
                    if m.graphs and b.graph is None:
                        m.capture(b)
                for b in made:
                    m.give(b)
        torch.cuda.synchronize()


class ServingEngine:
    """Continuous batching over ``slots`` streams within ``hbm_budget`` bytes of request
    buffers."""

    def __init__(self, cache, slots=8, hbm_budget=48 << 30):
        self.cache = cache
        self.slots = [torch.cuda.Stream(device=cache.device) for _ in range(int(slots))]
        self.budget = int(hbm_budget)
        self.peak_bytes = 0
        self.peak_inflight = 0

    def run(self, requests, keep_results=True, poll_sleep=20e-6):
        """Serve ``requests`` (sorted by arrival); returns them with ``admitted``, ``done``
        and (when ``keep_results``) ``y`` set.  Times are seconds from the start."""
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

    def _loop(self, requests, keep_results, poll_sleep, copier):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        free = list(range(len(self.slots)))
        inflight = []
        copying = []
        used = 0
        t0 = time.perf_counter()

        def finish(r):
            nonlocal used
            r.done = time.perf_counter() - t0
            r.model.give(r.bufs)
            used -= r.model.request_bytes
            free.append(r.slot)
            r.bufs = r.event = r.copy = None

        while pending or waiting or inflight or copying:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: arrival order, a free slot and room in the budget
            while waiting and free:
                r = waiting[0]
                model = self.cache.get(r.n, r.order)
                if used + model.request_bytes > self.budget:
                    break
                waiting.popleft()
                r.model = model
                r.slot = free.pop()
                r.bufs = model.take()
                used += model.request_bytes
                r.admitted = time.perf_counter() - t0
                r.event = model.launch(r, r.bufs, self.slots[r.slot])
                inflight.append(r)
                self.peak_bytes = max(self.peak_bytes, used)
                self.peak_inflight = max(self.peak_inflight, len(inflight))
            # completions: the result to its own memory on the copy thread, then the slot
            still = []
            for r in inflight:
                if r.event.query():
                    if keep_results and r.bufs.y_host.numel() > INLINE_RESULT:
                        r.copy = copier.submit(r.bufs.y_host.clone)
                        copying.append(r)
                    else:
                        if keep_results:
                            r.y = r.bufs.y_host.clone()
                        finish(r)
                else:
                    still.append(r)
            inflight = still
            still = []
            for r in copying:
                if r.copy.done():
                    r.y = r.copy.result()
                    finish(r)
                else:
                    still.append(r)
            copying = still
            if not waiting and (inflight or copying):
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
    ev = model.launch(req, bufs, stream)
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
    return {"requests": len(done), "throughput_req_s": len(done) / span,
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


def bench_sample(seconds=20.0, rate=None, slots=8, seed=0):
    """A bounded serving run for the benchmark line: requests at ``rate`` per second (by
    default about half the measured capacity) for about ``seconds``; returns latency and
    throughput figures and the error against FP64."""
    cache = ModelCache("fp32")
    cache.warm(buffers=slots)
    eng = ServingEngine(cache, slots=slots)
    cap = capacity(eng, seed=seed + 1)
    rate = rate or 0.5 * cap
    count = max(32, min(2000, int(rate * seconds)))
    reqs = poisson_requests(count, rate, seed=seed)
    eng.run(reqs)
    out = summary(reqs)
    out.update({"rate_req_s": rate, "capacity_req_s": cap, "slots": slots,
                "precision": "fp32 contraction and values",
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
    a = ap.parse_args(argv)
    cache = ModelCache("fp32", graphs=not a.eager)
    cache.warm(buffers=a.slots)
    eng = ServingEngine(cache, slots=a.slots)
    cap = capacity(eng, seed=7)
    print(json.dumps({"capacity_req_s": cap}), flush=True)
    for k, f in enumerate(float(x) for x in a.rates.split(",")):
        rate = f if a.absolute else f * cap
        reqs = poisson_requests(a.requests, rate, seed=5 + k)
        eng.run(reqs)
        out = summary(reqs)
        out.update({"rate_req_s": rate, "max_rel_err_vs_fp64": error_vs_fp64(cache, reqs)})
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
