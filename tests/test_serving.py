"""The serving form: continuous batching of assemble-and-apply requests.

A served product is checked against the oracle (FP64), the precision form against FP64
(relative 2-norm error at most 1e-6), every result of a mixed stream against the same request
served alone (bit-identical: the kernels are deterministic), and admission against its HBM
budget.
"""

import numpy as np
import pytest

from hippymfem_amd.serving import SIZES, poisson_requests, summary


def test_poisson_requests():
    reqs = poisson_requests(2000, 50.0, seed=3)
    t = np.array([r.arrival for r in reqs])
    assert len(reqs) == 2000 and np.all(np.diff(t) >= 0)
    assert abs(np.diff(t).mean() * 50.0 - 1.0) < 0.1
    assert {r.n for r in reqs} <= set(SIZES) and {r.order for r in reqs} == {1, 2}
    assert abs(np.mean([r.order == 2 for r in reqs]) - 0.5) < 0.05
    again = poisson_requests(2000, 50.0, seed=3)
    assert [(r.n, r.order, r.seed) for r in again] == [(r.n, r.order, r.seed) for r in reqs]


def test_summary_by_model():
    """Percentiles over the stream and per problem, from the requests' own clocks."""
    reqs = poisson_requests(400, 100.0, seed=4)
    for k, r in enumerate(reqs):
        r.done = r.arrival + 1e-3 * r.n / 8 * (1.0 + (k % 7) / 70)     # latency grows with n
    out = summary(reqs)
    lat = np.array([r.done - r.arrival for r in reqs])
    assert out["requests"] == 400
    assert abs(out["p99_ms"] - 1e3 * np.percentile(lat, 99)) < 1e-9
    assert sum(v["requests"] for v in out["by_model"].values()) == 400
    keys = list(out["by_model"])
    assert keys[0] == "64^3 P2" and keys[-1] == "8^3 P1"
    big = [r.done - r.arrival for r in reqs if (r.n, r.order) == (64, 2)]
    assert abs(out["by_model"]["64^3 P2"]["p50_ms"] - 1e3 * np.percentile(big, 50)) < 1e-9
    assert out["by_model"]["64^3 P2"]["p99_ms"] > out["by_model"]["8^3 P1"]["p99_ms"]


@pytest.mark.gpu
def test_served_field_has_population_std_one(gpu):
    """The served parameter field is normalized as the benchmarks' field is
    (``workloads.smoothed_noise``: ``np.std``, divisor n)."""
    from hippymfem_amd.serving import device_field

    for n in (4, 9):
        f = device_field(n, 77, gpu).cpu().numpy()
        assert f.size == (n + 1) ** 3
        assert abs(f.mean()) < 1e-13
        assert abs(f.std() - 1.0) < 1e-13, f.std()           # population deviation
        assert abs(f.std(ddof=1) - 1.0) > 1e-4                # not the unbiased one


@pytest.mark.gpu
def test_served_product_matches_oracle(gpu):
    from oracle import fe_oracle as O

    from hippymfem_amd import ADJOINT, STATE
    from hippymfem_amd.serving import (ModelCache, Request, ServingEngine, device_direction,
                                       device_field)

    cache = ModelCache("fp64")
    eng = ServingEngine(cache, slots=2)
    for order in (1, 2):
        r = Request(0, 4, order, seed=123 + order)
        eng.run([r])
        P = O.Problem(4, order=order)
        P.m = P.Vm.lattice_field(device_field(4, r.seed, gpu).cpu().numpy())
        x = device_direction(P.Vu.ndof, r.seed, gpu).cpu().numpy()
        ip, ix, d = P.assembled(ADJOINT, STATE)
        y = O.csr_matvec(ip, ix, d, x)
        err = np.abs(r.y.numpy() - y).max() / np.abs(y).max()
        print("order %d served vs oracle %.2e" % (order, err), flush=True)
        assert err <= 1e-13


@pytest.mark.gpu
def test_alone_equals_in_the_mix(gpu):
    import torch

    from hippymfem_amd.serving import ModelCache, Request, ServingEngine

    cache = ModelCache("fp32")
    eng = ServingEngine(cache, slots=4)
    spec = [(8, 1, 11), (8, 2, 12), (16, 1, 13), (16, 2, 14), (24, 2, 15), (8, 2, 16),
            (16, 2, 17), (4, 1, 18)]
    mix = [Request(k, n, o, s, arrival=0.0005 * k) for k, (n, o, s) in enumerate(spec)]
    eng.run(mix)
    assert eng.peak_inflight > 1                      # the requests did overlap
    for r in mix:
        alone = Request(100 + r.rid, r.n, r.order, r.seed)
        ServingEngine(cache, slots=1).run([alone])
        assert torch.equal(alone.y, r.y), "request %d differs alone and in the mix" % r.rid


@pytest.mark.gpu
def test_precision_form_error_and_admission(gpu):
    import torch

    from hippymfem_amd.serving import ModelCache, ServingEngine, reference_product

    cache = ModelCache("fp32")
    reqs = poisson_requests(12, 2000.0, seed=5, sizes=(8, 16, 24))
    largest = max(cache.get(r.n, r.order).request_bytes for r in reqs)
    eng = ServingEngine(cache, slots=8, hbm_budget=int(1.5 * largest))
    eng.run(reqs)
    assert all(r.y is not None for r in reqs)
    assert eng.peak_bytes <= int(1.5 * largest)
    worst = 0.0
    for r in reqs:
        y64 = reference_product(cache, r, "fp64")
        worst = max(worst, float(torch.linalg.vector_norm(r.y - y64) /
                                 torch.linalg.vector_norm(y64)))
    print("fp32 served vs fp64: %.2e" % worst, flush=True)
    assert worst <= 1e-6
    with pytest.raises(ValueError):
        ServingEngine(cache, slots=2, hbm_budget=1024).run(reqs[:1])


@pytest.mark.gpu
def test_graph_replays_equal_eager(gpu):
    """Every replay of a request's captured graph gives the bits of the same request run
    eagerly, with other eager work and synchronizations between the replays."""
    import torch

    from hippymfem_amd.serving import ModelCache, Request, ServingEngine, device_field

    eager = ModelCache("fp32", graphs=False)
    graphs = ModelCache("fp32", graphs=True)
    for n, order in ((8, 2), (16, 1)):
        seeds = [31, 32, 31, 33]
        ref = {}
        for s in set(seeds):
            r = Request(0, n, order, s)
            ServingEngine(eager, slots=1).run([r])
            ref[s] = r.y
        eng = ServingEngine(graphs, slots=1)
        for k, s in enumerate(seeds):
            r = Request(k, n, order, s)
            eng.run([r])
            assert graphs.get(n, order).pool[0].graph is not None
            assert torch.equal(r.y, ref[s]), (n, order, s)
            device_field(n, 1000 + k, gpu)            # eager work between the replays
            torch.cuda.synchronize()


@pytest.mark.gpu
def test_batched_requests_equal_alone(gpu):
    """Requests of one small problem arriving together are served by batched launches (a
    model of k disjoint copies of the mesh); each result is the bits of the request alone."""
    import torch

    from hippymfem_amd.serving import ModelCache, Request, ServingEngine

    cache = ModelCache("fp32")
    cache.warm(sizes=(8, 16), buffers=2)
    eng = ServingEngine(cache, slots=2)
    reqs = [Request(k, 8 if k % 3 else 16, 1 + k % 2, 500 + k, arrival=0.0) for k in range(14)]
    eng.run(reqs)
    assert max(r.batch for r in reqs) > 1
    assert eng.launches < len(reqs)
    single = ModelCache("fp32", batching=False)
    for r in reqs:
        alone = Request(100 + r.rid, r.n, r.order, r.seed)
        ServingEngine(single, slots=1).run([alone])
        assert torch.equal(alone.y, r.y), "request %d (batch of %d) differs alone" % (r.rid,
                                                                                       r.batch)


@pytest.mark.gpu
def test_small_requests_pass_a_large_one(gpu):
    """Size-aware admission: with the large problem's slots held, small requests that arrive
    after a large one are admitted before it."""
    from hippymfem_amd.serving import ModelCache, Request, ServingEngine

    cache = ModelCache("fp32")
    cache.warm(sizes=(8, 48), orders=(2,), buffers=2)
    assert cache.get(48, 2).cost > cache.get(8, 2).cost
    eng = ServingEngine(cache, slots=4, large=0.5 * cache.get(48, 2).cost, large_slots=1)
    reqs = [Request(0, 48, 2, 1, 0.0), Request(1, 48, 2, 2, 0.0)]
    reqs += [Request(2 + k, 8, 2, 10 + k, 0.0) for k in range(4)]
    eng.run(reqs)
    second = max(reqs[:2], key=lambda r: r.admitted)
    assert all(r.admitted <= second.admitted for r in reqs[2:])
