"""``hm_serving_field_noalloc`` (the serving field with its workspace in the library's static
device array) against ``hm_serving_field`` (the caller's workspace), through the C ABI: the
same bits on two calls in a row, the second on a smaller lattice (the array keeps the first
call's contents), with and without a dof permutation, and the size limit refused."""

import numpy as np
import pytest
import torch

import hippymfem_amd._lib as L

pytestmark = pytest.mark.gpu


def _field(seed, n1, passes, to, noalloc, device):
    sd = torch.tensor([int(seed)], dtype=torch.int64, device=device)
    out = torch.full((n1 ** 3,), np.nan, dtype=torch.float64, device=device)
    st = L.stream_handle(None)
    if noalloc:
        rc = L.lib().hm_serving_field_noalloc(L.ptr(sd), n1, passes, L.ptr(to), L.ptr(out), st)
        L.check(rc, "field (no allocation)")
    else:
        work = torch.full((2 * n1 ** 3 + 2,), np.nan, dtype=torch.float64, device=device)
        rc = L.lib().hm_serving_field(L.ptr(sd), n1, passes, L.ptr(work), L.ptr(to), L.ptr(out), st)
        L.check(rc, "field")
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("permuted", [False, True])
def test_noalloc_field_equals_workspace_field(gpu, permuted):
    cases = [(7, L.SERVING_FIELD_NOALLOC_MAX_N1, 3), (11, 17, 2)]   # the second call smaller
    for seed, n1, passes in cases:
        print("case seed=%d n1=%d passes=%d permuted=%s" % (seed, n1, passes, permuted),
              flush=True)
        to = None
        if permuted:
            to = torch.as_tensor(np.random.default_rng(seed).permutation(n1 ** 3),
                                 dtype=torch.int64, device=gpu)
        ref = _field(seed, n1, passes, to, False, gpu)
        got = _field(seed, n1, passes, to, True, gpu)
        assert torch.equal(got, ref), (seed, n1)
        assert bool(torch.isfinite(got).all())
        assert abs(float(got.mean())) < 1e-12
        assert abs(float(got.std(unbiased=False)) - 1.0) < 1e-12


def test_noalloc_field_refuses_larger_lattice(gpu):
    n1 = L.SERVING_FIELD_NOALLOC_MAX_N1 + 1
    sd = torch.tensor([3], dtype=torch.int64, device=gpu)
    out = torch.empty(n1 ** 3, dtype=torch.float64, device=gpu)
    rc = L.lib().hm_serving_field_noalloc(L.ptr(sd), n1, 3, None, L.ptr(out),
                                         L.stream_handle(None))
    assert rc == L.HM_ERR_ARG
    assert "exceeds" in L.lib().hm_last_error().decode()
