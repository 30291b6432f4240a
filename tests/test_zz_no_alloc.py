"""The allocation-free delayed-CGS2 step (inch_dcgs2_step_noalloc, static scratch) gives the
same results as inch_dcgs2_step with a caller workspace, bit for bit, on two calls in a row
(the second smaller), called through the C ABI; J beyond its scratch is rejected."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(len_, J, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    V = torch.rand((J + 1, len_), generator=g, dtype=torch.float64, device="cuda") - 0.5
    u = torch.rand(len_, generator=g, dtype=torch.float64, device="cuda") - 0.5
    beta = torch.rand(J, generator=g, dtype=torch.float64, device="cuda") + 0.5
    return V, u, beta


def _run(entry, V, u, beta, J):
    from inching_amd import ops
    from inching_amd._lib import call, lib
    V, u = V.clone(), u.clone()
    alpha = torch.zeros(J, dtype=torch.float64, device="cuda")
    bout = torch.zeros(1, dtype=torch.float64, device="cuda")
    vnext = torch.zeros_like(u)
    args = [u.numel(), J, ops.P(V), V.stride(0), ops.P(u), ops.P(alpha), ops.P(beta), ops.P(bout), ops.P(vnext)]
    if entry == "inch_dcgs2_step":
        work = torch.empty(int(lib().inch_dcgs2_work_size(J)), dtype=torch.float64, device="cuda")
        args.append(ops.P(work))
    call(entry, *args, ops.stream_ptr())
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (V, u, alpha, bout, vnext)]


def test_noalloc_dcgs2_equals_workspace_form(cuda):
    for len_, J, seed in ((200003, 37, 1), (50001, 9, 2)):     # the second call is smaller
        print("case: dcgs2 noalloc len=%d J=%d" % (len_, J), flush=True)
        V, u, beta = _case(len_, J, seed)
        ref = _run("inch_dcgs2_step", V, u, beta, J)
        got = _run("inch_dcgs2_step_noalloc", V, u, beta, J)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        assert np.isfinite(ref[3]).all() and ref[3][0] > 0


def test_noalloc_dcgs2_rejects_large_basis(cuda):
    from inching_amd import ops
    from inching_amd._lib import InchingError, call
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    with pytest.raises(InchingError) as e:
        call("inch_dcgs2_step_noalloc", 8, 257, ops.P(x), 8, ops.P(x), ops.P(x), ops.P(x), ops.P(x), None,
             ops.stream_ptr())
    assert e.value.status == -3 and "257" in str(e.value)
