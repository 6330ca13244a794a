"""The no-allocation entry point (heffte_mi355_execute_c2c_noalloc) against the allocating one.

The no-allocation call runs the same passes with the static per-GPU scratch as the workspace,
so its results must equal those of heffte_mi355_forward/backward_c2c (and _z2z) bit for bit.
Every call goes through the C ABI.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HMI_ERROR_INVALID_ARGUMENT = 2
HMI_ERROR_DEVICE = 4


def _plan(hmi, size, in_order, out_order):
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], in_order)
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], out_order)
    return hmi.fft3d(hmi.backend.mi355, inbox, outbox)


def _data(torch, cuda, n, dtype, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    return torch.from_numpy(x).to(dtype).to(cuda)


def _noalloc(hmi, plan, direction, dbl, batch, x, y, scaling):
    return hmi.lib.heffte_mi355_execute_c2c_noalloc(plan._h, direction, 1 if dbl else 0, batch,
                                                    ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                                    scaling)


def _allocating(hmi, plan, direction, dbl, batch, x, y, ws, scaling):
    name = ("forward_" if direction < 0 else "backward_") + ("z2z" if dbl else "c2c")
    fn = getattr(hmi.lib, "heffte_mi355_" + name)
    return fn(plan._h, batch, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
              ctypes.c_void_p(ws.data_ptr()) if ws is not None else None, scaling)


# two calls in a row on one GPU, the second with a smaller input (and other orders, so its passes
# use the scratch in another layout)
CASES = [((48, 40, 36), (0, 1, 2), (2, 0, 1), 2), ((20, 16, 12), (1, 0, 2), (0, 2, 1), 1)]


@pytest.mark.parametrize("dbl", [False, True], ids=["c2c", "z2z"])
@pytest.mark.parametrize("direction", [-1, 1], ids=["forward", "backward"])
@pytest.mark.parametrize("in_place", [True, False], ids=["in_place", "out_of_place"])
def test_noalloc_matches_allocating_entry_point(hmi, cuda, dbl, direction, in_place):
    import torch
    dtype = torch.complex128 if dbl else torch.complex64
    for i, (size, oi, oo, batch) in enumerate(CASES):
        plan = _plan(hmi, size, oi, oo)
        # an in-place call with a permuted output order runs an intermediate in the workspace
        assert plan.size_workspace() > 0, "the case must use a workspace"
        plan.prepare(double_precision=dbl)
        n = batch * size[0] * size[1] * size[2]
        x = _data(torch, cuda, n, dtype, 10 + i)
        ws = torch.empty(batch * plan.size_workspace(), dtype=dtype, device=cuda)
        outs = []
        for call in ("caller_workspace", "null_workspace", "noalloc"):
            src = x.clone()
            dst = src if in_place else torch.full_like(x, 7.0)
            if call == "noalloc":
                rc = _noalloc(hmi, plan, direction, dbl, batch, src, dst, 1)
            else:
                rc = _allocating(hmi, plan, direction, dbl, batch, src, dst,
                                 ws if call == "caller_workspace" else None, 1)
            assert rc == 0, (call, hmi.last_error())
            outs.append(dst)
        torch.cuda.synchronize()
        assert torch.equal(outs[2], outs[0]), f"case {i}: no-allocation result differs"
        assert torch.equal(outs[1], outs[0])


def test_noalloc_in_place_and_round_trip(hmi, cuda):
    import torch
    size, batch = (32, 24, 40), 3
    plan = _plan(hmi, size, (0, 1, 2), (1, 2, 0))
    plan.prepare()
    n = batch * size[0] * size[1] * size[2]
    x = _data(torch, cuda, n, torch.complex64, 3)
    ref = x.clone()
    ws = torch.empty(batch * plan.size_workspace(), dtype=torch.complex64, device=cuda)
    assert _allocating(hmi, plan, -1, False, batch, ref, ref, ws, 2) == 0, hmi.last_error()
    y = x.clone()
    assert _noalloc(hmi, plan, -1, False, batch, y, y, 2) == 0, hmi.last_error()
    torch.cuda.synchronize()
    assert torch.equal(y, ref)
    assert _noalloc(hmi, plan, 1, False, batch, y, y, 2) == 0, hmi.last_error()
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x)).item()
    assert err <= 5e-6, err


def test_noalloc_rejects_workspace_beyond_scratch(hmi, cuda):
    import torch
    # 256 x 256 x 160 complex<float> in place with a permuted output: an 84 MB workspace, more
    # than the 64 MiB scratch (the arrays are full size, so nothing could run out of bounds)
    size = (256, 256, 160)
    plan = _plan(hmi, size, (0, 1, 2), (2, 1, 0))
    plan.prepare()
    assert plan.size_workspace() * 8 > (64 << 20)
    x = torch.full((size[0] * size[1] * size[2],), 3.0, dtype=torch.complex64, device=cuda)
    rc = _noalloc(hmi, plan, -1, False, 1, x, x, 0)
    assert rc == HMI_ERROR_INVALID_ARGUMENT
    assert "static scratch" in hmi.last_error()
    torch.cuda.synchronize()
    assert torch.all(x == 3.0), "a refused call must enqueue nothing"
    del x
    # a plan that fits with one transform is refused for a batch whose workspace does not
    small = _plan(hmi, (64, 64, 64), (0, 1, 2), (2, 1, 0))
    small.prepare()
    y = torch.full((40 * 64 ** 3,), 3.0, dtype=torch.complex64, device=cuda)
    rc = _noalloc(hmi, small, -1, False, 40, y, y, 0)
    assert rc == HMI_ERROR_INVALID_ARGUMENT and "static scratch" in hmi.last_error()
    torch.cuda.synchronize()
    assert torch.all(y == 3.0), "a refused call must enqueue nothing"


def test_noalloc_needs_prepared_tables(hmi, cuda):
    import torch
    plan = _plan(hmi, (24, 20, 16), (0, 1, 2), (2, 0, 1))
    n = 24 * 20 * 16
    x = _data(torch, cuda, n, torch.complex64, 5)
    y = torch.full_like(x, 3.0)
    rc = _noalloc(hmi, plan, -1, False, 1, x, y, 0)
    assert rc == HMI_ERROR_DEVICE and "plan_prepare" in hmi.last_error()
    torch.cuda.synchronize()
    assert torch.all(y == 3.0), "a refused call must enqueue nothing"
    # single-precision tables do not serve a double-precision call
    plan.prepare(double_precision=False)
    xd = x.to(torch.complex128)
    yd = torch.full_like(xd, 3.0)
    rc = _noalloc(hmi, plan, -1, True, 1, xd, yd, 0)
    assert rc == HMI_ERROR_DEVICE
    torch.cuda.synchronize()
    assert torch.all(yd == 3.0)


def test_noalloc_argument_checks(hmi):
    plan = _plan(hmi, (8, 8, 8), (0, 1, 2), (0, 1, 2))
    f = hmi.lib.heffte_mi355_execute_c2c_noalloc
    assert f(None, -1, 0, 1, None, None, 0) == 3
    assert f(plan._h, 0, 0, 1, None, None, 0) == HMI_ERROR_INVALID_ARGUMENT
    assert f(plan._h, -1, 0, 1, None, None, 5) == HMI_ERROR_INVALID_ARGUMENT
    r2c = hmi.fft3d_r2c(hmi.backend.mi355, hmi.box3d([0, 0, 0], [7, 7, 7]), hmi.box3d([0, 0, 0], [4, 7, 7]), 0)
    assert f(r2c._h, -1, 0, 1, None, None, 0) == HMI_ERROR_INVALID_ARGUMENT
