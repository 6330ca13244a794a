"""GPU parity of the fused convolution (forward x, forward y, pointwise product folded into
the first backward pass, backward) against the FP64 oracle.

The oracle's convolve3d composes its pinned transforms as benchmarks/convolution.cpp:86-95
does and is itself pinned against a brute-force circular convolution.  Bounds: relative
L2 <= 1e-13 for complex<double>, <= 5e-6 for complex<float>.  At 256^3 (too large for the
oracle in seconds) the checks are the size-independent ones: a delta at the origin is the
identity, a shifted delta shifts the input, and the product commutes.
"""
import numpy as np
import pytest

import fftcase as fc

pytestmark = pytest.mark.gpu

CDT = {"complex64": np.complex64, "complex128": np.complex128}


def make_plan(hmi, size, inorder, outorder):
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], inorder)
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], outorder)
    return hmi.fft3d(hmi.backend.mi355, inbox, outbox)


def oracle_conv(oracle, size, inorder, outorder, x, y, scale, batch=1):
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], inorder)
    bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], outorder)
    return oracle.convolve3d(bi, bo, x, y, scale, batch)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("size", [(6, 7, 8), (16, 16, 16), (43, 76, 24), (1, 16, 1), (11, 13, 1), (64, 32, 8)])
def test_convolve_vs_oracle(hmi, cuda, oracle, dtype, size):
    import torch
    n = int(np.prod(size))
    for inorder, outorder in [((0, 1, 2), (0, 1, 2)), ((0, 1, 2), (2, 1, 0)), ((1, 2, 0), (0, 2, 1))]:
        plan = make_plan(hmi, size, inorder, outorder)
        x, y = fc.uniform_complex(n, 3), fc.uniform_complex(n, 4)
        tx, ty = fc.to_gpu(x, CDT[dtype]), fc.to_gpu(y, CDT[dtype])
        out = torch.empty_like(tx)
        ws = torch.empty(plan.size_convolve_workspace(), dtype=tx.dtype, device=cuda)
        plan.convolve(tx, ty, out, ws, hmi.scale.full)
        ref = oracle_conv(oracle, size, inorder, outorder, x, y, 1.0 / n)
        err = fc.rel_l2(fc.to_host(out), ref)
        assert err <= fc.TOL[dtype], (size, inorder, outorder, err)
        # inputs are untouched, and scale.none gives N times the convolution
        assert np.array_equal(fc.to_host(tx), x.astype(CDT[dtype]).astype(np.complex128))
        plan.convolve(tx, ty, out, None, hmi.scale.none)
        err = fc.rel_l2(fc.to_host(out), ref * n)
        assert err <= fc.TOL[dtype], (size, inorder, outorder, "none", err)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_convolve_in_place_and_batch(hmi, cuda, oracle, dtype):
    import torch
    size, batch = (8, 9, 10), 3
    n = int(np.prod(size))
    plan = make_plan(hmi, size, (0, 1, 2), (2, 0, 1))
    x = np.concatenate([fc.uniform_complex(n, 20 + b) for b in range(batch)])
    y = np.concatenate([fc.uniform_complex(n, 30 + b) for b in range(batch)])
    ref = oracle_conv(oracle, size, (0, 1, 2), (2, 0, 1), x, y, 1.0 / n, batch)
    ws = torch.empty(batch * plan.size_convolve_workspace(), dtype=fc.to_gpu(x[:1], CDT[dtype]).dtype,
                     device=cuda)
    # output aliases x
    tx, ty = fc.to_gpu(x, CDT[dtype]), fc.to_gpu(y, CDT[dtype])
    plan.convolve(tx, ty, tx, ws, hmi.scale.full, batch=batch)
    assert fc.rel_l2(fc.to_host(tx), ref) <= fc.TOL[dtype]
    # output aliases y
    tx, ty = fc.to_gpu(x, CDT[dtype]), fc.to_gpu(y, CDT[dtype])
    plan.convolve(tx, ty, ty, ws, hmi.scale.full, batch=batch)
    assert fc.rel_l2(fc.to_host(ty), ref) <= fc.TOL[dtype]
    # x and y the same array (the squared spectrum of benchmarks/convolution.cpp:93)
    tx = fc.to_gpu(x, CDT[dtype])
    out = torch.empty_like(tx)
    plan.convolve(tx, tx, out, None, hmi.scale.full, batch=batch)
    ref2 = oracle_conv(oracle, size, (0, 1, 2), (2, 0, 1), x, x, 1.0 / n, batch)
    assert fc.rel_l2(fc.to_host(out), ref2) <= fc.TOL[dtype]


def test_convolve_matches_unfused_sequence(hmi, cuda):
    """The fused call against forward, a torch product and backward through the same plan."""
    import torch
    size = (32, 24, 20)
    n = int(np.prod(size))
    plan = make_plan(hmi, size, (0, 1, 2), (1, 2, 0))
    tx = fc.to_gpu(fc.uniform_complex(n, 5), np.complex128)
    ty = fc.to_gpu(fc.uniform_complex(n, 6), np.complex128)
    fx, fy = torch.empty_like(tx), torch.empty_like(ty)
    plan.forward(tx, fx)
    plan.forward(ty, fy)
    fx.mul_(fy)
    ref = torch.empty_like(tx)
    plan.backward(fx, ref, hmi.scale.full)
    out = torch.empty_like(tx)
    plan.convolve(tx, ty, out)
    assert fc.rel_l2(fc.to_host(out), fc.to_host(ref)) <= 1e-14


@pytest.mark.parametrize("form", ["fp32", "split_fp16", "split_bf16"])
def test_convolve_split_precision(hmi, cuda, oracle, form):
    """A 512-point axis: the forward passes run the plan's precision form; the fused backward
    pass keeps the vector-ALU kernel."""
    import torch
    size = (512, 8, 6)
    n = int(np.prod(size))
    plan = make_plan(hmi, size, (0, 1, 2), (0, 1, 2))
    plan.set_precision(getattr(hmi.precision, form))
    x, y = fc.uniform_complex(n, 8), fc.uniform_complex(n, 9)
    tx, ty = fc.to_gpu(x, np.complex64), fc.to_gpu(y, np.complex64)
    out = torch.empty_like(tx)
    plan.convolve(tx, ty, out)
    ref = oracle_conv(oracle, size, (0, 1, 2), (0, 1, 2), x, y, 1.0 / n)
    err = fc.rel_l2(fc.to_host(out), ref)
    assert err <= 5e-6, (form, err)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_convolve_256_cubed_properties(hmi, cuda, dtype):
    import torch
    size = (256, 256, 256)
    n = 256 ** 3
    plan = make_plan(hmi, size, (0, 1, 2), (2, 1, 0))
    g = torch.Generator(device=cuda).manual_seed(1)
    rdt = torch.float32 if dtype == "complex64" else torch.float64
    x = torch.complex(torch.rand(n, generator=g, device=cuda, dtype=rdt) - 0.5,
                      torch.rand(n, generator=g, device=cuda, dtype=rdt) - 0.5)
    y = torch.complex(torch.rand(n, generator=g, device=cuda, dtype=rdt) - 0.5,
                      torch.rand(n, generator=g, device=cuda, dtype=rdt) - 0.5)
    tol = 5e-6 if dtype == "complex64" else 1e-13
    delta = torch.zeros_like(x)
    delta[0] = 1.0
    out = torch.empty_like(x)
    plan.convolve(x, delta, out)
    assert float(torch.linalg.norm(out - x) / torch.linalg.norm(x)) <= tol
    # delta at (i0, i1, i2) = (3, 1, 2): the output is x shifted by that vector (input order (0,1,2))
    delta.zero_()
    delta[3 + 256 * (1 + 256 * 2)] = 1.0
    plan.convolve(x, delta, out)
    shifted = torch.roll(x.view(256, 256, 256), shifts=(2, 1, 3), dims=(0, 1, 2)).reshape(-1)
    assert float(torch.linalg.norm(out - shifted) / torch.linalg.norm(x)) <= tol
    out2 = torch.empty_like(x)
    plan.convolve(x, y, out)
    plan.convolve(y, x, out2)
    assert float(torch.linalg.norm(out - out2) / torch.linalg.norm(out)) <= tol


def test_convolve_rejections(hmi, cuda):
    import torch
    size = (8, 8, 8)
    inbox = hmi.box3d([0, 0, 0], [7, 7, 7])
    r2c = hmi.fft3d_r2c(hmi.backend.mi355, inbox, inbox.r2c(0), 0)
    x = torch.zeros(512, dtype=torch.complex64, device=cuda)
    with pytest.raises(hmi.heffte_input_error):
        r2c.convolve(x, x, x)
    plan = make_plan(hmi, size, (0, 1, 2), (0, 1, 2))
    with pytest.raises(hmi.heffte_input_error):
        plan.convolve(x, x, x, scaling=3)
    with pytest.raises(hmi.heffte_input_error):
        plan.convolve(x, x.to(torch.complex128), x)
    with pytest.raises(hmi.heffte_input_error):
        plan.convolve(x[:100], x, x)
    with pytest.raises(hmi.heffte_input_error):
        plan.convolve(x, x, x, workspace=torch.zeros(4, dtype=torch.complex64, device=cuda))
    assert r2c.size_convolve_workspace() == -1


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_convolve_spectrum(hmi, cuda, oracle, dtype):
    """A precomputed spectrum (forward of y, scale none) gives the same result as convolve(x, y),
    also in place on x and with batch = 2."""
    import torch
    size, batch = (12, 10, 9), 2
    n = int(np.prod(size))
    plan = make_plan(hmi, size, (0, 1, 2), (2, 1, 0))
    x = np.concatenate([fc.uniform_complex(n, 40 + b) for b in range(batch)])
    y = np.concatenate([fc.uniform_complex(n, 50 + b) for b in range(batch)])
    tx, ty = fc.to_gpu(x, CDT[dtype]), fc.to_gpu(y, CDT[dtype])
    yhat = torch.empty_like(ty)
    plan.forward(ty, yhat, batch=batch)
    out = torch.empty_like(tx)
    plan.convolve_spectrum(tx, yhat, out, batch=batch)
    ref = oracle_conv(oracle, size, (0, 1, 2), (2, 1, 0), x, y, 1.0 / n, batch)
    assert fc.rel_l2(fc.to_host(out), ref) <= fc.TOL[dtype]
    ws = torch.empty(batch * max(1, plan.size_workspace()), dtype=tx.dtype, device=cuda)
    plan.convolve_spectrum(tx, yhat, tx, ws, hmi.scale.full, batch=batch)
    assert fc.rel_l2(fc.to_host(tx), ref) <= fc.TOL[dtype]
    with pytest.raises(hmi.heffte_input_error):
        plan.convolve_spectrum(ty, yhat, yhat, batch=batch)
