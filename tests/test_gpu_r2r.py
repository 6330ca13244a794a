"""GPU parity of rtransform<cos / sin / cos1> (DCT-II/III, DST-II/III, DCT-I on every axis).

Known answers: test/test_cos.cpp:36-80 (2x3x4 world, iota input; forward golden, backward
with scale::full recovers the input) and test/test_units_nompi.cpp:288-333 (1D DCT-II/DST-II of
1..4 and 1..5, embedded here in boxes whose other two axes have length 1, where each length-1
axis multiplies by 2).  Random inputs are compared with the FP64 oracle for every order pair,
odd/prime/power-of-two lengths, batches and in-place calls.  DCT-I: the golden of
test/test_cos.cpp:48-50 and the oracle (REDFT00 forward, 2 REDFT00 backward, scale::full =
1/(64 prod(n-1))).
"""
import numpy as np
import pytest

import fftcase as fc

pytestmark = pytest.mark.gpu

RDT = {"float32": (np.float32, 5e-6), "float64": (np.float64, 1e-13)}


def run_r2r(hmi, cuda, kind, size, inorder, outorder, x, dtype, direction="forward", scaling=0, batch=1,
            in_place=False, use_ws=True):
    import torch
    rdt, _ = RDT[dtype]
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], inorder)
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], outorder)
    plan = hmi.rtransform(hmi.backend.mi355, inbox, outbox, getattr(hmi.kind, kind))
    tx = torch.from_numpy(np.ascontiguousarray(x).astype(rdt)).to(cuda)
    ty = tx if in_place else torch.empty_like(tx)
    ws = torch.empty(max(1, batch * plan.size_workspace()), dtype=tx.dtype, device=cuda) if use_ws else None
    if direction == "forward":
        plan.forward_buffered(tx, ty, ws, scaling, batch) if use_ws else plan.forward(tx, ty, scaling, batch)
    else:
        plan.backward_buffered(tx, ty, ws, scaling, batch) if use_ws else plan.backward(tx, ty, scaling, batch)
    return fc.to_host(ty), plan


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("kind", ["cos", "sin", "cos1"])
def test_known_answers_3d(hmi, cuda, golden, dtype, kind):
    x = np.arange(1.0, 25.0)
    y, plan = run_r2r(hmi, cuda, kind, (2, 3, 4), (0, 1, 2), (0, 1, 2), x, dtype)
    ref = np.array(golden["r2r_3d"][kind])
    tol = 1e-2 if dtype == "float32" else 1e-9  # absolute, on values up to 2.4e3 (reference tolerance scale)
    assert np.abs(y - ref).max() <= tol
    back, _ = run_r2r(hmi, cuda, kind, (2, 3, 4), (0, 1, 2), (0, 1, 2), y, dtype, "backward", hmi.scale.full)
    assert np.abs(back - x).max() <= (1e-3 if dtype == "float32" else 1e-11)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("name", ["cos_len4", "cos_len5", "sin_len4", "sin_len5"])
def test_known_answers_1d(hmi, cuda, golden, dtype, name):
    g = golden["r2r_1d"][name]
    x = np.array(g["input"])
    n = x.size
    kind = name[:3]
    y, _ = run_r2r(hmi, cuda, kind, (n, 1, 1), (0, 1, 2), (0, 1, 2), x, dtype)
    # the two length-1 axes each multiply by 2 (DCT-II / DST-II of one sample is 2 x0)
    ref = 4.0 * np.array(g["output"])
    assert np.abs(y - ref).max() <= (1e-4 if dtype == "float32" else 1e-11) * np.abs(ref).max()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("kind", ["cos", "sin"])
@pytest.mark.parametrize("size", [(8, 9, 10), (16, 12, 5), (64, 32, 17), (43, 2, 7), (256, 8, 4)])
def test_vs_oracle(hmi, cuda, oracle, dtype, kind, size):
    rdt, tol = RDT[dtype]
    rng = np.random.default_rng(sum(size))
    fn = oracle.r2r_cos if kind == "cos" else oracle.r2r_sin
    for inorder, outorder, in_place in [((0, 1, 2), (0, 1, 2), False), ((0, 1, 2), (0, 1, 2), True),
                                        ((2, 0, 1), (1, 2, 0), False), ((1, 0, 2), (2, 1, 0), True)]:
        n = int(np.prod(size))
        x = rng.standard_normal(n)
        y, plan = run_r2r(hmi, cuda, kind, size, inorder, outorder, x, dtype, in_place=in_place)
        bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], inorder)
        bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], outorder)
        ref = fn(bi, bo, x, -1)
        assert fc.rel_l2(y, ref) <= tol, (inorder, outorder, in_place)
        # backward (2x DCT-III / DST-III) with scale::full returns the input
        z, _ = run_r2r(hmi, cuda, kind, size, inorder, outorder, y, dtype, "backward", hmi.scale.full,
                       in_place=in_place)
        assert fc.rel_l2(z, x) <= tol, (inorder, outorder, in_place)
        refb = fn(bo, bi, ref, +1, 1.0 / (64 * n))
        assert fc.rel_l2(z, refb) <= tol


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_batch_symmetric_scale(hmi, cuda, oracle, dtype):
    rdt, tol = RDT[dtype]
    size, batch = (12, 10, 8), 3
    n = int(np.prod(size))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(batch * n)
    y, _ = run_r2r(hmi, cuda, "cos", size, (0, 1, 2), (2, 1, 0), x, dtype, scaling=hmi.scale.symmetric,
                   batch=batch, use_ws=False)
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], (0, 1, 2))
    bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], (2, 1, 0))
    ref = oracle.r2r_cos(bi, bo, x, -1, 1.0 / np.sqrt(64.0 * n), batch)
    assert fc.rel_l2(y, ref) <= tol


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("size", [(9, 17, 5), (33, 65, 3), (257, 5, 6), (12, 7, 2), (1025, 3, 2), (129, 129, 9)])
def test_cos1_vs_oracle(hmi, cuda, oracle, dtype, size):
    """DCT-I: kernel FFT lengths 2(n-1) from 2 to 2048, compile-time and runtime-plan passes."""
    rdt, tol = RDT[dtype]
    rng = np.random.default_rng(sum(size) + 1)
    for inorder, outorder, in_place in [((0, 1, 2), (0, 1, 2), False), ((0, 1, 2), (0, 1, 2), True),
                                        ((2, 0, 1), (1, 2, 0), False)]:
        n = int(np.prod(size))
        x = rng.standard_normal(n)
        y, plan = run_r2r(hmi, cuda, "cos1", size, inorder, outorder, x, dtype, in_place=in_place)
        bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], inorder)
        bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], outorder)
        assert fc.rel_l2(y, oracle.r2r_cos1(bi, bo, x, -1)) <= tol, (inorder, outorder, in_place)
        z, _ = run_r2r(hmi, cuda, "cos1", size, inorder, outorder, y, dtype, "backward", hmi.scale.full,
                       in_place=in_place)
        assert fc.rel_l2(z, x) <= tol, (inorder, outorder, in_place)


def test_cos1_rejects_single_point_axis(hmi):
    b = hmi.box3d([0, 0, 0], [7, 0, 7])
    with pytest.raises(Exception):
        hmi.rtransform(hmi.backend.mi355, b, b, hmi.kind.cos1)
