"""GPU parity of fft3d_r2c (real <-> half spectrum) against the FP64 oracle.

Cases: the 1D known answers of test/test_units_nompi.cpp:336-367 (2x3x4 box,
input 1..24, r2c along each axis, backward divided by the length), the 3D
sizes of test/test_fft3d_r2c.cpp:286-312 style boxes along every r2c axis and
order, odd lengths, and the c2r round trip.
"""
import numpy as np
import pytest

import fftcase as fc

pytestmark = pytest.mark.gpu

RDT = {"float32": (np.float32, 5e-6), "float64": (np.float64, 1e-13)}


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_known_answers_2x3x4(hmi, cuda, golden, dtype, axis):
    """A 3D r2c on a box with one non-trivial axis is the 1D r2c executor."""
    import torch
    rdt, tol = RDT[dtype]
    full = [1, 1, 1]
    full[axis] = (2, 3, 4)[axis]
    # run the reference 2x3x4 lines one axis at a time: embed each line in a 3D box
    x = np.array(golden["box_2x3x4"]["input"], dtype=np.float64).reshape(4, 3, 2).transpose(2, 1, 0)
    size = (2, 3, 4)
    hs = list(size)
    hs[axis] = size[axis] // 2 + 1
    out = np.empty(hs, dtype=np.complex128)
    inbox = hmi.box3d([0, 0, 0], [f - 1 for f in full])
    plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, inbox.r2c(axis), axis)
    for idx in np.ndindex(*[s for d, s in enumerate(size) if d != axis]):
        sl = list(idx)
        sl.insert(axis, slice(None))
        line = torch.from_numpy(x[tuple(sl)].astype(rdt)).to(cuda)
        res = torch.empty(hs[axis], dtype=torch.complex64 if rdt == np.float32 else torch.complex128, device=cuda)
        plan.forward(line, res)
        out[tuple(sl)] = fc.to_host(res)
        back = torch.empty_like(line)
        plan.backward(res, back)
        assert np.abs(fc.to_host(back) / size[axis] - x[tuple(sl)]).max() < (1e-4 if rdt == np.float32 else 1e-12)
    ref = np.array([complex(a, b) for a, b in golden["r2c_1d"][f"axis{axis}"]])
    got = fc.flat(out, (0, 1, 2))
    assert np.abs(got - ref).max() < (5e-4 if rdt == np.float32 else 1e-11)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("size", [(6, 7, 8), (7, 8, 6), (16, 10, 9), (64, 32, 16), (33, 20, 12), (1024, 8, 6),
                                  (128, 96, 24)])
@pytest.mark.parametrize("r2c_dir", [0, 1, 2])
def test_r2c_c2r(hmi, cuda, oracle, dtype, size, r2c_dir):
    import torch
    rdt, tol = RDT[dtype]
    for inorder, outorder in [((0, 1, 2), (0, 1, 2)), ((0, 1, 2), (2, 0, 1)), ((1, 2, 0), (1, 0, 2))]:
        inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], inorder)
        outbox = inbox.r2c(r2c_dir)
        outbox = hmi.box3d(outbox.low, outbox.high, outorder)
        plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, outbox, r2c_dir)
        rng = np.random.default_rng(sum(size) + r2c_dir)
        x = rng.standard_normal(inbox.count())
        tx = torch.from_numpy(x.astype(rdt)).to(cuda)
        ty = torch.empty(outbox.count(), dtype=torch.complex64 if rdt == np.float32 else torch.complex128,
                         device=cuda)
        ws = torch.empty(max(1, plan.size_workspace()), dtype=ty.dtype, device=cuda)
        plan.forward_buffered(tx, ty, ws)
        bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], inorder)
        bo = oracle.Box.make(outbox.low, outbox.high, outorder)
        ref = oracle.fft3d_r2c(bi, bo, r2c_dir, x)
        assert fc.rel_l2(fc.to_host(ty), ref) <= tol, (inorder, outorder)
        back = torch.empty_like(tx)
        plan.backward_buffered(ty, back, ws, hmi.scale.full)
        assert fc.rel_l2(fc.to_host(back), x) <= tol, (inorder, outorder)
        # c2r against the oracle on a Hermitian-consistent spectrum (the forward result)
        ref_b = oracle.fft3d_c2r(bi, bo, r2c_dir, fc.to_host(ty), 1.0)
        back2 = torch.empty_like(tx)
        plan.backward(ty, back2)
        assert fc.rel_l2(fc.to_host(back2), ref_b) <= tol


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("n0", [8192, 4096])
def test_r2c_long_axis_across_columns(hmi, cuda, oracle, dtype, n0):
    """r2c along an 8192- or 4096-point axis that is not the fastest of the input: the r2c pass
    (a complex FFT of n0 / 2 points) reads its lines in column chunks of few lines (4096 points:
    4 lines for float, 2 for double, the tiles that share cache lines numbered in groups per XCD:
    one full group of 32 / two of 64 workgroups and a partial group over the 156 lines; 2048
    points: 8 / 4 lines); the c2r pass of the backward transform writes the same way."""
    import torch
    rdt, tol = RDT[dtype]
    size, inorder = (n0, 12, 13), (1, 0, 2)
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], inorder)
    outbox = inbox.r2c(0)
    plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, outbox, 0)
    x = np.random.default_rng(31).standard_normal(inbox.count())
    tx = torch.from_numpy(x.astype(rdt)).to(cuda)
    ty = torch.empty(outbox.count(), dtype=torch.complex64 if rdt == np.float32 else torch.complex128, device=cuda)
    plan.forward(tx, ty)
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], inorder)
    bo = oracle.Box.make(outbox.low, outbox.high, inorder)
    ref = oracle.fft3d_r2c(bi, bo, 0, x)
    assert fc.rel_l2(fc.to_host(ty), ref) <= tol
    ref_b = oracle.fft3d_c2r(bi, bo, 0, fc.to_host(ty), 1.0)
    back = torch.empty_like(tx)
    plan.backward(ty, back)
    assert fc.rel_l2(fc.to_host(back), ref_b) <= tol


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_r2c_large_round_trip(hmi, cuda, dtype):
    """1024 x 1024 x 64 real (the r2c along the first, contiguous axis of the 1024 x 1024 x 512
    configuration, fewer planes): c2r(r2c(x)) / N == x and Parseval on the half spectrum."""
    import torch
    rdt, tol = RDT[dtype]
    size = (1024, 1024, 64)
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size])
    outbox = inbox.r2c(0)
    plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, outbox, 0)
    g = torch.Generator(device=cuda)
    g.manual_seed(2)
    x = torch.randn(inbox.count(), dtype=torch.float32 if rdt == np.float32 else torch.float64, device=cuda,
                    generator=g)
    y = torch.empty(outbox.count(), dtype=torch.complex64 if rdt == np.float32 else torch.complex128, device=cuda)
    ws = torch.empty(max(1, plan.size_workspace()), dtype=y.dtype, device=cuda)
    plan.forward_buffered(x, y, ws)
    z = torch.empty_like(x)
    plan.backward_buffered(y, z, ws, hmi.scale.full)
    err = (torch.linalg.vector_norm((z - x).double()) / torch.linalg.vector_norm(x.double())).item()
    assert err <= tol, err
    # Parseval: sum |x|^2 = (|X_0|^2 + 2 sum_{0<k<n/2} |X_k|^2 + |X_{n/2}|^2) / N along the halved axis
    Y = y.view(size[2], size[1], size[0] // 2 + 1)
    p = (Y.abs() ** 2).double()
    tot = p[..., 0].sum() + p[..., -1].sum() + 2 * p[..., 1:-1].sum()
    lhs = (x.double() ** 2).sum() * inbox.count()
    assert abs((tot - lhs) / lhs).item() <= 10 * tol


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_r2c_c2r_unaligned_real_arrays(hmi, cuda, oracle, dtype):
    """Real arrays starting one element past an aligned address: the contiguous real-data pass
    must fall back from paired (8/16-byte) accesses to single elements."""
    import torch
    rdt, tol = RDT[dtype]
    size = (64, 12, 10)
    n = int(np.prod(size))
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size])
    outbox = inbox.r2c(0)
    plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, outbox, 0)
    x = np.random.default_rng(11).standard_normal(n)
    tdt = torch.float32 if dtype == "float32" else torch.float64
    cdt = torch.complex64 if dtype == "float32" else torch.complex128
    buf = torch.zeros(n + 1, dtype=tdt, device=cuda)
    buf[1:] = torch.from_numpy(x).to(tdt)
    tx = buf[1:]  # odd element offset
    ty = torch.empty(outbox.count(), dtype=cdt, device=cuda)
    plan.forward(tx, ty)
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size])
    ref = oracle.fft3d_r2c(bi, oracle.Box.make((0, 0, 0), [size[0] // 2, size[1] - 1, size[2] - 1]), 0, x)
    assert fc.rel_l2(fc.to_host(ty), ref) <= tol
    out = torch.zeros(n + 1, dtype=tdt, device=cuda)
    plan.backward(ty, out[1:], hmi.scale.full)
    assert fc.rel_l2(fc.to_host(out[1:]), x) <= tol
    assert float(out[0]) == 0.0
