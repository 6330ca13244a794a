"""GPU parity of the complex-to-complex 3D transform (fft3d) against the FP64 oracle.

Bounds: relative L2 <= 1e-13 for complex<double>, <= 5e-6 for complex<float>
(forward and round trip).  Cases follow heFFTe's own tests: the np=1 sizes of
test/test_fft3d_np1.cpp:17-21 (6x7x8, 7x8x6), the batch case of
test/test_fft3d.h:505-572 (8x9x10, batch 5), the cross-backend size
43x76x24 of test/test_units_nompi.cpp:673-809, every in/out order pair, in-place
and out-of-place calls, caller and internal workspace, and the real-input /
real-output overloads of test_fft3d_arrays (test/test_fft3d.h:397-503).
"""
import numpy as np
import pytest

import fftcase as fc

pytestmark = pytest.mark.gpu

CDT = {"complex64": np.complex64, "complex128": np.complex128}


def run_c2c(hmi, cuda, size, inorder, outorder, dtype, seed=0, batch=1, in_place=False, use_ws=True,
            scaling=0, direction="forward"):
    import torch
    n = int(np.prod(size))
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], inorder)
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], outorder)
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    xs = [fc.uniform_complex(n, seed + b) for b in range(batch)]
    x = np.concatenate(xs)  # stored in the order of the source box
    tx = fc.to_gpu(x, CDT[dtype])
    ty = tx if in_place else torch.empty(batch * n, dtype=tx.dtype, device=cuda)
    ws = torch.empty(max(1, batch * plan.size_workspace()), dtype=tx.dtype, device=cuda) if use_ws else None
    if direction == "forward":
        plan.forward_buffered(tx, ty, ws, scaling, batch=batch) if use_ws else plan.forward(tx, ty, scaling, batch)
    else:
        plan.backward_buffered(tx, ty, ws, scaling, batch=batch) if use_ws else plan.backward(tx, ty, scaling, batch)
    y = fc.to_host(ty)
    return x, y, plan


def oracle_c2c(oracle, size, inorder, outorder, x, sign, scale, batch=1):
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], inorder)
    bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], outorder)
    return oracle.fft3d_c2c(bi, bo, x, sign, scale, batch)


def test_config1_64cubed_double(hmi, cuda, oracle):
    """64^3 complex<double>, input order (0,1,2), output (2,1,0), U[-1,1] seed 0;
    forward scale::none then backward scale::full."""
    import torch
    size = (64, 64, 64)
    n = 64 ** 3
    inbox = hmi.box3d([0, 0, 0], [63, 63, 63], [0, 1, 2])
    outbox = hmi.box3d([0, 0, 0], [63, 63, 63], [2, 1, 0])
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    x = fc.uniform_complex(n, 0)
    tx = fc.to_gpu(x, np.complex128)
    ty = torch.empty_like(tx)
    ws = torch.empty(max(1, plan.size_workspace()), dtype=tx.dtype, device=cuda)
    plan.forward_buffered(tx, ty, ws, hmi.scale.none)
    ref = oracle_c2c(oracle, size, (0, 1, 2), (2, 1, 0), x, -1, 1.0)
    err = fc.rel_l2(fc.to_host(ty), ref)
    print(f"64^3 z2z forward rel L2 = {err:.3e}")
    assert err <= 1e-13
    tz = torch.empty_like(tx)
    plan.backward_buffered(ty, tz, ws, hmi.scale.full)
    err_rt = fc.rel_l2(fc.to_host(tz), x)
    print(f"64^3 z2z round trip rel L2 = {err_rt:.3e}")
    assert err_rt <= 1e-13


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("size", [(6, 7, 8), (7, 8, 6), (43, 76, 24), (1, 1, 1), (1, 16, 1), (32, 1, 12)])
def test_reference_sizes(hmi, cuda, oracle, dtype, size):
    for inorder, outorder in [((0, 1, 2), (0, 1, 2)), ((0, 1, 2), (2, 1, 0)), ((1, 2, 0), (0, 2, 1))]:
        x, y, _ = run_c2c(hmi, cuda, size, inorder, outorder, dtype, seed=11)
        ref = oracle_c2c(oracle, size, inorder, outorder, x, -1, 1.0)
        err = fc.rel_l2(y, ref)
        assert err <= fc.TOL[dtype], (size, inorder, outorder, err)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("oi", range(6))
@pytest.mark.parametrize("oo", range(6))
def test_all_order_pairs(hmi, cuda, oracle, dtype, oi, oo):
    size = (16, 12, 10)
    inorder, outorder = fc.ORDERS[oi], fc.ORDERS[oo]
    for in_place in (False, True):
        x, y, _ = run_c2c(hmi, cuda, size, inorder, outorder, dtype, seed=oi * 6 + oo, in_place=in_place)
        ref = oracle_c2c(oracle, size, inorder, outorder, x, -1, 1.0)
        err = fc.rel_l2(y, ref)
        assert err <= fc.TOL[dtype], (inorder, outorder, in_place, err)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_batch_backward_scales(hmi, cuda, oracle, dtype):
    """8x9x10, batch 5, all three scalings, backward direction, internal workspace."""
    size = (8, 9, 10)
    n = 720
    for scaling, s in ((0, 1.0), (1, 1.0 / n), (2, 1.0 / np.sqrt(n))):
        x, y, _ = run_c2c(hmi, cuda, size, (0, 1, 2), (2, 0, 1), dtype, seed=5, batch=5, use_ws=False,
                          scaling=scaling, direction="backward")
        ref = oracle_c2c(oracle, size, (2, 0, 1), (0, 1, 2), x, +1, s, batch=5)
        assert fc.rel_l2(y, ref) <= fc.TOL[dtype], scaling


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("size", [(64, 64, 64), (128, 32, 256), (512, 8, 4), (100, 45, 27), (96, 80, 63)])
def test_fast_and_generic_lengths(hmi, cuda, oracle, dtype, size):
    for inorder, outorder in [((0, 1, 2), (1, 2, 0)), ((2, 1, 0), (0, 1, 2))]:
        for in_place in (False, True):
            x, y, _ = run_c2c(hmi, cuda, size, inorder, outorder, dtype, seed=3, in_place=in_place)
            ref = oracle_c2c(oracle, size, inorder, outorder, x, -1, 1.0)
            err = fc.rel_l2(y, ref)
            assert err <= fc.TOL[dtype], (size, inorder, outorder, in_place, err)


@pytest.mark.parametrize("dtype,size", [("complex64", (8192, 2, 2)), ("complex64", (10240, 1, 2)),
                                        ("complex128", (5120, 2, 1)), ("complex64", (1021, 3, 2)),
                                        ("complex128", (2, 1021, 3)), ("complex64", (4096, 4, 3))])
def test_long_and_prime_lengths(hmi, cuda, oracle, dtype, size):
    """Lengths up to the single-pass LDS capacity (2 n element sizes <= 160 KiB: 10240 complex<float>,
    5120 complex<double>) and a large prime (1021: direct DFT stage of the runtime plan)."""
    for inorder, outorder in [((0, 1, 2), (0, 1, 2)), ((0, 1, 2), (2, 0, 1))]:
        x, y, _ = run_c2c(hmi, cuda, size, inorder, outorder, dtype, seed=8)
        ref = oracle_c2c(oracle, size, inorder, outorder, x, -1, 1.0)
        assert fc.rel_l2(y, ref) <= fc.TOL[dtype], (size, inorder, outorder)


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("size", [(1024, 20, 16), (2048, 12, 8), (16, 1024, 3)])
def test_wide_tiles_of_long_lines(hmi, cuda, oracle, dtype, size):
    """Transposing transforms with 1024- and 2048-point axes: a 1024-point pass whose load or store
    walks the column axis uses the widest tile (16 x 8 B or 8 x 16 B column chunks); the
    2048-point axis (whose widest tiles have 64-byte chunks) is scheduled on contiguous lines."""
    for inorder, outorder in [((0, 1, 2), (1, 2, 0)), ((2, 1, 0), (0, 1, 2)), ((1, 0, 2), (2, 0, 1))]:
        x, y, _ = run_c2c(hmi, cuda, size, inorder, outorder, dtype, seed=6)
        ref = oracle_c2c(oracle, size, inorder, outorder, x, -1, 1.0)
        assert fc.rel_l2(y, ref) <= fc.TOL[dtype], (size, inorder, outorder)


@pytest.mark.parametrize("dtype,size,tiles", [
    ("complex64", (4096, 16, 8), ("radices<16, 16, 16>, 16, 1,", "radices<16, 16, 16>, 16, 4,")),
    ("complex128", (4096, 8, 8), ("radices<16, 16, 16>, 16, 1,", "radices<16, 16, 16>, 16, 2,")),
    ("complex64", (2048, 16, 8), ("radices<8, 16, 16>, 16, 2,", "radices<8, 16, 16>, 16, 8,"))])
def test_long_axis_keeps_contiguous_lines(hmi, cuda, oracle, dtype, size, tiles):
    """A transposing transform whose long axis would run few-line tiles with column chunks
    narrower than a 128-byte line (4096 complex<float>: 4 lines of 32 bytes) is scheduled with
    that axis on contiguous lines on both sides; the short axes carry the transposition.  The
    schedule shows in the kernels the passes launch; the result is checked against the oracle."""
    import torch
    contiguous, narrow = tiles
    n = int(np.prod(size))
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], (0, 1, 2))
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], (1, 2, 0))
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    x = fc.uniform_complex(n, 21)
    tx = fc.to_gpu(x, CDT[dtype])
    ty = torch.empty(n, dtype=tx.dtype, device=cuda)
    plan.set_timing(True)
    plan.forward(tx, ty)
    names = list(plan.timing())
    plan.set_timing(False)
    assert any(contiguous in k for k in names), names
    assert not any(narrow in k for k in names), names
    ref = oracle_c2c(oracle, size, (0, 1, 2), (1, 2, 0), x, -1, 1.0)
    assert fc.rel_l2(fc.to_host(ty), ref) <= fc.TOL[dtype]


@pytest.mark.parametrize("rdt,cdt", [(np.float32, np.complex64), (np.float64, np.complex128)])
def test_real_input_and_real_output(hmi, cuda, oracle, rdt, cdt):
    """fft3d::forward(real in, complex out) and backward(complex in, real out),
    including the in-place complex -> real call of test_fft3d_arrays."""
    import torch
    size = (12, 10, 16)
    n = int(np.prod(size))
    inbox = hmi.box3d([0, 0, 0], [11, 9, 15], (0, 1, 2))
    outbox = hmi.box3d([0, 0, 0], [11, 9, 15], (2, 1, 0))
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    rng = np.random.default_rng(8)
    xr = rng.uniform(-1, 1, n)
    t_in = torch.from_numpy(xr.astype(rdt)).to(cuda)
    t_out = torch.empty(n, dtype=torch.complex64 if cdt == np.complex64 else torch.complex128, device=cuda)
    plan.forward(t_in, t_out)
    bi = oracle.Box.make((0, 0, 0), (11, 9, 15), (0, 1, 2))
    bo = oracle.Box.make((0, 0, 0), (11, 9, 15), (2, 1, 0))
    ref = oracle.fft3d_real_c2c(bi, bo, xr)
    tol = fc.TOL["complex64" if cdt == np.complex64 else "complex128"]
    assert fc.rel_l2(fc.to_host(t_out), ref) <= tol
    # complex -> real (truncation of the imaginary part), out of place and in place
    back = torch.empty(n, dtype=t_in.dtype, device=cuda)
    plan.backward(t_out, back, hmi.scale.full)
    assert fc.rel_l2(fc.to_host(back), xr) <= tol
    buf = t_out.clone()
    plan.backward(buf, buf.view(t_in.dtype)[:n], hmi.scale.full)
    assert fc.rel_l2(fc.to_host(buf.view(t_in.dtype)[:n]), xr) <= tol


def test_stream_binding(hmi, cuda, oracle):
    """The plan runs on the bound stream (fft3d(gpu_stream, ...))."""
    import torch
    size = (32, 32, 32)
    s = torch.cuda.Stream()
    inbox = hmi.box3d([0, 0, 0], [31, 31, 31])
    plan = hmi.fft3d(hmi.backend.mi355, inbox, inbox)
    plan.set_stream(s)
    x = fc.uniform_complex(32 ** 3, 4)
    tx = fc.to_gpu(x, np.complex128)
    ty = torch.empty_like(tx)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        plan.forward(tx, ty)
    s.synchronize()
    ref = oracle_c2c(oracle, size, (0, 1, 2), (0, 1, 2), x, -1, 1.0)
    assert fc.rel_l2(ty.cpu().numpy(), ref) <= 1e-13


def test_pass_timing_report(hmi, cuda, oracle):
    """Per-pass HIP-event timing: one record per pass launch, named by its kernel; results are
    unchanged while timing is on, and switching it off drops the records."""
    import torch
    size = (64, 48, 40)
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], [0, 1, 2])
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], [1, 2, 0])
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    x = fc.uniform_complex(int(np.prod(size)), 6)
    tx = fc.to_gpu(x, np.complex64)
    ty = torch.empty_like(tx)
    ws = torch.empty(plan.size_workspace(), dtype=tx.dtype, device=cuda)
    plan.set_timing(True)
    for _ in range(3):
        plan.forward_buffered(tx, ty, ws)
    rep = plan.timing()
    assert sum(v[0] for v in rep.values()) == 9, rep  # 3 calls x 3 passes
    assert all(v[1] > 0 for v in rep.values())
    assert all(k.startswith(("fft_uniform<float", "fft_generic<float")) for k in rep)
    ref = oracle_c2c(oracle, size, (0, 1, 2), (1, 2, 0), x, -1, 1.0)
    assert fc.rel_l2(ty.cpu().numpy(), ref) <= 5e-6
    plan.set_timing(False)
    plan.forward_buffered(tx, ty, ws)
    assert plan.timing() == {}


@pytest.mark.parametrize("dtype,size,orders", [
    ("complex64", (16384, 4, 2), ((0, 1, 2), (0, 1, 2))),
    ("complex64", (16384, 4, 2), ((0, 1, 2), (2, 1, 0))),
    ("complex64", (3, 12288, 2), ((1, 2, 0), (0, 1, 2))),
    ("complex64", (11025, 2, 2), ((0, 1, 2), (1, 0, 2))),
    ("complex128", (8192, 3, 2), ((0, 1, 2), (0, 1, 2))),
    ("complex128", (2, 3, 8192), ((2, 0, 1), (1, 2, 0))),
    ("complex128", (5376, 2, 2), ((0, 1, 2), (2, 0, 1))),
])
def test_split_lengths(hmi, cuda, oracle, dtype, size, orders):
    """Lengths beyond one LDS tile (complex<float> > 10240, complex<double> > 5120) run as two
    passes, n = n1 n2 (16384 = 128 x 128, 12288 = 96 x 128, 11025 = 105 x 105 on runtime-plan
    steps, 8192 = 64 x 128, 5376 = 42 x 128): forward against the oracle out of and in place, a
    batch of 2, and the backward transform with scale::full."""
    inorder, outorder = orders
    tol = fc.TOL[dtype]
    for in_place in (False, True):
        x, y, _ = run_c2c(hmi, cuda, size, inorder, outorder, dtype, seed=12, in_place=in_place)
        ref = oracle_c2c(oracle, size, inorder, outorder, x, -1, 1.0)
        assert fc.rel_l2(y, ref) <= tol, (in_place, fc.rel_l2(y, ref))
    n = int(np.prod(size))
    # a plan from outorder to inorder, run backward: data in inorder -> outorder
    x, y, _ = run_c2c(hmi, cuda, size, outorder, inorder, dtype, seed=13, batch=2, scaling=1, direction="backward")
    ref = oracle_c2c(oracle, size, inorder, outorder, x, +1, 1.0 / n, batch=2)
    assert fc.rel_l2(y, ref) <= tol


def test_split_length_real_input_and_r2c(hmi, cuda, oracle):
    """A split axis behind the real-input overload of fft3d and as a complex axis of fft3d_r2c."""
    import torch
    size = (16384, 3, 2)
    n = int(np.prod(size))
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], (0, 1, 2))
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], (2, 1, 0))
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    xr = np.random.default_rng(14).uniform(-1, 1, n)
    t_out = torch.empty(n, dtype=torch.complex64, device=cuda)
    plan.forward(torch.from_numpy(xr.astype(np.float32)).to(cuda), t_out)
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], (0, 1, 2))
    bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], (2, 1, 0))
    assert fc.rel_l2(fc.to_host(t_out), oracle.fft3d_real_c2c(bi, bo, xr)) <= 5e-6
    # r2c along axis 1 of a box whose axis 0 is split
    size = (12288, 6, 2)
    rin = hmi.box3d([0, 0, 0], [s - 1 for s in size])
    rout = rin.r2c(1)
    rplan = hmi.fft3d_r2c(hmi.backend.mi355, rin, rout, 1)
    x = np.random.default_rng(15).standard_normal(rin.count())
    tx = torch.from_numpy(x).to(cuda)
    ty = torch.empty(rout.count(), dtype=torch.complex128, device=cuda)
    rplan.forward(tx, ty)
    ref = oracle.fft3d_r2c(oracle.Box.make((0, 0, 0), [s - 1 for s in size]),
                           oracle.Box.make(rout.low, rout.high), 1, x)
    assert fc.rel_l2(fc.to_host(ty), ref) <= 1e-13
    back = torch.empty_like(tx)
    rplan.backward(ty, back, hmi.scale.full)
    assert fc.rel_l2(fc.to_host(back), x) <= 1e-13
