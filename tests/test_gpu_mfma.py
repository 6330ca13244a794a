"""GPU parity of the split-precision MFMA form of the 512-point complex<float> passes
(HMI_PRECISION_SPLIT_FP16: two FP16 terms, 3 products; HMI_PRECISION_SPLIT_BF16: three BF16
terms, 6 products; FP32 accumulation, a power-of-two scale per 32- or 16-point DFT) against the
FP64 oracle, bound 5e-6 relative L2, including inputs with 6 decades of dynamic range, partial
tiles (a line count that is not a multiple of 16) and batches."""
import numpy as np
import pytest

import fftcase as fc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["split_fp16", "split_bf16"])
@pytest.mark.parametrize("size,inorder,outorder", [((512, 24, 16), (0, 1, 2), (0, 1, 2)),
                                                   ((512, 24, 16), (0, 1, 2), (2, 0, 1)),
                                                   ((16, 512, 20), (1, 2, 0), (0, 1, 2)),
                                                   ((8, 20, 512), (2, 1, 0), (1, 0, 2)),
                                                   ((512, 512, 4), (0, 1, 2), (1, 2, 0)),
                                                   ((512, 7, 3), (0, 1, 2), (1, 0, 2)),
                                                   ((5, 512, 3), (1, 0, 2), (0, 2, 1))])
def test_split_forms_vs_oracle(hmi, cuda, oracle, form, size, inorder, outorder):
    import torch
    n = int(np.prod(size))
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], inorder)
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], outorder)
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    plan.set_precision(getattr(hmi.precision, form))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], inorder)
    bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], outorder)
    for in_place in (False, True):
        tx = fc.to_gpu(x, np.complex64)
        ty = tx if in_place else torch.empty_like(tx)
        plan.forward(tx, ty)
        ref = oracle.fft3d_c2c(bi, bo, x.astype(np.complex64).astype(np.complex128), -1, 1.0)
        err = fc.rel_l2(fc.to_host(ty), ref)
        print(form, size, inorder, outorder, in_place, f"forward rel L2 {err:.2e}")
        assert err <= 5e-6
        tz = torch.empty_like(ty)
        plan.backward(ty, tz, hmi.scale.full)
        assert fc.rel_l2(fc.to_host(tz), x) <= 5e-6


@pytest.mark.parametrize("form", ["split_fp16", "split_bf16"])
def test_split_forms_dynamic_range(hmi, cuda, oracle, form):
    """Elements whose magnitudes span 6 decades (1e-3 .. 1e3): the per-DFT scaling keeps
    both split terms in range."""
    import torch
    size = (512, 32, 8)
    n = int(np.prod(size))
    rng = np.random.default_rng(11)
    mag = 10.0 ** rng.uniform(-3, 3, n)
    x = mag * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    box = hmi.box3d([0, 0, 0], [s - 1 for s in size])
    plan = hmi.fft3d(hmi.backend.mi355, box, box)
    plan.set_precision(getattr(hmi.precision, form))
    tx = fc.to_gpu(x, np.complex64)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    b = oracle.Box.make((0, 0, 0), [s - 1 for s in size])
    ref = oracle.fft3d_c2c(b, b, x.astype(np.complex64).astype(np.complex128), -1, 1.0)
    err = fc.rel_l2(fc.to_host(ty), ref)
    print(form, f"6-decade forward rel L2 {err:.2e}")
    assert err <= 5e-6


def test_split_form_ignored_for_double_and_other_lengths(hmi, cuda, oracle):
    """complex<double> and lengths other than 512 keep the FP64 / FP32 vector passes."""
    import torch
    size = (64, 512, 6)
    n = int(np.prod(size))
    box = hmi.box3d([0, 0, 0], [s - 1 for s in size])
    plan = hmi.fft3d(hmi.backend.mi355, box, box)
    plan.set_precision(hmi.precision.split_fp16)
    x = fc.uniform_complex(n, 1)
    tx = fc.to_gpu(x, np.complex128)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    b = oracle.Box.make((0, 0, 0), [s - 1 for s in size])
    assert fc.rel_l2(fc.to_host(ty), oracle.fft3d_c2c(b, b, x, -1, 1.0)) <= 1e-13


def power_law_field(size, seed, slope=-5.0 / 3.0):
    """Complex Gaussian random field with power spectrum |k|^slope (x fastest, flattened)."""
    rng = np.random.default_rng(seed)
    n0, n1, n2 = size
    k0 = np.fft.fftfreq(n0) * n0
    k1 = np.fft.fftfreq(n1) * n1
    k2 = np.fft.fftfreq(n2) * n2
    kk = np.sqrt(k2[:, None, None] ** 2 + k1[None, :, None] ** 2 + k0[None, None, :] ** 2)
    kk[0, 0, 0] = 1.0
    amp = kk ** (slope / 2.0)
    white = rng.standard_normal((n2, n1, n0)) + 1j * rng.standard_normal((n2, n1, n0))
    return np.fft.ifftn(white * amp).reshape(-1)


@pytest.mark.parametrize("form", ["fp32", "split_fp16", "split_bf16"])
def test_power_law_field(hmi, cuda, oracle, form):
    """Gaussian random field with a |k|^-5/3 spectrum (seed 3): relative L2 within 5e-6 and the
    maximum error relative to the largest output magnitude stated."""
    import torch
    size = (512, 64, 16)
    x = power_law_field(size, 3)
    box = hmi.box3d([0, 0, 0], [s - 1 for s in size])
    plan = hmi.fft3d(hmi.backend.mi355, box, box)
    plan.set_precision(getattr(hmi.precision, form))
    tx = fc.to_gpu(x, np.complex64)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    b = oracle.Box.make((0, 0, 0), [s - 1 for s in size])
    ref = oracle.fft3d_c2c(b, b, x.astype(np.complex64).astype(np.complex128), -1, 1.0)
    y = fc.to_host(ty)
    err = fc.rel_l2(y, ref)
    mx = np.abs(y - ref).max() / np.abs(ref).max()
    print(form, f"|k|^-5/3 field forward rel L2 {err:.2e}, max {mx:.2e}")
    assert err <= 5e-6
    assert mx <= 5e-6


@pytest.mark.parametrize("form", ["split_fp16", "split_bf16"])
def test_split_forms_batch(hmi, cuda, oracle, form):
    """batch = 3 transforms of 512 x 9 x 5 (45 lines per axis pass and transform: partial
    tiles), out-of-place forward and in-place backward."""
    import torch
    size, batch = (512, 9, 5), 3
    n = int(np.prod(size))
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], (0, 1, 2))
    outbox = hmi.box3d([0, 0, 0], [s - 1 for s in size], (2, 0, 1))
    plan = hmi.fft3d(hmi.backend.mi355, inbox, outbox)
    plan.set_precision(getattr(hmi.precision, form))
    rng = np.random.default_rng(21)
    x = rng.standard_normal(batch * n) + 1j * rng.standard_normal(batch * n)
    xr = x.astype(np.complex64).astype(np.complex128)
    tx = fc.to_gpu(x, np.complex64)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty, batch=batch)
    bi = oracle.Box.make((0, 0, 0), [s - 1 for s in size], (0, 1, 2))
    bo = oracle.Box.make((0, 0, 0), [s - 1 for s in size], (2, 0, 1))
    ref = oracle.fft3d_c2c(bi, bo, xr, -1, 1.0, batch)
    assert fc.rel_l2(fc.to_host(ty), ref) <= 5e-6
    plan.backward(ty, ty, hmi.scale.full, batch=batch)
    assert fc.rel_l2(fc.to_host(ty), xr) <= 5e-6
