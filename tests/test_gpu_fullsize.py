"""Full-size checks at the headline configurations, through properties that do not need the
CPU oracle (too slow at these sizes): round trips, Parseval, linearity, a plane wave landing in
one bin, and agreement of the complex<float> forms with the complex<double> path on the GPU
(itself pinned against the oracle at smaller sizes).  Bounds: 5e-6 relative L2 for
complex<float> and the split-precision forms, 1e-13 for complex<double>.

Workloads: 512^3 c2c, inbox order (0,1,2) -> outbox order (1,2,0), in place with a caller
workspace, scale::symmetric (complex<float> and complex<double>); r2c 1024x1024x512 float
along axis 0; DCT-II/III 256^3 double; 1,000 mixed-size c2c transforms (sides from the smooth
sizes 16..128)."""
import math

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N = 512
SMOOTH = [16, 20, 24, 27, 28, 32, 40, 45, 48, 56, 63, 64, 80, 96, 105, 112, 128]


def rel(a, b):
    import torch
    return (torch.linalg.vector_norm((a - b).to(torch.complex128)) /
            torch.linalg.vector_norm(b.to(torch.complex128))).item()


def c2c_plan(hmi):
    inbox = hmi.box3d([0, 0, 0], [N - 1] * 3, [0, 1, 2])
    outbox = hmi.box3d([0, 0, 0], [N - 1] * 3, [1, 2, 0])
    return hmi.fft3d(hmi.backend.mi355, inbox, outbox)


@pytest.fixture(scope="module")
def field512(cuda):
    """N(0,1) real and imaginary parts (torch generator seed 1), complex<double>."""
    import torch
    g = torch.Generator(device=cuda)
    g.manual_seed(1)
    return torch.randn(N ** 3, dtype=torch.complex128, device=cuda, generator=g) * math.sqrt(2.0)


@pytest.fixture(scope="module")
def spectrum512(hmi, cuda, field512):
    """forward(x) in complex<double> (scale none), outbox order (1,2,0)."""
    import torch
    plan = c2c_plan(hmi)
    y = torch.empty_like(field512)
    plan.forward(field512, y)
    torch.cuda.synchronize()
    return y


def test_512_double_round_trip_and_parseval(hmi, cuda, field512, spectrum512):
    import torch
    plan = c2c_plan(hmi)
    x = field512.clone()
    ws = torch.empty(max(1, plan.size_workspace()), dtype=x.dtype, device=cuda)
    plan.forward_buffered(x, x, ws, hmi.scale.symmetric)
    plan.backward_buffered(x, x, ws, hmi.scale.symmetric)
    assert rel(x, field512) <= 1e-13
    lhs = (spectrum512.abs() ** 2).sum().item()
    rhs = (field512.abs() ** 2).sum().item() * N ** 3
    assert abs(lhs - rhs) / rhs <= 1e-12


@pytest.mark.parametrize("form", ["fp32", "split_fp16", "split_bf16"])
def test_512_float_forms_vs_double(hmi, cuda, field512, spectrum512, form):
    """Each complex<float> form against the complex<double> transform of the same (rounded)
    input, then the in-place symmetric round trip of the headline step."""
    import torch
    plan = c2c_plan(hmi)
    plan.set_precision(getattr(hmi.precision, form))
    x32 = field512.to(torch.complex64)
    y = torch.empty_like(x32)
    ws = torch.empty(max(1, plan.size_workspace()), dtype=x32.dtype, device=cuda)
    plan.forward_buffered(x32, y, ws, hmi.scale.none)
    ref = c2c_plan(hmi)
    y64 = torch.empty_like(field512)
    ref.forward(x32.to(torch.complex128), y64)
    err = rel(y, y64)
    print(form, f"forward rel L2 vs complex<double> {err:.2e}")
    assert err <= 5e-6
    z = x32.clone()
    plan.forward_buffered(z, z, ws, hmi.scale.symmetric)
    plan.backward_buffered(z, z, ws, hmi.scale.symmetric)
    assert rel(z, x32) <= 5e-6


def test_512_float_linearity_and_plane_wave(hmi, cuda, field512):
    import torch
    plan = c2c_plan(hmi)
    x = field512.to(torch.complex64)
    g = torch.Generator(device=cuda)
    g.manual_seed(7)
    w = torch.randn(N ** 3, dtype=torch.complex64, device=cuda, generator=g)
    fx, fw, fs = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
    plan.forward(x, fx)
    plan.forward(w, fw)
    plan.forward(2.5 * x - 1j * w, fs)
    assert rel(fs, 2.5 * fx - 1j * fw) <= 5e-6
    # plane wave exp(2 pi i (k . r) / N) -> N^3 at bin k, zero elsewhere (outbox order (1,2,0):
    # index = i1 + N (i2 + N i0))
    k = (3, 17, 250)
    i = torch.arange(N, device=cuda, dtype=torch.float64)
    ph = (k[0] * i[None, None, :] + k[1] * i[None, :, None] + k[2] * i[:, None, None]) / N
    pw = torch.exp(2j * math.pi * ph).reshape(-1).to(torch.complex64)  # inbox order (0,1,2)
    f = torch.empty_like(pw)
    plan.forward(pw, f)
    peak = k[1] + N * (k[2] + N * k[0])
    assert abs(f[peak].item() - N ** 3) <= 1e-4 * N ** 3
    f[peak] = 0
    assert torch.linalg.vector_norm(f).item() <= 1e-5 * N ** 3


def plane_wave(torch, cuda, k, dtype, sign=1):
    """exp(sign 2 pi i (k . r) / N) on the N^3 grid in order (0,1,2) (axis 0 fastest)."""
    i = torch.arange(N, device=cuda, dtype=torch.float64)
    ph = (k[0] * i[None, None, :] + k[1] * i[None, :, None] + k[2] * i[:, None, None]) / N
    return torch.exp(sign * 2j * math.pi * ph).reshape(-1).to(dtype)


def test_512_double_plane_waves_forward_and_backward(hmi, cuda):
    """complex<double> at full size: exp(+2 pi i k.r / N) forwards to N^3 in the one bin k of the
    outbox order (1,2,0) (a conjugated spectrum would put it at -k, a wrong output permutation at
    another index), and a single bin backwards to exp(+2 pi i k.r / N) in the inbox order."""
    import torch
    plan = c2c_plan(hmi)
    k = (5, 211, 37)
    pw = plane_wave(torch, cuda, k, torch.complex128)
    f = torch.empty_like(pw)
    plan.forward(pw, f)
    peak = k[1] + N * (k[2] + N * k[0])
    assert abs(f[peak].item() - N ** 3) <= 1e-11 * N ** 3
    f[peak] = 0
    assert torch.linalg.vector_norm(f).item() <= 1e-11 * N ** 3
    # the backward transform (no scale) of a unit value at bin k is the plane wave itself
    f.zero_()
    f[peak] = 1.0
    z = torch.empty_like(f)
    plan.backward(f, z)
    assert rel(z, pw) <= 1e-13


CFG3_SIZE = (1024, 1024, 512)


def cfg3_terms():
    """The 16 plane-wave cosines of the real-transform configuration: integer wave vectors
    uniform in [0, 32)^3 and amplitudes U[0, 1] from numpy seed 2 (same draw order as bench.py)."""
    rng = np.random.default_rng(2)
    terms = []
    for _ in range(16):
        kv = rng.integers(0, 32, 3)
        terms.append((tuple(int(v) for v in kv), float(rng.uniform())))
    return terms


def cfg3_field(torch, cuda, dtype, with_noise=True):
    """sum_m a_m cos(2 pi k_m . r / n) + 1e-3 N(0,1) on 1024 x 1024 x 512 (axis 0 fastest),
    built in float64 and rounded to dtype; returns (field, sum of squared noise)."""
    n0, n1, n2 = CFG3_SIZE
    i = torch.arange(n0, device=cuda, dtype=torch.float64) / n0
    j = torch.arange(n1, device=cuda, dtype=torch.float64) / n1
    kk = torch.arange(n2, device=cuda, dtype=torch.float64) / n2
    x = torch.zeros((n2, n1, n0), device=cuda, dtype=torch.float64)
    for kv, amp in cfg3_terms():
        x += amp * torch.cos(2 * math.pi * (kv[2] * kk[:, None, None] + kv[1] * j[None, :, None] +
                                            kv[0] * i[None, None, :]))
    noise2 = 0.0
    if with_noise:
        g = torch.Generator(device=cuda)
        g.manual_seed(2)
        nz = 1e-3 * torch.randn(x.shape, device=cuda, dtype=torch.float64, generator=g)
        noise2 = (nz * nz).sum().item()
        x += nz
        del nz
    return x.reshape(-1).to(dtype), noise2


@pytest.mark.parametrize("rdt", ["float32", "float64"])
def test_r2c_cfg3_closed_form_spectrum(hmi, cuda, rdt):
    """r2c along axis 0 of the configuration's plane-wave field: each cosine a cos(2 pi k.r / n)
    puts a N / 2 at bin k of the half spectrum (513 x 1024 x 512, axis 0 fastest) and, when
    k_0 = 0, also at its mirror (0, -k_1, -k_2) (k = 0: a N at the origin); everything else is the
    transform of the noise, whose energy over the half spectrum (planes k_0 = 0 and 512 once, the
    others twice) is N times the noise energy (Parseval)."""
    import torch
    dt = getattr(torch, rdt)
    cdt = torch.complex64 if rdt == "float32" else torch.complex128
    n0, n1, n2 = CFG3_SIZE
    ntot = n0 * n1 * n2
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in CFG3_SIZE])
    outbox = inbox.r2c(0)
    plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, outbox, 0)
    x, noise2 = cfg3_field(torch, cuda, dt)
    y = torch.empty(outbox.count(), dtype=cdt, device=cuda)
    plan.forward(x, y)
    del x
    h0 = n0 // 2 + 1
    expect = {}
    for kv, amp in cfg3_terms():
        for sgn in (1, -1):
            b = ((sgn * kv[0]) % n0, (sgn * kv[1]) % n1, (sgn * kv[2]) % n2)
            if b[0] < h0:
                idx = b[0] + h0 * (b[1] + n1 * b[2])
                expect[idx] = expect.get(idx, 0.0) + amp * ntot / 2
    # the noise adds about 1e-3 sqrt(N) (~23) to every bin, the float32 rounding far less
    tol = 2e-6 * ntot
    for idx, v in expect.items():
        assert abs(y[idx].item() - v) <= tol, (idx, y[idx].item(), v)
        y[idx] -= v
    r = y.view(n2, n1, h0)
    w = torch.full((h0,), 2.0, dtype=torch.float64, device=cuda)
    w[0] = w[n0 // 2] = 1.0
    energy = ((r.real.double() ** 2 + r.imag.double() ** 2) * w).sum().item()
    assert abs(energy - ntot * noise2) <= 0.01 * ntot * noise2, (energy, ntot * noise2)


@pytest.mark.parametrize("rdt", ["float32", "float64"])
def test_r2c_sine_wave_sign(hmi, cuda, rdt):
    """sin(2 pi k.r / n) with k_0 > 0: the half spectrum holds -i N / 2 at bin k and nothing else
    (a conjugated spectrum would hold +i N / 2)."""
    import torch
    dt = getattr(torch, rdt)
    cdt = torch.complex64 if rdt == "float32" else torch.complex128
    n0, n1, n2 = CFG3_SIZE
    ntot = n0 * n1 * n2
    k = (7, 300, 41)
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in CFG3_SIZE])
    outbox = inbox.r2c(0)
    plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, outbox, 0)
    i = torch.arange(n0, device=cuda, dtype=torch.float64) / n0
    j = torch.arange(n1, device=cuda, dtype=torch.float64) / n1
    kk = torch.arange(n2, device=cuda, dtype=torch.float64) / n2
    x = torch.sin(2 * math.pi * (k[2] * kk[:, None, None] + k[1] * j[None, :, None] + k[0] * i[None, None, :]))
    x = x.reshape(-1).to(dt)
    y = torch.empty(outbox.count(), dtype=cdt, device=cuda)
    plan.forward(x, y)
    h0 = n0 // 2 + 1
    idx = k[0] + h0 * (k[1] + n1 * k[2])
    tol = (2e-6 if rdt == "float32" else 1e-12) * ntot
    assert abs(y[idx].item() - (-0.5j * ntot)) <= tol, y[idx].item()
    y[idx] = 0
    assert torch.linalg.vector_norm(y).item() <= tol * 10


def test_r2c_1024x1024x512_float(hmi, cuda):
    """r2c along axis 0 and back (scale::full) and agreement with the double transform."""
    import torch
    size = (1024, 1024, 512)
    inbox = hmi.box3d([0, 0, 0], [s - 1 for s in size])
    outbox = inbox.r2c(0)
    plan = hmi.fft3d_r2c(hmi.backend.mi355, inbox, outbox, 0)
    g = torch.Generator(device=cuda)
    g.manual_seed(2)
    x = torch.randn(inbox.count(), dtype=torch.float32, device=cuda, generator=g)
    y = torch.empty(outbox.count(), dtype=torch.complex64, device=cuda)
    plan.forward(x, y)
    y64 = torch.empty(outbox.count(), dtype=torch.complex128, device=cuda)
    plan.forward(x.double(), y64)
    assert rel(y, y64) <= 5e-6
    del y64
    z = torch.empty_like(x)
    plan.backward(y, z, hmi.scale.full)
    assert (torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x)).item() <= 5e-6


def test_dct_256_double_round_trip(hmi, cuda):
    import torch
    n = 256
    box = hmi.box3d([0, 0, 0], [n - 1] * 3)
    plan = hmi.rtransform(hmi.backend.mi355, box, box, hmi.kind.cos)
    g = torch.Generator(device=cuda)
    g.manual_seed(2)
    x = torch.randn(n ** 3, dtype=torch.float64, device=cuda, generator=g)
    y, z = torch.empty_like(x), torch.empty_like(x)
    plan.forward(x, y)
    plan.backward(y, z, hmi.scale.full)
    assert (torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x)).item() <= 1e-13
    # Parseval of REDFT10 (y_k = 2 sum_j x_j cos(pi (j + 1/2) k / n)) per axis:
    # y_0^2 / 2 + sum_{k>0} y_k^2 = 2 n sum_j x_j^2, so in 3D sum w y^2 = (2 n)^3 sum x^2 with
    # weight 1/2 for each zero index
    w = torch.ones(n, dtype=torch.float64, device=cuda)
    w[0] = 0.5
    ww = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    lhs = (ww * y * y).sum().item()
    rhs = (x * x).sum().item() * (2 * n) ** 3
    assert abs(lhs - rhs) / rhs <= 1e-12


def test_batch_1000_mixed_round_trip(hmi, cuda):
    """The 1,000-transform batch (seed 4 sides, U[-1,1] values): batched forward and backward
    (scale::full) returns the input; every transform's forward equals the plain fft3d plan's
    complex<double> forward of the same input (a separate kernel path: the fft3d passes), to the
    complex<float> bound; the fused x-y phase (all 1000 transforms) and the separate phases agree."""
    import torch
    rng = np.random.default_rng(4)
    sizes = [tuple(int(v) for v in rng.choice(SMOOTH, 3)) for _ in range(1000)]
    counts = [a * b * c for a, b, c in sizes]
    plan = hmi.batch_plan(sizes)
    g = torch.Generator(device=cuda)
    g.manual_seed(4)
    x = (torch.rand(sum(counts), dtype=torch.complex64, device=cuda, generator=g) * 2 - (1 + 1j)).contiguous()
    y = torch.empty_like(x)
    plan.forward(x, y)
    assert plan.fused_count(False) == 1000
    offs = np.cumsum([0] + counts)[:-1]
    singles = {}
    worst = 0.0
    for idx in range(len(sizes)):
        s, o, c = sizes[idx], int(offs[idx]), counts[idx]
        if s not in singles:
            b = hmi.box3d([0, 0, 0], [v - 1 for v in s])
            singles[s] = hmi.fft3d(hmi.backend.mi355, b, b)
        ref = torch.empty(c, dtype=torch.complex128, device=cuda)
        singles[s].forward(x[o:o + c].to(torch.complex128), ref)
        e = rel(y[o:o + c], ref)
        assert e <= 5e-6, (idx, s, e)
        worst = max(worst, e)
    print(f"worst per-transform rel L2 vs fft3d complex<double>: {worst:.2e}")
    plan.set_fused(False)
    ys = torch.empty_like(x)
    plan.forward(x, ys)
    assert rel(ys, y) <= 5e-6
    plan.set_fused(True)
    z = torch.empty_like(x)
    plan.backward(y, z, hmi.scale.full)
    assert rel(z, x) <= 5e-6
