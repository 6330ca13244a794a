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
    (scale::full) returns the input, and each transform's forward equals the plain fft3d plan."""
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
    offs = np.cumsum([0] + counts)[:-1]
    for idx in (0, 1, 499, 998, 999):
        s, o, c = sizes[idx], int(offs[idx]), counts[idx]
        b = hmi.box3d([0, 0, 0], [v - 1 for v in s])
        single = hmi.fft3d(hmi.backend.mi355, b, b)
        ref = torch.empty(c, dtype=torch.complex128, device=cuda)
        single.forward(x[o:o + c].to(torch.complex128), ref)
        assert rel(y[o:o + c], ref) <= 5e-6, (idx, s)
    z = torch.empty_like(x)
    plan.backward(y, z, hmi.scale.full)
    assert rel(z, x) <= 5e-6
