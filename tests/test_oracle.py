"""Pins the CPU oracle (oracle/heffte_oracle.c) before it is trusted as the checker.

Each oracle routine is checked against something it does not share code with:
  * heFFTe's own known-answer vectors (tests/golden/reference_known_answers.json,
    transcribed from test/test_units_nompi.cpp and test/test_cos.cpp);
  * a brute-force O(n^2) DFT in long double with exact index reduction;
  * numpy.fft (pocketfft) in FP64;
  * closed forms of the DCT/DST and DFT special cases (impulse, constant,
    single plane wave), Parseval and linearity.
"""
import numpy as np
import pytest

TOL64 = 1e-13


def box(oracle, low, high, order=(0, 1, 2)):
    return oracle.Box.make(low, high, order)


def canonical(arr, size):
    """flat array in order (0,1,2) -> numpy array indexed [i0, i1, i2]."""
    return np.asarray(arr).reshape(size[2], size[1], size[0]).transpose(2, 1, 0)


def flat(a, order):
    """numpy array indexed [i0, i1, i2] -> flat array stored in `order`."""
    # slowest axis first in C order
    return np.ascontiguousarray(a.transpose(order[2], order[1], order[0])).ravel()


def cx(pairs):
    return np.array([complex(a, b) for a, b in pairs])


# ------------------------------------------------------------ known answers

@pytest.mark.parametrize("axis", [0, 1, 2])
def test_1d_c2c_known_answers(oracle, golden, axis):
    """test/test_units_nompi.cpp:209-262: one 1D transform per line along `axis`."""
    x = canonical(np.array(golden["box_2x3x4"]["input"], dtype=np.complex128), (2, 3, 4))
    y = np.empty_like(x)
    for idx in np.ndindex(*[s for d, s in enumerate((2, 3, 4)) if d != axis]):
        sl = list(idx)
        sl.insert(axis, slice(None))
        y[tuple(sl)] = oracle.fft1d(x[tuple(sl)])
    ref = cx(golden["c2c_1d"][f"axis{axis}"])
    assert np.abs(flat(y, (0, 1, 2)) - ref).max() < 1e-11
    # backward of the known answer returns n * input
    back = np.empty_like(y)
    for idx in np.ndindex(*[s for d, s in enumerate((2, 3, 4)) if d != axis]):
        sl = list(idx)
        sl.insert(axis, slice(None))
        back[tuple(sl)] = oracle.fft1d(y[tuple(sl)], +1)
    assert np.abs(back / (2, 3, 4)[axis] - x).max() < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_1d_r2c_known_answers(oracle, golden, axis):
    """test/test_units_nompi.cpp:336-367: r2c along `axis` keeps n/2+1 entries."""
    size = (2, 3, 4)
    x = canonical(np.array(golden["box_2x3x4"]["input"], dtype=np.float64), size)
    hs = list(size)
    hs[axis] = size[axis] // 2 + 1
    y = np.empty(hs, dtype=np.complex128)
    for idx in np.ndindex(*[s for d, s in enumerate(size) if d != axis]):
        sl = list(idx)
        sl.insert(axis, slice(None))
        y[tuple(sl)] = oracle.fft1d(x[tuple(sl)].astype(np.complex128))[:hs[axis]]
    ref = cx(golden["r2c_1d"][f"axis{axis}"])
    assert np.abs(flat(y, (0, 1, 2)) - ref).max() < 1e-11


@pytest.mark.parametrize("name,kind", [("cos_len4", 0), ("cos_len5", 0), ("sin_len4", 1), ("sin_len5", 1)])
def test_1d_r2r_known_answers(oracle, golden, name, kind):
    """test/test_units_nompi.cpp:288-333 (DCT-II / DST-II, GPU/stock 4N round trip)."""
    g = golden["r2r_1d"][name]
    x = np.array(g["input"])
    y = oracle.r2r_1d(x, kind, -1)
    assert np.abs(y - np.array(g["output"])).max() < 1e-11
    back = oracle.r2r_1d(y, kind, +1) / (golden["r2r_1d"]["roundtrip_factor_per_length"] * x.size)
    assert np.abs(back - x).max() < 1e-12


@pytest.mark.parametrize("kind", ["cos", "sin", "cos1"])
def test_3d_r2r_known_answers(oracle, golden, kind):
    """test/test_cos.cpp:36-80: forward golden, backward with scale::full = 1/(64 N)
    (DCT-I: 1/(64 (n0-1)(n1-1)(n2-1)), include/heffte_fft3d.h:643-646)."""
    b = box(oracle, (0, 0, 0), (1, 2, 3))
    x = np.arange(1.0, 25.0)
    fn = {"cos": oracle.r2r_cos, "sin": oracle.r2r_sin, "cos1": oracle.r2r_cos1}[kind]
    y = fn(b, b, x, -1)
    assert np.abs(y - np.array(golden["r2r_3d"][kind])).max() < 1e-9
    back = fn(b, b, y, +1, 1.0 / (64 * 24) if kind != "cos1" else 1.0 / (64 * 1 * 2 * 3))
    assert np.abs(back - x).max() < 1e-11


@pytest.mark.parametrize("order", [(1, 2, 0), (2, 1, 0), (0, 2, 1)])
def test_transpose_known_answers(oracle, golden, order):
    """test/test_units_nompi.cpp:577-637: bit-exact in-node transposes."""
    src = box(oracle, (0, 0, 0), (1, 2, 3))
    dst = box(oracle, (0, 0, 0), (1, 2, 3), order)
    x = np.arange(1.0, 25.0)
    y = oracle.transpose(src, dst, x)
    ref = np.array(golden["transpose"]["order_" + "".join(map(str, order))])
    assert np.array_equal(y, ref)
    assert np.array_equal(oracle.transpose(dst, src, y), x)  # "flip back" case


# ------------------------------------------------------- brute force / numpy

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 27, 43, 64, 97, 100, 105, 128, 243, 256])
def test_fft1d_vs_long_double_dft(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    for sign in (-1, 1):
        ref = oracle.dft1d_ld(x, sign)
        assert oracle.rel_l2(oracle.fft1d(x, sign), ref) < 2e-15 * max(1.0, np.log2(n))
        nref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
        assert oracle.rel_l2(ref, nref) < 1e-14


def test_long_double_dft_closed_forms(oracle):
    """impulse at j0 -> exp(-2 pi i j0 k / n); constant -> n delta; plane wave -> one bin."""
    n = 12
    for j0 in range(n):
        x = np.zeros(n, complex)
        x[j0] = 1.0
        y = oracle.dft1d_ld(x)
        k = np.arange(n)
        assert np.abs(y - np.exp(-2j * np.pi * ((j0 * k) % n) / n)).max() < 4e-15
    y = oracle.dft1d_ld(np.ones(n, complex))
    assert abs(y[0] - n) < 1e-14 and np.abs(y[1:]).max() < 1e-14
    k0 = 5
    y = oracle.dft1d_ld(np.exp(2j * np.pi * k0 * np.arange(n) / n))
    assert abs(y[k0] - n) < 1e-13 and np.abs(np.delete(y, k0)).max() < 1e-13


ORDERS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


@pytest.mark.parametrize("size", [(6, 7, 8), (7, 8, 6), (8, 9, 10), (43, 76, 24), (1, 5, 1), (16, 1, 9)])
@pytest.mark.parametrize("oi,oo", [(0, 0), (0, 5), (0, 3), (4, 2)])
def test_fft3d_c2c_vs_numpy(oracle, size, oi, oo):
    rng = np.random.default_rng(7)
    inorder, outorder = ORDERS[oi], ORDERS[oo]
    a = rng.standard_normal(size) + 1j * rng.standard_normal(size)
    bi = box(oracle, (0, 0, 0), [s - 1 for s in size], inorder)
    bo = box(oracle, (0, 0, 0), [s - 1 for s in size], outorder)
    y = oracle.fft3d_c2c(bi, bo, flat(a, inorder))
    assert oracle.rel_l2(y, flat(np.fft.fftn(a), outorder)) < TOL64
    z = oracle.fft3d_c2c(bo, bi, y, +1, 1.0 / a.size)
    assert oracle.rel_l2(z, flat(a, inorder)) < TOL64


def test_fft3d_c2c_batch_and_scale(oracle):
    rng = np.random.default_rng(3)
    size, batch = (8, 9, 10), 5
    bi = box(oracle, (0, 0, 0), (7, 8, 9))
    bo = box(oracle, (0, 0, 0), (7, 8, 9), (2, 1, 0))
    xs = [rng.standard_normal(size) + 1j * rng.standard_normal(size) for _ in range(batch)]
    x = np.concatenate([flat(a, (0, 1, 2)) for a in xs])
    s = 1.0 / np.sqrt(np.prod(size))
    y = oracle.fft3d_c2c(bi, bo, x, -1, s, batch)
    ref = np.concatenate([flat(np.fft.fftn(a) * s, (2, 1, 0)) for a in xs])
    assert oracle.rel_l2(y, ref) < TOL64


def test_fft3d_invariants(oracle):
    """Parseval and linearity of the 3D transform (no reference formula involved)."""
    rng = np.random.default_rng(11)
    size = (12, 10, 9)
    b = box(oracle, (0, 0, 0), [s - 1 for s in size])
    x = rng.standard_normal(b.count()) + 1j * rng.standard_normal(b.count())
    z = rng.standard_normal(b.count()) + 1j * rng.standard_normal(b.count())
    fx, fz = oracle.fft3d_c2c(b, b, x), oracle.fft3d_c2c(b, b, z)
    assert abs(np.vdot(fx, fx).real / b.count() - np.vdot(x, x).real) < 1e-12 * np.vdot(x, x).real
    f = oracle.fft3d_c2c(b, b, 2.5 * x - 1j * z)
    assert oracle.rel_l2(f, 2.5 * fx - 1j * fz) < TOL64


@pytest.mark.parametrize("size,r2c_dir", [((6, 7, 8), 0), ((6, 7, 8), 1), ((6, 7, 8), 2), ((9, 4, 5), 0),
                                          ((16, 1, 1), 0), ((43, 76, 24), 1)])
def test_r2c_c2r_vs_numpy(oracle, size, r2c_dir):
    rng = np.random.default_rng(5)
    a = rng.standard_normal(size)
    order = (1, 2, 0)
    bi = box(oracle, (0, 0, 0), [s - 1 for s in size], order)
    bo = bi.r2c(r2c_dir)
    y = oracle.fft3d_r2c(bi, bo, r2c_dir, flat(a, order))
    full = np.fft.fftn(a)
    sl = [slice(None)] * 3
    sl[r2c_dir] = slice(0, size[r2c_dir] // 2 + 1)
    assert oracle.rel_l2(y, flat(full[tuple(sl)], order)) < TOL64
    back = oracle.fft3d_c2r(bi, bo, r2c_dir, y, 1.0 / a.size)
    assert oracle.rel_l2(back, flat(a, order)) < TOL64
    # real-input c2c keeps the full spectrum
    yf = oracle.fft3d_real_c2c(bi, bi, flat(a, order))
    assert oracle.rel_l2(yf, flat(full, order)) < TOL64


def dct2_closed(x):
    n = x.size
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    return 2.0 * (np.cos(np.pi * k * (2 * j + 1) / (2 * n)) @ x)


def dct3x2_closed(x):
    # 2 * FFTW REDFT01: 2 * (x0 + 2 sum_{k>=1} x_k cos(pi k (2n+1) / 2N))
    n = x.size
    j = np.arange(n)[:, None]
    k = np.arange(1, n)[None, :]
    return 2.0 * (x[0] + 2.0 * (np.cos(np.pi * k * (2 * j + 1) / (2 * n)) @ x[1:]))


def dst2_closed(x):
    n = x.size
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    return 2.0 * (np.sin(np.pi * (k + 1) * (2 * j + 1) / (2 * n)) @ x)


def dst3x2_closed(x):
    # 2 * FFTW RODFT01: 2 * ((-1)^j x_{N-1} + 2 sum_{k<N-1} x_k sin(pi (k+1)(2j+1) / 2N))
    n = x.size
    j = np.arange(n)
    k = np.arange(n - 1)[None, :]
    return 2.0 * ((-1.0) ** j * x[-1] + 2.0 * (np.sin(np.pi * (k + 1) * (2 * j[:, None] + 1) / (2 * n)) @ x[:-1]))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 13, 16, 30, 64])
def test_r2r_1d_closed_forms(oracle, n):
    rng = np.random.default_rng(n + 100)
    x = rng.standard_normal(n)
    assert np.abs(oracle.r2r_1d(x, 0, -1) - dct2_closed(x)).max() < 1e-12 * n
    assert np.abs(oracle.r2r_1d(x, 0, +1) - dct3x2_closed(x)).max() < 1e-12 * n
    assert np.abs(oracle.r2r_1d(x, 1, -1) - dst2_closed(x)).max() < 1e-12 * n
    assert np.abs(oracle.r2r_1d(x, 1, +1) - dst3x2_closed(x)).max() < 1e-12 * n


def redft00_closed(x):
    """DCT-I: y[k] = x[0] + (-1)^k x[N-1] + 2 sum_{j=1}^{N-2} x[j] cos(pi j k / (N-1))."""
    n = x.size
    k = np.arange(n)[:, None]
    j = np.arange(1, n - 1)[None, :]
    return x[0] + (-1.0) ** np.arange(n) * x[-1] + 2.0 * (np.cos(np.pi * j * k / (n - 1)) @ x[1:-1])


@pytest.mark.parametrize("n", [2, 3, 4, 5, 9, 17, 33, 65])
def test_cos1_1d_closed_form(oracle, n):
    """DCT-I forward = REDFT00, backward = 2 REDFT00; round trip 4 (N-1)."""
    x = np.random.default_rng(n + 7).standard_normal(n)
    y = oracle.r2r_1d(x, 2, -1)
    assert np.abs(y - redft00_closed(x)).max() < 1e-12 * n
    assert np.abs(oracle.r2r_1d(x, 2, +1) - 2.0 * redft00_closed(x)).max() < 1e-12 * n
    assert np.abs(oracle.r2r_1d(y, 2, +1) / (4 * (n - 1)) - x).max() < 1e-13 * n


def test_cos1_3d_separable_and_rejects_short_axes(oracle):
    rng = np.random.default_rng(19)
    size = (5, 3, 9)
    a = rng.standard_normal(size)
    ref = np.apply_along_axis(redft00_closed, 0, a)
    ref = np.apply_along_axis(redft00_closed, 1, ref)
    ref = np.apply_along_axis(redft00_closed, 2, ref)
    bi = box(oracle, (0, 0, 0), (4, 2, 8), (1, 2, 0))
    bo = box(oracle, (0, 0, 0), (4, 2, 8), (2, 0, 1))
    y = oracle.r2r_cos1(bi, bo, flat(a, (1, 2, 0)))
    assert oracle.rel_l2(y, flat(ref, (2, 0, 1))) < TOL64
    b1 = box(oracle, (0, 0, 0), (4, 0, 8))
    with pytest.raises(Exception):
        oracle.r2r_cos1(b1, b1, np.zeros(45))


def test_r2r_3d_separable(oracle):
    """The 3D cosine transform is the 1D transform along each axis (closed form)."""
    rng = np.random.default_rng(9)
    size = (5, 6, 7)
    a = rng.standard_normal(size)
    ref = np.apply_along_axis(dct2_closed, 0, a)
    ref = np.apply_along_axis(dct2_closed, 1, ref)
    ref = np.apply_along_axis(dct2_closed, 2, ref)
    bi = box(oracle, (0, 0, 0), (4, 5, 6), (2, 0, 1))
    bo = box(oracle, (0, 0, 0), (4, 5, 6), (0, 1, 2))
    y = oracle.r2r_cos(bi, bo, flat(a, (2, 0, 1)))
    assert oracle.rel_l2(y, flat(ref, (0, 1, 2))) < TOL64


def test_planted_mistake_is_caught(oracle):
    """A transposed operand (conjugate sign) or a dropped twiddle would fail the checks above."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal(12) + 1j * rng.standard_normal(12)
    good = oracle.fft1d(x)
    assert oracle.rel_l2(oracle.fft1d(x, +1), oracle.dft1d_ld(x)) > 0.1
    assert oracle.rel_l2(good, oracle.dft1d_ld(x)) < 1e-14


def _circular_convolution(a, b):
    """Brute force: c[n] = sum_m a[m] b[(n - m) mod N] over the three axes (no transform)."""
    c = np.zeros(a.shape, dtype=np.complex128)
    for m in np.ndindex(a.shape):
        c += a[m] * np.roll(b, m, axis=(0, 1, 2))
    return c


@pytest.mark.parametrize("size,oi,oo", [((4, 3, 5), 0, 0), ((5, 4, 3), 0, 5), ((2, 6, 1), 3, 1)])
def test_convolve3d_vs_brute_force(oracle, size, oi, oo):
    """backward(forward(x) * forward(y)) / N is the circular convolution; with scale 1 it is
    N times it; the product is commutative and a delta at the origin is the identity."""
    rng = np.random.default_rng(13)
    inorder, outorder = ORDERS[oi], ORDERS[oo]
    a = rng.standard_normal(size) + 1j * rng.standard_normal(size)
    b = rng.standard_normal(size) + 1j * rng.standard_normal(size)
    bi = box(oracle, (0, 0, 0), [s - 1 for s in size], inorder)
    bo = box(oracle, (0, 0, 0), [s - 1 for s in size], outorder)
    ref = _circular_convolution(a, b)
    c = oracle.convolve3d(bi, bo, flat(a, inorder), flat(b, inorder), 1.0 / a.size)
    assert oracle.rel_l2(c, flat(ref, inorder)) < TOL64
    c1 = oracle.convolve3d(bi, bo, flat(b, inorder), flat(a, inorder), 1.0)
    assert oracle.rel_l2(c1, flat(ref * a.size, inorder)) < TOL64
    delta = np.zeros(size, dtype=np.complex128)
    delta[0, 0, 0] = 1.0
    d = oracle.convolve3d(bi, bo, flat(a, inorder), flat(delta, inorder), 1.0 / a.size)
    assert oracle.rel_l2(d, flat(a, inorder)) < TOL64
    # a shifted delta shifts the input (a wrong sign in either transform would shift the other way)
    delta = np.zeros(size, dtype=np.complex128)
    delta[1 % size[0], 1 % size[1], 0] = 1.0
    d = oracle.convolve3d(bi, bo, flat(a, inorder), flat(delta, inorder), 1.0 / a.size)
    assert oracle.rel_l2(d, flat(np.roll(a, (1, 1), axis=(0, 1)), inorder)) < TOL64


def test_convolve3d_batch(oracle):
    rng = np.random.default_rng(17)
    size, batch = (3, 4, 2), 3
    bi = box(oracle, (0, 0, 0), [s - 1 for s in size])
    bo = box(oracle, (0, 0, 0), [s - 1 for s in size], (2, 0, 1))
    xs = [rng.standard_normal(size) + 1j * rng.standard_normal(size) for _ in range(batch)]
    ys = [rng.standard_normal(size) + 1j * rng.standard_normal(size) for _ in range(batch)]
    x = np.concatenate([flat(a, (0, 1, 2)) for a in xs])
    y = np.concatenate([flat(a, (0, 1, 2)) for a in ys])
    c = oracle.convolve3d(bi, bo, x, y, 1.0 / 24, batch)
    ref = np.concatenate([flat(_circular_convolution(a, b), (0, 1, 2)) for a, b in zip(xs, ys)])
    assert oracle.rel_l2(c, ref) < TOL64
