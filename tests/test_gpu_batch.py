"""GPU parity of the batched mixed-size transforms (persistent work-list kernel + hipGraph)
against the FP64 oracle, transform by transform.  Sizes are drawn from the smooth lengths
{16, 20, 24, 27, 28, 32, 40, 45, 48, 56, 63, 64, 80, 96, 105, 112, 128} (radices 2, 3, 5, 7).
The captured graph is replayed several times with eager work in between and every replay
is checked."""
import numpy as np
import pytest

import fftcase as fc

pytestmark = pytest.mark.gpu

SMOOTH = [16, 20, 24, 27, 28, 32, 40, 45, 48, 56, 63, 64, 80, 96, 105, 112, 128]
TOL = {"complex64": 5e-6, "complex128": 1e-13}


def draw_sizes(count, seed, choices):
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(choices, 3)) for _ in range(count)]


def oracle_each(oracle, sizes, offsets, x, sign, scaling):
    out = []
    for s, off in zip(sizes, offsets):
        n = int(np.prod(s))
        b = oracle.Box.make((0, 0, 0), [v - 1 for v in s])
        sc = 1.0 if scaling == 0 else (1.0 / n if scaling == 1 else 1.0 / np.sqrt(n))
        out.append(oracle.fft3d_c2c(b, b, x[off:off + n], sign, sc))
    return out


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_vs_oracle(hmi, cuda, oracle, dtype, graph):
    import torch
    sizes = draw_sizes(30, 4, [16, 20, 24, 27, 28, 32, 40, 45, 48, 56, 63, 64])
    plan = hmi.batch_plan(sizes)
    plan.set_graph(graph)
    n = plan.size_input()
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in sizes])[:-1]
    x = fc.uniform_complex(n, 4)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    tx = torch.from_numpy(x).to(cuda).to(tdt)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    y = fc.to_host(ty)
    refs = oracle_each(oracle, sizes, offsets, x, -1, 0)
    for s, off, r in zip(sizes, offsets, refs):
        assert fc.rel_l2(y[off:off + r.size], r) <= TOL[dtype], s
    # backward with full scaling, in place, returns the input
    plan.backward(ty, ty, hmi.scale.full)
    assert fc.rel_l2(fc.to_host(ty), x) <= TOL[dtype]


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("autotune", [False, True])
def test_batch_length_105_tiles(hmi, cuda, oracle, dtype, autotune):
    """Length 105 = 3 * 5 * 7 on every axis, with line counts that are not multiples of the tile's
    lines (16 / 8): the staged compile-time tile (untimed default) and the tuner's choice."""
    import torch
    sizes = [(105, 7, 3), (5, 105, 9), (3, 2, 105), (105, 105, 2), (105, 1, 1), (1, 105, 105)]
    plan = hmi.batch_plan(sizes)
    plan.set_autotune(autotune)
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in sizes])[:-1]
    x = fc.uniform_complex(plan.size_input(), 11)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    tx = torch.from_numpy(x).to(cuda).to(tdt)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    y = fc.to_host(ty)
    for s, off, r in zip(sizes, offsets, oracle_each(oracle, sizes, offsets, x, -1, 0)):
        assert fc.rel_l2(y[off:off + r.size], r) <= TOL[dtype], s
    tz = torch.empty_like(tx)
    plan.backward(ty, tz, hmi.scale.full)
    assert fc.rel_l2(fc.to_host(tz), x) <= TOL[dtype]


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_batch_all_smooth_lengths_and_offsets(hmi, cuda, oracle, dtype):
    """Every smooth length on every axis, output at shuffled, gapped offsets, symmetric scale."""
    import torch
    sizes = [(SMOOTH[i], SMOOTH[(i + 5) % 17], SMOOTH[(i + 11) % 17]) for i in range(17)]
    counts = [int(np.prod(s)) for s in sizes]
    in_off = list(np.cumsum([0] + counts)[:-1])
    order = np.random.default_rng(1).permutation(len(sizes))
    out_off = [0] * len(sizes)
    pos = 7  # leave gaps between the outputs
    for i in order:
        out_off[i] = pos
        pos += counts[i] + 3
    plan = hmi.batch_plan(sizes, in_off, out_off)
    x = fc.uniform_complex(plan.size_input(), 9)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    tx = torch.from_numpy(x).to(cuda).to(tdt)
    ty = torch.zeros(plan.size_output(), dtype=tdt, device=cuda)
    plan.forward(tx, ty, hmi.scale.symmetric)
    y = fc.to_host(ty)
    refs = oracle_each(oracle, sizes, in_off, x, -1, 2)
    for s, off, r in zip(sizes, out_off, refs):
        assert fc.rel_l2(y[off:off + r.size], r) <= TOL[dtype], s
    # the gaps were not written
    mask = np.ones(y.size, bool)
    for off, c in zip(out_off, counts):
        mask[off:off + c] = False
    assert np.all(y[mask] == 0)


def test_graph_replays(hmi, cuda, oracle):
    """Replays of the captured graph, with eager work and synchronisations between them."""
    import torch
    sizes = draw_sizes(12, 7, [16, 24, 32, 48, 64])
    plan = hmi.batch_plan(sizes)
    x = fc.uniform_complex(plan.size_input(), 2)
    tx = torch.from_numpy(x).to(cuda).to(torch.complex128)
    ty = torch.empty_like(tx)
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in sizes])[:-1]
    refs = oracle_each(oracle, sizes, offsets, x, -1, 1)
    for rep in range(4):
        ty.fill_(float(rep))
        junk = torch.randn(1 << 16, device=cuda).sum()  # eager work between replays
        torch.cuda.synchronize()
        plan.forward(tx, ty, hmi.scale.full)
        y = fc.to_host(ty)
        for s, off, r in zip(sizes, offsets, refs):
            assert fc.rel_l2(y[off:off + r.size], r) <= 1e-13, (rep, s)
        del junk


def test_batch_matches_single_plans(hmi, cuda):
    """The batched result equals, to rounding, the per-transform fft3d plans."""
    import torch
    sizes = draw_sizes(8, 3, SMOOTH)
    plan = hmi.batch_plan(sizes)
    x = fc.uniform_complex(plan.size_input(), 5)
    tx = torch.from_numpy(x).to(cuda).to(torch.complex64)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    off = 0
    for s in sizes:
        n = int(np.prod(s))
        b = hmi.box3d([0, 0, 0], [v - 1 for v in s])
        p = hmi.fft3d(hmi.backend.mi355, b, b)
        out = torch.empty(n, dtype=tx.dtype, device=cuda)
        p.forward(tx[off:off + n], out)
        assert fc.rel_l2(fc.to_host(ty[off:off + n]), fc.to_host(out).astype(np.complex128)) <= 5e-6
        off += n


def test_autotune_cache_and_parity(hmi, cuda, oracle, tmp_path, monkeypatch):
    """The tuner times the candidate tiles of every side length on the first call, writes one
    cache line per (length, precision), and the tuned plan matches the oracle; a second plan
    reads the cache and gives the same result bit for bit; untimed defaults also pass."""
    import torch
    cache = tmp_path / "sub" / "tuning.txt"
    monkeypatch.setenv("HMI_TUNE_CACHE", str(cache))
    sizes = [(20, 48, 105), (48, 20, 32), (105, 105, 16)]
    counts = [int(np.prod(s)) for s in sizes]
    offsets = np.cumsum([0] + counts)[:-1]
    plan = hmi.batch_plan(sizes)
    x = fc.uniform_complex(plan.size_input(), 21)
    tx = torch.from_numpy(x).to(cuda).to(torch.complex64)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    y = fc.to_host(ty)
    for s, off, r in zip(sizes, offsets, oracle_each(oracle, sizes, offsets, x, -1, 0)):
        assert fc.rel_l2(y[off:off + r.size], r) <= 5e-6, s
    lines = cache.read_text().splitlines()
    assert lines[0].startswith("# heffte_mi355 batch tuning, build ")
    rows = [line.split() for line in lines[1:] if line.strip()]
    tuned = {(int(r[0]), int(r[1])) for r in rows}
    assert {(20, 0), (48, 0), (105, 0), (32, 0), (16, 0)} <= tuned
    for r in rows:
        assert r[2] in ("0", "1") and int(r[3]) >= (0 if r[2] == "1" else 1)
    plan2 = hmi.batch_plan(sizes)
    ty2 = torch.empty_like(tx)
    plan2.forward(tx, ty2)
    assert torch.equal(ty, ty2)
    plan3 = hmi.batch_plan(sizes)
    plan3.set_autotune(False)
    ty3 = torch.empty_like(tx)
    plan3.forward(tx, ty3)
    assert fc.rel_l2(fc.to_host(ty3), y.astype(np.complex128)) <= 5e-6


def test_autotune_cache_from_another_build_is_ignored(hmi, cuda, oracle, tmp_path, monkeypatch):
    """A cache file whose first line names another build of the library is not read: its choice
    (a runtime plan of 3 lines for length 20, never a candidate) is replaced by a timed one and
    the file rewritten under this build's header."""
    import torch
    cache = tmp_path / "tuning.txt"
    cache.write_text("# heffte_mi355 batch tuning, build Jan  1 2000 00:00:00 / Jan  1 2000 00:00:00\n"
                     "20 0 0 3 2,2,5\n")
    monkeypatch.setenv("HMI_TUNE_CACHE", str(cache))
    sizes = [(20, 20, 20)]
    plan = hmi.batch_plan(sizes)
    x = fc.uniform_complex(plan.size_input(), 22)
    tx = torch.from_numpy(x).to(cuda).to(torch.complex64)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    ref = oracle_each(oracle, sizes, [0], x, -1, 0)[0]
    assert fc.rel_l2(fc.to_host(ty), ref) <= 5e-6
    lines = cache.read_text().splitlines()
    assert lines[0].startswith("# heffte_mi355 batch tuning, build ") and "2000" not in lines[0]
    assert "20 0 0 3 2,2,5" not in lines[1:]
    assert any(line.split()[:2] == ["20", "0"] for line in lines[1:])


# elements per thread of the compile-time tile of each side length (0: the staged tile of
# length 105, register class 0)
TILE_E = {16: 16, 20: 10, 24: 12, 27: 9, 28: 14, 32: 8, 40: 10, 45: 15, 48: 12, 56: 14, 63: 21, 64: 8,
          80: 10, 96: 12, 112: 14, 128: 16, 105: 0}


def fused_eligible(size, dtype):
    """Whether the fused x-y phase takes a transform: both x and y lengths have compile-time tiles;
    register class = the larger of the two (E <= 12: 0, E <= 16: 1, else 2); complex<float>: a
    256-thread build for every class and planes of at most 10240 elements, a 512-thread build for
    planes up to 20480; complex<double>: class 0, planes up to 5120 (a plane counts n1 rows of
    n0 | 1 elements)."""
    if size[0] not in TILE_E or size[1] not in TILE_E:
        return False
    e0, e1 = TILE_E[size[0]], TILE_E[size[1]]
    cls = max(0 if e <= 12 else (1 if e <= 16 else 2) for e in (e0, e1))
    plane = (size[0] | 1) * size[1]  # LDS rows padded to an odd length
    if dtype == "complex64":
        return plane <= 20480
    return cls == 0 and plane <= 5120


# every build of the fused phase, z-plane groups with a remainder, columns with 3 threads per line
# (27, 45, 63 on y), the staged length 105 on x and y, transforms left to the separate phases
# (length 1, planes beyond the budget, the 14-21-element tiles in complex<double>); z lengths with
# compile-time tiles (the column-block z phase, blocks with a remainder) and without
FUSED_SIZES = [(16, 16, 33), (27, 45, 5), (63, 64, 7), (20, 63, 4), (128, 128, 3), (96, 112, 2), (128, 80, 2),
               (105, 16, 4), (16, 105, 4), (63, 128, 3), (40, 24, 48), (112, 63, 2), (1, 32, 16), (64, 1, 8),
               (45, 27, 9), (56, 28, 6), (32, 128, 5), (105, 105, 3), (105, 64, 2), (20, 105, 7), (105, 128, 2),
               (16, 16, 128), (27, 45, 63), (20, 24, 105), (105, 105, 105), (128, 96, 112), (1, 32, 64), (3, 5, 80)]


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_fused_xy_phase(hmi, cuda, oracle, dtype):
    """Fused x-y phase against the oracle transform by transform (forward), the round trip
    (backward, full scale, in place), and the separate-phase path of the same plan; the number of
    transforms it takes follows the eligibility rule above."""
    import torch
    sizes = FUSED_SIZES
    plan = hmi.batch_plan(sizes)
    plan.set_autotune(False)
    dbl = dtype == "complex128"
    assert plan.fused_count(dbl) == -1
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in sizes])[:-1]
    x = fc.uniform_complex(plan.size_input(), 31)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    tx = torch.from_numpy(x).to(cuda).to(tdt)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty)
    assert plan.fused_count(dbl) == sum(fused_eligible(s, dtype) for s in sizes)
    y = fc.to_host(ty)
    for s, off, r in zip(sizes, offsets, oracle_each(oracle, sizes, offsets, x, -1, 0)):
        assert fc.rel_l2(y[off:off + r.size], r) <= TOL[dtype], s
    plan.set_fused(False)
    ts = torch.empty_like(tx)
    plan.forward(tx, ts)
    assert fc.rel_l2(fc.to_host(ts), y) <= TOL[dtype]
    plan.set_fused(True)
    plan.backward(ty, ty, hmi.scale.full)
    assert fc.rel_l2(fc.to_host(ty), x) <= TOL[dtype]


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_batch_random_sides(hmi, cuda, oracle, dtype):
    """80 transforms with sides drawn from 1..40 (primes, lengths without compile-time tiles, sides
    of 1) mixed with 20 of table lengths: the fused x-y, column-block z and separate phases in one
    batch, against the oracle transform by transform, forward and backward (symmetric scale)."""
    import torch
    rng = np.random.default_rng(12)
    sizes = [tuple(int(v) for v in rng.integers(1, 41, 3)) for _ in range(80)]
    sizes += [tuple(int(v) for v in rng.choice([16, 20, 24, 27, 28, 32, 40], 3)) for _ in range(20)]
    order = rng.permutation(len(sizes))
    sizes = [sizes[i] for i in order]
    plan = hmi.batch_plan(sizes)
    offsets = np.cumsum([0] + [int(np.prod(s)) for s in sizes])[:-1]
    x = fc.uniform_complex(plan.size_input(), 13)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    tx = torch.from_numpy(x).to(cuda).to(tdt)
    ty = torch.empty_like(tx)
    plan.forward(tx, ty, hmi.scale.symmetric)
    y = fc.to_host(ty)
    for s, off, r in zip(sizes, offsets, oracle_each(oracle, sizes, offsets, x, -1, 2)):
        assert fc.rel_l2(y[off:off + r.size], r) <= TOL[dtype], s
    tz = torch.empty_like(tx)
    plan.backward(ty, tz, hmi.scale.symmetric)
    z = fc.to_host(tz)
    for s, off in zip(sizes, offsets):
        n = int(np.prod(s))
        assert fc.rel_l2(z[off:off + n], x[off:off + n]) <= TOL[dtype], s
