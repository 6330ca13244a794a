"""CPU checks of the C ABI (no kernel launches): the library loads, exports every
entry point include/heffte_mi355.h declares, and the host-side plan logic
accepts and rejects boxes with heFFTe's status codes."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "heffte_mi355.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(heffte_mi355_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(hmi):
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(hmi.lib, n), f"{n} is declared in include/heffte_mi355.h but not exported"
    assert sorted(hmi.C_ABI_SYMBOLS) == names


def test_library_has_no_cpu_fallback_symbols(hmi):
    """The product library links no FFT library: the transforms are its own HIP kernels."""
    out = os.popen(f"ldd {hmi.LIBRARY_PATH}").read()
    for banned in ("rocfft", "hipfft", "fftw", "torch", "liboracle"):
        assert banned not in out, out


def LL(v):
    return (ctypes.c_longlong * 3)(*v)


def I3(v):
    return (ctypes.c_int * 3)(*v)


def create(hmi, lo, hi, io, oo):
    h = ctypes.c_void_p()
    rc = hmi.lib.heffte_mi355_plan_create(LL(lo), LL(hi), I3(io), LL(lo), LL(hi), I3(oo), None, ctypes.byref(h))
    return rc, h


def test_plan_create_and_sizes(hmi):
    rc, h = create(hmi, (0, 0, 0), (63, 63, 63), (0, 1, 2), (2, 1, 0))
    assert rc == 0, hmi.last_error()
    assert hmi.lib.heffte_mi355_size_inbox(h) == 64 ** 3
    assert hmi.lib.heffte_mi355_size_outbox(h) == 64 ** 3
    ws = hmi.lib.heffte_mi355_size_workspace(h)
    # one full intermediate array at most (heFFTe's own would be 3 N for this plan)
    assert 0 <= ws <= 64 ** 3
    assert hmi.lib.heffte_mi355_get_kind(h) == 0
    assert hmi.lib.heffte_mi355_plan_destroy(h) == 0


def test_plan_rejects_bad_boxes(hmi):
    # inbox and outbox must hold the same indexes on one rank
    h = ctypes.c_void_p()
    rc = hmi.lib.heffte_mi355_plan_create(LL((0, 0, 0)), LL((7, 7, 7)), I3((0, 1, 2)), LL((0, 0, 0)),
                                          LL((7, 7, 6)), I3((0, 1, 2)), None, ctypes.byref(h))
    assert rc == 2 and "same indexes" in hmi.last_error()
    # order must be a permutation
    rc, h = create(hmi, (0, 0, 0), (7, 7, 7), (0, 0, 2), (0, 1, 2))
    assert rc == 2
    # empty box
    rc, h = create(hmi, (0, 0, 0), (7, -1, 7), (0, 1, 2), (0, 1, 2))
    assert rc == 2


def test_r2c_plan_boxes(hmi):
    inbox = hmi.box3d([0, 0, 0], [15, 9, 7])
    p = hmi.fft3d_r2c(hmi.backend.mi355, inbox, inbox.r2c(1), 1)
    assert p.size_inbox() == 16 * 10 * 8 and p.size_outbox() == 16 * 6 * 8
    with pytest.raises(hmi.heffte_input_error):
        hmi.fft3d_r2c(hmi.backend.mi355, inbox, inbox.r2c(0), 1)
    with pytest.raises(hmi.heffte_input_error):
        hmi.fft3d_r2c(hmi.backend.mi355, inbox, inbox.r2c(0), 3)


def test_invalid_plan_and_arguments(hmi):
    assert hmi.lib.heffte_mi355_plan_destroy(None) == 3
    assert hmi.lib.heffte_mi355_size_inbox(None) == -1
    rc, h = create(hmi, (0, 0, 0), (3, 3, 3), (0, 1, 2), (0, 1, 2))
    # scale outside {0, 1, 2}
    assert hmi.lib.heffte_mi355_forward_c2c(h, 1, ctypes.c_void_p(16), ctypes.c_void_p(16), None, 7) == 2
    # null arrays are rejected before any launch
    assert hmi.lib.heffte_mi355_forward_c2c(h, 1, None, None, None, 0) == 2
    # r2r call on a complex plan
    assert hmi.lib.heffte_mi355_forward_d2d(h, 1, ctypes.c_void_p(16), ctypes.c_void_p(16), None, 0) == 2
    assert hmi.lib.heffte_mi355_plan_destroy(h) == 0


def test_communicator_must_have_one_rank(hmi):
    b = hmi.box3d([0, 0, 0], [3, 3, 3])
    with pytest.raises(hmi.heffte_input_error):
        hmi.fft3d(hmi.backend.mi355, b, b, comm=2)
    hmi.fft3d(hmi.backend.mi355, b, b, comm=1)


def test_default_options(hmi):
    class opts(ctypes.Structure):
        _fields_ = [("use_reorder", ctypes.c_int), ("algorithm", ctypes.c_int), ("use_pencils", ctypes.c_int),
                    ("use_gpu_aware", ctypes.c_int)]
    o = opts()
    assert hmi.lib.heffte_mi355_set_default_options(ctypes.byref(o)) == 0
    assert o.use_reorder == 0 and o.use_pencils == 1


def test_cos1_plan_rules(hmi):
    """DCT-I plans need two points per axis; the workspace covers the real intermediates."""
    b = hmi.box3d([0, 0, 0], [8, 16, 4])
    p = hmi.rtransform(hmi.backend.mi355, b, b, hmi.kind.cos1)
    assert p.size_inbox() == p.size_outbox() == 9 * 17 * 5
    assert p.size_workspace() >= 0
    with pytest.raises(hmi.heffte_input_error):
        flat = hmi.box3d([0, 0, 0], [8, 0, 4])
        hmi.rtransform(hmi.backend.mi355, flat, flat, hmi.kind.cos1)


def test_proc_setup_and_split_world(hmi):
    """Host geometry helpers (python/heffte.py:428-460): one rank owns the world; larger grids
    partition it exactly, with the remainder on the first boxes."""
    world = hmi.box3d([0, 0, 0], [511, 511, 511])
    assert list(hmi.proc_setup(world, 1)) == [1, 1, 1]
    assert sorted(hmi.proc_setup(world, 8).tolist()) == [2, 2, 2]
    assert list(hmi.proc_setup(hmi.box3d([0, 0, 0], [1023, 1023, 3]), 4)) == [2, 2, 1]
    one = hmi.split_world(world, [1, 1, 1])
    assert len(one) == 1 and one[0].size == world.size
    w2 = hmi.box3d([2, 0, 1], [11, 6, 5], (2, 0, 1))
    parts = hmi.split_world(w2, [3, 2, 1])
    assert len(parts) == 6 and sum(p.count() for p in parts) == w2.count()
    assert parts[0].size == (4, 4, 5) and parts[1].size == (3, 4, 5) and parts[3].size == (4, 3, 5)
    assert parts[0].low.tolist() == [2, 0, 1] and list(parts[5].high) == [11, 6, 5]
    assert all(list(p.order) == [2, 0, 1] for p in parts)
    covered = set()
    for p in parts:
        for i in range(p.low[0], p.high[0] + 1):
            for j in range(p.low[1], p.high[1] + 1):
                for k in range(p.low[2], p.high[2] + 1):
                    assert (i, j, k) not in covered
                    covered.add((i, j, k))
    assert len(covered) == w2.count()


def test_length_beyond_single_pass_capacity(hmi):
    """A line that does not fit the 160 KiB LDS of one workgroup (2 n element sizes) runs as two
    passes n = n1 n2, with a step buffer of one array in the workspace; a length with no such
    split (a prime beyond the capacity) is refused at plan creation with a message, not run."""
    ok = hmi.box3d([0, 0, 0], [10239, 0, 0])
    hmi.fft3d(hmi.backend.mi355, ok, ok)
    n = 16384 * 4 * 2
    split = hmi.box3d([0, 0, 0], [16383, 3, 1])  # 16384 = 128 x 128
    p = hmi.fft3d(hmi.backend.mi355, split, split)
    # the step buffer, and an intermediate array for the in-place call with real input
    assert p.size_workspace() == 2 * n
    permuted = hmi.box3d([0, 0, 0], [16383, 3, 1], [2, 1, 0])
    p = hmi.fft3d(hmi.backend.mi355, split, permuted)
    assert p.size_workspace() == 2 * n  # an intermediate array and the step buffer
    big = hmi.box3d([0, 0, 0], [10246, 0, 0])  # 10247 is prime
    with pytest.raises(hmi.heffte_input_error) as e:
        hmi.fft3d(hmi.backend.mi355, big, big)
    assert "LDS" in str(e.value)


def test_batch_plan_creation_rules(hmi):
    """Batch plans are created on the host: sizes and offsets are checked, extents reported."""
    p = hmi.batch_plan([(16, 20, 24), (7, 1, 3)])
    assert p.size_input() == 16 * 20 * 24 + 21 and p.size_output() == p.size_input()
    q = hmi.batch_plan([(4, 4, 4), (2, 2, 2)], in_offsets=[100, 0], out_offsets=[0, 64])
    assert q.size_input() == 164 and q.size_output() == 72
    for bad in ([], [(0, 4, 4)], [(4, -1, 4)]):
        with pytest.raises(hmi.heffte_input_error):
            hmi.batch_plan(bad)
    with pytest.raises(hmi.heffte_input_error):
        hmi.batch_plan([(4, 4, 4)], in_offsets=[-1])
    with pytest.raises(hmi.heffte_input_error):
        hmi.batch_plan([(20000, 2, 2)])  # a side beyond the single-pass LDS capacity


def test_batch_c_abi_rejects_bad_handles(hmi):
    import ctypes
    lib = hmi.lib
    assert lib.heffte_mi355_batch_set_graph(None, 1) != 0
    assert lib.heffte_mi355_batch_set_autotune(None, 1) != 0
    assert lib.heffte_mi355_batch_size_input(None) == -1
    h = ctypes.c_void_p()
    assert lib.heffte_mi355_batch_create(1, None, None, None, ctypes.byref(h)) != 0
    assert "size" in hmi.last_error()


def test_batch_aliasing_rules(hmi):
    """Output ranges of a batch must be disjoint (checked at creation); a call with input ==
    output needs equal offsets per transform, and otherwise the input and output ranges must
    not overlap (checked before anything is launched, so these run without a GPU)."""
    lib = hmi.lib
    with pytest.raises(hmi.heffte_input_error) as e:
        hmi.batch_plan([(4, 4, 4), (2, 2, 2)], in_offsets=[0, 64], out_offsets=[0, 60])
    assert "overlap" in str(e.value)
    with pytest.raises(hmi.heffte_input_error) as e:
        hmi.batch_plan([(4, 4, 4), (2, 2, 2)], in_offsets=[0])  # shorter than the list of sizes
    assert "offsets" in str(e.value)
    # in == out with different offsets per transform: rejected
    p = hmi.batch_plan([(4, 4, 4), (2, 2, 2)], in_offsets=[8, 72], out_offsets=[0, 64])
    base = 1 << 40  # a fake device address: the check fails before any GPU call
    assert lib.heffte_mi355_batch_forward_c2c(p._h, ctypes.c_void_p(base), ctypes.c_void_p(base), 0) == 2
    assert "equal input and output offsets" in hmi.last_error()
    # distinct arrays whose element ranges overlap: rejected
    assert lib.heffte_mi355_batch_forward_c2c(p._h, ctypes.c_void_p(base), ctypes.c_void_p(base + 8 * 16), 0) == 2
    assert "overlap" in hmi.last_error()
    assert lib.heffte_mi355_batch_backward_z2z(p._h, ctypes.c_void_p(base + 16 * 40), ctypes.c_void_p(base), 0) == 2
    assert lib.heffte_mi355_batch_graph_captures(p._h) == 0
    assert lib.heffte_mi355_batch_graph_captures(None) == -1


def test_prepare_rejects_bad_handles(hmi):
    assert hmi.lib.heffte_mi355_plan_prepare(None, 0) == 3


def test_matrix_core_pass_count(hmi):
    """Which passes the split-precision forms move to the matrix cores (host planning only): the
    three 512-point passes of the 512^3 c2c transform; none of a 1024-point axis or a 256^3 box;
    none under FP32."""
    b = hmi.box3d([0, 0, 0], [511, 511, 511])
    out = hmi.box3d([0, 0, 0], [511, 511, 511], [1, 2, 0])
    plan = hmi.fft3d(hmi.backend.mi355, b, out)
    assert plan.matrix_core_passes(-1) == (0, 3)
    plan.set_precision(hmi.precision.split_fp16)
    assert plan.matrix_core_passes(-1) == (3, 3)
    assert plan.matrix_core_passes(+1, in_place=True) == (3, 3)
    plan.set_precision(hmi.precision.split_bf16)
    assert plan.matrix_core_passes(+1) == (3, 3)
    mixed = hmi.fft3d(hmi.backend.mi355, hmi.box3d([0, 0, 0], [1023, 511, 255]),
                      hmi.box3d([0, 0, 0], [1023, 511, 255]))
    mixed.set_precision(hmi.precision.split_fp16)
    m, t = mixed.matrix_core_passes(-1)
    assert (m, t) == (1, 3)
    with pytest.raises(hmi.heffte_input_error):
        plan.matrix_core_passes(0)
