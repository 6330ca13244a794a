"""Python binding of libheffte_mi355.so, the MI355X single-rank 3D FFT backend.

It mirrors heFFTe's ctypes binding (python/heffte.py: backend, scale, box3d,
fft3d(), fft3d_r2c(), plan.forward/backward and their _buffered variants) with
torch tensors on the GPU as the arrays.  torch is only the allocator here:
every transform runs in the hand-written HIP kernels of the library, and there
is no fallback -- if the library is missing, importing this module raises.

    import torch, heffte_amd as heffte
    box = heffte.box3d([0, 0, 0], [63, 63, 63])
    out = heffte.box3d([0, 0, 0], [63, 63, 63], [2, 1, 0])
    plan = heffte.fft3d(heffte.backend.mi355, box, out)
    y = torch.empty(plan.size_outbox(), dtype=torch.complex128, device="cuda")
    plan.forward(x, y, heffte.scale.none)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# HEFFTE_MI355_LIBRARY selects another build of the same library (e.g. to compare two builds)
LIBRARY_PATH = os.environ.get("HEFFTE_MI355_LIBRARY") or os.path.join(_HERE, "libheffte_mi355.so")


class heffte_input_error(Exception):
    """Raised on invalid input, as python/heffte.py:17-33 does."""

    def __init__(self, message):
        super().__init__(message)
        self.message = message


def _load():
    # torch ships its own libamdhip64.so.7; loading torch first makes the library
    # bind to that same HIP runtime (one runtime per process, shared device memory).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(
            f"{LIBRARY_PATH} is missing: build it with `python heffte_amd/build.py` "
            "(there is no CPU or PyTorch fallback for the transforms)")
    lib = ctypes.CDLL(LIBRARY_PATH)
    LL3 = ctypes.POINTER(ctypes.c_longlong)
    I3 = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    plan_p = ctypes.c_void_p
    lib.heffte_mi355_plan_create.argtypes = [LL3, LL3, I3, LL3, LL3, I3, vp, ctypes.POINTER(plan_p)]
    lib.heffte_mi355_plan_create_r2c.argtypes = [LL3, LL3, I3, LL3, LL3, I3, ctypes.c_int, vp, ctypes.POINTER(plan_p)]
    lib.heffte_mi355_plan_create_r2r.argtypes = [ctypes.c_int, LL3, LL3, I3, LL3, LL3, I3, vp, ctypes.POINTER(plan_p)]
    lib.heffte_mi355_plan_destroy.argtypes = [plan_p]
    for f in ("heffte_mi355_size_inbox", "heffte_mi355_size_outbox", "heffte_mi355_size_workspace"):
        getattr(lib, f).argtypes = [plan_p]
        getattr(lib, f).restype = ctypes.c_longlong
    lib.heffte_mi355_get_kind.argtypes = [plan_p]
    lib.heffte_mi355_set_stream.argtypes = [plan_p, vp]
    lib.heffte_mi355_set_precision.argtypes = [plan_p, ctypes.c_int]
    lib.heffte_mi355_plan_prepare.argtypes = [plan_p, ctypes.c_int]
    for f in ("forward_c2c", "forward_z2z", "forward_s2c", "forward_d2z", "backward_c2c", "backward_z2z",
              "backward_c2s", "backward_z2d", "forward_s2s", "forward_d2d", "backward_s2s", "backward_d2d"):
        fn = getattr(lib, "heffte_mi355_" + f)
        fn.argtypes = [plan_p, ctypes.c_int, vp, vp, vp, ctypes.c_int]
        fn.restype = ctypes.c_int
    lib.heffte_mi355_plan_set_timing.argtypes = [plan_p, ctypes.c_int]
    lib.heffte_mi355_plan_timing_report.argtypes = [plan_p]
    lib.heffte_mi355_plan_timing_report.restype = ctypes.c_char_p
    lib.heffte_mi355_execute_c2c_noalloc.argtypes = [plan_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp,
                                                     ctypes.c_int]
    lib.heffte_mi355_execute_c2c_noalloc.restype = ctypes.c_int
    lib.heffte_mi355_reshape.argtypes = [ctypes.c_int, LL3, LL3, I3, I3, ctypes.c_int, vp, vp, vp]
    lib.heffte_mi355_size_convolve_workspace.argtypes = [plan_p]
    lib.heffte_mi355_size_convolve_workspace.restype = ctypes.c_longlong
    for f in ("convolve_c2c", "convolve_z2z", "convolve_spectrum_c2c", "convolve_spectrum_z2z"):
        fn = getattr(lib, "heffte_mi355_" + f)
        fn.argtypes = [plan_p, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int]
        fn.restype = ctypes.c_int
    lib.heffte_mi355_batch_create.argtypes = [ctypes.c_int, I3, LL3, LL3, ctypes.POINTER(plan_p)]
    lib.heffte_mi355_batch_destroy.argtypes = [plan_p]
    lib.heffte_mi355_batch_set_stream.argtypes = [plan_p, vp]
    lib.heffte_mi355_batch_set_graph.argtypes = [plan_p, ctypes.c_int]
    lib.heffte_mi355_batch_set_autotune.argtypes = [plan_p, ctypes.c_int]
    lib.heffte_mi355_batch_set_fused.argtypes = [plan_p, ctypes.c_int]
    lib.heffte_mi355_plan_matrix_core_passes.argtypes = [plan_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                         ctypes.POINTER(ctypes.c_int)]
    lib.heffte_mi355_batch_fused_count.argtypes = [plan_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    for f in ("heffte_mi355_batch_size_input", "heffte_mi355_batch_size_output", "heffte_mi355_batch_graph_captures"):
        getattr(lib, f).argtypes = [plan_p]
        getattr(lib, f).restype = ctypes.c_longlong
    for f in ("forward_c2c", "forward_z2z", "backward_c2c", "backward_z2z"):
        fn = getattr(lib, "heffte_mi355_batch_" + f)
        fn.argtypes = [plan_p, vp, vp, ctypes.c_int]
        fn.restype = ctypes.c_int
    lib.heffte_mi355_last_error.restype = ctypes.c_char_p
    lib.heffte_mi355_set_default_options.argtypes = [vp]
    return lib


lib = _load()

# exported symbols of include/heffte_mi355.h (checked by the tests)
C_ABI_SYMBOLS = (
    "heffte_mi355_set_default_options", "heffte_mi355_plan_create", "heffte_mi355_plan_create_r2c",
    "heffte_mi355_plan_create_r2r", "heffte_mi355_plan_destroy", "heffte_mi355_size_inbox",
    "heffte_mi355_size_outbox", "heffte_mi355_size_workspace", "heffte_mi355_get_kind", "heffte_mi355_set_stream",
    "heffte_mi355_forward_c2c", "heffte_mi355_forward_z2z", "heffte_mi355_forward_s2c", "heffte_mi355_forward_d2z",
    "heffte_mi355_backward_c2c", "heffte_mi355_backward_z2z", "heffte_mi355_backward_c2s",
    "heffte_mi355_backward_z2d", "heffte_mi355_forward_s2s", "heffte_mi355_forward_d2d",
    "heffte_mi355_backward_s2s", "heffte_mi355_backward_d2d", "heffte_mi355_reshape", "heffte_mi355_last_error",
    "heffte_mi355_batch_create", "heffte_mi355_batch_destroy", "heffte_mi355_batch_set_stream",
    "heffte_mi355_batch_set_graph", "heffte_mi355_batch_size_input", "heffte_mi355_batch_size_output",
    "heffte_mi355_batch_forward_c2c", "heffte_mi355_batch_forward_z2z", "heffte_mi355_batch_backward_c2c",
    "heffte_mi355_batch_backward_z2z", "heffte_mi355_set_precision", "heffte_mi355_batch_set_autotune",
    "heffte_mi355_size_convolve_workspace", "heffte_mi355_convolve_c2c", "heffte_mi355_convolve_z2z",
    "heffte_mi355_convolve_spectrum_c2c", "heffte_mi355_convolve_spectrum_z2z", "heffte_mi355_plan_prepare",
    "heffte_mi355_batch_graph_captures", "heffte_mi355_execute_c2c_noalloc", "heffte_mi355_plan_set_timing",
    "heffte_mi355_plan_timing_report", "heffte_mi355_batch_set_fused", "heffte_mi355_batch_fused_count",
    "heffte_mi355_plan_matrix_core_passes",
)


class backend:
    """Backend ids (python/heffte.py:110-116); this library provides mi355 only."""
    mi355 = 12
    valid = [12]


class scale:
    none = 0
    full = 1
    symmetric = 2


class precision:
    """Precision form of the complex<float> 512-point passes (HMI_PRECISION_*)."""
    fp32 = 0
    split_fp16 = 1
    split_bf16 = 2


class kind:
    c2c = 0
    r2c = 1
    cos = 2
    sin = 3
    cos1 = 4  # DCT-I, the convention of backend::rocfft_cos1


class box3d:
    """heffte::box3d: inclusive low/high indexes and the (fast, mid, slow) order."""

    def __init__(self, clow, chigh, corder=(0, 1, 2)):
        self.low = np.array(clow, dtype=np.int64)
        self.high = np.array(chigh, dtype=np.int64)
        self.order = np.array(corder, dtype=np.int32)
        if self.low.size != 3 or self.high.size != 3 or self.order.size != 3:
            raise heffte_input_error("box3d needs three low, high and order entries")
        if sorted(self.order.tolist()) != [0, 1, 2]:
            raise heffte_input_error("box3d order must be a permutation of (0, 1, 2)")

    @property
    def size(self):
        return tuple(int(h - l + 1) for l, h in zip(self.low, self.high))

    def count(self):
        s = self.size
        return 0 if min(s) <= 0 else s[0] * s[1] * s[2]

    def osize(self, i):
        return self.size[self.order[i]]

    def r2c(self, dim):
        """heffte_geometry.h:96-104."""
        high = self.high.copy()
        high[dim] = self.low[dim] + self.size[dim] // 2
        return box3d(self.low, high, self.order)

    def _c(self):
        LL = ctypes.c_longlong * 3
        I3 = ctypes.c_int * 3
        return LL(*self.low.tolist()), LL(*self.high.tolist()), I3(*self.order.tolist())


def last_error():
    return lib.heffte_mi355_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise heffte_input_error(f"{what} failed with code {rc}: {last_error()}")


def _ptr(t):
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    raise heffte_input_error("arrays must be torch tensors on the GPU or raw device addresses")


def _dtype_name(t):
    return str(t.dtype).replace("torch.", "")


class fft_plan:
    """A plan object; create it with fft3d(), fft3d_r2c() or rtransform()."""

    def __init__(self, handle, k):
        self._h = handle
        self.kind = k
        self.use_r2c = (k == kind.r2c)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and lib is not None:
            lib.heffte_mi355_plan_destroy(h)
            self._h = None

    def size_inbox(self):
        return lib.heffte_mi355_size_inbox(self._h)

    def size_outbox(self):
        return lib.heffte_mi355_size_outbox(self._h)

    def size_workspace(self):
        return lib.heffte_mi355_size_workspace(self._h)

    def set_precision(self, form):
        """precision.fp32 (default), precision.split_fp16 or precision.split_bf16."""
        _check(lib.heffte_mi355_set_precision(self._h, int(form)), "set_precision")

    def matrix_core_passes(self, direction=-1, in_place=False):
        """(passes on the matrix cores, all passes) of the complex<float> transform in direction
        -1 (forward) or +1 (backward) under the current precision form."""
        m, t = ctypes.c_int(), ctypes.c_int()
        _check(lib.heffte_mi355_plan_matrix_core_passes(self._h, int(direction), 1 if in_place else 0,
                                                        ctypes.byref(m), ctypes.byref(t)), "matrix_core_passes")
        return m.value, t.value

    def prepare(self, double_precision=False):
        """Uploads the device tables of every transform variant in one precision now, so that
        later calls (e.g. under stream capture) upload nothing."""
        _check(lib.heffte_mi355_plan_prepare(self._h, 1 if double_precision else 0), "prepare")

    def set_timing(self, enable=True):
        """Starts (or stops) recording a hipEvent around every pass launch of this plan's calls;
        either drops the records so far."""
        _check(lib.heffte_mi355_plan_set_timing(self._h, 1 if enable else 0), "set_timing")

    def timing(self):
        """Per pass kernel since set_timing(True): {name: (launches, total ms)}; waits for the
        recorded events."""
        txt = lib.heffte_mi355_plan_timing_report(self._h)
        if txt is None:
            raise heffte_input_error(f"timing report failed: {last_error()}")
        out = {}
        for line in txt.decode().splitlines():
            name, launches, ms = line.split("\t")
            out[name] = (int(launches), float(ms))
        return out

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream, a raw hipStream_t address, or None (default stream)."""
        s = getattr(stream, "cuda_stream", stream)
        _check(lib.heffte_mi355_set_stream(self._h, ctypes.c_void_p(s) if s else None), "set_stream")

    def _call(self, direction, inarray, outarray, workspace, scaling, batch):
        if scaling not in (0, 1, 2):
            raise heffte_input_error(f"{direction}() called with invalid scaling")
        ti, to = _dtype_name(inarray), _dtype_name(outarray)
        if self.kind in (kind.cos, kind.sin, kind.cos1):
            table = {("float32", "float32"): "s2s", ("float64", "float64"): "d2d"}
        elif direction == "forward":
            table = {("float32", "complex64"): "s2c", ("float64", "complex128"): "d2z"}
            if not self.use_r2c:
                table.update({("complex64", "complex64"): "c2c", ("complex128", "complex128"): "z2z"})
        else:
            table = {("complex64", "float32"): "c2s", ("complex128", "float64"): "z2d"}
            if not self.use_r2c:
                table.update({("complex64", "complex64"): "c2c", ("complex128", "complex128"): "z2z"})
        key = table.get((ti, to))
        if key is None:
            raise heffte_input_error(f"{direction}() called with incompatible types {ti} -> {to}")
        nin, nout = (self.size_inbox(), self.size_outbox()) if direction == "forward" else (
            self.size_outbox(), self.size_inbox())
        if inarray.numel() < batch * nin or outarray.numel() < batch * nout:
            raise heffte_input_error(f"{direction}() called with invalid array size")
        if workspace is not None and workspace.numel() * workspace.element_size() < \
                batch * self.size_workspace() * (16 if key in ("z2z", "d2z", "z2d") else
                                                 8 if key in ("c2c", "s2c", "c2s", "d2d") else 4):
            raise heffte_input_error(f"{direction}() called with insufficient workspace size")
        fn = getattr(lib, f"heffte_mi355_{direction}_{key}")
        _check(fn(self._h, batch, _ptr(inarray), _ptr(outarray), _ptr(workspace), scaling), direction)

    def forward(self, inarray, outarray, scaling=scale.none, batch=1):
        self._call("forward", inarray, outarray, None, scaling, batch)

    def backward(self, inarray, outarray, scaling=scale.none, batch=1):
        self._call("backward", inarray, outarray, None, scaling, batch)

    def forward_buffered(self, inarray, outarray, workspace, scaling=scale.none, batch=1):
        self._call("forward", inarray, outarray, workspace, scaling, batch)

    def backward_buffered(self, inarray, outarray, workspace, scaling=scale.none, batch=1):
        self._call("backward", inarray, outarray, workspace, scaling, batch)

    def size_convolve_workspace(self):
        return lib.heffte_mi355_size_convolve_workspace(self._h)

    def convolve(self, x, y, outarray, workspace=None, scaling=scale.full, batch=1):
        """outarray = backward(forward(x) * forward(y)) with the backward scaling; scale.full
        gives the circular convolution of x and y (the sequence of benchmarks/convolution.cpp:86-95
        in one call, the product folded into the first backward pass).  c2c plans only;
        x, y and outarray are complex arrays in the inbox layout, outarray may alias x or y."""
        if self.kind != kind.c2c:
            raise heffte_input_error("convolve() needs a c2c plan")
        if scaling not in (0, 1, 2):
            raise heffte_input_error("convolve() called with invalid scaling")
        names = {_dtype_name(x), _dtype_name(y), _dtype_name(outarray)}
        if len(names) != 1 or names.copy().pop() not in ("complex64", "complex128"):
            raise heffte_input_error(f"convolve() called with incompatible types {sorted(names)}")
        key = "c2c" if _dtype_name(x) == "complex64" else "z2z"
        n = batch * self.size_inbox()
        if x.numel() < n or y.numel() < n or outarray.numel() < n:
            raise heffte_input_error("convolve() called with invalid array size")
        if workspace is not None and workspace.numel() * workspace.element_size() < \
                batch * self.size_convolve_workspace() * (8 if key == "c2c" else 16):
            raise heffte_input_error("convolve() called with insufficient workspace size")
        fn = getattr(lib, f"heffte_mi355_convolve_{key}")
        _check(fn(self._h, batch, _ptr(x), _ptr(y), _ptr(outarray), _ptr(workspace), scaling), "convolve")

    def convolve_spectrum(self, x, spectrum, outarray, workspace=None, scaling=scale.full, batch=1):
        """outarray = backward(forward(x) * spectrum) for a spectrum = forward(y) (scale.none,
        outbox layout) computed once by the caller: one forward and one backward transform per
        call.  outarray may alias x, not spectrum; workspace as for forward_buffered."""
        if self.kind != kind.c2c:
            raise heffte_input_error("convolve_spectrum() needs a c2c plan")
        if scaling not in (0, 1, 2):
            raise heffte_input_error("convolve_spectrum() called with invalid scaling")
        names = {_dtype_name(x), _dtype_name(spectrum), _dtype_name(outarray)}
        if len(names) != 1 or names.copy().pop() not in ("complex64", "complex128"):
            raise heffte_input_error(f"convolve_spectrum() called with incompatible types {sorted(names)}")
        key = "c2c" if _dtype_name(x) == "complex64" else "z2z"
        if x.numel() < batch * self.size_inbox() or spectrum.numel() < batch * self.size_outbox() or \
                outarray.numel() < batch * max(self.size_inbox(), self.size_outbox()):
            raise heffte_input_error("convolve_spectrum() called with invalid array size")
        if workspace is not None and workspace.numel() * workspace.element_size() < \
                batch * self.size_workspace() * (8 if key == "c2c" else 16):
            raise heffte_input_error("convolve_spectrum() called with insufficient workspace size")
        fn = getattr(lib, f"heffte_mi355_convolve_spectrum_{key}")
        _check(fn(self._h, batch, _ptr(x), _ptr(spectrum), _ptr(outarray), _ptr(workspace), scaling),
               "convolve_spectrum")


def _check_backend(backend_tag):
    if backend_tag not in backend.valid:
        raise heffte_input_error("Invalid backend, use heffte_amd.backend.mi355")


def _check_comm(comm):
    if comm is not None:
        size = comm.Get_size() if hasattr(comm, "Get_size") else int(comm)
        if size != 1:
            raise heffte_input_error("this backend runs one rank per transform (communicator of size 1)")


def fft3d(backend_tag, inbox, outbox, comm=None):
    """python/heffte.py fft3d(); comm may be None or any communicator of size 1."""
    _check_backend(backend_tag)
    _check_comm(comm)
    h = ctypes.c_void_p()
    il, ih, io = inbox._c()
    ol, oh, oo = outbox._c()
    _check(lib.heffte_mi355_plan_create(il, ih, io, ol, oh, oo, None, ctypes.byref(h)), "fft3d")
    return fft_plan(h, kind.c2c)


def fft3d_r2c(backend_tag, inbox, outbox, r2c_direction, comm=None):
    _check_backend(backend_tag)
    _check_comm(comm)
    if r2c_direction not in (0, 1, 2):
        raise heffte_input_error("fft3d_r2c() called with invalid r2c_direction, must use 0, 1, or 2")
    h = ctypes.c_void_p()
    il, ih, io = inbox._c()
    ol, oh, oo = outbox._c()
    _check(lib.heffte_mi355_plan_create_r2c(il, ih, io, ol, oh, oo, r2c_direction, None, ctypes.byref(h)),
           "fft3d_r2c")
    return fft_plan(h, kind.r2c)


def rtransform(backend_tag, inbox, outbox, transform=kind.cos, comm=None):
    """heffte::rtransform<backend::*_cos / *_sin / *_cos1> (transform = kind.cos, kind.sin, kind.cos1)."""
    _check_backend(backend_tag)
    _check_comm(comm)
    h = ctypes.c_void_p()
    il, ih, io = inbox._c()
    ol, oh, oo = outbox._c()
    _check(lib.heffte_mi355_plan_create_r2r(transform, il, ih, io, ol, oh, oo, None, ctypes.byref(h)),
           "rtransform")
    return fft_plan(h, transform)


def proc_setup(world, num_procs):
    """Processor grid (i, j, k), i j k = num_procs, whose sub-boxes have the least surface
    (python/heffte.py:428-448 with the integer semantics of proc_setup_min_surface,
    include/heffte_geometry.h:643-686).  This backend runs one rank: proc_setup(world, 1)
    is (1, 1, 1); larger grids only describe how a caller would split the world."""
    if world.count() <= 0 or num_procs < 1:
        raise heffte_input_error("proc_setup() needs a non-empty world and at least one rank")
    if num_procs == 1:
        return np.array([1, 1, 1], dtype=np.int32)
    n = world.size

    def surface(g):
        b = [n[d] // g[d] for d in range(3)]
        return b[0] * b[1] + b[1] * b[2] + b[2] * b[0]

    best, best_surface = [1, 1, 1], None
    for i in range(1, min(num_procs, n[0]) + 1):
        if num_procs % i:
            continue
        jmax = min(num_procs // i, n[1])
        for j in range(1, jmax + 1):
            if jmax % j:
                continue
            k = num_procs // (i * j)
            if k <= n[2] and i * j * k == num_procs:
                s = surface((i, j, k))
                if best_surface is None or s < best_surface:
                    best, best_surface = [i, j, k], s
    return np.array(best, dtype=np.int32)


def split_world(world, proc_grid):
    """heffte::split_world (include/heffte_geometry.h:410-431): the boxes of a processor grid,
    fast index first; the first (size % grid) boxes along an axis get one extra index."""
    g = [int(v) for v in proc_grid]
    n = world.size

    def edge(d, i):
        return int(world.low[d]) + i * (n[d] // g[d]) + min(i, n[d] % g[d])

    return [box3d([edge(0, i), edge(1, j), edge(2, k)], [edge(0, i + 1) - 1, edge(1, j + 1) - 1, edge(2, k + 1) - 1],
                  world.order)
            for k in range(g[2]) for j in range(g[1]) for i in range(g[0])]


def reshape(inbox, out_order, inarray, outarray, batch=1, stream=None):
    """Bit-exact in-node reshape of inarray (stored in inbox.order) into out_order."""
    il, ih, io = inbox._c()
    oo = (ctypes.c_int * 3)(*[int(v) for v in out_order])
    s = getattr(stream, "cuda_stream", stream)
    _check(lib.heffte_mi355_reshape(inarray.element_size(), il, ih, io, oo, batch, _ptr(inarray), _ptr(outarray),
                                    ctypes.c_void_p(s) if s else None), "reshape")


class batch_plan:
    """Independent c2c transforms of different sizes, each a box in order (0, 1, 2).

    sizes: list of (nx, ny, nz); offsets in complex elements (None = packed back to back)."""

    def __init__(self, sizes, in_offsets=None, out_offsets=None):
        sizes = [tuple(int(v) for v in s) for s in sizes]
        if any(len(s) != 3 for s in sizes):
            raise heffte_input_error("every transform of a batch has three sides")
        for name, offs in (("in_offsets", in_offsets), ("out_offsets", out_offsets)):
            if offs is not None and len(offs) != len(sizes):
                raise heffte_input_error(f"{name} holds {len(offs)} offsets for {len(sizes)} transforms")
        flat = [v for s in sizes for v in s]
        S = (ctypes.c_int * len(flat))(*flat)
        io = (ctypes.c_longlong * len(sizes))(*in_offsets) if in_offsets is not None else None
        oo = (ctypes.c_longlong * len(sizes))(*out_offsets) if out_offsets is not None else None
        h = ctypes.c_void_p()
        _check(lib.heffte_mi355_batch_create(len(sizes), S, io, oo, ctypes.byref(h)), "batch_plan")
        self._h = h
        self.sizes = sizes

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and lib is not None:
            lib.heffte_mi355_batch_destroy(h)
            self._h = None

    def size_input(self):
        return lib.heffte_mi355_batch_size_input(self._h)

    def size_output(self):
        return lib.heffte_mi355_batch_size_output(self._h)

    def set_stream(self, stream):
        s = getattr(stream, "cuda_stream", stream)
        _check(lib.heffte_mi355_batch_set_stream(self._h, ctypes.c_void_p(s) if s else None), "set_stream")

    def set_graph(self, enable):
        _check(lib.heffte_mi355_batch_set_graph(self._h, 1 if enable else 0), "set_graph")

    def graph_captures(self):
        """Number of hipGraph captures made so far (a replayed call adds none)."""
        return lib.heffte_mi355_batch_graph_captures(self._h)

    def set_autotune(self, enable):
        """Time the tile candidates of each side length on the first call (default) or not."""
        _check(lib.heffte_mi355_batch_set_autotune(self._h, 1 if enable else 0), "set_autotune")

    def set_fused(self, enable):
        """Run the x and y axes of eligible transforms in one fused pass (default) or not."""
        _check(lib.heffte_mi355_batch_set_fused(self._h, 1 if enable else 0), "set_fused")

    def fused_count(self, double_precision=False):
        """Transforms the fused x-y phase takes (-1 before the first call of that precision)."""
        c = ctypes.c_int()
        _check(lib.heffte_mi355_batch_fused_count(self._h, 1 if double_precision else 0, ctypes.byref(c)),
               "fused_count")
        return c.value

    def _run(self, direction, inarray, outarray, scaling):
        t = _dtype_name(inarray)
        if t != _dtype_name(outarray) or t not in ("complex64", "complex128"):
            raise heffte_input_error("batch transforms take complex64 or complex128 arrays of one type")
        if inarray.numel() < self.size_input() or outarray.numel() < self.size_output():
            raise heffte_input_error(f"{direction}() called with invalid array size")
        key = "c2c" if t == "complex64" else "z2z"
        fn = getattr(lib, f"heffte_mi355_batch_{direction}_{key}")
        _check(fn(self._h, _ptr(inarray), _ptr(outarray), scaling), direction)

    def forward(self, inarray, outarray, scaling=scale.none):
        self._run("forward", inarray, outarray, scaling)

    def backward(self, inarray, outarray, scaling=scale.none):
        self._run("backward", inarray, outarray, scaling)
