"""Benchmark of the single-rank 3D FFT path on one MI355X.

Workload (the headline configuration): c2c 512 x 512 x 512 complex<float>,
input box order (0,1,2), output box order (1,2,0), in place with a
caller-provided workspace, forward + backward with scale::symmetric, values iid
N(0,1).  One step = one forward + backward pair, as heFFTe's speed3d harness
times it (benchmarks/speed3d.h:162-191, 228-229).

Prints one JSON line with ms per pair, 5 N log2 N GFLOP/s, the roofline of the
dominant kernel (the 512-point pass; algorithmic bytes = read + write of the
whole array per launch), the error against the FP64 path, the rocFFT time of
the same transform (torch.fft on this GPU, for comparison) and the CPU baseline
(the FP64 oracle on the host cores).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "512³ c2c fwd+bwd ms & %HBM roofline; split-FP16 vs FP32 at ≤5e-6 rel L2"
HBM_PEAK_GBS = 8000.0
# HBM bytes per pass launch from rocprofv3 FETCH_SIZE / WRITE_SIZE, each scaled by the value the
# same counter reads for a plain 1 GiB copy (profiles/fft_uniform_512_c64_hbm_traffic.md)
MEASURED_TRAFFIC = {(512, "complex64"): (2147835904, "profiles/fft_uniform_512_c64_hbm_traffic.md")}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    # untimed pairs before the timed ones: the first ~10 pairs on a fresh process run ~2 % slower
    # than the steady state (clock and power ramp)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--n", type=int, default=512, help="cube side")
    p.add_argument("--dtype", default="complex64", choices=["complex64", "complex128"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-rocfft", action="store_true", help="skip the rocFFT comparison")
    p.add_argument("--no-error", action="store_true", help="skip the FP64 error check")
    p.add_argument("--no-batched", action="store_true", help="skip the 1000-transform mixed-size batch leg")
    p.add_argument("--no-split", action="store_true", help="skip the split-precision MFMA forms")
    p.add_argument("--no-double", action="store_true", help="skip the 512^3 complex<double> leg")
    p.add_argument("--no-secondary", action="store_true",
                   help="skip the r2c 1024x1024x512, DCT 256^3 and 512^3 reshape legs")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a plain 1 GiB device copy (reference for the HBM byte counters)")
    return p.parse_args()


def pass_times(plan, step, steps, torch, bytes_per_launch, ms_step):
    """Average launch duration of each pass kernel in a step.  The library records HIP events on
    the plan's stream around every pass launch over `steps` further calls of step(); each kernel's
    share of that event-timed total, applied to ms_step (the time of one step in the timed region,
    whose launches run back to back with no events between them), gives its average duration
    there.  The events alone add the few microseconds of each marker to every launch
    (avg_ms_events)."""
    plan.set_timing(True)
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    rep = plan.timing()
    plan.set_timing(False)
    total_all = sum(total for _, total in rep.values())
    out = {}
    for name, (launches, total) in rep.items():
        avg = ms_step * (total / total_all) / (launches / steps)
        out[name] = {"launches": launches, "total_ms_events": round(total, 4),
                     "avg_ms_events": round(total / launches, 5), "avg_ms": round(avg, 5),
                     "GBps": round(bytes_per_launch / (avg * 1e-3) / 1e9, 1),
                     "frac": round(bytes_per_launch / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def seed1_input(n):
    """n^3 complex<double> values with iid N(0, 1) real and imaginary parts (numpy seed 1)."""
    import numpy as np
    rng = np.random.default_rng(1)
    return rng.standard_normal(2 * n ** 3).view(np.complex128)


def cpu_baseline(n, x=None, y_gpu=None):
    """The FP64 oracle (a CPU restatement of heFFTe's path) on the host cores:
    one forward + backward pair of the same transform, full size.  When the complex<double> GPU
    leg ran on the same input (x, seed 1), its forward output y_gpu is checked against the
    oracle's forward output here."""
    import numpy as np
    from oracle import pyoracle
    threads = min(16, os.cpu_count() or 1)
    pyoracle.set_threads(threads)
    if n > 512:
        return None
    m = n
    b = pyoracle.Box.make((0, 0, 0), (m - 1, m - 1, m - 1), (0, 1, 2))
    bo = pyoracle.Box.make((0, 0, 0), (m - 1, m - 1, m - 1), (1, 2, 0))
    if x is None:
        x = seed1_input(m)
    s = 1.0 / math.sqrt(m ** 3)
    t0 = time.perf_counter()
    y = pyoracle.fft3d_c2c(b, bo, x, -1, s)
    z = pyoracle.fft3d_c2c(bo, b, y, +1, s)
    t1 = time.perf_counter()
    res = {"value": round((t1 - t0) * 1e3, 1), "unit": "ms per fwd+bwd pair", "cores": threads, "kind": "port",
           "sample": f"one full {m}^3 forward+backward pair of the FP64 oracle (complex<double>, "
                     f"OpenMP, {threads} threads), (0,1,2)->(1,2,0), scale::symmetric"}
    if y_gpu is not None:
        for key, yg in y_gpu.items():
            res[f"gpu_{key}_forward_rel_l2_vs_oracle"] = float(np.linalg.norm(yg - y) / np.linalg.norm(y))
    del y, z
    return res


def double_leg(heffte, torch, dev, plan, args, x_np):
    """512^3 complex<double> (the second element type of the headline configuration): the same
    boxes and protocol (in place, caller workspace, forward + backward scale::symmetric), input
    iid N(0, 1) (numpy seed 1).  Per-pass times from the library's HIP events; the forward outputs
    of the same input in complex<double> and complex<float> go to the CPU leg, which checks them
    against the FP64 oracle."""
    n = args.n
    N = n ** 3
    bytes_per_launch = 2.0 * N * 16
    data = torch.from_numpy(x_np).to(dev)
    ws = torch.empty(max(1, plan.size_workspace()), dtype=torch.complex128, device=dev)

    def step():
        plan.forward_buffered(data, data, ws, heffte.scale.symmetric)
        plan.backward_buffered(data, data, ws, heffte.scale.symmetric)

    for _ in range(args.warmup):
        step()
    ms = _timed(torch, step, args.steps)
    passes = pass_times(plan, step, args.steps, torch, bytes_per_launch, ms)
    dom = max(passes, key=lambda k: passes[k]["total_ms_events"])
    res = {"workload": "c2c 512x512x512 complex<double>, inbox order (0,1,2) -> outbox order (1,2,0), in place, "
                       "caller workspace, forward+backward scale::symmetric, iid N(0,1) (numpy seed 1)",
           "ms_per_pair": round(ms, 4), "gflops_5nlog2n": round(2 * 5.0 * N * math.log2(N) / (ms * 1e-3) / 1e9, 1),
           "dominant_kernel": dom, "dominant_avg_ms": passes[dom]["avg_ms"], "dominant_frac": passes[dom]["frac"],
           "passes": passes, "bytes_per_launch": bytes_per_launch}
    data.copy_(torch.from_numpy(x_np))
    plan.forward_buffered(data, data, ws, heffte.scale.symmetric)
    y = {"complex_double": data.cpu().numpy()}
    del data, ws
    torch.cuda.empty_cache()
    d32 = torch.from_numpy(x_np).to(dev).to(torch.complex64)
    ws32 = torch.empty(max(1, plan.size_workspace()), dtype=torch.complex64, device=dev)
    plan.forward_buffered(d32, d32, ws32, heffte.scale.symmetric)
    y["complex_float"] = d32.cpu().numpy().astype("complex128")
    del d32, ws32
    torch.cuda.empty_cache()
    return res, y


SMOOTH = [16, 20, 24, 27, 28, 32, 40, 45, 48, 56, 63, 64, 80, 96, 105, 112, 128]


def batched_leg(heffte, torch, dev, steps):
    """1,000 independent c2c complex<float> transforms, each side drawn from the smooth sizes
    (seed 4, sides independent), values U[-1,1]: the batched work-list kernels (fused x-y phase,
    column-block z phase) replayed from a hipGraph vs a per-transform launch loop of the plain
    fft3d plans.  One step = forward + backward (scale::full) of the whole batch."""
    import numpy as np
    rng = np.random.default_rng(4)
    sizes = [tuple(int(v) for v in rng.choice(SMOOTH, 3)) for _ in range(1000)]
    counts = [a * b * c for a, b, c in sizes]
    total = sum(counts)
    plan = heffte.batch_plan(sizes)
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    x = (torch.rand(total, dtype=torch.complex64, device=dev, generator=g) * 2 - (1 + 1j)).contiguous()
    y = torch.empty_like(x)

    def timed(fn, reps, warm=1):
        for _ in range(warm):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def batched():
        plan.forward(x, y)
        plan.backward(y, y, heffte.scale.full)

    ms_batched = timed(batched, steps, warm=3)
    yb = y.clone()
    # per-transform loop of the plain kernels (one fft3d plan per shape, 3 launches per transform)
    plans = {}
    offs = np.cumsum([0] + counts)[:-1]
    ws = torch.empty(max(counts), dtype=torch.complex64, device=dev)
    for s in set(sizes):
        b = heffte.box3d([0, 0, 0], [v - 1 for v in s])
        plans[s] = heffte.fft3d(heffte.backend.mi355, b, b)
    views = [(plans[s], x[o:o + c], y[o:o + c]) for s, o, c in zip(sizes, offs, counts)]

    def loop():
        for p, xi, yi in views:
            p.forward_buffered(xi, yi, ws)
        for p, xi, yi in views:
            p.backward_buffered(yi, yi, ws, heffte.scale.full)

    ms_loop = timed(loop, max(2, steps // 4))
    # both return the input (round trip); report the larger deviation
    err_b = (torch.linalg.vector_norm(yb - x) / torch.linalg.vector_norm(x)).item()
    err_l = (torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x)).item()
    # 3-pass model: one read and one write of every element per axis and direction; the fused
    # phases move 2 passes' worth (x and y in one, z in the other)
    model = 2 * 3 * 2 * 8 * total
    moved = 2 * 2 * 2 * 8 * total
    return {"transforms": 1000, "elements": int(total), "distinct_shapes": len(plans),
            "ms_per_pair_batched_graph": round(ms_batched, 3), "ms_per_pair_launch_loop": round(ms_loop, 3),
            "speedup": round(ms_loop / ms_batched, 2),
            "hbm_frac_3pass_model": round(model / (ms_batched * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "hbm_GBps_moved": round(moved / (ms_batched * 1e-3) / 1e9, 1),
            "round_trip_rel_l2": max(err_b, err_l)}


def _timed(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def plane_waves(torch, dev, shape, seed, dtype):
    """16 plane-wave cosines (integer wave vectors in [0, 32)^3, amplitudes U[0, 1]) plus
    1e-3 N(0, 1) noise on a grid of shape (n0, n1, n2), x fastest."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n0, n1, n2 = shape
    i = torch.arange(n0, device=dev, dtype=torch.float64) / n0
    j = torch.arange(n1, device=dev, dtype=torch.float64) / n1
    k = torch.arange(n2, device=dev, dtype=torch.float64) / n2
    x = torch.zeros((n2, n1, n0), device=dev, dtype=dtype)
    for _ in range(16):
        kv = rng.integers(0, 32, 3)
        amp = float(rng.uniform())
        ph = (2 * math.pi * (kv[2] * k[:, None, None] + kv[1] * j[None, :, None])).to(dtype)
        x += amp * torch.cos(ph + (2 * math.pi * kv[0] * i[None, None, :]).to(dtype))
        del ph
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    x += 1e-3 * torch.randn(x.shape, device=dev, dtype=dtype, generator=g)
    return x.reshape(-1)


def r2c_leg(heffte, torch, dev, steps):
    """r2c along axis 0 of a 1024 x 1024 x 512 float box, in and out order (0,1,2), out of
    place: one real-data pass (2N reals -> N/2+1 bins) and two complex passes."""
    n0, n1, n2 = 1024, 1024, 512
    inbox = heffte.box3d([0, 0, 0], [n0 - 1, n1 - 1, n2 - 1])
    outbox = inbox.r2c(0)
    plan = heffte.fft3d_r2c(heffte.backend.mi355, inbox, outbox, 0)
    x = plane_waves(torch, dev, (n0, n1, n2), 2, torch.float32)
    y = torch.empty(outbox.count(), dtype=torch.complex64, device=dev)
    z = torch.empty_like(x)
    ws = torch.empty(max(1, plan.size_workspace()), dtype=torch.complex64, device=dev)
    ms_f = _timed(torch, lambda: plan.forward_buffered(x, y, ws), steps)
    ms_b = _timed(torch, lambda: plan.backward_buffered(y, z, ws, heffte.scale.full), steps)
    nr, nc = inbox.count(), outbox.count()
    bytes_f = 4 * nr + 8 * nc + 2 * 2 * 8 * nc
    rt = (torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x)).item()
    del z, ws
    x64 = x.to(torch.float64)
    del x
    ws64 = torch.empty(max(1, plan.size_workspace()), dtype=torch.complex128, device=dev)
    y64 = torch.empty(nc, dtype=torch.complex128, device=dev)
    plan.forward_buffered(x64, y64, ws64)
    ms_f64 = _timed(torch, lambda: plan.forward_buffered(x64, y64, ws64), max(2, steps // 2))
    err = (torch.linalg.vector_norm(y.to(torch.complex128) - y64) / torch.linalg.vector_norm(y64)).item()
    del x64, y64, ws64, y
    torch.cuda.empty_cache()
    return {"workload": "r2c 1024x1024x512 float -> 513x1024x512 complex<float>, axis 0, order (0,1,2), "
                        "out of place, 16 plane waves + 1e-3 noise",
            "ms_forward": round(ms_f, 4), "ms_backward": round(ms_b, 4),
            "forward_hbm_GBps": round(bytes_f / (ms_f * 1e-3) / 1e9, 1),
            "forward_hbm_frac": round(bytes_f / (ms_f * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "forward_bytes": bytes_f, "ms_forward_double": round(ms_f64, 4),
            "rel_l2_vs_fp64": err, "round_trip_rel_l2": rt}


def precision_leg(heffte, torch, dev, plan, n):
    """Forward error of each precision form against the FP64 path on the precision-form inputs:
    a complex Gaussian random field with power spectrum |k|^-5/3 (seed 3) and the same field
    with 6 decades of dynamic range (each element times 10^u, u uniform in [-3, 3], seed 3).
    Relative L2 and the largest error relative to the largest output magnitude."""
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    white = torch.randn((n, n, n), dtype=torch.complex128, device=dev, generator=g)
    k = torch.fft.fftfreq(n, d=1.0 / n, device=dev, dtype=torch.float64)
    kk = torch.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    kk[0, 0, 0] = 1.0
    field = torch.fft.ifftn(white * kk ** (-5.0 / 6.0)).reshape(-1)
    del white, kk
    u = torch.rand(field.shape, dtype=torch.float64, device=dev, generator=g) * 6.0 - 3.0
    cases = {"power_law_k-5/3": field, "six_decades": field * torch.pow(10.0, u)}
    del u
    ws = torch.empty(max(1, plan.size_workspace()), dtype=torch.complex64, device=dev)
    ws64 = torch.empty(max(1, plan.size_workspace()), dtype=torch.complex128, device=dev)
    res = {}
    for name, x64 in cases.items():
        x32 = x64.to(torch.complex64)
        x64 = x32.to(torch.complex128)  # the reference sees the same rounded input
        y64 = torch.empty_like(x64)
        plan.set_precision(heffte.precision.fp32)
        plan.forward_buffered(x64, y64, ws64, heffte.scale.none)
        den, peak = torch.linalg.vector_norm(y64).item(), y64.abs().max().item()
        res[name] = {}
        for form in ("fp32", "split_fp16", "split_bf16"):
            plan.set_precision(getattr(heffte.precision, form))
            y = torch.empty_like(x32)
            plan.forward_buffered(x32, y, ws, heffte.scale.none)
            d = y.to(torch.complex128) - y64
            res[name][form] = {"rel_l2": torch.linalg.vector_norm(d).item() / den,
                               "max_rel": d.abs().max().item() / peak}
            del y, d
        del x32, x64, y64
    plan.set_precision(heffte.precision.fp32)
    del cases, field, ws, ws64
    torch.cuda.empty_cache()
    return res


def reshape_leg(heffte, torch, dev, steps):
    """The bit-exact in-node reshape alone (heFFTe's transpose between pencil orders, here
    permute_kernel): 512^3 complex<float> from order (0,1,2) to (1,2,0) and to (2,1,0), out of
    place.  In the fused path this permutation costs no pass of its own; timed here as what an
    unfused reshape would add.  Beside it, a plain device copy of the same 1 GiB array (torch's
    copy kernel): the rate a read + write of every element reaches without any reordering, the
    practical ceiling the passes are compared with (below the 8 TB/s peak)."""
    n = 512
    box = heffte.box3d([0, 0, 0], [n - 1] * 3)
    x = torch.randn(n ** 3, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    ms = _timed(torch, lambda: y.copy_(x), steps)
    res = {"plain_copy": {"ms": round(ms, 4), "hbm_GBps": round(2 * 8 * n ** 3 / (ms * 1e-3) / 1e9, 1)}}
    for order in ((1, 2, 0), (2, 1, 0)):
        ms = _timed(torch, lambda: heffte.reshape(box, order, x, y), steps)
        res["to_" + "".join(map(str, order))] = {"ms": round(ms, 4),
                                                 "hbm_GBps": round(2 * 8 * n ** 3 / (ms * 1e-3) / 1e9, 1)}
    del x, y
    torch.cuda.empty_cache()
    return res


def convolve_leg(heffte, torch, dev, steps):
    """Convolution through the transform (benchmarks/convolution.cpp:86-95: forward, pointwise
    product, backward) on 512^3 complex<float>, inbox order (0,1,2), outbox (1,2,0).  Unfused =
    the reference's sequence through the plan with the product as a separate torch kernel;
    fused = plan.convolve (product folded into the first backward pass).  'self' squares one
    spectrum as the reference benchmark does; 'pair' convolves two arrays."""
    n = 512
    inbox = heffte.box3d([0, 0, 0], [n - 1] * 3, [0, 1, 2])
    outbox = heffte.box3d([0, 0, 0], [n - 1] * 3, [1, 2, 0])
    plan = heffte.fft3d(heffte.backend.mi355, inbox, outbox)
    x = plane_waves(torch, dev, (n, n, n), 3, torch.complex64)
    y = plane_waves(torch, dev, (n, n, n), 4, torch.complex64)
    out, fy = torch.empty_like(x), torch.empty_like(x)
    ws = torch.empty(max(1, plan.size_convolve_workspace()), dtype=torch.complex64, device=dev)
    wsf = ws[:max(1, plan.size_workspace())]

    def self_unfused():
        plan.forward_buffered(x, out, wsf, heffte.scale.full)
        out.mul_(out)
        plan.backward_buffered(out, out, wsf)

    def pair_unfused():
        plan.forward_buffered(x, out, wsf)
        plan.forward_buffered(y, fy, wsf)
        out.mul_(fy)
        plan.backward_buffered(out, out, wsf, heffte.scale.full)

    res = {"workload": "convolution 512^3 complex<float>, order (0,1,2) -> (1,2,0), plane waves + noise"}
    ms_su = _timed(torch, self_unfused, steps)
    ref = out.clone()
    ms_sf = _timed(torch, lambda: plan.convolve(x, x, out, ws, heffte.scale.none), steps)
    # the reference scales the forward transform (1/N twice through the square); same work
    d_self = (torch.linalg.vector_norm(out / float(n ** 3) ** 2 - ref) / torch.linalg.vector_norm(ref)).item()
    ms_pu = _timed(torch, pair_unfused, steps)
    ref.copy_(out)
    ms_pf = _timed(torch, lambda: plan.convolve(x, y, out, ws, heffte.scale.full), steps)
    d_pair = (torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref)).item()
    # the same with the spectrum of y computed once (a filter applied to many inputs)
    plan.forward_buffered(y, fy, wsf)
    ms_ps = _timed(torch, lambda: plan.convolve_spectrum(x, fy, out, wsf, heffte.scale.full), steps)
    d_spec = (torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref)).item()
    res.update({"self_ms_unfused": round(ms_su, 4), "self_ms_fused": round(ms_sf, 4),
                "self_speedup": round(ms_su / ms_sf, 3), "self_rel_diff": d_self,
                "pair_ms_unfused": round(ms_pu, 4), "pair_ms_fused": round(ms_pf, 4),
                "pair_speedup": round(ms_pu / ms_pf, 3), "pair_rel_diff": d_pair,
                "pair_ms_precomputed_spectrum": round(ms_ps, 4), "precomputed_rel_diff": d_spec})
    del x, y, out, fy, ws, wsf, ref
    torch.cuda.empty_cache()
    return res


def dct_leg(heffte, torch, dev, steps):
    """rtransform<cos> (DCT-II forward, DCT-III backward) of a 256^3 double box, order (0,1,2),
    out of place: three real-to-real passes per direction."""
    n = 256
    box = heffte.box3d([0, 0, 0], [n - 1] * 3)
    plan = heffte.rtransform(heffte.backend.mi355, box, box, heffte.kind.cos)
    x = plane_waves(torch, dev, (n, n, n), 2, torch.float64)
    y, z = torch.empty_like(x), torch.empty_like(x)
    ws = torch.empty(max(1, plan.size_workspace()), dtype=torch.float64, device=dev)
    ms_f = _timed(torch, lambda: plan.forward_buffered(x, y, ws), steps)
    ms_b = _timed(torch, lambda: plan.backward_buffered(y, z, ws, heffte.scale.full), steps)
    bytes_f = 3 * 2 * 8 * n ** 3
    rt = (torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x)).item()
    del x, y, z, ws
    torch.cuda.empty_cache()
    return {"workload": "DCT-II/III 256^3 double, order (0,1,2), out of place, 16 plane waves + 1e-3 noise",
            "ms_forward": round(ms_f, 4), "ms_backward": round(ms_b, 4),
            "forward_hbm_GBps": round(bytes_f / (ms_f * 1e-3) / 1e9, 1),
            "forward_hbm_frac": round(bytes_f / (ms_f * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "forward_bytes": bytes_f, "round_trip_rel_l2": rt}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only (single-rank path); "
                                   "multi-GPU runs are not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import heffte_amd as heffte

    dev = torch.device("cuda:0")
    n = args.n
    N = n ** 3
    cdt = torch.complex64 if args.dtype == "complex64" else torch.complex128
    esize = 8 if args.dtype == "complex64" else 16
    inbox = heffte.box3d([0, 0, 0], [n - 1] * 3, [0, 1, 2])
    outbox = heffte.box3d([0, 0, 0], [n - 1] * 3, [1, 2, 0])
    plan = heffte.fft3d(heffte.backend.mi355, inbox, outbox)

    g = torch.Generator(device=dev)
    g.manual_seed(1)
    data = torch.randn(N, dtype=cdt, device=dev, generator=g) * math.sqrt(2.0)  # re, im ~ N(0,1)
    ws = torch.empty(max(1, plan.size_workspace()), dtype=cdt, device=dev)
    x0 = data.clone() if not args.no_error else None

    def step():
        plan.forward_buffered(data, data, ws, heffte.scale.symmetric)
        plan.backward_buffered(data, data, ws, heffte.scale.symmetric)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    ms_pair = ms_total / args.steps

    # every launch of a pair is an n-point pass (3 forward + 3 backward) reading and writing the
    # whole array; the passes differ in their store layout (column chunks or contiguous lines).
    # Each pass kernel is timed with HIP events on the plan's stream over the same number of pairs
    # (a second timed region, so that the events do not perturb the first); the dominant kernel
    # is the one with the largest total time.
    launches_per_pair = 6
    bytes_per_launch = 2.0 * N * esize  # read + write of the array
    passes = pass_times(plan, step, args.steps, torch, bytes_per_launch, ms_pair)
    dominant = max(passes, key=lambda k: passes[k]["total_ms_events"])
    avg_launch_ms = passes[dominant]["avg_ms"]
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    gflops = 2 * 5.0 * N * math.log2(N) / (ms_pair * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(ms_pair, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_pair, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.dtype == "complex64" else "f64",
        "data": "synthetic (iid N(0,1) real and imaginary parts, torch generator seed 1)",
        "config": {"workload": f"c2c {n}x{n}x{n} {'complex<float>' if esize == 8 else 'complex<double>'}, "
                               "inbox order (0,1,2) -> outbox order (1,2,0), in place, caller workspace, "
                               "forward+backward scale::symmetric",
                   "n": n, "batch": 1, "parallelism": "single rank, 1 GPU"},
        "gflops_5nlog2n": round(gflops, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((n, args.dtype), (None,))[0],
                     "traffic_source": MEASURED_TRAFFIC.get((n, args.dtype), (None, None))[1],
                     "kernel": dominant, "avg_launch_ms": round(avg_launch_ms, 5),
                     "bytes_per_launch": bytes_per_launch,
                     "timing": "the timed region's pair time split by each kernel's share of the time "
                               "between HIP events recorded around each pass launch on the plan's stream "
                               f"({args.steps} more pairs)",
                     "passes": passes},
    }

    # split-precision MFMA forms of the same transform (HMI_PRECISION_SPLIT_FP16 / _BF16)
    if args.dtype == "complex64" and n == 512 and not args.no_split:
        forms = {}
        for name, form in (("split_fp16", heffte.precision.split_fp16), ("split_bf16", heffte.precision.split_bf16)):
            plan.set_precision(form)
            for _ in range(args.warmup):
                step()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(args.steps):
                step()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / args.steps
            fp = pass_times(plan, step, args.steps, torch, bytes_per_launch, ms)
            dom = max(fp, key=lambda k: fp[k]["total_ms_events"])
            entry = {"ms_per_pair": round(ms, 4), "speedup_vs_fp32": round(ms_pair / ms, 3),
                     "hbm_frac_pair_over_6": round(bytes_per_launch / (ms / launches_per_pair * 1e-3) / 1e9 /
                                                   HBM_PEAK_GBS, 4),
                     "dominant_kernel": dom, "dominant_frac": fp[dom]["frac"], "passes": fp}
            if x0 is not None:
                yf = torch.empty_like(x0)
                plan.forward_buffered(x0, yf, ws, heffte.scale.none)
                entry["_y"] = yf
            forms[name] = entry
        plan.set_precision(heffte.precision.fp32)
        out["split_precision"] = forms

    if args.calibrate_copy:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE per dispatch
        # calibrate the counters read for the FFT passes in the same profiled run
        src = torch.empty(N, dtype=cdt, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = N * esize
        del src, dst

    if not args.no_error:
        # error of one forward transform against the FP64 path (itself pinned to the oracle by the tests)
        y = torch.empty_like(x0)
        plan.forward_buffered(x0, y, ws, heffte.scale.none)
        ws64 = torch.empty(max(1, plan.size_workspace()), dtype=torch.complex128, device=dev)
        x64 = x0.to(torch.complex128)
        del x0
        y64 = torch.empty_like(x64)
        plan.forward_buffered(x64, y64, ws64, heffte.scale.none)
        torch.cuda.synchronize()
        num = torch.linalg.vector_norm(y.to(torch.complex128) - y64).item()
        den = torch.linalg.vector_norm(y64).item()
        out["rel_l2_vs_fp64"] = num / den
        for entry in out.get("split_precision", {}).values():
            yf = entry.pop("_y", None)
            if yf is not None:
                entry["rel_l2_vs_fp64"] = torch.linalg.vector_norm(yf.to(torch.complex128) - y64).item() / den
                del yf
        del y, x64, y64, ws64
        torch.cuda.empty_cache()
        if args.dtype == "complex64" and n == 512 and not args.no_split:
            out["precision_inputs"] = precision_leg(heffte, torch, dev, plan, n)

    if not args.no_rocfft:
        try:
            x = torch.randn(n, n, n, dtype=cdt, device=dev, generator=g)
            for _ in range(2):
                torch.fft.ifftn(torch.fft.fftn(x, norm="ortho"), norm="ortho")
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = max(3, args.steps // 2)
            e0.record()
            for _ in range(reps):
                torch.fft.ifftn(torch.fft.fftn(x, norm="ortho"), norm="ortho")
            e1.record()
            torch.cuda.synchronize()
            out["rocfft_torch_fftn_ms_per_pair"] = round(e0.elapsed_time(e1) / reps, 4)
            del x
            torch.cuda.empty_cache()
        except Exception as e:  # pragma: no cover
            out["rocfft_torch_fftn_ms_per_pair"] = f"unavailable: {e}"

    if not args.no_batched and n == 512:
        del data, ws
        torch.cuda.empty_cache()
        out["batched_mixed_1000"] = batched_leg(heffte, torch, dev, max(3, args.steps))
        torch.cuda.empty_cache()
        data = ws = None

    if not args.no_secondary and n == 512:
        del data, ws
        torch.cuda.empty_cache()
        out["r2c_1024x1024x512_f32"] = r2c_leg(heffte, torch, dev, max(3, args.steps // 2))
        out["dct_256_f64"] = dct_leg(heffte, torch, dev, max(3, args.steps // 2))
        out["reshape_512_c64"] = reshape_leg(heffte, torch, dev, max(3, args.steps // 2))
        out["convolve_512_c64"] = convolve_leg(heffte, torch, dev, max(3, args.steps // 2))
        data = ws = None

    x_np = y_gpu = None
    if not args.no_double and n == 512 and args.dtype == "complex64":
        del data, ws
        torch.cuda.empty_cache()
        x_np = seed1_input(n)
        out["c2c_512_complex_double"], y_gpu = double_leg(heffte, torch, dev, plan, args, x_np)
        data = ws = None

    if not args.no_cpu:
        del data, ws
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline(n, x_np, y_gpu)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
