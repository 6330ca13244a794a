"""Shared helpers of the GPU parity tests: seeded inputs, layout conversion and
the oracle call that a GPU result is compared with."""
import numpy as np

ORDERS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
TOL = {"complex64": 5e-6, "complex128": 1e-13, "float32": 5e-6, "float64": 1e-13}


def flat(a, order):
    """numpy array indexed [i0, i1, i2] -> flat array stored in box order `order`."""
    return np.ascontiguousarray(a.transpose(order[2], order[1], order[0])).ravel()


def unflat(v, size, order):
    """inverse of flat()."""
    shape = (size[order[2]], size[order[1]], size[order[0]])
    a = np.asarray(v).reshape(shape)
    inv = np.argsort([order[2], order[1], order[0]])
    return a.transpose(inv)


def uniform_complex(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


def to_gpu(x, dtype):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x).astype(dtype))
    return t.to("cuda:0")


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.complex128 if t.is_complex() else np.float64)


def rel_l2(y, ref):
    y = np.asarray(y).ravel()
    ref = np.asarray(ref).ravel()
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(y - ref) / den) if den > 0 else float(np.linalg.norm(y))
