"""Bit-exact GPU reshape (in-node transpose) on integer-valued payloads.

Known answers: test/test_units_nompi.cpp:577-637 (2x3x4 box to orders (1,2,0),
(2,1,0), (0,2,1) and back); general case: the index-encoding payloads of
test/test_reshape3d.cpp:47-67 (every entry holds its own global index) checked
against the oracle's permutation for all order pairs, element sizes 4/8/16
bytes, ragged extents and batches.
"""
import numpy as np
import pytest

import fftcase as fc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [(1, 2, 0), (2, 1, 0), (0, 2, 1)])
def test_reference_transposes(hmi, cuda, golden, order):
    import torch
    box = hmi.box3d([0, 0, 0], [1, 2, 3], (0, 1, 2))
    x = torch.arange(1, 25, dtype=torch.float64, device=cuda)
    y = torch.empty_like(x)
    hmi.reshape(box, order, x, y)
    ref = np.array(golden["transpose"]["order_" + "".join(map(str, order))])
    assert np.array_equal(fc.to_host(y), ref)
    back = torch.empty_like(x)
    hmi.reshape(hmi.box3d([0, 0, 0], [1, 2, 3], order), (0, 1, 2), y, back)
    assert torch.equal(back, x)


@pytest.mark.parametrize("dtype", ["int32", "int64", "complex128"])
@pytest.mark.parametrize("oi", range(6))
@pytest.mark.parametrize("oo", range(6))
def test_index_payload_all_orders(hmi, cuda, oracle, dtype, oi, oo):
    import torch
    inorder, outorder = fc.ORDERS[oi], fc.ORDERS[oo]
    low, high = (3, -2, 5), (3 + 36, -2 + 40, 5 + 32)  # 37 x 43 x 33, ragged against the 32-wide tiles
    batch = 2
    box = hmi.box3d(low, high, inorder)
    n = box.count()
    payload = np.arange(batch * n, dtype=np.int64) * 7 + 1
    if dtype == "int32":
        x = torch.from_numpy(payload.astype(np.int32)).to(cuda)
    elif dtype == "int64":
        x = torch.from_numpy(payload).to(cuda)
    else:
        x = torch.from_numpy(payload.astype(np.float64) + 1j * (-payload.astype(np.float64))).to(cuda)
    y = torch.empty_like(x)
    hmi.reshape(box, outorder, x, y, batch=batch)
    ob_in = oracle.Box.make(low, high, inorder)
    ob_out = oracle.Box.make(low, high, outorder)
    host = x.cpu().numpy()
    ref = np.concatenate([oracle.transpose(ob_in, ob_out, host[b * n:(b + 1) * n]) for b in range(batch)])
    assert np.array_equal(y.cpu().numpy(), ref)


def test_reshape_rejects_overlap_and_bad_sizes(hmi, cuda):
    import torch
    box = hmi.box3d([0, 0, 0], [3, 3, 3])
    x = torch.zeros(64, dtype=torch.float32, device=cuda)
    with pytest.raises(hmi.heffte_input_error):
        hmi.reshape(box, (2, 1, 0), x, x)
    h = torch.zeros(64, dtype=torch.int16, device=cuda)
    with pytest.raises(hmi.heffte_input_error):
        hmi.reshape(box, (2, 1, 0), h, torch.zeros_like(h))
