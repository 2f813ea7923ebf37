"""The CPU oracle of the learned preconditioner (oracle/nn_oracle.py), pinned
without the GPU: its convolution, pooling and upsampling against PyTorch's
float32 operators, its backward pass against central finite differences of
the float64 network, its BF16 rounding against torch's bfloat16 cast, and its
parameter layout against the library's count."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from bevy_eulerian_fluid_amd import native
from oracle import nn_oracle as O


def test_param_count_matches_library():
    assert O.param_count() == native.Network.param_count()
    assert 0.2e6 < O.param_count() < 0.35e6


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(10000) * 10.0 ** rng.integers(-8, 8, 10000)).astype(np.float32)
    x[:4] = [1.00390625, 1.01171875, -2.0, 0.0]  # ties to even
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(O.bf16(x), ref)


def test_conv_pool_up_match_torch():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 12, 10, 16))
    w = rng.standard_normal((8, 16, 3, 3))
    b = rng.standard_normal(8)
    ours = O.conv3x3(x, w, b)
    ref = F.conv2d(torch.from_numpy(x).permute(0, 3, 1, 2).float(), torch.from_numpy(w).float(),
                   torch.from_numpy(b).float(), padding=1).permute(0, 2, 3, 1).double().numpy()
    assert np.allclose(ours, ref, rtol=1e-4, atol=1e-4)
    xt = torch.from_numpy(x).permute(0, 3, 1, 2)
    assert np.allclose(O.pool2(x), F.avg_pool2d(xt, 2).permute(0, 2, 3, 1).numpy())
    assert np.allclose(O.up2(x), F.interpolate(xt, scale_factor=2, mode="nearest").permute(0, 2, 3, 1).numpy())


def _planes(rng, n, h, w):
    p = np.zeros((n, 5, h, w))
    p[:, 0] = rng.standard_normal((n, h, w))
    p[:, 1] = rng.uniform(0, 1, (n, h, w))
    p[:, 2] = rng.uniform(-1, 1, (n, h, w))
    p[:, 3] = rng.uniform(0, 1, (n, h, w)) < 0.7
    p[:, 4] = rng.standard_normal((n, h, w))
    return p


@pytest.mark.parametrize("seed", [3, 4])
def test_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    params = rng.standard_normal(O.param_count()) * 0.1
    planes = _planes(rng, 1, 16, 16)
    loss, grads, _ = O.loss_and_grads(params, planes, round_=False)
    # one weight and one bias of every layer, plus a few random entries
    idx = []
    off = 0
    for ca, cb, co, *_ in O.NET:
        nw = co * (ca + cb) * 9
        idx += [off + int(rng.integers(0, min(nw, 9 * co * 3))), off + nw + int(rng.integers(0, co))]
        off += nw + co
    idx += list(rng.integers(0, O.param_count(), 10))
    eps = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (O.loss_and_grads(pp, planes, False)[0] - O.loss_and_grads(pm, planes, False)[0]) / (2 * eps)
        assert abs(fd - grads[i]) <= 2e-4 * max(1.0, abs(fd)) + 1e-7, (i, fd, grads[i])
