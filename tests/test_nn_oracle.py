"""CPU checks of the learned preconditioner's oracle (oracle/nn_oracle.py).

The network has no counterpart in the reference, so the oracle is pinned by
an independent library (torch's CPU conv3d / conv_transpose3d and autograd),
by adjointness, by the loss's closed form, and by finite differences.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from oracle import nn_oracle as O  # noqa: E402

torch = pytest.importorskip("torch")


def _torch_weight(W):
    """W[tap][ci][co] -> torch conv3d weight [co][ci][kd][kh][kw]."""
    return torch.from_numpy(np.ascontiguousarray(W.reshape(3, 3, 3, W.shape[1], W.shape[2]).transpose(4, 3, 0, 1, 2)))


@pytest.mark.parametrize("shape", [(2, 5, 6, 7), (1, 8, 8, 8), (1, 3, 9, 4)])
def test_conv_matches_torch_conv3d(shape):
    rng = np.random.default_rng(1)
    x = rng.normal(size=shape + (16,))
    W = rng.normal(size=(27, 16, 16))
    ref = torch.nn.functional.conv3d(torch.from_numpy(x).permute(0, 4, 1, 2, 3), _torch_weight(W), padding=1)
    out = O.conv(x, W)
    np.testing.assert_allclose(out, ref.permute(0, 2, 3, 4, 1).numpy(), rtol=1e-10, atol=1e-10)


def test_flipped_filter_is_the_adjoint_conv():
    """conv with flip_transpose(W) = torch conv_transpose3d = adjoint of conv."""
    rng = np.random.default_rng(2)
    dz = rng.normal(size=(2, 6, 5, 7, 16))
    W = rng.normal(size=(27, 16, 16))
    ref = torch.nn.functional.conv_transpose3d(torch.from_numpy(dz).permute(0, 4, 1, 2, 3), _torch_weight(W), padding=1)
    out = O.conv(dz, O.flip_transpose(W))
    np.testing.assert_allclose(out, ref.permute(0, 2, 3, 4, 1).numpy(), rtol=1e-10, atol=1e-10)
    x = rng.normal(size=dz.shape)
    assert np.isclose(np.sum(O.conv(x, W) * dz), np.sum(x * out), rtol=1e-12)


def test_weight_gradient_matches_autograd():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 5, 4, 6, 16))
    dz = rng.normal(size=(2, 5, 4, 6, 16))
    W = rng.normal(size=(27, 16, 16))
    wt = _torch_weight(W).clone().requires_grad_(True)
    y = torch.nn.functional.conv3d(torch.from_numpy(x).permute(0, 4, 1, 2, 3), wt, padding=1)
    (y * torch.from_numpy(dz).permute(0, 4, 1, 2, 3)).sum().backward()
    g = wt.grad.numpy()  # [co][ci][kd][kh][kw]
    ref = g.transpose(2, 3, 4, 1, 0).reshape(27, 16, 16)
    np.testing.assert_allclose(O.conv_dw(x, dz), ref, rtol=1e-10, atol=1e-10)


def test_pool_up_and_their_adjoints():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, 5, 6, 3, 16))
    p = O.pool(x)
    assert p.shape == (2, 3, 3, 2, 16)
    # explicit loop: missing children count as zero, divide by 8
    for (b, z, y, xx) in [(0, 0, 0, 0), (1, 2, 2, 1), (0, 2, 1, 1)]:
        s = np.zeros(16)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    fz, fy, fx = 2 * z + dz, 2 * y + dy, 2 * xx + dx
                    if fz < 5 and fy < 6 and fx < 3:
                        s += x[b, fz, fy, fx]
        np.testing.assert_allclose(p[b, z, y, xx], s / 8, rtol=1e-13)
    c = rng.normal(size=p.shape)
    assert np.isclose(np.sum(p * c), np.sum(x * O.pool_back(c, x.shape[1:4])), rtol=1e-12)
    u = O.up(c, x.shape[1:4])
    assert u.shape == x.shape and np.array_equal(u[1, 4, 5, 2], c[1, 2, 2, 1])
    assert np.isclose(np.sum(u * x), np.sum(c * O.up_back(x, c.shape[1:4])), rtol=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40], np.float32)
    r = O.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0           # tie -> even
    assert r[2] == np.float32(1.0 + 2 ** -7)      # above the tie -> up
    assert r[3] == -2.5
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_loss_closed_form():
    """loss = 0 when y is any non-zero multiple of the error direction, in (0,1) otherwise."""
    rng = np.random.default_rng(5)
    e = rng.normal(size=(2, 6, 7, 5))
    r = O.k_apply(e)  # K e = r  ->  e = K^-1 r
    G = np.einsum("bzyx,bzyx->b", e, r)
    for s in (1.0, 3.0, -2.0):
        loss, dy = O.loss_and_dy(s * e, r, G)
        assert abs(loss) < 1e-12
        assert np.max(np.abs(dy)) < 1e-10
    loss, _ = O.loss_and_dy(rng.normal(size=e.shape), r, G)
    assert 0.0 < loss < 1.0


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    p = O.init_params(7).astype(np.float64)
    p[O.LAYER_PARAMS * np.arange(O.N_LAYERS)[:, None] + O.TAPS * 256 + np.arange(16)] = rng.normal(0, 0.1, (14, 16))
    r = rng.normal(size=(2, 8, 8, 8))
    G = np.array([3.0, 5.0])
    loss, grad = O.loss_grad(p, r, G, rounding=False)
    idx = []
    for l, (cin, cout, *_rest) in enumerate(O.NET):
        base = l * O.LAYER_PARAMS
        for _ in range(2):
            t, ci, co = rng.integers(27), rng.integers(cin), rng.integers(cout)
            idx.append(base + t * 256 + ci * 16 + co)
        idx.append(base + O.TAPS * 256 + rng.integers(cout))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (O.loss_grad(pp, r, G, rounding=False)[0] - O.loss_grad(pm, r, G, rounding=False)[0]) / (2 * eps)
        assert np.isclose(grad[i], fd, rtol=2e-4, atol=1e-9), (i, grad[i], fd)
    # entries outside the channel counts carry no gradient
    lift = grad[:27 * 256].reshape(27, 16, 16)
    assert np.all(lift[:, 1:, :] == 0)
    out = grad[13 * O.LAYER_PARAMS:13 * O.LAYER_PARAMS + 27 * 256].reshape(27, 16, 16)
    assert np.all(out[:, :, 1:] == 0)


def test_rounded_forward_stays_close_to_exact():
    """BF16 storage changes the output by well under a per cent of its size."""
    rng = np.random.default_rng(8)
    p = O.init_params(9)
    r = rng.normal(size=(1, 12, 10, 9)).astype(np.float32)
    y32, _ = O.forward(p, r, rounding=True)
    y64, _ = O.forward(p.astype(np.float64), r.astype(np.float64), rounding=False)
    assert np.linalg.norm(y32 - y64) / np.linalg.norm(y64) < 2e-2


def test_exact_out_skip_changes_only_the_output_skip_rounding():
    """exact_out_skip leaves every activation as it is and, without rounding,
    the output too; with rounding the output moves by about the BF16 rounding
    of the skip sum (0.4 % per value at most), not more."""
    rng = np.random.default_rng(10)
    p = O.init_params(11)
    r = rng.normal(size=(1, 11, 9, 14)).astype(np.float32)
    y0, a0 = O.forward(p, r, rounding=True)
    y1, a1 = O.forward(p, r, rounding=True, exact_out_skip=True)
    for u, v in zip(a0, a1):
        assert np.array_equal(u, v)
    d = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
    assert 0.0 < d < 1e-2
    z0, _ = O.forward(p.astype(np.float64), r.astype(np.float64), rounding=False)
    z1, _ = O.forward(p.astype(np.float64), r.astype(np.float64), rounding=False, exact_out_skip=True)
    assert np.array_equal(z0, z1)


def test_output_layer_matches_forward():
    """output_layer from the forward's own activations reproduces the forward's
    output (rounded skip sum) and its exact_out_skip variant."""
    rng = np.random.default_rng(12)
    p = O.init_params(13)
    r = rng.normal(size=(1, 10, 7, 12)).astype(np.float32)
    y, acts = O.forward(p, r, rounding=True)
    assert np.allclose(O.output_layer(p, acts[0], acts[12]), y, rtol=0, atol=1e-6 * np.abs(y).max())
    ye, _ = O.forward(p, r, rounding=True, exact_out_skip=True)
    assert np.allclose(O.output_layer(p, acts[0], acts[12], exact_skip=True), ye, rtol=0, atol=1e-6 * np.abs(y).max())
