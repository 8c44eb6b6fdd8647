"""Warm-start network: the FP64 reference is pinned by finite differences (CPU); the HIP forward/backward kernels
match the reference (GPU); training on the GPU lowers the loss."""
import numpy as np
import pytest

from impressionfea_amd.mlp import init_params, layer_shapes


def test_reference_gradient_finite_difference():
    import mlp_ref as R
    shapes = layer_shapes(5, 8, 2, 3)
    rng = np.random.default_rng(0)
    p = init_params(1, 5, 8, 2, 3).astype(np.float64) * 3
    x = rng.standard_normal((7, 5))
    y = rng.standard_normal((7, 3))
    loss, g = R.loss_grad(p, shapes, x, y, bf16=False)
    for i in rng.choice(p.size, 25, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_grad(pp, shapes, x, y, False)[0] - R.loss_grad(pm, shapes, x, y, False)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9)


def test_bf16_rounding_ties_to_even():
    import mlp_ref as R
    one = np.float32(1.0)
    step = np.float32(2.0 ** -8)  # bf16 spacing at 1.0 is 2^-7, so 1 + 2^-8 is a tie
    assert R.to_bf16(np.array([one + step]))[0] == 1.0
    assert R.to_bf16(np.array([np.float32(1.0 + 3 * 2.0 ** -8)]))[0] == 1.0 + 2.0 ** -6


@pytest.mark.gpu
def test_device_forward_backward_match_reference():
    import mlp_ref as R
    from impressionfea_amd.mlp import MlpDevice
    shapes = layer_shapes()
    rng = np.random.default_rng(3)
    p = init_params(2) * np.float32(3.0)
    net = MlpDevice(p)
    x = rng.standard_normal((1000, 11)).astype(np.float32)
    y = rng.standard_normal((1000, 3)).astype(np.float32)
    yh = net.forward(x)
    yr = R.forward(p, shapes, x)[0]
    # Both sides round the same operands to BF16, but from activations that differ in the last FP32 bits (FP32
    # accumulation on the device, FP64 in the reference): an activation within those bits of a BF16 rounding boundary
    # rounds to the neighbouring BF16 value on one side (a 2^-8 relative step), so a few outputs differ by up to a few
    # such steps of one term while almost all agree to the FP32 accumulation error.
    err = np.abs(yh - yr) / np.abs(yr).max()
    flipped = float(np.mean(err > 1e-5))
    print(f"forward: median {np.median(err):.2e}, max {err.max():.2e}, outputs beyond 1e-5: {100 * flipped:.2f} %")
    assert np.median(err) <= 1e-6 and flipped <= 0.02 and err.max() <= 4e-3
    loss, g = net.loss_grad(x, y)
    lr, gr = R.loss_grad(p, shapes, x, y)
    gerr = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print(f"loss {loss:.8e} vs {lr:.8e}; gradient relative error {gerr:.2e}")
    assert loss == pytest.approx(lr, rel=1e-5)
    assert gerr <= 1e-4
    # bit-reproducible reductions
    loss2, g2 = net.loss_grad(x, y)
    assert loss2 == loss and np.array_equal(g, g2)


@pytest.mark.gpu
def test_device_training_reduces_loss():
    from impressionfea_amd.mlp import MlpDevice
    rng = np.random.default_rng(5)
    x = rng.standard_normal((20000, 11)).astype(np.float32)
    w = rng.standard_normal((11, 3)).astype(np.float32) * 0.3
    y = np.tanh(x @ w).astype(np.float32)
    net = MlpDevice()
    net.set_dataset(x, y)
    first = net.train_epoch(rng.permutation(x.shape[0]).astype(np.int32), 2048, 2e-3)
    for _ in range(10):
        last = net.train_epoch(rng.permutation(x.shape[0]).astype(np.int32), 2048, 2e-3)
    assert last < 0.2 * first
