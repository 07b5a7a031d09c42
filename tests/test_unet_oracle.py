"""The U-Net oracle pinned: central differences of its FP64 loss, BF16 rounding
known answers, and the adjoint identities of pooling / upsampling / conv."""
import numpy as np

from oracle import unet as U


def _params(seed):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    p["out"] = (rng.uniform(-0.1, 0.1, p["out"][0].shape).astype(np.float32), np.zeros(1, np.float32))
    return {k: (w, rng.uniform(-0.05, 0.05, b.shape).astype(np.float32)) for k, (w, b) in p.items()}


def test_gradients_match_central_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((8, 8, 8)).astype(np.float32)
    t = rng.standard_normal((8, 8, 8)).astype(np.float32)
    worst = U.grad_check(_params(1), x, t, ["e1a", "e2b", "ba", "bb", "d3a", "d2a", "d1b", "out"], n_probe=2,
                         eps=1e-5)
    assert worst < 1e-3


def test_bf16_rounding_known_answers():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)                    # BF16 spacing at 1
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0],
                    dtype=np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 * 2.0 ** -7, 1.0 + ulp, -2.5, 0.0], dtype=np.float32)
    assert np.array_equal(U.bf16(vals), want)
    assert U.bf16(one) == one


def test_adjoints():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((3, 8, 8, 8))
    c = rng.standard_normal((3, 4, 4, 4))
    assert np.isclose(np.sum(U.pool2(a) * c), np.sum(a * U.pool2_bwd(c)))
    assert np.isclose(np.sum(U.up2(c) * a), np.sum(c * U.up2_bwd(a)))
    w = rng.standard_normal((5, 3, 27))
    dz = rng.standard_normal((5, 8, 8, 8))
    dx, dw = U.conv3d_bwd(a, w, dz)
    assert np.isclose(np.sum(U.conv3d(a, w, np.zeros(5)) * dz), np.sum(a * dx))
    # dw: linearity in w
    assert np.isclose(np.sum(U.conv3d(a, w, np.zeros(5)) * dz), np.sum(w * dw))


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 4, 5, 6))
    w = rng.standard_normal((3, 2, 27))
    b = rng.standard_normal(3)
    y = U.conv3d(x, w, b)
    co, z, yy, xx = 1, 0, 4, 3
    s = b[co]
    for t, (dz, dy, dx) in enumerate([(a, bb, c) for a in (-1, 0, 1) for bb in (-1, 0, 1) for c in (-1, 0, 1)]):
        zz, y2, x2 = z + dz, yy + dy, xx + dx
        if 0 <= zz < 4 and 0 <= y2 < 5 and 0 <= x2 < 6:
            s += np.dot(w[co, :, t], x[:, zz, y2, x2])
    assert np.isclose(y[co, z, yy, xx], s)


def test_adam_matches_closed_form_first_step():
    w = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    w1, m, v = U.adam_step(w, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # bias-corrected first step moves every weight by lr * sign(g)
    assert np.allclose(w1, w - 0.1 * np.sign(g), atol=1e-7)
