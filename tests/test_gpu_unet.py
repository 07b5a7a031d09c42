"""U-Net kernels against the CPU oracle (oracle/unet.py) on the same BF16
inputs.  Convolutions: FP32 accumulation vs the oracle's FP64 sum, so 1e-5
relative.  Whole network: the only differences are FP32 accumulation order and
the BF16 roundings it flips (1 BF16 ulp = 0.4 %), which compound through up to
14 layers: outputs within 1e-2 and every gradient within 5e-2 relative L2,
cosine similarity > 0.998."""
import ctypes as C

import numpy as np
import pytest

from oracle import unet as U

pytestmark = pytest.mark.gpu


def _bf16_tensor(a):
    """float32 array (already BF16-representable) -> torch int16 raw bits on the device."""
    import torch
    raw = np.ascontiguousarray(U.bf16(a)).view(np.uint32) >> 16
    return torch.from_numpy(raw.astype(np.uint16).view(np.int16)).to("cuda")


def _bits_to_f32(t):
    raw = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return raw.view(np.float32)


@pytest.mark.parametrize("cin,cout,ldx,ldy", [(16, 16, 16, 16), (1, 16, 1, 16), (112, 48, 112, 48), (16, 1, 16, 1),
                                              (32, 64, 48, 80), (64, 32, 64, 32)])
@pytest.mark.parametrize("mode", [0, 1])
def test_conv3d_fwd(gpu, cin, cout, ldx, ldy, mode):
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    D, H, W = 6, 8, 10
    print(f"conv fwd cin={cin} cout={cout} ld={ldx},{ldy} mode={mode}", flush=True)
    rng = np.random.default_rng(cin * 100 + cout)
    x = U.bf16(rng.standard_normal((D * H * W, ldx)).astype(np.float32))
    w = U.bf16(rng.standard_normal((cout, 27, cin)).astype(np.float32) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    X = _bf16_tensor(x)
    Wt = _bf16_tensor(w)
    B = torch.from_numpy(b).to("cuda")
    if mode == 1:
        Y = torch.zeros(D * H * W * ldy, dtype=torch.int16, device="cuda")
    else:
        Y = torch.zeros(D * H * W * ldy, dtype=torch.float32, device="cuda")
    _lib.call("pc_unet_conv3d_fwd", ptr(X), ldx, cin, ptr(Wt), ptr(B), cout, D, H, W, mode, 0.1, ptr(Y), ldy,
              stream())
    torch.cuda.synchronize()
    xo = x[:, :cin].T.reshape(cin, D, H, W)
    ref = U.conv3d(xo, w.transpose(0, 2, 1), b).reshape(cout, -1).T    # [V][cout]
    if mode == 1:
        got = _bits_to_f32(Y).reshape(-1, ldy)[:, :cout]
        want = U.bf16(U.leaky(ref).astype(np.float32))
        # BF16 outputs: equal except where FP32 accumulation flips a rounding (1 ulp)
        assert np.max(np.abs(got - want) / (np.abs(want) + 1e-3)) < 1.6e-2
    else:
        got = Y.cpu().numpy().reshape(-1, ldy)[:, :cout]
        assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    if ldy > cout:   # channels beyond cout untouched
        rest = Y.cpu().numpy().reshape(-1, ldy)[:, cout:]
        assert not np.any(rest)


@pytest.mark.parametrize("cin,cout", [(16, 16), (1, 16), (48, 16), (16, 1), (64, 64)])
def test_conv3d_wgrad(gpu, cin, cout):
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    D, H, W = 8, 8, 8
    print(f"wgrad cin={cin} cout={cout}", flush=True)
    rng = np.random.default_rng(7 + cin + cout)
    x = U.bf16(rng.standard_normal((D * H * W, cin)).astype(np.float32))
    dz = U.bf16(rng.standard_normal((D * H * W, cout)).astype(np.float32))
    nb = C.c_size_t()
    _lib.call("pc_unet_wgrad_workspace", cout, cin, C.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    dW = torch.empty(cout * 27 * cin, dtype=torch.float32, device="cuda")
    dZt, Xt = _bf16_tensor(dz), _bf16_tensor(x)     # keep both alive until the kernel has run
    _lib.call("pc_unet_conv3d_wgrad", ptr(dZt), cout, cout, ptr(Xt), cin, cin, D, H, W, ptr(dW), ptr(work),
              nb.value, stream())
    torch.cuda.synchronize()
    _, ref = U.conv3d_bwd(x.T.reshape(cin, D, H, W), np.zeros((cout, cin, 27)), dz.T.reshape(cout, D, H, W))
    got = dW.cpu().numpy().reshape(cout, 27, cin).transpose(0, 2, 1)
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed)
    p["out"] = (rng.uniform(-0.1, 0.1, p["out"][0].shape).astype(np.float32), np.zeros(1, np.float32))
    return {k: (w, rng.uniform(-0.05, 0.05, b.shape).astype(np.float32)) for k, (w, b) in p.items()}


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def _cos(a, b):
    a, b = np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))


def test_unet_forward_backward_vs_oracle(gpu):
    import torch
    from parsec_python_amd.learned import UNet
    print("unet fwd/bwd 16^3", flush=True)
    p = _rand_params(5)
    rng = np.random.default_rng(6)
    D = 16
    x = rng.uniform(0, 1, (D, D, D)).astype(np.float32)
    t = (0.1 * rng.standard_normal((D, D, D))).astype(np.float32)
    net = UNet(p)
    xt = torch.from_numpy(x).to("cuda")
    tt = torch.from_numpy(t).to("cuda")
    y = net.forward(xt)
    loss = net.loss_grad(y, tt)
    net.backward(net.bufs["dy"].view_as(y))
    torch.cuda.synchronize()
    ref = U.Net(p)
    y_ref = ref.forward(x)
    l_ref, dy_ref = ref.loss(y_ref, t)
    g_ref = ref.backward(dy_ref)
    assert _rel(y.cpu().numpy(), y_ref) < 1e-2
    assert abs(float(loss.item()) - l_ref) < 1e-2 * l_ref
    g = net.grads()
    for name in g_ref:
        assert _rel(g[name][0], g_ref[name][0]) < 5e-2, name
        assert _cos(g[name][0], g_ref[name][0]) > 0.998, name
        assert _rel(g[name][1], g_ref[name][1]) < 5e-2, name


def test_adam_matches_oracle(gpu):
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    rng = np.random.default_rng(8)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n)
    v = np.zeros(n)
    W = torch.from_numpy(w.copy()).to("cuda")
    M = torch.zeros(n, device="cuda")
    Vv = torch.zeros(n, device="cuda")
    wb = torch.empty(n, dtype=torch.int16, device="cuda")
    wr = w.astype(np.float64)
    for step in range(1, 4):
        g = rng.standard_normal(n).astype(np.float32)
        G = torch.from_numpy(g).to("cuda")
        _lib.call("pc_unet_adam", ptr(W), ptr(G), ptr(M), ptr(Vv), n, 1e-2, 0.9, 0.999, 1e-8, 1e-4, step, ptr(wb),
                  stream())
        wr, m, v = U.adam_step(wr, g, m, v, step, 1e-2, wd=1e-4)
    torch.cuda.synchronize()
    assert np.allclose(W.cpu().numpy(), wr, rtol=1e-5, atol=1e-6)
    assert np.array_equal(_bits_to_f32(wb), U.bf16(W.cpu().numpy()))


def test_training_reduces_loss(gpu):
    import torch
    from parsec_python_amd.learned import UNet
    rng = np.random.default_rng(9)
    D = 16
    x = torch.from_numpy(rng.uniform(0, 1, (D, D, D)).astype(np.float32)).to("cuda")
    t = 0.3 * x * x - 0.1
    net = UNet(seed=3)
    losses = [float(net.train_step(x, t, lr=1e-3).item()) for _ in range(40)]
    # the FP64 oracle trained the same way reaches 0.17x the initial loss
    assert losses[-1] < 0.35 * losses[0]
