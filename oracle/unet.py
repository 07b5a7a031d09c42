"""CPU restatement of the learned density initialiser (test infrastructure).

The reference has no network or training code: its ML start only consumes a
density file (src/V_ion/pseudoDiag_ML4Den_poisson.py:392-540, restated in
``oracle/scf.py``'s ``rho_init``).  This module fixes the network the HIP
kernels implement and computes its forward pass, loss and gradients in NumPy
(FP64 accumulation), rounding to BF16 at exactly the points where the device
path stores BF16, so the device gradients can be compared with it.

Network (a 4-level 3D U-Net, 3x3x3 convolutions, zero padding 1):
    enc   e1a 1->16, e1b 16->16 | pool | e2a 16->32, e2b 32->32 | pool |
          e3a 32->48, e3b 48->48 | pool | ba 48->64, bb 64->64
    dec   up(64) ++ e3 -> d3a 112->48, d3b 48->48 | up(48) ++ e2 -> d2a 80->32,
          d2b 32->32 | up(32) ++ e1 -> d1a 48->16, d1b 16->16 | out 16->1
Every conv but ``out`` is followed by leaky ReLU (slope 0.1).  pool = 2x2x2
average, up = 2x nearest-neighbour, ++ = channel concatenation (upsampled
channels first).  Weights W[co][ci][t], tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).

Precision contract (device and oracle): weights used in the convolutions are
the BF16 rounding of the FP32 master weights; every activation that feeds a
convolution (network input, leaky-ReLU outputs, pooled and upsampled tensors)
is stored in BF16; conv outputs are FP32 (FP32 accumulation on the device);
the output layer is FP32.  Backward: the conv-output gradient dZ is FP32; the
bias gradient sums dZ; dZ is rounded to BF16 for the weight gradient and the
input gradient, which accumulate in FP32.

Pinning: ``grad_check`` compares every gradient with central differences of
the loss in an all-FP64 variant of the same network (``bf16=False``); the tests
run it on small volumes.
"""
from __future__ import annotations

import numpy as np

LEVELS = (16, 32, 48, 64)
SLOPE = 0.1


def layer_specs():
    c1, c2, c3, c4 = LEVELS
    return [("e1a", 1, c1), ("e1b", c1, c1), ("e2a", c1, c2), ("e2b", c2, c2), ("e3a", c2, c3), ("e3b", c3, c3),
            ("ba", c3, c4), ("bb", c4, c4), ("d3a", c4 + c3, c3), ("d3b", c3, c3), ("d2a", c3 + c2, c2),
            ("d2b", c2, c2), ("d1a", c2 + c1, c1), ("d1b", c1, c1), ("out", c1, 1)]


def init_params(seed: int = 0):
    """He-uniform weights (fan-in 27*cin), zero biases; the output layer starts at zero."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, ci, co in layer_specs():
        if name == "out":
            w = np.zeros((co, ci, 27), np.float32)
        else:
            lim = np.sqrt(6.0 / (27 * ci))
            w = rng.uniform(-lim, lim, (co, ci, 27)).astype(np.float32)
        params[name] = (w, np.zeros(co, np.float32))
    return params


def bf16(x):
    """Round FP32 values to the nearest BF16 (ties to even), returned as FP32."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def _taps():
    return [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv3d(x, w, b):
    """x [ci,D,H,W], w [co,ci,27], b [co] -> [co,D,H,W] (FP64 accumulation)."""
    ci, D, H, W = x.shape
    xp = np.pad(x.astype(np.float64), ((0, 0), (1, 1), (1, 1), (1, 1)))
    out = np.zeros((w.shape[0], D, H, W))
    for t, (dz, dy, dx) in enumerate(_taps()):
        xs = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        out += np.tensordot(w[:, :, t].astype(np.float64), xs, axes=(1, 0))
    return out + b.astype(np.float64)[:, None, None, None]


def conv3d_bwd(x, w, dz):
    """Gradients of conv3d: (dx [ci,...], dw [co,ci,27]); dz [co,D,H,W]."""
    ci, D, H, W = x.shape
    xp = np.pad(x.astype(np.float64), ((0, 0), (1, 1), (1, 1), (1, 1)))
    dxp = np.zeros_like(xp)
    dw = np.zeros(w.shape)
    dz = dz.astype(np.float64)
    for t, (a, bb, c) in enumerate(_taps()):
        sl = (slice(None), slice(1 + a, 1 + a + D), slice(1 + bb, 1 + bb + H), slice(1 + c, 1 + c + W))
        dw[:, :, t] = np.tensordot(dz, xp[sl], axes=([1, 2, 3], [1, 2, 3]))
        dxp[sl] += np.tensordot(w[:, :, t].astype(np.float64), dz, axes=(0, 0))
    return dxp[:, 1:-1, 1:-1, 1:-1], dw


def leaky(z):
    return np.where(z > 0, z, SLOPE * z)


def pool2(x):
    c, D, H, W = x.shape
    return x.reshape(c, D // 2, 2, H // 2, 2, W // 2, 2).astype(np.float64).mean(axis=(2, 4, 6))


def pool2_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up2_bwd(dy):
    c, D, H, W = dy.shape
    return dy.reshape(c, D // 2, 2, H // 2, 2, W // 2, 2).sum(axis=(2, 4, 6))


class Net:
    """Forward + backward of the U-Net on one volume (D, H, W multiples of 8)."""

    def __init__(self, params, use_bf16: bool = True):
        self.p = params
        self.use_bf16 = use_bf16
        r = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
        self.r = r
        self.wq = {k: r(w) for k, (w, _) in params.items()}

    def _q(self, a):
        return self.r(a) if self.use_bf16 else np.asarray(a, np.float64)

    def _f32(self, a):
        return np.asarray(a, np.float32) if self.use_bf16 else np.asarray(a, np.float64)

    def conv(self, name, a_in, act=True):
        z = self._f32(conv3d(a_in, self.wq[name], self.p[name][1]))
        self.cache[name] = a_in
        if not act:
            return z
        return self._q(leaky(z))

    def forward(self, x):
        self.cache = {}
        self.acts = {}
        a = self._q(np.asarray(x)[None])
        e1 = self.conv("e1b", self.conv("e1a", a))
        p1 = self._q(pool2(e1))
        e2 = self.conv("e2b", self.conv("e2a", p1))
        p2 = self._q(pool2(e2))
        e3 = self.conv("e3b", self.conv("e3a", p2))
        p3 = self._q(pool2(e3))
        bb = self.conv("bb", self.conv("ba", p3))
        c3 = np.concatenate([up2(bb), e3], axis=0)
        d3 = self.conv("d3b", self.conv("d3a", c3))
        c2 = np.concatenate([up2(d3), e2], axis=0)
        d2 = self.conv("d2b", self.conv("d2a", c2))
        c1 = np.concatenate([up2(d2), e1], axis=0)
        d1 = self.conv("d1b", self.conv("d1a", c1))
        y = self.conv("out", d1, act=False)
        self.acts = dict(e1=e1, e2=e2, e3=e3, bb=bb, d3=d3, d2=d2, d1=d1)
        # outputs of each activated conv (for the leaky-ReLU derivative)
        self.outs = {"e1a": self.cache["e1b"], "e1b": e1, "e2a": self.cache["e2b"], "e2b": e2,
                     "e3a": self.cache["e3b"], "e3b": e3, "ba": self.cache["bb"], "bb": bb,
                     "d3a": self.cache["d3b"], "d3b": d3, "d2a": self.cache["d2b"], "d2b": d2,
                     "d1a": self.cache["d1b"], "d1b": d1}
        return y[0]

    def loss(self, y, t):
        d = np.asarray(y, np.float64) - np.asarray(t, np.float64)
        return float(np.mean(d * d)), self._f32(2.0 * d / d.size)

    def backward(self, dy):
        """dy [D,H,W] (FP32) -> dict name -> (dW, db)."""
        grads = {}

        def conv_bwd(name, dz, act_out=None):
            # dz: gradient at the conv output (after the leaky ReLU derivative)
            db = np.asarray(dz, np.float64).sum(axis=(1, 2, 3))
            dzq = self._q(dz)
            dx, dw = conv3d_bwd(self.cache[name], self.wq[name], dzq)
            grads[name] = (dw, db)
            return self._f32(dx)

        def act_bwd(name, da):
            a = self.outs[name]
            return self._f32(np.where(a > 0, da, SLOPE * da))

        g = conv_bwd("out", np.asarray(dy)[None])
        g = conv_bwd("d1b", act_bwd("d1b", g))
        g = conv_bwd("d1a", act_bwd("d1a", g))
        c2 = LEVELS[1]
        g_up, g_e1 = g[:c2], g[c2:]
        g = self._f32(up2_bwd(g_up))
        g = conv_bwd("d2b", act_bwd("d2b", g))
        g = conv_bwd("d2a", act_bwd("d2a", g))
        c3 = LEVELS[2]
        g_up, g_e2 = g[:c3], g[c3:]
        g = self._f32(up2_bwd(g_up))
        g = conv_bwd("d3b", act_bwd("d3b", g))
        g = conv_bwd("d3a", act_bwd("d3a", g))
        c4 = LEVELS[3]
        g_up, g_e3 = g[:c4], g[c4:]
        g = self._f32(up2_bwd(g_up))
        g = conv_bwd("bb", act_bwd("bb", g))
        g = conv_bwd("ba", act_bwd("ba", g))
        g = self._f32(pool2_bwd(g) + g_e3)
        g = conv_bwd("e3b", act_bwd("e3b", g))
        g = conv_bwd("e3a", act_bwd("e3a", g))
        g = self._f32(pool2_bwd(g) + g_e2)
        g = conv_bwd("e2b", act_bwd("e2b", g))
        g = conv_bwd("e2a", act_bwd("e2a", g))
        g = self._f32(pool2_bwd(g) + g_e1)
        g = conv_bwd("e1b", act_bwd("e1b", g))
        conv_bwd("e1a", act_bwd("e1a", g))
        return grads


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam update (bias-corrected), FP64 reference of the device kernel."""
    g = np.asarray(g, np.float64) + wd * np.asarray(w, np.float64)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return np.asarray(w, np.float64) - lr * mh / (np.sqrt(vh) + eps), m, v


def grad_check(params, x, t, names, n_probe=3, eps=1e-3, seed=0):
    """Central differences of the FP64 network's loss for a few weights of each
    named layer against its analytic gradient; returns the worst relative error."""
    rng = np.random.default_rng(seed)
    net = Net(params, use_bf16=False)
    y = net.forward(x)
    _, dy = net.loss(y, t)
    g = net.backward(dy)
    worst = 0.0
    for name in names:
        w, b = params[name]
        for _ in range(n_probe):
            idx = tuple(rng.integers(0, s) for s in w.shape)
            w0 = w[idx]
            vals = []
            for s in (+1, -1):
                pp = {k: (v[0].astype(np.float64).copy(), v[1].astype(np.float64)) for k, v in params.items()}
                pp[name][0][idx] = w0 + s * eps
                n2 = Net(pp, use_bf16=False)
                vals.append(n2.loss(n2.forward(x), t)[0])
            num = (vals[0] - vals[1]) / (2 * eps)
            ana = g[name][0][idx]
            worst = max(worst, abs(num - ana) / max(abs(num), abs(ana), 1e-12))
    return worst
