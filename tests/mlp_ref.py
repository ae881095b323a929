"""numpy restatement of the learned warm start's network, for checking the HIP
kernels (closures-2d_amd/csrc/mnc_mlp.hip): MLP 15 -> 128 -> 128 -> 128 -> 15,
exact (erf) GELU, mean squared error on standardised targets.

With bf16=True every GEMM operand is rounded to bf16 (round to nearest even)
as the matrix-core kernels do; sums are float64 here (FP32 on the GPU)."""
import numpy as np
from scipy.special import erf

DIMS = (15, 128, 128, 128, 15)
NPARAM = 37007


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    """[(W, b)] * 4 (W[in][out]) and the standardisation (in_mean, in_std, out_mean, out_std)."""
    p = np.asarray(params, dtype=np.float64)
    layers, off = [], 0
    for i in range(4):
        a, b = DIMS[i], DIMS[i + 1]
        W = p[off:off + a * b].reshape(a, b)
        off += a * b
        bias = p[off:off + b]
        off += b
        layers.append((W, bias))
    assert off == NPARAM
    st = p[NPARAM:NPARAM + 60].reshape(4, 15) if p.size >= NPARAM + 60 else None
    return layers, st


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def standardise(params, us, alpha=None):
    _, st = unpack(params)
    x = (np.asarray(us) - st[0]) / st[1]
    t = None if alpha is None else (np.asarray(alpha) - st[2]) / st[3]
    return x, t


def forward(params, x, bf=False):
    layers, _ = unpack(params)
    r = bf16 if bf else (lambda v: v)
    zs, acts = [], [x]
    h = x
    for i, (W, b) in enumerate(layers):
        z = r(h) @ r(W) + b
        zs.append(z)
        h = gelu(z) if i < 3 else z
        acts.append(h)
    return zs, acts


def loss_and_grad(params, x, t, bf=False):
    """MSE = mean((y - t)^2) over rows and the 15 outputs, and its gradient as a
    flat vector in the parameter layout."""
    layers, _ = unpack(params)
    r = bf16 if bf else (lambda v: v)
    zs, acts = forward(params, x, bf)
    y = acts[-1]
    n = y.size
    loss = float(((y - t) ** 2).sum() / n)
    d = 2.0 * (y - t) / n
    grads = [None] * 4
    for i in range(3, -1, -1):
        W, _ = layers[i]
        grads[i] = (r(acts[i]).T @ r(d), d.sum(axis=0))
        if i > 0:
            d = (r(d) @ r(W).T) * gelu_grad(zs[i - 1])
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat
