"""CPU f64 restatement of the warm-start network (tanh MLP, mean-squared-error
loss, Adam). TEST INFRASTRUCTURE ONLY: the checker the HIP kernels of
split_operator_amd/csrc/so_mlp.hip are compared against; the product never
imports it.

The network has no counterpart in the split_operator crate. With `bf=True`
every GEMM operand is rounded to BF16 (round to nearest even) exactly where
the GPU rounds it — activations, weights and back-propagated gradients on
their way into a product — and stored activations and gradients are rounded
to FP32 where the GPU stores FP32. Sums run in f64 (the GPU sums in FP32, so
the two agree to FP32 round-off). With `bf=False` it is the plain f64
network, whose gradients the tests pin by central differences.
"""
import numpy as np


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as f64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(a, on):
    return np.asarray(a, dtype=np.float32).astype(np.float64) if on else np.asarray(a, dtype=np.float64)


def forward(params, x, bf=True):
    """Activations [x, Z1, ..., ZL]; ZL is the raw output (no base)."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    zs = [np.asarray(x, dtype=np.float64)]
    L = len(params)
    for l, (w, b) in enumerate(params):
        y = q(zs[-1]) @ q(w).T
        y = _f32(y, bf) + np.asarray(b, dtype=np.float64)
        y = _f32(y, bf)
        if l < L - 1:
            y = _f32(np.tanh(y), bf)
        zs.append(y)
    return zs


def loss_and_grads(params, x, target, base, bf=True, scale=1.0):
    """Loss (1/batch) sum ((base + scale net(x) - target) / scale)^2 (sum over outputs, mean over samples)
    and its gradients [(gW, gb), ...] per layer."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    zs = forward(params, x, bf)
    L = len(params)
    d = _f32((np.asarray(base, dtype=np.float64) + scale * zs[-1] - np.asarray(target, dtype=np.float64)) / scale, bf)
    rows = d.shape[0]
    loss = float(np.sum(d * d) / rows)
    g = _f32(d * (2.0 / rows), bf)
    grads = [None] * L
    for l in range(L - 1, -1, -1):
        if l < L - 1:
            g = _f32(g * (1.0 - zs[l + 1] * zs[l + 1]), bf)
        gw = _f32(q(g).T @ q(zs[l]), bf)
        gb = _f32(g.sum(axis=0), bf)
        grads[l] = (gw, gb)
        if l > 0:
            g = _f32(q(g) @ q(params[l][0]), bf)
    return loss, grads


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) with bias correction; returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    p = p - lr * (m / (1.0 - b1 ** t)) / (np.sqrt(v / (1.0 - b2 ** t)) + eps)
    return p, m, v
