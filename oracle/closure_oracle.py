"""TEST INFRASTRUCTURE ONLY: numpy restatement of the closure network's
forward and backward passes (the checker of bastr_closure_apply and
bastr_closure_loss_grad).  Imported only by tests/ and bench.py's checks.

The network (36 -> 64 -> 64 -> 2, GELU in its tanh form, Hendrycks and
Gimpel 2016) and its features (the 3x3 periodic stencil of u1, u2, theta,
omega, feature f = 4*pos + field, pos = 3*(dy+1) + (dx+1)) are those of
include/bastr_closure.h.  With emulate=True the arithmetic rounds to BF16
exactly where the device kernels do (weights W1, W2, W3; the scaled features;
the layer-1 activations; the backward operands dZ2 and dZ1, the bias gradients
b1, b2 summing those rounded values) and computes everything else in FP64;
with emulate=False it is plain FP64 (the form the finite-difference checks
pin).
"""
import numpy as np

NF, NH, NO = 36, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NO * NH
NPARAM = OFF_B3 + NO


def bf16(x):
    """float32 values rounded to BF16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def gelu(z):
    t = np.tanh(np.sqrt(2.0 / np.pi) * (z + 0.044715 * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_d(z):
    t = np.tanh(np.sqrt(2.0 / np.pi) * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * np.sqrt(2.0 / np.pi) * (1.0 + 3 * 0.044715 * z * z)


def unpack(p):
    p = np.asarray(p)
    W1 = p[OFF_W1:OFF_B1].reshape(NH, NF)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(NH, NH)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(NO, NH)
    b3 = p[OFF_B3:]
    return W1, b1, W2, b2, W3, b3


def features(fields, scales, emulate=True):
    """fields [4, n, n] (index [field, x, y]) -> X [n*n, 36], points in Fortran
    order (x fastest)."""
    f = np.asarray(fields, dtype=np.float32 if emulate else np.float64)
    n = f.shape[1]
    cols = []
    for pos in range(9):
        dx, dy = pos % 3 - 1, pos // 3 - 1
        sh = np.roll(np.roll(f, -dx, axis=1), -dy, axis=2)  # sh[c, x, y] = f[c, x+dx, y+dy]
        for c in range(4):
            v = sh[c].reshape(n * n, order="F")
            if emulate:
                inv = np.float32(1.0) / np.float32(scales[c])
                v = bf16(v * inv)
            else:
                v = v / scales[c]
            cols.append(v.astype(np.float64))
    return np.stack(cols, axis=1)


def targets(target, scales, emulate=True):
    t = np.asarray(target, dtype=np.float32 if emulate else np.float64)
    n = t.shape[1]
    if emulate:
        return np.stack([(t[c].reshape(n * n, order="F") / np.float32(scales[4 + c])).astype(np.float64)
                         for c in range(2)], axis=1)
    return np.stack([t[c].reshape(n * n, order="F") / scales[4 + c] for c in range(2)], axis=1)


def forward(params, X, emulate=True):
    W1, b1, W2, b2, W3, b3 = [np.asarray(a, dtype=np.float64) for a in unpack(params)]
    if emulate:
        W1, W2, W3 = (bf16(W).astype(np.float64) for W in (W1, W2, W3))
    z1 = X @ W1.T + b1
    H1 = gelu(z1)
    if emulate:
        H1 = bf16(H1).astype(np.float64)
    z2 = H1 @ W2.T + b2
    a2 = gelu(z2)
    out = a2 @ W3.T + b3
    return out, dict(z1=z1, H1=H1, z2=z2, a2=a2, W1=W1, W2=W2, W3=W3)


def loss_grad(params, X, T, emulate=True):
    """L = sum (out - T)^2 / (2 P) and its gradient (NPARAM, FP64)."""
    out, c = forward(params, X, emulate)
    P = X.shape[0]
    e = out - T
    loss = float(np.sum(e * e) / (2 * P))
    g = e / P
    dW3 = g.T @ c["a2"]
    db3 = g.sum(axis=0)
    dz2 = (g @ c["W3"]) * gelu_d(c["z2"])
    if emulate:
        dz2 = bf16(dz2).astype(np.float64)
    dW2 = dz2.T @ c["H1"]
    db2 = dz2.sum(axis=0)
    dz1 = (dz2 @ c["W2"]) * gelu_d(c["z1"])
    if emulate:
        dz1 = bf16(dz1).astype(np.float64)
    dW1 = dz1.T @ X
    db1 = dz1.sum(axis=0)
    grad = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return loss, grad
