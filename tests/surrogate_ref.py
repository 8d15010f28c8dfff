"""CPU reference of the tail-parameter MLP (128-128-128-16, three outputs
used): forward, loss and backward, with an option to round operands to
bfloat16 exactly where the device kernels do (hicops_amd/csrc/k_mlp.hip)."""
import numpy as np

F = H = 128
O = 16
OUT = 3
oW1 = 0
ob1 = oW1 + H * F
oW2 = ob1 + H
ob2 = oW2 + H * H
oW3 = ob2 + H
ob3 = oW3 + O * H
NPARAM = ob3 + O


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to the nearest bfloat16 (ties to even), back as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    return (p[oW1:ob1].reshape(H, F), p[ob1:oW2], p[oW2:ob2].reshape(H, H), p[ob2:oW3],
            p[oW3:ob3].reshape(O, H), p[ob3:ob3 + O])


def forward_backward(p, X, T, emulate_bf16=True, dtype=np.float32):
    """Loss sum_r w_r sum_{j<3} (y - t)^2 / sum w and its gradient (same layout as p)."""
    r = bf16 if emulate_bf16 else (lambda a: a)
    p = p.astype(dtype)
    X = X.astype(dtype)
    T = T.astype(dtype)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    w = T[:, 3]
    inv = 1.0 / w.sum()
    Xb, W1b, W2b, W3b = r(X), r(W1), r(W2), r(W3)
    h1 = np.maximum(Xb @ W1b.T + b1, 0)
    h1b = r(h1)
    h2 = np.maximum(h1b @ W2b.T + b2, 0)
    h2b = r(h2)
    y = h2b @ W3b.T + b3
    d = np.zeros_like(y)
    d[:, :OUT] = y[:, :OUT] - T[:, :OUT]
    loss = float((w[:, None] * d * d).sum() * inv)
    dy = 2.0 * w[:, None] * d * inv
    g = np.zeros(NPARAM, dtype)
    gW1, gb1, gW2, gb2, gW3, gb3 = unpack(g)
    gW3[:OUT] = (r(dy).T @ h2b)[:OUT]
    gb3[:OUT] = dy.sum(0)[:OUT]
    dh2 = (r(dy) @ W3b) * (h2b > 0)
    gW2[:] = r(dh2).T @ h1b
    gb2[:] = dh2.sum(0)
    dh1 = (r(dh2) @ W2b) * (h1b > 0)
    gW1[:] = r(dh1).T @ Xb
    gb1[:] = dh1.sum(0)
    return loss, g, y
