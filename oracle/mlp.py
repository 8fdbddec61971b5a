"""numpy restatement of the per-contact warm-start network (ostrich_amd/csrc/fb_mlp.hip).

TEST INFRASTRUCTURE ONLY (the checker of the GPU network): 12 -> 64 -> 64 -> 3, ReLU,
BF16 weights and activations (round to nearest even at the same points as the GPU),
sums in float64, loss mean_rows sum_c (y_c - t_c)^2.  The learned warm start has no
counterpart in the reference, so the restatement is pinned by its own tests: finite
differences of the loss in float64 without rounding, and the GPU forward/backward.
"""
import numpy as np

IN, INP, H, OUT, OUTP = 12, 16, 64, 3, 16
OW1 = 0
OB1 = OW1 + INP * H
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H * OUTP
NPARAM = OB3 + OUTP


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, np.float32)
    return (p[OW1:OB1].reshape(INP, H), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3].reshape(H, OUTP), p[OB3:NPARAM])


def forward(p, x, rounding=True):
    """Returns (y [N,3], cache) in float64."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = x.shape[0]
    X = np.zeros((n, INP)); X[:, :IN] = x
    X = r(X).astype(np.float64)
    W1r, W2r, W3r = (r(W).astype(np.float64) for W in (W1, W2, W3))
    H1 = r(np.maximum((X @ W1r).astype(np.float32) + b1, 0)).astype(np.float64)
    H2 = r(np.maximum((H1 @ W2r).astype(np.float32) + b2, 0)).astype(np.float64)
    Y = (H2 @ W3r).astype(np.float32) + b3
    return Y[:, :OUT].astype(np.float64), (X, H1, H2, W1r, W2r, W3r)


def loss_and_grad(p, x, t, rounding=True):
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    y, (X, H1, H2, W1r, W2r, W3r) = forward(p, x, rounding)
    n = x.shape[0]
    diff = y - t
    loss = float((diff ** 2).sum() / n)
    dY = np.zeros((n, OUTP)); dY[:, :OUT] = 2 * diff / n
    dY = r(dY).astype(np.float64)
    g = np.zeros(NPARAM)
    g[OW3:OB3] = (H2.T @ dY).reshape(-1)
    g[OB3:NPARAM] = dY.sum(0)
    dZ2 = r(np.where(H2 > 0, dY @ W3r.T, 0)).astype(np.float64)
    g[OW2:OB2] = (H1.T @ dZ2).reshape(-1)
    g[OB2:OW3] = dZ2.sum(0)
    dZ1 = r(np.where(H1 > 0, dZ2 @ W2r.T, 0)).astype(np.float64)
    gW1 = X.T @ dZ1
    gW1[IN:] = 0
    g[OW1:OB1] = gW1.reshape(-1)
    g[OB1:OW2] = dZ1.sum(0)
    g[OW3:OB3].reshape(H, OUTP)[:, OUT:] = 0
    g[OB3 + OUT:NPARAM] = 0
    return loss, g


def init_params(seed=0):
    """He-normal weights, zero biases, zero padding rows/columns."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    W1 = np.zeros((INP, H), np.float32); W1[:IN] = rng.normal(0, np.sqrt(2 / IN), (IN, H))
    W3 = np.zeros((H, OUTP), np.float32); W3[:, :OUT] = rng.normal(0, np.sqrt(1 / H), (H, OUT))
    p[OW1:OB1] = W1.reshape(-1)
    p[OW2:OB2] = rng.normal(0, np.sqrt(2 / H), (H, H)).astype(np.float32).reshape(-1)
    p[OW3:OB3] = W3.reshape(-1)
    return p
