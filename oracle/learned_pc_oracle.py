"""CPU oracle (numpy, float64) of the learned block-Jacobi correction: the
features, the three 3x3 convolutions, the training loss and its analytic
weight gradient.  Test infrastructure only: the product runs the HIP
kernels of csrc/learned_pc.hip; this module is the checker.

The network has no counterpart in the reference; its definition is the
one stated in include/llg/learned_pc.h:
    f = s*gamma*[u, m x u, (m.u) m, 0 x 7]      (16 channels, zero off-mask)
    h1 = C1 f, h2 = C2 h1, o = C3 h2            (3x3 taps, W[tap][cin][cout])
    loss = sum_active |s u + o[:3] - s x|^2 / ncell
BF16 rounding (round to nearest even) is applied where the device stores
BF16: features, weights, h1, h2 and the output gradient, so the oracle sees
the device's inputs to every MFMA.
"""
import numpy as np

C_IN, C1, C2, C3 = 16, 32, 32, 16


def bf16(a):
    """round float32 values to bfloat16 (nearest even), returned as float64"""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split(w):
    """flat FP32 weight vector -> [W1 (9,16,32), W2 (9,32,32), W3 (9,32,16)]"""
    sizes = [(9, C_IN, C1), (9, C1, C2), (9, C2, C3)]
    out, o = [], 0
    for s in sizes:
        n = int(np.prod(s))
        out.append(np.asarray(w[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def shift(a, dy, dx, periodic):
    """b[y, x] = a[y + dy, x + dx] (zero outside unless periodic); a: (ny, ng, C)"""
    if periodic:
        return np.roll(np.roll(a, -dy, axis=0), -dx, axis=1)
    b = np.zeros_like(a)
    ny, ng = a.shape[:2]
    ys = slice(max(0, -dy), min(ny, ny - dy))
    xs = slice(max(0, -dx), min(ng, ng - dx))
    ys2 = slice(max(0, dy), min(ny, ny + dy))
    xs2 = slice(max(0, dx), min(ng, ng + dx))
    b[ys, xs] = a[ys2, xs2]
    return b


def conv(a, W, periodic):
    """out[y, x, co] = sum_t sum_ci a[(y, x) + d_t, ci] W[t, ci, co]"""
    out = np.zeros(a.shape[:2] + (W.shape[2],))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out += shift(a, dy, dx, periodic) @ W[t]
    return out


def conv_adjoint(g, W, periodic):
    """input gradient of conv: taps flipped, W transposed"""
    out = np.zeros(g.shape[:2] + (W.shape[1],))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out += shift(g, -dy, -dx, periodic) @ W[t].T
    return out


def wgrad(a, g, periodic):
    """dW[t] = sum_cells a[cell + d_t]^T g[cell]"""
    dW = np.zeros((9, a.shape[2], g.shape[2]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        sa = shift(a, dy, dx, periodic)
        dW[t] = sa.reshape(-1, a.shape[2]).T @ g.reshape(-1, g.shape[2])
    return dW


def features(u, y, active, sg, ng, ny):
    n = ng * ny
    u = u.reshape(3, n)
    m = y.reshape(3, n)
    mu = (m * u).sum(0)
    v = np.stack([u[0], u[1], u[2],
                  m[1] * u[2] - m[2] * u[1], m[2] * u[0] - m[0] * u[2], m[0] * u[1] - m[1] * u[0],
                  mu * m[0], mu * m[1], mu * m[2]], axis=1) * sg
    v = v * (active.reshape(n, 1) != 0)
    f = np.zeros((n, C_IN))
    f[:, :9] = v
    return f.reshape(ny, ng, C_IN)


def forward(w, u, y, active, gamma, scale, ng, ny, periodic, rounding=True):
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, W3 = [r(x) for x in split(w)]
    f = r(features(u, y, active, scale * gamma, ng, ny))
    h1 = r(conv(f, W1, periodic))
    h2 = r(conv(h1, W2, periodic))
    o = conv(h2, W3, periodic)
    return dict(f=f, h1=h1, h2=h2, o=o, W=(W1, W2, W3))


def loss_grad(w, u, y, x, active, gamma, scale, ng, ny, periodic, rounding=True):
    """loss and flat weight gradient, as the device computes them"""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    st = forward(w, u, y, active, gamma, scale, ng, ny, periodic, rounding)
    n = ng * ny
    on = (active.reshape(ny, ng, 1) != 0)
    pred = scale * np.moveaxis(u.reshape(3, ny, ng), 0, -1) + st["o"][..., :3]
    d = (pred - scale * np.moveaxis(x.reshape(3, ny, ng), 0, -1)) * on
    loss = float((d ** 2).sum() / n)
    g3 = np.zeros((ny, ng, C3))
    g3[..., :3] = r(2.0 * d / n)
    W1, W2, W3 = st["W"]
    dW3 = wgrad(st["h2"], g3, periodic)
    g2 = r(conv_adjoint(g3, W3, periodic))
    dW2 = wgrad(st["h1"], g2, periodic)
    g1 = r(conv_adjoint(g2, W2, periodic))
    dW1 = wgrad(st["f"], g1, periodic)
    return loss, np.concatenate([dW1.ravel(), dW2.ravel(), dW3.ravel()]), st
