"""CPU restatement of the learned G1 extrapolator (test infrastructure only).

The network and features of g12kba_amd/csrc/extrap.hip, in numpy: the
features of one upper-triangle element of G1, the forward pass
A1 = tanh(X W1 + b1), A2 = tanh(A1 W2 + b2), A3 = tanh(A2 W3 + b3),
Y = A3 W4 + b4 + X Ws, the training loss (squared residual up to 1, linear
beyond) and its gradient.  With
bf=True every matrix-product operand is rounded to BF16 (round to nearest even)
as the GPU stages it, and products accumulate in float64 (the GPU: FP32);
bf=False is the exact float64 network the finite-difference checks pin.
The extrapolator has no counterpart in the reference.
"""
import numpy as np


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def dims(hist=64, horizon=16, hidden=320):
    return dict(hist=hist, horizon=horizon, din=2 * hist + 2 * (hist + horizon), dout=2 * horizon, H=hidden)


def layout(d):
    """(name, shape) of the parameter blocks in storage order."""
    din, H, dout = d["din"], d["H"], d["dout"]
    return [("W1", (din, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
            ("W4", (H, dout)), ("b4", (dout,)), ("Ws", (din, dout))]


def unpack(flat, d):
    out, o = {}, 0
    for name, shape in layout(d):
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(blocks, d):
    return np.concatenate([np.asarray(blocks[name], dtype=np.float64).reshape(-1) for name, _ in layout(d)])


def _mm(a, b, bf):
    if bf:
        a, b = bf16(a).astype(np.float64), bf16(b).astype(np.float64)
    return a @ b


def forward(flat, X, d, bf=True):
    P = unpack(flat, d)
    X = np.asarray(X, dtype=np.float64)
    A1 = np.tanh(_mm(X, P["W1"], bf) + P["b1"])
    A2 = np.tanh(_mm(A1, P["W2"], bf) + P["b2"])
    A3 = np.tanh(_mm(A2, P["W3"], bf) + P["b3"])
    Y = _mm(A3, P["W4"], bf) + P["b4"] + _mm(X, P["Ws"], bf)
    return Y, (X, A1, A2, A3)


def loss_grad(flat, X, T, d, bf=True):
    """Loss averaged over all outputs and its gradient (storage layout): per
    output r^2 for |r| <= 1 and 2|r| - 1 beyond, r = y - target."""
    P = unpack(flat, d)
    Y, (X, A1, A2, A3) = forward(flat, X, d, bf)
    T = np.asarray(T, dtype=np.float64)
    diff = Y - T
    a = np.abs(diff)
    loss = float(np.mean(np.where(a <= 1, diff ** 2, 2 * a - 1)))
    dY = 2.0 * np.clip(diff, -1, 1) / diff.size
    g = {}
    g["W4"] = _mm(A3.T, dY, bf)
    g["Ws"] = _mm(X.T, dY, bf)
    g["b4"] = dY.sum(0)
    D3 = _mm(dY, P["W4"].T, bf) * (1 - A3 ** 2)
    g["W3"] = _mm(A2.T, D3, bf)
    g["b3"] = D3.sum(0)
    D2 = _mm(D3, P["W3"].T, bf) * (1 - A2 ** 2)
    g["W2"] = _mm(A1.T, D2, bf)
    g["b2"] = D2.sum(0)
    D1 = _mm(D2, P["W2"].T, bf) * (1 - A1 ** 2)
    g["W1"] = _mm(X.T, D1, bf)
    g["b1"] = D1.sum(0)
    return loss, pack(g, d)


def upper_pairs(N):
    return [(p, q) for p in range(N) for q in range(p, N)]


def features(hist, w, Ns, Nq, d, dt):
    """Features of every upper-triangle element from `hist` (steps x N x N
    complex, oldest first) and the quench shift w (hist + horizon values).
    Returns X (n_el x din), ref (n_el complex), scale (n_el)."""
    H, F = d["hist"], d["horizon"]
    hist = np.asarray(hist)[-H:]
    N = hist.shape[1]
    w = np.asarray(w, dtype=np.float64)[:H + F]
    pairs = upper_pairs(N)
    X = np.zeros((len(pairs), d["din"]))
    ref = np.zeros(len(pairs), np.complex128)
    scale = np.zeros(len(pairs))
    for e, (p, q) in enumerate(pairs):
        h = hist[:, p, q]
        m = (1.0 if p % Ns < Nq else 0.0) - (1.0 if q % Ns < Nq else 0.0)
        dd = w * m
        r = h[-1]
        s = max(np.abs(h - r).max(), 4.0 * dt * np.abs(dd).max() * abs(r), 1e-6 * max(1.0, abs(r)))
        z = (h - r) / s
        X[e, 0:2 * H:2], X[e, 1:2 * H:2] = z.real, z.imag
        f = -1j * dd * dt * r / s
        X[e, 2 * H::2], X[e, 2 * H + 1::2] = f.real, f.imag
        ref[e], scale[e] = r, s
    return X, ref, scale


def predict(flat, hist, w, Ns, Nq, d, dt, bf=True):
    """One block: the next `horizon` matrices (lower triangle by G^dagger = -G)."""
    X, ref, scale = features(hist, w, Ns, Nq, d, dt)
    Y, _ = forward(flat, X.astype(np.float32), d, bf)
    N = np.asarray(hist).shape[1]
    F = d["horizon"]
    out = np.zeros((F, N, N), np.complex128)
    for e, (p, q) in enumerate(upper_pairs(N)):
        g = ref[e] + scale[e] * (Y[e, 0::2] + 1j * Y[e, 1::2])
        out[:, p, q] = g
        if p != q:
            out[:, q, p] = -np.conj(g)
    return out
