"""CPU checks of the extrapolator oracle (oracle/mlp.py): the analytic
gradient against central finite differences, BF16 rounding known answers and
the feature normalisation on inputs with known features."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40], np.float32)
    y = mlp.bf16(x)
    # ulp(1) = 2^-7 in BF16: 1 + 2^-8 is a tie (to even: 1), 1 + 3 2^-8 rounds up to 1 + 2^-6
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == np.float32(1 + 2 ** -6) and y[3] == np.float32(1 + 2 ** -7)
    assert y[4] == -2.5
    assert np.isfinite(y[5]) and abs(y[5] / 3.0e38 - 1) < 2 ** -8
    assert (mlp.bf16(y) == y).all()  # idempotent


def test_gradient_matches_finite_differences():
    d = mlp.dims(hist=4, horizon=2, hidden=8)
    rng = np.random.default_rng(0)
    n = sum(int(np.prod(s)) for _, s in mlp.layout(d))
    flat = rng.standard_normal(n) * 0.3
    X = rng.standard_normal((7, d["din"]))
    T = rng.standard_normal((7, d["dout"])) * 2  # residuals on both sides of 1
    loss, g = mlp.loss_grad(flat, X, T, d, bf=False)
    eps = 1e-6
    for i in rng.choice(n, size=60, replace=False):
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        num = (mlp.loss_grad(fp, X, T, d, bf=False)[0] - mlp.loss_grad(fm, X, T, d, bf=False)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-7 + 1e-6 * abs(num), (i, num, g[i])
    # every parameter block receives gradient
    o = 0
    for name, shape in mlp.layout(d):
        k = int(np.prod(shape))
        assert np.abs(g[o:o + k]).max() > 0, name
        o += k


def test_features_known_inputs():
    d = mlp.dims(hist=8, horizon=4, hidden=8)
    N, dt = 3, 0.01
    t = np.arange(8) * dt
    hist = np.zeros((8, N, N), np.complex128)
    hist[:, 0, 1] = np.exp(-2j * t)            # varying element
    hist[:, 1, 2] = 0.5j                       # constant element
    w = np.zeros(12)
    X, ref, scale = mlp.features(hist, w, 3, 1, d, dt)
    pairs = mlp.upper_pairs(N)
    e01, e12 = pairs.index((0, 1)), pairs.index((1, 2))
    assert ref[e01] == hist[-1, 0, 1] and scale[e01] == np.abs(hist[:, 0, 1] - hist[-1, 0, 1]).max()
    z = X[e01, 0:16:2] + 1j * X[e01, 1:16:2]
    assert abs(np.abs(z).max() - 1.0) < 1e-12 and z[-1] == 0 and (X[e01, 16:] == 0).all()
    assert scale[e12] == 1e-6 and (X[e12] == 0).all()
    # drive: sites < Nq = 1 carry w; element (0, 1) sees d = w, (1, 2) sees 0
    w = np.full(12, 2.0)
    X, ref, scale = mlp.features(hist, w, 3, 1, d, dt)
    r = hist[-1, 0, 1]
    assert scale[e01] >= 4 * dt * 2.0 * abs(r)
    f = -1j * 2.0 * dt * r / scale[e01]
    assert np.allclose(X[e01, 16::2], f.real) and np.allclose(X[e01, 17::2], f.imag)
    assert (X[e12, 16:] == 0).all()
