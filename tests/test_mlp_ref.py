"""The numpy cost-model reference against central finite differences and a
closed form, so that it can serve as the check of the HIP kernels."""
import numpy as np

from mlp_ref import IN, PARAMS, forward, init_params, loss_and_grad, unpack


def test_parameter_count():
    assert PARAMS == 5825


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = init_params(rng)
    p[IN * 64:IN * 64 + 64] = rng.normal(0, 0.1, 64)  # nonzero b1 so ReLU kinks are off the grid
    X = rng.normal(0, 1, (37, IN))
    y = rng.normal(0, 1, 37)
    _, g = loss_and_grad(p, X, y)
    idx = rng.choice(PARAMS, 200, replace=False)
    idx = np.concatenate([idx, [PARAMS - 1, 0, IN * 64]])
    h = 1e-6
    for i in idx:
        d = np.zeros(PARAMS)
        d[i] = h
        num = (loss_and_grad(p + d, X, y)[0] - loss_and_grad(p - d, X, y)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_linear_region_closed_form():
    """With every unit active the network is affine: out = W3 (W2 (W1 x + b1)
    + b2) + b3."""
    rng = np.random.default_rng(4)
    p = np.zeros(PARAMS)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    o = 0
    p[o:o + 64 * IN] = rng.uniform(0.01, 0.1, 64 * IN); o += 64 * IN
    p[o:o + 64] = 1.0; o += 64
    p[o:o + 64 * 64] = rng.uniform(0.01, 0.1, 64 * 64); o += 64 * 64
    p[o:o + 64] = 1.0; o += 64
    p[o:o + 64] = rng.normal(0, 1, 64); o += 64
    p[o] = 0.5
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = rng.uniform(0, 1, (5, IN))
    out, _ = forward(p, X)
    want = ((X @ W1.T + b1) @ W2.T + b2) @ W3 + b3
    np.testing.assert_allclose(out, want, rtol=1e-12)
