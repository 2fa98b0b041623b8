"""Numpy restatement of the superblock cost model (24-64-64-1 perceptron, ReLU)
for checking the HIP kernels: forward pass, and the gradient of the mean
squared error over standardised targets. Parameter layout: W1[64][24], b1[64],
W2[64][64], b2[64], W3[64], b3 (5825 floats). Test infrastructure only."""
import numpy as np

IN, H = 24, 64
PARAMS = H * IN + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, Xs):
    """Standardised features -> standardised prediction, with activations."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = Xs @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    out = h2 @ W3 + b3
    return out, (h1, h2)


def loss_and_grad(p, Xs, ys):
    """Mean squared error over n samples and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = Xs.shape[0]
    out, (h1, h2) = forward(p, Xs)
    e = out - ys
    loss = float(np.mean(e * e))
    d3 = 2.0 * e / n
    gW3 = h2.T @ d3
    gb3 = d3.sum()
    d2 = np.outer(d3, W3) * (h2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (h1 > 0)
    gW1 = d1.T @ Xs
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g


def standardise(X, y, mean, dev, y_mean, y_dev):
    return (np.asarray(X, np.float64) - mean) / dev, (np.asarray(y, np.float64) - y_mean) / y_dev


def predict_us(p, X, mean, dev, y_mean, y_dev):
    Xs = (np.asarray(X, np.float64) - mean) / dev
    out, _ = forward(p, Xs)
    return out * y_dev + y_mean


def init_params(rng):
    p = np.zeros(PARAMS)
    W1, b1, W2, b2, W3, b3 = (None,) * 6
    p[:H * IN] = rng.normal(0, np.sqrt(2 / IN), H * IN)
    o = H * IN + H
    p[o:o + H * H] = rng.normal(0, np.sqrt(2 / H), H * H)
    o += H * H + H
    p[o:o + H] = rng.normal(0, np.sqrt(2 / H), H)
    return p
