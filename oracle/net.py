"""TEST INFRASTRUCTURE ONLY: numpy restatement of the path network.

The network of the learned-path form (include/pct_net.h): 12 inputs -> 64 tanh
-> 64 tanh -> 8 linear outputs, parameters in one FP32 vector (W1[12][64],
b1[64], W2[64][64], b2[64], W3[64][8], b3[8]). The device kernels multiply in
BF16 (round to nearest even) with FP32 accumulation; `bf16=True` rounds the
operands of every product the same way here (accumulation in FP64), so the
device results agree to FP32 accumulation error. `bf16=False` is the plain
FP64 network, on which the backward pass is pinned against finite
differences.

Loss: sum over samples and outputs of (y - target)^2 / (8 n).
"""
import numpy as np

IN, HID, OUT = 12, 64, 8
OFF = dict(W1=0, b1=768, W2=832, b2=4928, W3=4992, b3=5504)
P = 5512


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return (p[0:768].reshape(IN, HID), p[768:832], p[832:4928].reshape(HID, HID), p[4928:4992],
            p[4992:5504].reshape(HID, OUT), p[5504:5512])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)


def bf16(a):
    """round float32 values to bfloat16 (nearest, ties to even), returned as float64"""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    u = ((u + r) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def _q(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def init_params(seed=0):
    """Xavier-uniform weights, zero biases"""
    rng = np.random.default_rng(seed)

    def xavier(fi, fo):
        lim = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-lim, lim, size=(fi, fo))
    return pack(xavier(IN, HID), np.zeros(HID), xavier(HID, HID), np.zeros(HID), xavier(HID, OUT), np.zeros(OUT))


def forward(params, X, bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(params)
    X = np.asarray(X, dtype=np.float32)
    z1 = _q(X, bf16) @ _q(W1, bf16) + b1
    a1 = np.tanh(z1).astype(np.float32) if bf16 else np.tanh(z1)
    z2 = _q(a1, bf16) @ _q(W2, bf16) + b2
    a2 = np.tanh(z2).astype(np.float32) if bf16 else np.tanh(z2)
    z3 = _q(a2, bf16) @ _q(W3, bf16) + b3
    return dict(X=X, a1=np.asarray(a1, np.float64), a2=np.asarray(a2, np.float64), y=z3)


def loss(params, X, Y, bf16=True):
    y = forward(params, X, bf16)["y"]
    return float(np.sum((y - np.asarray(Y, np.float64)) ** 2) / (OUT * len(X)))


def gradient(params, X, Y, bf16=True):
    """(grad[5512], loss) of the mean-squared error"""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    f = forward(params, X, bf16)
    n = len(X)
    r = f["y"] - np.asarray(Y, np.float64)
    dz3 = 2.0 * r / (OUT * n)
    if bf16:
        dz3 = dz3.astype(np.float32).astype(np.float64)
    a1, a2 = f["a1"], f["a2"]
    gW3 = _q(a2, bf16).T @ _q(dz3, bf16)
    gb3 = dz3.sum(0)
    da2 = _q(dz3, bf16) @ _q(W3, bf16).T
    dz2 = da2 * (1.0 - a2 * a2)
    if bf16:
        dz2 = dz2.astype(np.float32).astype(np.float64)
    gW2 = _q(a1, bf16).T @ _q(dz2, bf16)
    gb2 = dz2.sum(0)
    da1 = _q(dz2, bf16) @ _q(W2, bf16).T
    dz1 = da1 * (1.0 - a1 * a1)
    if bf16:
        dz1 = dz1.astype(np.float32).astype(np.float64)
    gW1 = _q(f["X"], bf16).T @ _q(dz1, bf16)
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return g, float(np.sum(r * r) / (OUT * n))


def adam_step(params, m, v, grad, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """one Adam update (t = 1, 2, ...), FP32 state like the device kernel"""
    p = np.asarray(params, np.float32).copy()
    m = (np.float32(beta1) * np.asarray(m, np.float32) + np.float32(1 - beta1) * np.asarray(grad, np.float32))
    v = (np.float32(beta2) * np.asarray(v, np.float32)
         + np.float32(1 - beta2) * np.asarray(grad, np.float32) * np.asarray(grad, np.float32))
    c1 = np.float32(1.0 / (1.0 - beta1 ** t))
    c2 = np.float32(1.0 / (1.0 - beta2 ** t))
    p = p - np.float32(lr) * (m * c1) / (np.sqrt(v * c2) + np.float32(eps))
    return p.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


def spline_path(out8, t_in, t_out, u2, dT, sigma):
    """t of the surrogate path (Hermite + clamped spline of the offsets) at
    normalised depths sigma, FP64, for a single history (reference for tests)"""
    from scipy.interpolate import CubicSpline
    kn = np.linspace(0.0, 1.0, 6)
    d = np.concatenate([[0.0], np.asarray(out8[:4], np.float64) * 0.01, [0.0]])
    s = np.asarray(sigma, np.float64)
    h01 = 3 * s * s - 2 * s ** 3
    h11 = s ** 3 - s * s
    return t_in + (t_out - t_in) * h01 + u2 * dT * h11 + CubicSpline(kn, d, bc_type="clamped")(s)
