"""CPU oracle, test infrastructure only: the superquadric warm-start network
(18 -> 64 -> 64 -> 3, tanh hidden layers) in numpy. The product path trains
and runs it on the GPU's BF16 matrix cores; this file is the checker.

forward() / gradient() reproduce the GPU kernels' rounding points exactly:
inputs, weights, hidden activations and the output / hidden-layer deltas are
rounded to BF16 (round to nearest even) before each matrix product, products
accumulate in FP32, biases, tanh and its derivative stay FP32. With
exact=True the same formulas run in FP64 without rounding (used to pin the
backward pass against finite differences).

Parameter vector (5571 floats): W1 [18][64], b1 [64], W2 [64][64], b2 [64],
W3 [64][3], b3 [3].
"""
import numpy as np

N_IN, N_HID, N_OUT = 18, 64, 3
SIZES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
         ("W3", (N_HID, N_OUT)), ("b3", (N_OUT,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(params):
    p = np.asarray(params)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SIZES])


def _r(x, exact):
    return x if exact else bf16(x)


def forward(params, X, exact=False):
    dt = np.float64 if exact else np.float32
    P = {k: v.astype(dt) for k, v in unpack(params).items()}
    Xb = _r(np.asarray(X, dtype=dt), exact)
    W1, W2, W3 = (_r(P[k], exact) for k in ("W1", "W2", "W3"))
    H1 = np.tanh(Xb @ W1 + P["b1"]).astype(dt)
    H1b = _r(H1, exact)
    H2 = np.tanh(H1b @ W2 + P["b2"]).astype(dt)
    H2b = _r(H2, exact)
    Y = (H2b @ W3 + P["b3"]).astype(dt)
    return Y, (Xb, W1, W2, W3, H1, H1b, H2, H2b)


def gradient(params, X, T, exact=False):
    """(flat gradient of mean_k |y_k - t_k|^2, that loss)."""
    dt = np.float64 if exact else np.float32
    Y, (Xb, W1, W2, W3, H1, H1b, H2, H2b) = forward(params, X, exact)
    B = Y.shape[0]
    E = Y - np.asarray(T, dtype=dt)
    loss = float((E.astype(np.float64) ** 2).sum() / B)
    dY = (dt(2.0) / dt(B)) * E
    dYb = _r(dY, exact)
    g = {"W3": H2b.T @ dYb, "b3": dY.sum(0)}
    dZ2 = (dYb @ W3.T) * (1 - H2 * H2)
    dZ2b = _r(dZ2, exact)
    g["W2"] = H1b.T @ dZ2b
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2b @ W2.T) * (1 - H1 * H1)
    dZ1b = _r(dZ1, exact)
    g["W1"] = Xb.T @ dZ1b
    g["b1"] = dZ1.sum(0)
    return pack(g).astype(dt), loss


def loss_exact(params, X, T):
    Y, _ = forward(np.asarray(params, dtype=np.float64), X, exact=True)
    return float(((Y - T) ** 2).sum() / Y.shape[0])
