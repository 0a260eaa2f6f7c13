"""FP64 NumPy statement of the learned-Smooth MLP -- TEST INFRASTRUCTURE ONLY.

Only tests/ import it.  Forward, loss and analytic backward of
28 -> 64 (SiLU) -> 64 (SiLU) -> 3 with the flat parameter layout the GPU uses
(W1[28][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]), plus finite-difference
checks that pin the backward, and a BF16 (round-to-nearest-even) emulation of the
inference path.
"""
import numpy as np

NIN, NH, NOUT = 28, 64, 3
SIZES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NOUT)), ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(w):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(w[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def normalize(X, T, norm):
    mu, isig, osc = norm
    return (np.asarray(X, np.float64) - mu) * isig, np.asarray(T, np.float64) / osc


def loss_and_grad(w, Xn, Tn):
    """Mean over rows and outputs of (y - t)^2 and its gradient (FP64)."""
    p = unpack(w)
    z1 = Xn @ p["W1"] + p["b1"]
    a1 = silu(z1)
    z2 = a1 @ p["W2"] + p["b2"]
    a2 = silu(z2)
    y = a2 @ p["W3"] + p["b3"]
    e = y - Tn
    n = e.size
    loss = float((e * e).sum() / n)
    dy = 2.0 * e / n
    g = {"W3": a2.T @ dy, "b3": dy.sum(0)}
    dz2 = (dy @ p["W3"].T) * silu_grad(z2)
    g["W2"] = a1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * silu_grad(z1)
    g["W1"] = Xn.T @ dz1
    g["b1"] = dz1.sum(0)
    return loss, np.concatenate([g[name].reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round FP32 values to BF16 (nearest even) and back to FP32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def predict_bf16(w, Xn):
    """The inference path: BF16 inputs, weights and hidden activations, FP32 sums."""
    p = {k: np.asarray(v, np.float32) for k, v in unpack(w).items()}
    x = bf16(np.asarray(Xn, np.float32))
    h1 = bf16(silu((x @ bf16(p["W1"])) + p["b1"]).astype(np.float32))
    h2 = bf16(silu((h1 @ bf16(p["W2"])) + p["b2"]).astype(np.float32))
    return (h2 @ bf16(p["W3"])) + p["b3"]
