"""CPU restatement of the relaxation surrogate network (3-64-64-64-3, tanh).

TEST INFRASTRUCTURE ONLY: the checker of the GPU BF16-MFMA forward and
backward kernels.  The network has no counterpart in the reference.

Two evaluations:
  forward(..., bf16=True)   the GPU numerics: layer-1 FP32, BF16 rounding of
                            every hidden activation and of W2, W3, FP64 sums
  forward(..., bf16=False)  exact FP64 (for finite-difference checks)
backward() is the analytic gradient of L = sum (y - t)^2 / (3 B) in FP64,
with the same rounding points as the GPU when bf16=True (the GPU also
rounds the back-propagated dz to BF16 before its MFMAs)."""
import numpy as np

SHAPES = [("W1", (64, 3)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (3, 64)), ("b4", (3,))]


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(w[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([d[n].ravel() for n, _ in SHAPES])


def forward(w, xn, bf16_round=True):
    """xn: normalised inputs [B, 3]; returns (yn [B, 3], activations)."""
    p = unpack(w)
    r = bf16 if bf16_round else (lambda a: np.asarray(a, dtype=np.float64))
    if bf16_round:
        x32 = np.asarray(xn, dtype=np.float32)
        z1 = (p["b1"].astype(np.float32)[None] + x32 @ p["W1"].astype(np.float32).T).astype(np.float64)
    else:
        z1 = p["b1"][None] + np.asarray(xn, dtype=np.float64) @ p["W1"].T
    h1 = r(np.tanh(z1))
    W2, W3 = (bf16(p["W2"]), bf16(p["W3"])) if bf16_round else (p["W2"], p["W3"])
    h2 = r(np.tanh(h1 @ W2.T + p["b2"][None]))
    h3 = r(np.tanh(h2 @ W3.T + p["b3"][None]))
    y = h3 @ p["W4"].T + p["b4"][None]
    return y, (np.asarray(xn, dtype=np.float64), h1, h2, h3, W2, W3)


def loss(w, xn, tn, bf16_round=False):
    y, _ = forward(w, xn, bf16_round)
    return np.sum((y - tn) ** 2) / (3 * len(xn))


def backward(w, xn, tn, bf16_round=True):
    """Gradient of loss() with respect to the packed parameters."""
    p = unpack(w)
    y, (x, h1, h2, h3, W2, W3) = forward(w, xn, bf16_round)
    r = bf16 if bf16_round else (lambda a: a)
    B = len(x)
    dy = 2.0 * (y - tn) / (3 * B)
    g = {"W4": dy.T @ h3, "b4": dy.sum(0)}
    dz3 = (dy @ p["W4"]) * (1 - h3 ** 2)
    g["b3"] = dz3.sum(0)
    dz3r = r(dz3)
    g["W3"] = dz3r.T @ h2
    dz2 = (dz3r @ W3) * (1 - h2 ** 2)
    g["b2"] = r(dz2).sum(0)
    dz2r = r(dz2)
    g["W2"] = dz2r.T @ h1
    dz1 = (dz2r @ W2) * (1 - h1 ** 2)
    dz1r = r(dz1)
    g["b1"] = dz1r.sum(0)
    g["W1"] = dz1r.T @ x
    return pack(g)
