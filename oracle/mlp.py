"""CPU oracle of the learned split proposer (test infrastructure only).

The network has no counterpart in the reference; this file restates what the
device kernels compute so the tests can pin them:

* ``node_features``  the 102 inputs of a node (centroid-bin counts and
                     extents), as the builder writes them;
* ``forward`` / ``loss_and_grads``  FP32 MLP 102 -> 128 -> 128 -> 96 with ReLU,
                     softmax cross-entropy, analytic gradients (checked in the
                     tests against central finite differences in float64);
* ``bf16``           round-to-nearest-even float32 -> bfloat16 (known answers
                     in the tests);
* ``infer_bf16``     the inference path: BF16 inputs, weights and hidden
                     activations, exact products summed in float64, rounded to
                     float32 (the device accumulates in FP32, so tests allow a
                     small tolerance).

Parameter vector layout (float32): W1[128][128] (inputs 102..127 zero),
b1[128], W2[128][128], b2[128], W3[96][128], b3[96].
"""
from __future__ import annotations

import numpy as np

IN, INP, H, OUT = 102, 128, 128, 96
SIZES = [("W1", (H, INP)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)),
         ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float32).ravel() for name, _ in SIZES])


def node_features(counts: np.ndarray, n: int, bmn, bmx, cext) -> np.ndarray:
    """counts: (3, 32) centroid-bin counts of a node of n objects; bmn/bmx its
    bounds; cext its centroid extents."""
    f = np.zeros(IN, dtype=np.float32)
    f[:96] = (np.log1p(np.asarray(counts, np.float32).ravel()) / np.float32(np.log1p(np.float32(n))))
    ext = np.asarray(bmx, np.float32) - np.asarray(bmn, np.float32)
    emax = ext.max()
    f[96:99] = ext / emax if emax > 0 else 0.0
    ce = np.asarray(cext, np.float32)
    f[99:102] = np.where(ext > 0, np.minimum(ce / np.where(ext > 0, ext, 1), 1.0), 0.0)
    return f


def _pad(X):
    X = np.asarray(X, dtype=np.float64)
    out = np.zeros((X.shape[0], INP), dtype=np.float64)
    out[:, :IN] = X[:, :IN]
    return out


def forward(p: np.ndarray, X: np.ndarray, dtype=np.float64):
    w = {k: v.astype(dtype) for k, v in unpack(np.asarray(p, dtype)).items()}
    x = _pad(X).astype(dtype)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0)
    z3 = h2 @ w["W3"].T + w["b3"]
    return x, h1, h2, z3


def loss_and_grads(p: np.ndarray, X: np.ndarray, y: np.ndarray, dtype=np.float64):
    """Mean softmax cross-entropy and its gradient in the parameter layout."""
    w = {k: v.astype(dtype) for k, v in unpack(np.asarray(p, dtype)).items()}
    x, h1, h2, z3 = forward(p, X, dtype)
    B = x.shape[0]
    z = z3 - z3.max(axis=1, keepdims=True)
    e = np.exp(z)
    s = e.sum(axis=1, keepdims=True)
    prob = e / s
    loss = float(np.mean(np.log(s[:, 0]) - z[np.arange(B), y]))
    d3 = prob.copy()
    d3[np.arange(B), y] -= 1.0
    d3 /= B
    g = {"W3": d3.T @ h2, "b3": d3.sum(axis=0)}
    d2 = (d3 @ w["W3"]) * (h2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(axis=0)
    d1 = (d2 @ w["W2"]) * (h1 > 0)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(axis=0)
    return loss, np.concatenate([np.asarray(g[name]).ravel() for name, _ in SIZES])


def loss_only(p: np.ndarray, X: np.ndarray, y: np.ndarray) -> float:
    _, _, _, z3 = forward(p, X, np.float64)
    z = z3 - z3.max(axis=1, keepdims=True)
    s = np.exp(z).sum(axis=1)
    return float(np.mean(np.log(s) - z[np.arange(z.shape[0]), y]))


def bf16(x) -> np.ndarray:
    """float32 -> bfloat16 (round to nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def infer_bf16(p: np.ndarray, X: np.ndarray) -> np.ndarray:
    w = unpack(np.asarray(p, np.float32))
    x = bf16(_pad(X).astype(np.float32)).astype(np.float64)
    W1, W2, W3 = (bf16(w[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    z1 = (x @ W1.T).astype(np.float32) + w["b1"]
    h1 = bf16(np.maximum(z1, 0)).astype(np.float64)
    z2 = (h1 @ W2.T).astype(np.float32) + w["b2"]
    h2 = bf16(np.maximum(z2, 0)).astype(np.float64)
    return (h2 @ W3.T).astype(np.float32) + w["b3"]
