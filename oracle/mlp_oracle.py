"""TEST ORACLE for the learned troubled-cell indicator (include/esdg_mlp.h).  Test infrastructure only.

The indicator has no counterpart in the reference; this module restates its definition in float64
numpy: the element features, the BF16-rounded forward pass and the weighted cross-entropy loss with
its gradient (straight-through through the BF16 roundings).  The gradient is pinned by central
finite differences of the loss with the roundings switched off (test_mlp_oracle.py); the modal
features by closed forms (a polynomial of degree k has no energy above shell k).
"""
from __future__ import annotations

import numpy as np

NIN, NH = 72, 32
OFF_W1 = 0
OFF_B1 = NIN * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def rbf(a) -> np.ndarray:
    """Round to nearest even BF16 (through float32), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[OFF_W1:OFF_B1].reshape(NIN, NH), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(NH, NH),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3], w[OFF_B3])


def features(q, b, V, Nelem):
    """[Nelem][72] features of a state q ([Nelem][3][64]) over bathymetry b ([Nelem][64])."""
    Q = np.asarray(q, dtype=np.float64).reshape(Nelem, 3, 64)
    h = Q[:, 0]
    B = np.asarray(b, dtype=np.float64).reshape(Nelem, 64)
    V = np.asarray(V, dtype=np.float64).reshape(8, 8)
    mean = h.mean(axis=1, keepdims=True)
    dev = h - mean
    s = np.abs(dev).max(axis=1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        x = np.where(s > 0, dev / np.where(s > 0, s, 1.0), 0.0)
    U = (B + h).reshape(Nelem, 8, 8)                     # U[e][j][i] = u at node j*8 + i
    M = np.einsum("ai,eji,bj->eba", V, U, V)             # M[e][k2][k1]
    k = np.arange(8)
    shell = np.maximum(k[:, None], k[None, :])
    E2 = M * M
    Ek = np.stack([(E2 * (shell == s_)).sum(axis=(1, 2)) for s_ in range(8)], axis=1)
    Et = Ek.sum(axis=1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.maximum(np.log10(Ek / np.where(Et > 0, Et, 1.0)), -16.0) / 8.0
    f = np.where(Et > 0, f, -2.0)
    return np.concatenate([x, f], axis=1)


def forward(w, X, exact=False):
    """z1, a1, z2, a2, z3, p for feature rows X; BF16 roundings unless exact."""
    W1, b1, W2, b2, w3, b3 = unpack(w)
    r = (lambda a: np.asarray(a, dtype=np.float64)) if exact else rbf
    xq = r(X)
    z1 = xq @ r(W1) + b1
    a1 = r(np.maximum(z1, 0.0))
    z2 = a1 @ r(W2) + b2
    a2 = r(np.maximum(z2, 0.0))
    z3 = a2 @ r(w3) + b3
    p = 1.0 / (1.0 + np.exp(-z3))
    return xq, z1, a1, z2, a2, z3, p


def loss_grad(w, X, y, pos_weight, exact=False):
    """Mean of -(pw y log p + (1 - y) log(1 - p)) and its gradient (straight-through roundings)."""
    W1, b1, W2, b2, w3, b3 = unpack(w)
    r = (lambda a: np.asarray(a, dtype=np.float64)) if exact else rbf
    xq, z1, a1, z2, a2, z3, p = forward(w, X, exact)
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    sp_neg = np.logaddexp(0.0, -z3)      # -log p
    sp_pos = np.logaddexp(0.0, z3)       # -log(1 - p)
    L = np.mean(pos_weight * y * sp_neg + (1 - y) * sp_pos)
    dz3 = ((1 - y) * p - pos_weight * y * (1 - p)) / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = a2.T @ dz3
    g[OFF_B3] = dz3.sum()
    dz2 = np.where(z2 > 0, np.outer(dz3, r(w3)), 0.0)
    g[OFF_W2:OFF_B2] = (a1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz1 = np.where(z1 > 0, dz2 @ r(W2).T, 0.0)
    g[OFF_W1:OFF_B1] = (xq.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return L, g
