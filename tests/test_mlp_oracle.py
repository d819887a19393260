"""CPU checks that pin the learned-indicator oracle (oracle/mlp_oracle.py).

- the loss gradient against central finite differences (roundings off: the formula itself);
- BF16 rounding against known answers (round to nearest, ties to even);
- the features against closed forms: constant water height gives zero nodal features, a polynomial
  surface of degree k has no modal energy above shell k, the nodal features span [-1, 1].
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case
from oracle import mlp_oracle as M


def random_problem(n=64, seed=0):
    rng = np.random.default_rng(seed)
    w = np.concatenate([rng.uniform(-0.3, 0.3, M.NIN * M.NH), rng.uniform(-0.1, 0.1, M.NH),
                        rng.uniform(-0.3, 0.3, M.NH * M.NH), rng.uniform(-0.1, 0.1, M.NH),
                        rng.uniform(-0.5, 0.5, M.NH), [0.1]])
    X = np.concatenate([rng.uniform(-1, 1, (n, 64)), rng.uniform(-2, 0, (n, 8))], axis=1)
    y = (rng.uniform(size=n) < 0.3).astype(float)
    return w, X, y


def test_gradient_matches_finite_differences():
    w, X, y = random_problem()
    L, g = M.loss_grad(w, X, y, pos_weight=3.0, exact=True)
    rng = np.random.default_rng(1)
    picks = np.concatenate([rng.choice(M.OFF_B1, 20, replace=False),
                            M.OFF_B1 + rng.choice(M.NH, 5, replace=False),
                            M.OFF_W2 + rng.choice(M.NH * M.NH, 20, replace=False),
                            M.OFF_B2 + rng.choice(M.NH, 5, replace=False),
                            M.OFF_W3 + rng.choice(M.NH, 8, replace=False), [M.OFF_B3]])
    h = 1e-6
    for k in picks:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (M.loss_grad(wp, X, y, 3.0, exact=True)[0] - M.loss_grad(wm, X, y, 3.0, exact=True)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_gradient_with_roundings_is_the_same_formula_at_rounded_values():
    """Straight-through: with BF16 roundings the gradient equals the exact-formula gradient taken
    at rounded weights and inputs when no activation rounding changes a ReLU's side."""
    w, X, y = random_problem(seed=3)
    Lr, gr = M.loss_grad(w, X, y, 2.0)
    assert np.isfinite(gr).all() and np.isfinite(Lr)
    # rounding perturbs the loss only at the BF16 level
    Le, _ = M.loss_grad(w, X, y, 2.0, exact=True)
    assert abs(Lr - Le) < 0.05 * abs(Le)


def test_bf16_rounding_known_answers():
    one = 1.0
    assert M.rbf([one])[0] == 1.0
    assert M.rbf([1.0 + 2.0 ** -9])[0] == 1.0                       # below half: down
    assert M.rbf([1.0 + 3 * 2.0 ** -9])[0] == 1.0 + 2.0 ** -7        # above half: up
    assert M.rbf([1.0 + 2.0 ** -8])[0] == 1.0                        # tie: to even (down)
    assert M.rbf([1.0 + 3 * 2.0 ** -8])[0] == 1.0 + 2.0 ** -6        # tie: to even (up)
    assert M.rbf([-2.5])[0] == -2.5


def test_features_closed_forms():
    c = Case.dam_break(N=7, nel=3, noise=0.0, seed=0)
    V = c["VdmInv"]
    Ne = c.Nelem
    # constant water height: nodal features vanish
    q = np.zeros((Ne, 3, 64))
    q[:, 0] = 1.5
    F = M.features(q.ravel(), c["b"], V, Ne)
    assert np.abs(F[:, :64]).max() == 0.0
    # a surface u = h + b of degree <= 2 in (xi, eta): no energy above shell 2
    xg = c["x_GL"]
    XI, ETA = np.meshgrid(xg, xg)                 # [j][i]
    u = 1.0 + 0.3 * XI - 0.2 * ETA ** 2 + 0.1 * XI * ETA
    q = np.zeros((Ne, 3, 64))
    b = np.zeros((Ne, 64))
    q[:, 0] = u.ravel()
    F = M.features(q.ravel(), b.ravel(), V, Ne)
    assert np.all(F[:, 64 + 3:] <= -1.5)          # energy ratio of shells 3..7 below 1e-12
    assert np.all(F[:, 64:67] > -2.0)
    # nodal features span [-1, 1] exactly at the extreme node
    q[:, 0] = u.ravel() + np.linspace(0, 1, 64)
    F = M.features(q.ravel(), b.ravel(), V, Ne)
    assert np.allclose(np.abs(F[:, :64]).max(axis=1), 1.0)
