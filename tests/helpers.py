"""Shared test helpers: cases, states and error norms."""
from __future__ import annotations

import numpy as np

from esdgsem_mpiocca_amd import Case, RunParams


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def perturbed_state(case: Case, seed=1, amp_h=0.05, amp_u=0.3):
    """A smooth-plus-noise wet state on the case's mesh (h > 0, non-zero momenta)."""
    rng = np.random.default_rng(seed)
    ngl2 = case.ngl ** 2
    q = case["q"].reshape(case.Nelem, 3, ngl2).copy()
    x = case["x"].reshape(case.Nelem, ngl2)
    y = case["y"].reshape(case.Nelem, ngl2)
    h = q[:, 0]
    h *= 1.0 + amp_h * (np.sin(1.3 * x + 0.2) * np.cos(0.7 * y) + 0.2 * rng.uniform(-1, 1, h.shape))
    q[:, 1] = h * (amp_u * np.cos(0.9 * y) + 0.1 * rng.uniform(-1, 1, h.shape))
    q[:, 2] = h * (amp_u * np.sin(1.1 * x) + 0.1 * rng.uniform(-1, 1, h.shape))
    return np.ascontiguousarray(q.ravel())


def small_cases():
    """(id, case, run params) tuples covering Cartesian/curved, N = 4/5/7, both ES fluxes."""
    out = []
    c = Case.gaussian_bump(N=4, nel=8)
    out.append(("cart_N4_es", c, RunParams(num_flux=0)))
    c = Case.dam_break(N=7, nel=6, noise=1e-3, seed=3)
    out.append(("curved_N7_esrot", c, RunParams(num_flux=3)))
    c = Case.dam_break(N=5, nel=5, noise=1e-3, seed=4)
    out.append(("curved_N5_es", c, RunParams(num_flux=0)))
    return out
