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


def oracle_residual(case, rp, q, av_visc=None):
    """Qt of one stage (gather, volume, numerical flux, surface lift [, BR1 + viscous terms]) built
    from the oracle kernels in the DGtimeloop order.  Returns Qt."""
    from oracle import binding as O
    p = O.params(case, rp)
    a = case.arrays
    Ne, Nf, ngl = case.Nelem, case.Nfaces, case.ngl
    fn, nn = Nf * ngl, Ne * ngl * ngl
    q = np.ascontiguousarray(q, dtype=np.float64)
    qL, qR, bL, bR = np.zeros(fn * 3), np.zeros(fn * 3), np.zeros(fn), np.zeros(fn)
    O.call("collect_edge_data_bottom", p, Nf, a["EdgeData"], a["b"], bL, bR)
    O.call("collect_edge_data", p, Nf, a["EdgeData"], q, a["nx"], a["ny"], bL, bR, qL, qR)
    Qt = np.zeros(nn * 3)
    O.call("volume_fluxdiff", p, Ne, a["Jac"], a["yXi"], a["yEta"], a["xXi"], a["xEta"], q, a["D"], a["Bx"],
           a["By"], Qt)
    SP = np.zeros(fn * 3)
    O.call("num_flux_es_rotated" if rp.num_flux == 3 else "num_flux_es", p, Nf, a["EdgeReversed"], a["nx"],
           a["ny"], a["scal"], qL, qR, bL, bR, SP)
    DB = np.zeros(fn)
    O.call("surface", p, Ne, a["Jac"], a["ElemEdgeMasterSlave"], a["ElemEdgeOrientation"], a["ElemToEdge"], SP,
           DB, DB, Qt)
    if av_visc is not None:
        visc = np.ascontiguousarray(av_visc, dtype=np.float64)
        SGX, SGY = np.zeros(fn * 2), np.zeros(fn * 2)
        O.call("num_fluxes_gradient", p, Nf, a["EdgeData"], a["nx"], a["ny"], a["scal"], qL, qR, SGX, SGY)
        gX, gY = np.zeros(nn * 2), np.zeros(nn * 2)
        O.call("calc_gradient", p, Ne, a["Jac"], a["yXi"], a["yEta"], a["xXi"], a["xEta"], q, a["Dhat"], gX, gY)
        O.call("surface_gradient", p, Ne, a["Jac"], a["ElemEdgeMasterSlave"], a["ElemEdgeOrientation"],
               a["ElemToEdge"], SGX, SGY, gX, gY)
        O.call("scale_gradient", p, Ne, q, gX, gY)
        VL, VR = np.zeros(Nf), np.zeros(Nf)
        XL, XR, YL, YR = (np.zeros(fn * 2) for _ in range(4))
        O.call("collect_edge_data_gradient", p, Nf, a["EdgeData"], gX, gY, visc, VL, VR, XL, XR, YL, YR)
        O.call("volume_viscose", p, Ne, a["Jac"], a["yXi"], a["yEta"], a["xXi"], a["xEta"], gX, gY, a["Dstrong"],
               visc, Qt)
        SPV = np.zeros(fn * 2)
        O.call("num_fluxes_viscose", p, Nf, a["EdgeData"], a["nx"], a["ny"], a["scal"], VL, VR, XL, XR, YL, YR, SPV)
        O.call("surface_visc", p, Ne, a["Jac"], a["ElemEdgeMasterSlave"], a["ElemEdgeOrientation"],
               a["ElemToEdge"], SPV, Qt)
    return Qt


def quad_weights(case):
    """w_i w_j J at every node (J = 1/Jac): the discrete volume measure."""
    w = case["w_GL"]
    ww = np.outer(w, w).ravel()            # index j*ngl + i  ->  w_j w_i
    return (ww[None, :] / case["Jac"].reshape(case.Nelem, -1))
