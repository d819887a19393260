"""Entropy budget of the GPU residual (the semi-discrete statement behind basis::calcTotalEntropy,
src/basis.cpp:970-1111, and calcESFlux.okl:111-200).

For the entropy S = sum w_i w_j J (h|u|^2/2 + g h^2/2 + g h b) with entropy variables
v = (g(h+b) - |u|^2/2, u, v), the rate of the stage residual Qt is dS/dt = sum w J v . Qt.  With the
EC volume flux and the EC surface flux (ES=0) it is zero on a periodic mesh; with the ES flux
(ES=1) it equals the summed face dissipation

    -1/2 sum_faces sum_k w_k scal_k [v]^T R |Lambda| R^T [v],

R |Lambda| R^T the eigen-structure of calcESFlux.okl, evaluated here in numpy from the face
traces the GPU gathered.  Both hold to roundoff; a dropped factor (1/2, scal, a weight) or a
wrong sign in any of the volume, surface or flux kernels breaks them.
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, RunParams
from helpers import perturbed_state

pytestmark = pytest.mark.gpu

G = 9.81


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


def gpu_residual(dev, case, ES, q):
    """Qt (gather, flux-differencing volume, ES/EC surface flux, lift) and the face traces."""
    rp = RunParams(num_flux=0, ES=ES, g=G)
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        a = case.arrays
        d = {k: ctx.array(a[k]) for k in ("EdgeData", "EdgeReversed", "b", "nx", "ny", "scal", "Jac", "yXi", "yEta",
                                          "xXi", "xEta", "D", "Bx", "By", "ElemEdgeMasterSlave",
                                          "ElemEdgeOrientation", "ElemToEdge", "x", "y")}
        Nf, Ne, ngl = case.Nfaces, case.Nelem, case.ngl
        fn = Nf * ngl
        dq = ctx.array(q)
        bL, bR, qL, qR = ctx.empty(fn), ctx.empty(fn), ctx.empty(fn * 3), ctx.empty(fn * 3)
        ctx.collect_edge_data_bottom(Nf, d["EdgeData"], d["b"], bL, bR)
        ctx.collect_edge_data(Nf, d["EdgeData"], dq, d["x"], d["y"], d["nx"], d["ny"], bL, bR, qL, qR, 0.0)
        Qt = ctx.empty(Ne * ngl * ngl * 3)
        ctx.volume_kernel(Ne, d["Jac"], d["yXi"], d["yEta"], d["xXi"], d["xEta"], dq, d["D"], d["Bx"], d["By"], Qt)
        SP = ctx.empty(fn * 3)
        ctx.calc_num_fluxes(Nf, d["EdgeData"], d["EdgeReversed"], d["nx"], d["ny"], d["scal"], qL, qR, bL, bR, SP)
        zero = ctx.empty(fn)
        ctx.surface_kernel(Ne, d["Jac"], d["ElemEdgeMasterSlave"], d["ElemEdgeOrientation"], d["ElemToEdge"], SP,
                           zero, zero, Qt)
        return Qt.download(), qL.download(), qR.download(), bL.download(), bR.download()
    finally:
        ctx.close()


def entropy_vars(h, hu, hv, b):
    u, v = hu / h, hv / h
    return np.stack([G * (h + b) - 0.5 * (u * u + v * v), u, v], -1)


@pytest.mark.parametrize("nel,N", [(8, 4), (6, 7)])
def test_entropy_rate_equals_face_dissipation(dev, nel, N):
    case = Case.gaussian_bump(N=N, nel=nel)
    assert (case["EdgeReversed"] == 1).all()        # both sides of every face list their nodes alike
    q = perturbed_state(case, seed=6, amp_u=0.5)
    Ne, ngl, Nf = case.Nelem, case.ngl, case.Nfaces
    w = case["w_GL"]
    wJ = (np.outer(w, w).ravel()[None, :] / case["Jac"].reshape(Ne, -1))
    Q = q.reshape(Ne, 3, -1)
    vq = entropy_vars(Q[:, 0], Q[:, 1], Q[:, 2], case["b"].reshape(Ne, -1))           # [e][node][3]

    rates = {}
    for ES in (0, 1):
        Qt, qL, qR, bL, bR = gpu_residual(dev, case, ES, q)
        T = Qt.reshape(Ne, 3, -1).transpose(0, 2, 1)
        rates[ES] = ((wJ[..., None] * vq * T).sum(), np.abs(wJ[..., None] * vq * T).sum())

    # the face dissipation from the traces (face arrays [f][v][k], nodes aligned on both sides)
    L, R = qL.reshape(Nf, 3, ngl), qR.reshape(Nf, 3, ngl)
    jv = entropy_vars(R[:, 0], R[:, 1], R[:, 2], bR.reshape(Nf, ngl)) - \
        entropy_vars(L[:, 0], L[:, 1], L[:, 2], bL.reshape(Nf, ngl))                   # [f][k][3]
    hL, hR = L[:, 0], R[:, 0]
    uA = 0.5 * (L[:, 1] / hL + R[:, 1] / hR)
    vA = 0.5 * (L[:, 2] / hL + R[:, 2] / hR)
    c = np.sqrt(0.5 * G * (hL + hR))
    nx, ny = case["nx"].reshape(Nf, ngl), case["ny"].reshape(Nf, ngl)
    un = uA * nx + vA * ny
    Rm = np.empty((Nf, ngl, 3, 3))
    Rm[..., 0, :] = np.stack([np.ones_like(uA), np.zeros_like(uA), np.ones_like(uA)], -1)
    Rm[..., 1, :] = np.stack([uA + nx * c, ny, uA - nx * c], -1)
    Rm[..., 2, :] = np.stack([vA + ny * c, nx, vA - ny * c], -1)
    lam = np.stack([np.abs(un + c) / (2 * G), 0.5 * (hL + hR) * np.abs(un), np.abs(un - c) / (2 * G)], -1)
    z = np.einsum("fkji,fkj->fki", Rm, jv)                                              # R^T [v]
    diss = -0.5 * (lam * z * z).sum(-1)                                                # [f][k]
    predicted = (w[None, :] * case["scal"].reshape(Nf, ngl) * diss).sum()

    dS0, scale0 = rates[0]
    dS1, scale1 = rates[1]
    assert abs(dS0) <= 1e-12 * scale0, (dS0, scale0)
    assert predicted < -1e-6 * scale1
    assert abs(dS1 - predicted) <= 1e-11 * scale1, (dS1, predicted, scale1)
