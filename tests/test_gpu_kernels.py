"""GPU parity of every C-ABI kernel against the oracle on identical inputs.

Each kernel is fed the same arrays on the host (oracle) and on the device (HIP, through the
C ABI); the inputs of downstream kernels are produced by the oracle so each kernel is checked in
isolation.  Tolerances: copies and index work are bit-exact.  Device code and oracle are both
compiled without multiply-add contraction, so the arithmetic kernels mostly agree bitwise; the
bounds stated per test leave room for the transcendental functions (sqrt is exact; log10 and sin
in the shock sensor may differ by an ulp).
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import RunParams
from helpers import max_rel, perturbed_state, rel_l2, small_cases
from oracle import binding as O

pytestmark = pytest.mark.gpu

CASES = small_cases()


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


class Setup:
    def __init__(self, dev, case, rp, av=False, eps0=0.0):
        rp = RunParams(**{**rp.__dict__, "eps0": eps0 if av else rp.eps0,
                          "sigma_min": -9.0, "sigma_max": -3.0})
        self.case, self.rp = case, rp
        self.p = O.params(case, rp)
        self.ctx = dev.Context(dev.make_params(case, rp))
        a = case.arrays
        self.h = {
            "Jac": a["Jac"], "Yxi": a["yXi"], "Yeta": a["yEta"], "Xxi": a["xXi"], "Xeta": a["xEta"],
            "Bx": a["Bx"], "By": a["By"], "B": a["b"], "D": a["D"], "Dhat": a["Dhat"],
            "Dstrong": a["Dstrong"], "VdmInv": a["VdmInv"], "nx": a["nx"], "ny": a["ny"],
            "scal": a["scal"], "EleSizes": a["EleSizes"], "GLw": a["w_GL"], "x": a["x"], "y": a["y"],
            "EdgeData": a["EdgeData"], "ER": a["EdgeReversed"], "E2E": a["ElemToEdge"],
            "EMS": a["ElemEdgeMasterSlave"], "EOR": a["ElemEdgeOrientation"]}
        self.d = {k: self.ctx.array(v) for k, v in self.h.items()}
        self.q = perturbed_state(case)
        self.Ne, self.Nf, self.ngl = case.Nelem, case.Nfaces, case.ngl

    def zeros(self, n):
        return np.zeros(n)

    def close(self):
        self.ctx.close()


@pytest.fixture(params=CASES, ids=[c[0] for c in CASES])
def S(request, dev):
    _, case, rp = request.param
    s = Setup(dev, case, rp)
    yield s
    s.close()


def test_collect_edge_data_bottom(S):
    fn = S.Nf * S.ngl
    bL, bR = np.zeros(fn), np.zeros(fn)
    O.call("collect_edge_data_bottom", S.p, S.Nf, S.h["EdgeData"], S.h["B"], bL, bR)
    dL, dR = S.ctx.empty(fn), S.ctx.empty(fn)
    S.ctx.collect_edge_data_bottom(S.Nf, S.d["EdgeData"], S.d["B"], dL, dR)
    assert np.array_equal(dL.download(), bL) and np.array_equal(dR.download(), bR)


@pytest.mark.parametrize("bc", [1, 2], ids=["inner_outer", "ocean"])
def test_collect_edge_data_boundary_variants(dev, bc):
    """InnerOuter.okl and OceanBoundary.okl exterior states (both ocean branches occur)."""
    from esdgsem_mpiocca_amd import Case
    case = Case.ocean_uplift(N=4, nel=6, amp=2.0)
    rp = RunParams(bc_kind=bc, h_zero=4000.0)
    s = Setup(dev, case, rp)
    try:
        fn = s.Nf * s.ngl
        bL, bR = np.zeros(fn), np.zeros(fn)
        O.call("collect_edge_data_bottom", s.p, s.Nf, s.h["EdgeData"], s.h["B"], bL, bR)
        qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
        O.call("collect_edge_data", s.p, s.Nf, s.h["EdgeData"], s.q, s.h["nx"], s.h["ny"], bL, bR, qL, qR)
        dq = s.ctx.array(s.q)
        dL, dR = s.ctx.empty(fn * 3), s.ctx.empty(fn * 3)
        dbL, dbR = s.ctx.array(bL), s.ctx.array(bR)
        s.ctx.collect_edge_data(s.Nf, s.d["EdgeData"], dq, s.d["x"], s.d["y"], s.d["nx"], s.d["ny"], dbL, dbR,
                                dL, dR, 0.0)
        assert np.array_equal(dL.download(), qL) and np.array_equal(dR.download(), qR)
    finally:
        s.close()


@pytest.mark.parametrize("bc,t", [(3, 0.37), (4, 123.0), (4, 400.0), (5, 0.37), (6, 0.37), (7, 0.37)],
                         ids=["convergence", "three_mound_early", "three_mound_late", "conv_test3",
                              "conv_test4", "conv_test5"])
def test_collect_edge_data_time_dependent(dev, bc, t):
    """Dirichlet_ConvTest{,3,4,5}.okl (manufactured exterior state) and 3MoundInflow.okl (walls,
    inflow at x = 0 switching at t = 300, outflow): the exterior states use the node coordinates and t."""
    from esdgsem_mpiocca_amd import Case
    from esdgsem_mpiocca_amd import setup_api as SA
    if bc in (3, 5, 6, 7):
        testcase = {3: 1, 5: 8, 6: 88, 7: 89}[bc]
        case = Case.convergence(N=4, nel=5, curved=True, testcase=testcase)
    else:
        case = Case(SA.CASE_CONSTANT, SA.MESH_CARTESIAN, 4, 6, 3, 0.0, 4.0, 0.0, 2.0, p=(1.0, 0.2, 0.1, 0))
    rp = RunParams(bc_kind=bc)
    s = Setup(dev, case, rp)
    try:
        fn = s.Nf * s.ngl
        bL, bR = np.zeros(fn), np.zeros(fn)
        O.call("collect_edge_data_bottom", s.p, s.Nf, s.h["EdgeData"], s.h["B"], bL, bR)
        qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
        O.call("collect_edge_data_xt", s.p, s.Nf, s.h["EdgeData"], s.q, s.h["x"], s.h["y"], s.h["nx"], s.h["ny"],
               bL, bR, t, qL, qR)
        dq = s.ctx.array(s.q)
        dL, dR = s.ctx.empty(fn * 3), s.ctx.empty(fn * 3)
        dbL, dbR = s.ctx.array(bL), s.ctx.array(bR)
        s.ctx.collect_edge_data(s.Nf, s.d["EdgeData"], dq, s.d["x"], s.d["y"], s.d["nx"], s.d["ny"], dbL, dbR,
                                dL, dR, t)
        assert np.array_equal(dL.download(), qL)
        gR = dR.download()
        # the manufactured state goes through cos/sin: device and host libm may differ in the last ulp
        assert np.abs(gR - qR).max() <= 4e-15 * np.abs(qR).max()
    finally:
        s.close()


def test_collect_edge_data_walls(S):
    fn = S.Nf * S.ngl
    qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
    O.call("collect_edge_data_walls", S.p, S.Nf, S.h["EdgeData"], S.q, S.h["nx"], S.h["ny"], qL, qR)
    dq = S.ctx.array(S.q)
    dL, dR = S.ctx.empty(fn * 3), S.ctx.empty(fn * 3)
    bl = S.ctx.empty(fn)
    S.ctx.collect_edge_data(S.Nf, S.d["EdgeData"], dq, S.d["x"], S.d["y"], S.d["nx"], S.d["ny"], bl, bl,
                            dL, dR, 0.0)
    gL, gR = dL.download(), dR.download()
    assert np.array_equal(gL, qL)            # interior traces are copies
    assert max_rel(gR, qR) <= 4e-16          # wall mirror: a few flops per node


def test_volume_kernel(S):
    nq = S.Ne * S.ngl ** 2 * 3
    Qt = np.zeros(nq)
    h = S.h
    O.call("volume_fluxdiff", S.p, S.Ne, h["Jac"], h["Yxi"], h["Yeta"], h["Xxi"], h["Xeta"], S.q, h["D"],
           h["Bx"], h["By"], Qt)
    d = S.d
    dq, dQt = S.ctx.array(S.q), S.ctx.empty(nq)
    S.ctx.volume_kernel(S.Ne, d["Jac"], d["Yxi"], d["Yeta"], d["Xxi"], d["Xeta"], dq, d["D"], d["Bx"],
                        d["By"], dQt)
    g = dQt.download()
    for v in range(3):
        a = g.reshape(S.Ne, 3, -1)[:, v]
        b = Qt.reshape(S.Ne, 3, -1)[:, v]
        assert rel_l2(a, b) <= 1e-13, (v, rel_l2(a, b))


@pytest.mark.parametrize("flux", [0, 1, 3, 4, 5])
def test_calc_num_fluxes(dev, flux):
    _, case, rp = CASES[1]
    rp = RunParams(**{**rp.__dict__, "num_flux": flux})
    S = Setup(dev, case, rp)
    try:
        fn = S.Nf * S.ngl
        qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
        O.call("collect_edge_data_walls", S.p, S.Nf, S.h["EdgeData"], S.q, S.h["nx"], S.h["ny"], qL, qR)
        bL, bR = np.zeros(fn), np.zeros(fn)
        O.call("collect_edge_data_bottom", S.p, S.Nf, S.h["EdgeData"], S.h["B"], bL, bR)
        SP = np.zeros(fn * 3)
        if flux == 5:
            # boundary bottom traces on both sides of 6.7 so that both branches are taken
            bL = np.where(np.arange(fn) % 2 == 0, bL + 5.0, bL + 8.0)
            O.call("num_flux_ocean", S.p, S.Nf, S.h["EdgeData"], S.h["ER"], S.h["nx"], S.h["ny"], S.h["scal"],
                   qL, qR, bL, bR, SP)
        else:
            name = {0: "num_flux_es", 1: "num_flux_lf", 3: "num_flux_es_rotated", 4: "num_flux_roe"}[flux]
            O.call(name, S.p, S.Nf, S.h["ER"], S.h["nx"], S.h["ny"], S.h["scal"], qL, qR, bL, bR, SP)
        dqL, dqR, dbL, dbR = (S.ctx.array(a) for a in (qL, qR, bL, bR))
        dSP = S.ctx.empty(fn * 3)
        S.ctx.calc_num_fluxes(S.Nf, S.d["EdgeData"], S.d["ER"], S.d["nx"], S.d["ny"], S.d["scal"], dqL, dqR,
                              dbL, dbR, dSP)
        assert rel_l2(dSP.download(), SP) <= 1e-14
    finally:
        S.close()


def test_volume_kernel_std(S):
    """FluxDifferencing = 0: standard strong-form volume term with D = Dstrong."""
    nq = S.Ne * S.ngl ** 2 * 3
    Qt = np.zeros(nq)
    h = S.h
    Ds = np.ascontiguousarray(h["Dstrong"])
    O.call("volume_std", S.p, S.Ne, h["Jac"], h["Yxi"], h["Yeta"], h["Xxi"], h["Xeta"], S.q, Ds, h["Bx"],
           h["By"], Qt)
    d = S.d
    dQt = S.ctx.empty(nq)
    S.ctx.volume_kernel_std(S.Ne, d["Jac"], d["Yxi"], d["Yeta"], d["Xxi"], d["Xeta"], S.ctx.array(S.q),
                            d["Dstrong"], d["Bx"], d["By"], dQt)
    assert rel_l2(dQt.download(), Qt) <= 1e-14


def test_surface_kernel(S):
    nq = S.Ne * S.ngl ** 2 * 3
    fn = S.Nf * S.ngl
    rng = np.random.default_rng(5)
    FS = rng.standard_normal(fn * 3)
    Qt0 = rng.standard_normal(nq)
    DB = np.zeros(fn)
    Qt = Qt0.copy()
    O.call("surface", S.p, S.Ne, S.h["Jac"], S.h["EMS"], S.h["EOR"], S.h["E2E"], FS, DB, DB, Qt)
    dQt = S.ctx.array(Qt0)
    dDB = S.ctx.array(DB)
    S.ctx.surface_kernel(S.Ne, S.d["Jac"], S.d["EMS"], S.d["EOR"], S.d["E2E"], S.ctx.array(FS), dDB, dDB, dQt)
    assert max_rel(dQt.download(), Qt) <= 1e-15


def test_find_lambda_max(S):
    lam = np.zeros(S.Ne)
    O.call("find_lambda_max", S.p, S.Ne, S.q, lam)
    dl = S.ctx.empty(S.Ne)
    S.ctx.find_lambda_max(S.Ne, S.ctx.array(S.q), dl)
    assert max_rel(dl.download(), lam) <= 4e-16
    dmax = S.ctx.empty(1)
    S.ctx.reduce_max(S.Ne, dl, dmax)
    assert dmax.download()[0] == dl.download().max()


def test_rk_kernels(S):
    nq = S.Ne * S.ngl ** 2 * 3
    rng = np.random.default_rng(7)
    Qt, g0, q0, qt = (rng.standard_normal(nq) for _ in range(4))
    A, Cc, dt = -0.4178904745, 0.3792103129, 1.3e-3
    g, q = g0.copy(), q0.copy()
    O.call("rk_ls_calc", S.p, S.Ne, Qt, A, g)
    O.call("rk_ls_update", S.p, S.Ne, Cc, dt, g, q)
    dg, dq = S.ctx.array(g0), S.ctx.array(q0)
    S.ctx.calc_rk(S.Ne, S.ctx.array(Qt), A, 0.0, dg)
    S.ctx.update_kernel_ls(S.Ne, Cc, dt, dg, dq)
    assert max_rel(dg.download(), g) <= 2e-16
    assert max_rel(dq.download(), q) <= 2e-16
    q2 = q0.copy()
    O.call("rk_ssp_update", S.p, S.Ne, 0.75, 0.25, 0.25, dt, Qt, qt, q2)
    dq2 = S.ctx.array(q0)
    S.ctx.update_kernel_ssp(S.Ne, 0.75, 0.25, 0.25, dt, S.ctx.array(Qt), S.ctx.array(qt), dq2)
    assert max_rel(dq2.download(), q2) <= 4e-16


def test_shock_capturing(dev):
    _, case, rp = CASES[1]
    S = Setup(dev, case, rp, av=True, eps0=0.01)
    try:
        visc = np.zeros(S.Ne)
        O.call("shock_capturing", S.p, S.Ne, S.q, S.h["B"], S.h["VdmInv"], visc, None)
        dv, dvp = S.ctx.empty(S.Ne), S.ctx.empty(S.Ne)
        S.ctx.shock_capturing(S.Ne, S.ctx.array(S.q), S.d["B"], S.d["VdmInv"], S.d["EleSizes"], dv, dvp)
        g = dv.download()
        assert np.abs(g - visc).max() <= 1e-12 * 0.01
        assert np.array_equal(g, dvp.download())
        assert (visc > 0).any() and (visc < 0.01).any()   # the sine ramp is exercised
    finally:
        S.close()


def test_pos_pres(S):
    ngl2 = S.ngl ** 2
    q = S.q.reshape(S.Ne, 3, ngl2).copy()
    rng = np.random.default_rng(11)
    # drive some nodes negative / below tolerance in a few elements
    for e in rng.choice(S.Ne, size=max(2, S.Ne // 5), replace=False):
        k = rng.choice(ngl2, size=3, replace=False)
        q[e, 0, k] = -0.3 * q[e, 0, k]
    q[0, 0, :] = -1.0                                  # negative mean -> -inf branch
    q = q.ravel()
    qo = q.copy()
    O.call("pos_pres", S.p, S.Ne, S.h["EleSizes"], S.h["GLw"], S.h["Jac"], qo)
    dq = S.ctx.array(q)
    S.ctx.pos_pres(S.Ne, S.d["EleSizes"], S.d["GLw"], S.d["Jac"], dq)
    g = dq.download()
    fin = np.isfinite(qo)
    assert np.array_equal(np.isfinite(g), fin)
    assert np.array_equal(g[~fin], qo[~fin])
    assert max_rel(g[fin], qo[fin]) <= 1e-15


def test_gradient_chain(dev):
    _, case, rp = CASES[1]
    S = Setup(dev, case, rp, av=True, eps0=0.01)
    try:
        Ne, Nf, ngl = S.Ne, S.Nf, S.ngl
        fn, nn = Nf * ngl, Ne * ngl * ngl
        h, d = S.h, S.d
        qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
        O.call("collect_edge_data_walls", S.p, Nf, h["EdgeData"], S.q, h["nx"], h["ny"], qL, qR)
        SGX, SGY = np.zeros(fn * 2), np.zeros(fn * 2)
        O.call("num_fluxes_gradient", S.p, Nf, h["EdgeData"], h["nx"], h["ny"], h["scal"], qL, qR, SGX, SGY)
        gX, gY = np.zeros(nn * 2), np.zeros(nn * 2)
        O.call("calc_gradient", S.p, Ne, h["Jac"], h["Yxi"], h["Yeta"], h["Xxi"], h["Xeta"], S.q, h["Dhat"], gX, gY)
        gX1, gY1 = gX.copy(), gY.copy()
        O.call("surface_gradient", S.p, Ne, h["Jac"], h["EMS"], h["EOR"], h["E2E"], SGX, SGY, gX1, gY1)
        gX2, gY2 = gX1.copy(), gY1.copy()
        O.call("scale_gradient", S.p, Ne, S.q, gX2, gY2)
        visc = np.linspace(0.0, 0.01, Ne)
        VL, VR = np.zeros(Nf), np.zeros(Nf)
        XL, XR, YL, YR = (np.zeros(fn * 2) for _ in range(4))
        O.call("collect_edge_data_gradient", S.p, Nf, h["EdgeData"], gX2, gY2, visc, VL, VR, XL, XR, YL, YR)
        Qt0 = np.random.default_rng(3).standard_normal(nn * 3)
        Qt = Qt0.copy()
        O.call("volume_viscose", S.p, Ne, h["Jac"], h["Yxi"], h["Yeta"], h["Xxi"], h["Xeta"], gX2, gY2,
               h["Dstrong"], visc, Qt)
        SPV = np.zeros(fn * 2)
        O.call("num_fluxes_viscose", S.p, Nf, h["EdgeData"], h["nx"], h["ny"], h["scal"], VL, VR, XL, XR, YL, YR,
               SPV)
        Qt2 = Qt.copy()
        O.call("surface_visc", S.p, Ne, h["Jac"], h["EMS"], h["EOR"], h["E2E"], SPV, Qt2)

        c = S.ctx
        dqL, dqR = c.array(qL), c.array(qR)
        dSGX, dSGY = c.empty(fn * 2), c.empty(fn * 2)
        c.calc_num_fluxes_gradient(Nf, d["EdgeData"], d["nx"], d["ny"], d["scal"], dqL, dqR, dSGX, dSGY)
        assert max_rel(dSGX.download(), SGX) <= 4e-16 and max_rel(dSGY.download(), SGY) <= 4e-16
        dgX, dgY = c.empty(nn * 2), c.empty(nn * 2)
        c.calc_gradient(Ne, d["Jac"], d["Yxi"], d["Yeta"], d["Xxi"], d["Xeta"], c.array(S.q), d["B"], d["Dhat"],
                        dgX, dgY)
        assert rel_l2(dgX.download(), gX) <= 1e-13 and rel_l2(dgY.download(), gY) <= 1e-13
        dgX, dgY = c.array(gX), c.array(gY)
        c.surface_kernel_gradient(Ne, d["Jac"], d["EMS"], d["EOR"], d["E2E"], c.array(SGX), c.array(SGY), dgX, dgY)
        assert max_rel(dgX.download(), gX1) <= 1e-15 and max_rel(dgY.download(), gY1) <= 1e-15
        dgX, dgY = c.array(gX1), c.array(gY1)
        c.scale_gradient(Ne, c.array(S.q), dgX, dgY)
        assert np.array_equal(dgX.download(), gX2) and np.array_equal(dgY.download(), gY2)
        dv = c.array(visc)
        dVL, dVR = c.empty(Nf), c.empty(Nf)
        dXL, dXR, dYL, dYR = (c.empty(fn * 2) for _ in range(4))
        c.collect_edge_data_gradient(Nf, d["EdgeData"], c.array(gX2), c.array(gY2), dv, dVL, dVR, dXL, dXR, dYL, dYR)
        for a, b in ((dVL, VL), (dVR, VR), (dXL, XL), (dXR, XR), (dYL, YL), (dYR, YR)):
            assert np.array_equal(a.download(), b)
        dQt = c.array(Qt0)
        c.volume_kernel_viscose(Ne, d["Jac"], d["Yxi"], d["Yeta"], d["Xxi"], d["Xeta"], c.array(gX2), c.array(gY2),
                                d["Dstrong"], dv, dQt)
        assert rel_l2(dQt.download(), Qt) <= 1e-14
        dSPV = c.empty(fn * 2)
        c.calc_num_fluxes_viscose(Nf, d["EdgeData"], d["nx"], d["ny"], d["scal"], c.array(VL), c.array(VR),
                                  c.array(XL), c.array(XR), c.array(YL), c.array(YR), dSPV)
        assert max_rel(dSPV.download(), SPV) <= 1e-15
        dQt = c.array(Qt)
        c.surface_kernel_visc(Ne, d["Jac"], d["EMS"], d["EOR"], d["E2E"], c.array(SPV), dQt)
        assert max_rel(dQt.download(), Qt2) <= 1e-15
    finally:
        S.close()
