"""GPU tests of the learned troubled-cell indicator (include/esdg_mlp.h) against the oracle
(oracle/mlp_oracle.py) and of its use in the solver's artificial-viscosity path.

Tolerances: features are FP32 values of FP64 sums taken in another order (1e-6 absolute).  The
network rounds activations to BF16 after FP32 accumulation; when an accumulated value sits within
an FP32 rounding of a BF16 tie, device and oracle may round it to neighbouring BF16 values, so
outputs are compared by quantiles (99.9% of rows within 1e-3, all within 2e-2).  Gradients are FP32
sums over the batch: within 2e-3 of each block's largest entry.
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, RunParams
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


def av_params():
    return RunParams(num_flux=3, artificial_viscosity=1, eps0=0.005, sigma_min=-4.5, sigma_max=-2.5, DFL=0.1)


@pytest.fixture(scope="module")
def state(dev):
    """A dam-break state after a few steps on a 10x10 N=7 curved mesh (fronts with viscosity)."""
    case = Case.dam_break(N=7, nel=10, noise=1e-3, seed=2, ratio=3.0, angle=0.5)
    rp = av_params()
    ctx = dev.Context(dev.make_params(case, rp))
    s = dev.Solver(ctx, case, rp, fused=1)
    s.run(6)
    q = s.get_q()
    s.close()
    ctx.close()
    return case, rp, q


def random_params(seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-0.3, 0.3, M.NIN * M.NH), rng.uniform(-0.1, 0.1, M.NH),
                           rng.uniform(-0.3, 0.3, M.NH * M.NH), rng.uniform(-0.1, 0.1, M.NH),
                           rng.uniform(-0.5, 0.5, M.NH), [0.05]]).astype(np.float32)


def test_features_match_oracle(dev, state):
    case, rp, q = state
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        dq, dB, dV = ctx.array(q), ctx.array(case["b"]), ctx.array(case["VdmInv"])
        df = ctx.empty(case.Nelem * 72, np.float32)
        dev.mlp_features(ctx, case.Nelem, dq, dB, dV, df)
        F = df.download().reshape(case.Nelem, 72)
    finally:
        ctx.close()
    Fo = M.features(q, case["b"], case["VdmInv"], case.Nelem)
    assert np.abs(F - Fo).max() <= 1e-6, np.abs(F - Fo).max()
    assert (Fo[:, 64:] > -2).any() and (Fo[:, :64] != 0).any()


def test_inference_matches_oracle(dev, state):
    """BF16 MFMA forward (also pins the MFMA operand/result lane layout)."""
    case, rp, q = state
    w = random_params(5)
    Fo = M.features(q, case["b"], case["VdmInv"], case.Nelem).astype(np.float32)
    rng = np.random.default_rng(6)
    Fr = np.concatenate([rng.uniform(-1, 1, (333, 64)), rng.uniform(-2, 0, (333, 8))], axis=1).astype(np.float32)
    X = np.concatenate([Fo, Fr])
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        net = dev.MLP(ctx, params=w)
        dX = ctx.array(X)
        dp = ctx.empty(len(X), np.float32)
        net.infer(dX, len(X), dp)
        p = dp.download()
        # fused features + inference equals features followed by inference
        dq, dB, dV = ctx.array(q), ctx.array(case["b"]), ctx.array(case["VdmInv"])
        dpf = ctx.empty(case.Nelem, np.float32)
        net.indicator(case.Nelem, dq, dB, dV, dpf)
        pf = dpf.download()
        df = ctx.empty(case.Nelem * 72, np.float32)
        dev.mlp_features(ctx, case.Nelem, dq, dB, dV, df)
        dp2 = ctx.empty(case.Nelem, np.float32)
        net.infer(df, case.Nelem, dp2)
        p2 = dp2.download()
        net.close()
    finally:
        ctx.close()
    po = M.forward(w.astype(np.float64), X.astype(np.float64))[-1]
    d = np.abs(p - po)
    assert np.quantile(d, 0.999) <= 1e-3 and d.max() <= 2e-2, (np.quantile(d, 0.999), d.max())
    assert np.array_equal(pf, p2)


def test_loss_gradient_matches_oracle(dev, state):
    case, rp, q = state
    w = random_params(7)
    rng = np.random.default_rng(8)
    X = np.concatenate([rng.uniform(-1, 1, (1000, 64)), rng.uniform(-2, 0, (1000, 8))], axis=1).astype(np.float32)
    y = (rng.uniform(size=1000) < 0.25).astype(np.float32)
    idx = rng.permutation(1000).astype(np.int32)[:700]
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        net = dev.MLP(ctx, params=w)
        dX, dy, di = ctx.array(X), ctx.array(y), ctx.array(idx)
        dg, dl = ctx.empty(M.NPARAM, np.float32), ctx.empty(1, np.float32)
        net.loss_grad(700, dX, dy, di, 4.0, dg, dl)
        g, L = dg.download(), dl.download()[0]
        net.close()
    finally:
        ctx.close()
    Lo, go = M.loss_grad(w.astype(np.float64), X[idx].astype(np.float64), y[idx].astype(np.float64), 4.0)
    assert abs(L - Lo) <= 1e-5 * abs(Lo), (L, Lo)
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)):
        scale = np.abs(go[a:b]).max()
        assert np.abs(g[a:b] - go[a:b]).max() <= 2e-3 * scale + 1e-9, (a, np.abs(g[a:b] - go[a:b]).max(), scale)


def test_training_learns_a_separable_rule(dev, state):
    """Adam on the device's own forward/backward kernels fits a linear rule on two modal features."""
    case, rp, q = state
    rng = np.random.default_rng(9)
    n = 20000
    X = np.concatenate([rng.uniform(-1, 1, (n, 64)), rng.uniform(-2, 0, (n, 8))], axis=1).astype(np.float32)
    y = (X[:, 71] + 0.5 * X[:, 70] > -1.2).astype(np.float32)
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        net = dev.MLP(ctx, seed=1)
        dX, dy = ctx.array(X), ctx.array(y)
        losses = []
        for it in range(600):
            idx = rng.integers(0, n, 1024).astype(np.int32)
            di = ctx.array(idx)
            losses.append(net.train_step(1024, dX, dy, di, 1.0, 5e-3, want_loss=(it % 50 == 0 or it == 599)))
            di.free()
        dp = ctx.empty(n, np.float32)
        net.infer(dX, n, dp)
        p = dp.download()
        net.close()
    finally:
        ctx.close()
    acc = np.mean((p > 0.5) == (y > 0.5))
    assert losses[-1] < 0.25 * losses[0], (losses[0], losses[-1])
    assert acc > 0.95, acc


def run_av(dev, case, rp, steps, indicator=None, lo=0.5, hi=0.5, audit=False, params=None):
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        s = dev.Solver(ctx, case, rp, fused=1)
        net = None
        if indicator is not None:
            net = dev.MLP(ctx, seed=3, params=params)
            s.set_indicator(net, lo, hi)
            s.indicator_audit(audit)
        s.run(steps)
        q = s.get_q()
        counts = s.indicator_counts() if indicator is not None else None
        s.close()
        if net:
            net.close()
    finally:
        ctx.close()
    return q, counts


def test_indicator_bands_reduce_to_known_schedules(dev):
    """p >= hi everywhere (troubled) or inside the band everywhere (the sensor decides) gives the
    modal-sensor result; p <= lo everywhere (smooth) gives the result without viscosity."""
    case = Case.dam_break(N=7, nel=8, noise=1e-3, seed=4, ratio=2.5, angle=0.3)
    rp = av_params()
    plain, _ = run_av(dev, case, rp, 8)
    troubled, c1 = run_av(dev, case, rp, 8, indicator=True, lo=-1.0, hi=-1.0)
    unsure, _ = run_av(dev, case, rp, 8, indicator=True, lo=-1.0, hi=2.0)
    assert np.array_equal(plain, troubled) and np.array_equal(plain, unsure)
    assert c1["viscous"] > 0 and c1["element_stages"] == 8 * 5 * case.Nelem
    assert c1["not_smooth"] == c1["element_stages"]
    no_av_rp = RunParams(**{**rp.__dict__, "artificial_viscosity": 0})
    no_av, _ = run_av(dev, case, no_av_rp, 8)
    smooth, c0 = run_av(dev, case, rp, 8, indicator=True, lo=2.0, hi=2.0, audit=True)
    assert np.array_equal(no_av, smooth)
    assert c0["viscous"] == 0 and c0["false_negatives"] == c0["sensor_positives"] > 0
    assert c0["not_smooth"] == 0
