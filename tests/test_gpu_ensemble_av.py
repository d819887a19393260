"""GPU tests of artificial viscosity in the batched ensemble (include/esdg_ensemble.h).

A scenario with artificial viscosity must come out of an ensemble bit for bit as the same scenario
run alone by a solver with the fused schedule (the schedule that runs the AV path, itself bitwise
equal to the reference kernel sequence), next to scenarios without viscosity, of other degrees,
with the limiter, with an early end time and sharing a mesh; likewise with the learned indicator
for the N = 7 members, whose counts must add up to the solo runs' counts.  Task order, launch shape,
graph replay and autotuning must not change a bit.
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, RunParams
from esdgsem_mpiocca_amd.indicator import dam_break_state

pytestmark = pytest.mark.gpu

STEPS = 10


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


def av(**kw):
    base = dict(num_flux=3, artificial_viscosity=1, eps0=0.005, sigma_min=-4.5, sigma_max=-2.5, DFL=0.1)
    return RunParams(**{**base, **kw})


def alone(dev, case, rp, q, steps, net_seed=None, lo=0.5, hi=0.5, fused=None, plot=False):
    ctx = dev.Context(dev.make_params(case, rp))
    if fused is None:
        fused = 1 if rp.artificial_viscosity else 2
    try:
        s = dev.Solver(ctx, case, rp, fused=fused)
        net = None
        if net_seed is not None:
            net = dev.MLP(ctx, seed=net_seed)
            s.set_indicator(net, lo, hi)
            s.indicator_audit(True)
        if q is not None:
            s.set_q(q)
        n, t = s.run(steps)
        out = s.get_q()
        counts = s.indicator_counts() if net else None
        if plot:
            counts = s.view("ViscForPlot", case.Nelem).download()
        s.close()
        if net:
            net.close()
    finally:
        ctx.close()
    return out, n, t, counts


def run_ensemble(dev, mem, steps, schedule=None, autotune=False, net_seed=None, lo=0.5, hi=0.5):
    case0, rp0, _ = mem[0]
    ctx = dev.Context(dev.make_params(case0, rp0))
    try:
        ens = dev.Ensemble(ctx, mem)
        net = None
        if net_seed is not None:
            net = dev.MLP(ctx, seed=net_seed)
            ens.set_indicator(net, lo, hi)
            ens.indicator_audit(True)
        if schedule:
            ens.set_schedule(**schedule)
        if autotune:
            ens.autotune(trial_steps=2)
        ens.run(steps)
        res = [(ens.get_q(i),) + ens.status(i)[:2] for i in range(len(mem))]
        counts = ens.indicator_counts() if net else None
        ens.close()
        if net:
            net.close()
    finally:
        ctx.close()
    return res, counts


def members(dev):
    d7 = Case.dam_break(N=7, nel=8, noise=1e-3, seed=4, ratio=2.5, angle=0.3)
    d4 = Case.dam_break(N=4, nel=12, noise=1e-3, seed=5, ratio=3.0, angle=1.2)
    d3 = Case.dam_break(N=3, nel=10, noise=1e-3, seed=6, ratio=2.0, angle=0.0)
    fl = Case.flood(N=4, nel=8, seed=3)
    m5 = Case.ocean_uplift(N=5, nel=6, amp=2.5, axis_a=70e3, axis_c=30e3, angle=2.0)
    rng = np.random.default_rng(11)
    q7b = dam_break_state(d7, 3.0, 2.0, 1e-3, rng)
    out = [
        (d7, av(), None),
        (m5, RunParams(num_flux=3, bc_kind=2, h_zero=4000.0), None),            # no viscosity
        (d4, av(num_flux=0, DFL=0.05), None),
        (d7, av(), q7b),                                                          # shares d7's mesh
        (fl, av(positivity=1, pos_pres_tol=1e-3, CFL=0.25), None),               # limiter + AV
        (d3, av(), None),
    ]
    # the N = 3 scenario stops early: T between its 4th and 5th step
    t4 = alone(dev, d3, av(), None, 4)[2]
    t5 = alone(dev, d3, av(), None, 5)[2]
    out[5] = (d3, av(T=0.5 * (t4 + t5)), None)
    return out


@pytest.fixture(scope="module")
def mem(dev):
    return members(dev)


@pytest.fixture(scope="module")
def ensemble_default(dev, mem):
    return run_ensemble(dev, mem, STEPS)[0]


def test_av_scenarios_equal_solo_fused_runs(dev, mem, ensemble_default):
    for i, (case, rp, q) in enumerate(mem):
        qe, te, ne = ensemble_default[i]
        qa, na, ta, _ = alone(dev, case, rp, q, STEPS)
        assert (ne, te) == (na, ta), i
        assert np.array_equal(qe, qa), i
    assert ensemble_default[5][2] < STEPS and ensemble_default[5][1] == mem[5][1].T


def test_viscosity_changes_the_result(dev, mem, ensemble_default):
    """The AV work is not a no-op on these members: without it the states differ."""
    plain = [(c, RunParams(**{**rp.__dict__, "artificial_viscosity": 0}), q) for c, rp, q in mem]
    res, _ = run_ensemble(dev, plain, STEPS)
    for i in (0, 2, 3):
        assert not np.array_equal(res[i][0], ensemble_default[i][0]), i


@pytest.mark.parametrize("schedule", [
    dict(order=1, blocks_per_cu=0, per_degree=0, xcd=1, graph=1),
    dict(order=2, blocks_per_cu=1, per_degree=1, xcd=0, graph=0),
    dict(order=0, blocks_per_cu=2, per_degree=0, xcd=0, graph=1),
    dict(order=1, blocks_per_cu=0, per_degree=1, xcd=1, graph=1, tile=2),
    dict(order=0, blocks_per_cu=0, per_degree=0, xcd=0, graph=0, tile=4),
], ids=["tile_xcd", "degree_persistent_perdeg_eager", "persistent2", "elemtile2_interleaved_xcd",
        "elemtile4_eager"])
def test_av_schedules_do_not_change_results(dev, mem, ensemble_default, schedule):
    res, _ = run_ensemble(dev, mem, STEPS, schedule=schedule)
    for i in range(len(mem)):
        assert res[i][1:] == ensemble_default[i][1:], i
        assert np.array_equal(res[i][0], ensemble_default[i][0]), i


def test_av_autotune_keeps_state(dev, mem, ensemble_default):
    res, _ = run_ensemble(dev, mem, STEPS, autotune=True)
    for i in range(len(mem)):
        assert res[i][1:] == ensemble_default[i][1:], i
        assert np.array_equal(res[i][0], ensemble_default[i][0]), i


@pytest.mark.parametrize("lo,hi", [(0.3, 0.7), (-1.0, -1.0), (2.0, 2.0)], ids=["band", "troubled", "smooth"])
def test_learned_members_equal_solo_runs(dev, mem, ensemble_default, lo, hi):
    """The learned indicator on the N = 7 members (0 and 3), the modal sensor on the others; the
    indicator counts of the ensemble are the sums of the solo runs' counts."""
    res, counts = run_ensemble(dev, mem, STEPS, net_seed=3, lo=lo, hi=hi)
    total = None
    for i, (case, rp, q) in enumerate(mem):
        learned = rp.artificial_viscosity and case.ngl == 8
        qa, na, ta, ca = alone(dev, case, rp, q, STEPS, net_seed=3 if learned else None, lo=lo, hi=hi)
        assert res[i][1:] == (ta, na), i
        assert np.array_equal(res[i][0], qa), i
        if learned:
            total = ca if total is None else {k: total[k] + ca[k] for k in total}
    assert counts == total
    assert counts["element_stages"] == 2 * STEPS * 5 * mem[0][0].Nelem
    if hi < 0.0:
        # every element troubled: the modal sensor's result
        for i in (0, 3):
            assert np.array_equal(res[i][0], ensemble_default[i][0]), i
        assert counts["viscous"] > 0
    if lo > 1.0:
        assert counts["viscous"] == 0 and counts["false_negatives"] == counts["sensor_positives"] > 0


def test_learned_autotune_keeps_counts(dev, mem):
    ref, c_ref = run_ensemble(dev, mem, STEPS, net_seed=3, lo=0.3, hi=0.7)
    res, c = run_ensemble(dev, mem, STEPS, net_seed=3, lo=0.3, hi=0.7, autotune=True)
    assert c == c_ref
    for i in range(len(mem)):
        assert res[i][1:] == ref[i][1:], i
        assert np.array_equal(res[i][0], ref[i][0]), i


@pytest.mark.parametrize("learned", [False, True], ids=["sensor", "learned"])
def test_single_kernel_schedule_with_av_equals_fused(dev, mem, learned):
    """The solver's single-kernel schedule (fused = 2, a batch of one) with artificial viscosity
    equals the fused schedule bit for bit, with the learned indicator and its counts too; without
    the indicator it leaves the final state's viscosity in ViscForPlot, as the fused schedule does."""
    for i in (0, 2, 4, 5):
        case, rp, q = mem[i]
        seed = 3 if learned and case.ngl == 8 else None
        if learned and seed is None:
            continue
        qa, na, ta, ca = alone(dev, case, rp, q, STEPS, net_seed=seed, lo=0.3, hi=0.7, fused=1)
        qb, nb, tb, cb = alone(dev, case, rp, q, STEPS, net_seed=seed, lo=0.3, hi=0.7, fused=2)
        assert (na, ta) == (nb, tb), i
        assert np.array_equal(qa, qb), i
        if learned:
            assert ca == cb, i
        else:
            pa = alone(dev, case, rp, q, STEPS, fused=1, plot=True)[3]
            pb = alone(dev, case, rp, q, STEPS, fused=2, plot=True)[3]
            assert np.array_equal(pa, pb), i
