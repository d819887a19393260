"""GPU parity of the benchmarked workloads against the oracle, at their own degree and physics and,
for the curved dam break, at its full size.

- The wet/dry flood of the 2048^2 run (N=7, 16-mode Fourier bathymetry, positivity limiter, modal
  sensor + artificial viscosity, CFL 0.1, dry tolerance 0.1) on a 12^2 mesh: both schedules against
  the oracle's DGtimeloop (src/deviceclass.cpp:897-1130, okl/Positivity/PosPres.okl:17-223).
- The learned indicator under the limiter: band limits that make every element troubled, or leave
  every element to the sensor, equal the plain method bit for bit.
- The fused schedule with its marked-element list launch on a mesh where that list spans many
  workgroups, bit for bit against the reference kernel sequence.
- The 512^2 curved dam break (N=7, AV) of the headline measurement, fused schedule with the
  device clock, three steps against the oracle.

Acceptance bar: relative L2 <= 1e-12 per field against the oracle.
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, indicator
from helpers import rel_l2
from oracle import binding as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


def flood_params(**kw):
    return indicator.av_run_params(**{"positivity": 1, "pos_pres_tol": 0.1, "CFL": 0.1, **kw})


def fields_rel_l2(case, a, b):
    A = a.reshape(case.Nelem, 3, -1)
    B = b.reshape(case.Nelem, 3, -1)
    return [rel_l2(A[:, v], B[:, v]) for v in range(3)]


def run(dev, case, rp, steps, fused, net_params=None, lo=0.5, hi=0.5, want=()):
    ctx = dev.Context(dev.make_params(case, rp))
    extra = {}
    try:
        s = dev.Solver(ctx, case, rp, fused=fused)
        net = None
        if net_params is not None:
            net = dev.MLP(ctx, seed=3) if isinstance(net_params, str) else dev.MLP(ctx, params=net_params)
            s.set_indicator(net, lo, hi)
        n, t = s.run(steps)
        q = s.get_q()
        for name in want:
            extra[name] = s.view(name, case.Nelem).download()
        if net:
            extra["counts"] = s.indicator_counts()
        s.close()
        if net:
            net.close()
    finally:
        ctx.close()
    return q, n, t, extra


@pytest.fixture(scope="module")
def flood12():
    case = Case.flood(N=7, nel=12, seed=0)
    rp = flood_params()
    qo, no, to, _ = O.timeloop(case, rp, 20)
    return case, rp, qo, no, to


@pytest.mark.parametrize("fused", [0, 1], ids=["reference_sequence", "fused"])
def test_flood_n7_av_limiter_matches_oracle(dev, flood12, fused):
    """The 2048^2 run's physics on a 12^2 mesh: wet and dry nodes, viscous elements, the limiter
    needed (without it the oracle's run reaches NaN), 20 steps against the oracle."""
    case, rp, qo, no, to = flood12
    qg, ng, tg, ex = run(dev, case, rp, 20, fused, want=("ViscPara",))
    assert ng == no == 20
    assert abs(tg - to) <= 1e-12 * to
    errs = fields_rel_l2(case, qg, qo)
    assert max(errs) <= 1e-12, errs
    Q = qg.reshape(case.Nelem, 3, -1)
    assert np.isfinite(Q).all() and (Q[:, 0] >= 0).all()
    assert (Q[:, 0] < rp.pos_pres_tol).sum() > 0.2 * Q[:, 0].size        # dry land remains
    assert (ex["ViscPara"] > 0).sum() >= 5                                # viscous elements


def test_flood_n7_without_limiter_fails_in_oracle(flood12):
    """The case exercises the limiter: the same run without it breaks down."""
    case, _, _, _, _ = flood12
    q, _, _, _ = O.timeloop(case, flood_params(positivity=0), 20)
    assert not np.isfinite(q).all() or (q.reshape(case.Nelem, 3, -1)[:, 0] < 0).any()


def test_learned_indicator_bands_under_limiter(dev, flood12):
    """The learned indicator with the limiter on: every element classed troubled (p >= hi) or
    every element left to the sensor (lo < p < hi) equals the plain method bit for bit; every
    element classed smooth equals the run without artificial viscosity (still limited)."""
    case, rp, qo, _, _ = flood12
    plain, _, tp, _ = run(dev, case, rp, 12, 1)
    troubled, _, tt, ex1 = run(dev, case, rp, 12, 1, net_params="seeded", lo=-1.0, hi=-1.0)
    unsure, _, tu, _ = run(dev, case, rp, 12, 1, net_params="seeded", lo=-1.0, hi=2.0)
    assert np.array_equal(plain, troubled) and np.array_equal(plain, unsure)
    assert tp == tt == tu
    c = ex1["counts"]
    assert c["viscous"] > 0 and c["not_smooth"] == c["element_stages"] == 12 * 5 * case.Nelem
    no_av, _, _, _ = run(dev, case, flood_params(artificial_viscosity=0), 12, 1)
    smooth, _, _, ex0 = run(dev, case, rp, 12, 1, net_params="seeded", lo=2.0, hi=2.0)
    assert np.array_equal(no_av, smooth)
    assert ex0["counts"]["viscous"] == 0
    qo12, _, _, _ = O.timeloop(case, rp, 12)
    assert max(fields_rel_l2(case, plain, qo12)) <= 1e-12


def test_learned_indicator_trained_under_limiter_stays_close(dev, flood12):
    """The trained network (with the sensor as fallback) under the limiter on the flood: a
    positive, finite state whose change against the plain method is of the size of the viscosity
    it withholds or adds."""
    params, _ = indicator.load_weights()
    if params is None:
        pytest.skip("no trained indicator weights")
    case, rp, _, _, _ = flood12
    plain, _, _, _ = run(dev, case, rp, 20, 1)
    learned, n, _, ex = run(dev, case, rp, 20, 1, net_params=params, lo=0.1, hi=0.9)
    assert n == 20
    Q = learned.reshape(case.Nelem, 3, -1)
    assert np.isfinite(Q).all() and (Q[:, 0] >= 0).all()
    assert rel_l2(Q[:, 0], plain.reshape(case.Nelem, 3, -1)[:, 0]) <= 1e-2
    assert ex["counts"]["element_stages"] == 20 * 5 * case.Nelem


def test_fused_list_launch_spans_many_workgroups_bitwise(dev):
    """A 96^2 N=7 dam break with a step across x = 5: the sensor gives about 200 elements along
    the front viscosity, and the face pass lists them with their neighbours, so the list launch on
    the second stream runs across many workgroups beside the unmarked launch, while most elements
    stay unmarked.  Five steps of the fused schedule equal the reference kernel sequence bit for
    bit."""
    case = Case.dam_break(N=7, nel=96, noise=1e-3, seed=7)
    rp = indicator.av_run_params()
    qa, na, ta, exa = run(dev, case, rp, 5, 0, want=("ViscPara",))
    qb, nb, tb, _ = run(dev, case, rp, 5, 1)
    marked = int((exa["ViscPara"] > 0).sum())
    assert 150 <= marked <= case.Nelem // 4, marked
    assert na == nb == 5 and ta == tb
    assert np.array_equal(qa, qb)


def test_c1_full_size_fused_matches_oracle(dev):
    """The headline workload at its full size: 512^2 curved dam break, N=7, modal sensor + AV,
    fused schedule (element list, second stream, device clock): three steps against the oracle."""
    case = Case.dam_break(N=7, nel=512, noise=1e-3, seed=0)
    rp = indicator.av_run_params()
    qo, no, to, _ = O.timeloop(case, rp, 3)
    qg, ng, tg, ex = run(dev, case, rp, 3, 1, want=("ViscPara",))
    assert ng == no == 3
    assert abs(tg - to) <= 1e-14 * to
    errs = fields_rel_l2(case, qg, qo)
    assert max(errs) <= 1e-12, errs
    assert (ex["ViscPara"] > 0).sum() > 64
