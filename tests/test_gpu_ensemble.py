"""GPU tests of the batched ensemble (include/esdg_ensemble.h).

Every scenario of an ensemble must come out bit for bit as the same scenario run alone by a
solver (single-kernel schedule, itself bitwise equal to the fused and reference schedules), and
within 1e-12 of the oracle; the task ordering, launch shape, graph replay and autotuning must not
change a bit.
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, RunParams
from esdgsem_mpiocca_amd.scenarios import uplift_state
from helpers import rel_l2
from oracle import binding as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


def members():
    """Mixed degrees and sizes, two scenarios sharing one mesh, walls + positivity, a short T."""
    ocean = RunParams(num_flux=3, bc_kind=2, h_zero=4000.0)
    m4 = Case.ocean_uplift(N=4, nel=6, amp=1.0)
    m7 = Case.ocean_uplift(N=7, nel=5, amp=4.0, axis_a=90e3, axis_c=60e3, angle=1.1)
    m5 = Case.ocean_uplift(N=5, nel=7, amp=2.5, axis_a=70e3, axis_c=30e3, angle=2.0)
    fl = Case.flood(N=4, nel=6, seed=3)
    out = [
        (m4, ocean, None),
        (m7, RunParams(num_flux=0, bc_kind=2, h_zero=4000.0), None),
        (m4, ocean, uplift_state(m4, 4.5, 30e3, 90e3, 0.3)),       # shares m4's mesh arrays
        (fl, RunParams(num_flux=3, positivity=1, pos_pres_tol=1e-3, CFL=0.25), None),
        (m5, ocean, None),
        (m5, RunParams(num_flux=3, bc_kind=1, CFL=0.4), uplift_state(m5, 1.0, 50e3, 50e3, 0.0)),
        # larger meshes so that persistent grids of one or two blocks per CU loop over tasks
        (Case.ocean_uplift(N=4, nel=64, amp=3.0), ocean, None),
        (Case.ocean_uplift(N=7, nel=48, amp=2.0, angle=0.7), ocean, None),
    ]
    # scenario 4 stops early: T between its 5th and 6th step
    t5 = alone(m5, ocean, None, 5)[2]
    t6 = alone(m5, ocean, None, 6)[2]
    short = RunParams(**{**ocean.__dict__, "T": 0.5 * (t5 + t6)})
    out[4] = (m5, short, None)
    return out


def alone(case, rp, q, steps, fused=2):
    import esdgsem_mpiocca_amd as E
    d = E.device()
    ctx = d.Context(d.make_params(case, rp))
    try:
        s = d.Solver(ctx, case, rp, fused=fused)
        if q is not None:
            s.set_q(q)
        n, t = s.run(steps)
        out = s.get_q()
        s.close()
    finally:
        ctx.close()
    return out, n, t


def run_ensemble(dev, mem, steps, schedule=None, autotune=False):
    case0, rp0, _ = mem[0]
    ctx = dev.Context(dev.make_params(case0, rp0))
    try:
        ens = dev.Ensemble(ctx, mem)
        if schedule:
            ens.set_schedule(**schedule)
        if autotune:
            ens.autotune(trial_steps=2)
        ens.run(steps)
        res = [(ens.get_q(i),) + ens.status(i)[:2] for i in range(len(mem))]
        ens.close()
    finally:
        ctx.close()
    return res


STEPS = 12


@pytest.fixture(scope="module")
def mem():
    return members()


@pytest.fixture(scope="module")
def ensemble_default(dev, mem):
    return run_ensemble(dev, mem, STEPS)


def test_ensemble_equals_scenarios_alone(dev, mem, ensemble_default):
    for i, (case, rp, q) in enumerate(mem):
        qe, te, ne = ensemble_default[i]
        qa, na, ta = alone(case, rp, q, STEPS)
        assert (ne, te) == (na, ta), i
        assert np.array_equal(qe, qa), i
    # the short scenario stopped at its T
    assert ensemble_default[4][2] < STEPS and ensemble_default[4][1] == mem[4][1].T


def test_ensemble_matches_oracle(dev, mem, ensemble_default):
    for i in (1, 2, 3, 5):
        case, rp, q = mem[i]
        c2 = case
        if q is not None:
            c2 = case.with_state(q)
        qo, no, to, _ = O.timeloop(c2, rp, STEPS)
        qe, te, ne = ensemble_default[i]
        assert ne == no and abs(te - to) <= 1e-14 * to
        A, B = qe.reshape(case.Nelem, 3, -1), qo.reshape(case.Nelem, 3, -1)
        for v in range(3):
            nb = np.linalg.norm(B[:, v])
            err = rel_l2(A[:, v], B[:, v]) if nb > 0 else np.abs(A[:, v]).max()
            assert err <= 1e-12, (i, v, err)


@pytest.mark.parametrize("schedule", [
    dict(order=1, blocks_per_cu=0, per_degree=0, xcd=1, graph=1),
    dict(order=2, blocks_per_cu=1, per_degree=1, xcd=0, graph=0),
    dict(order=1, blocks_per_cu=1, per_degree=0, xcd=1, graph=1),
    dict(order=0, blocks_per_cu=2, per_degree=0, xcd=0, graph=1),
    dict(order=0, blocks_per_cu=1, per_degree=1, xcd=1, graph=1),
    dict(order=0, blocks_per_cu=0, per_degree=1, xcd=0, graph=1, tile=2),
    dict(order=1, blocks_per_cu=0, per_degree=0, xcd=1, graph=0, tile=4),
    dict(order=2, blocks_per_cu=1, per_degree=0, xcd=0, graph=1, tile=4),
], ids=["tile_xcd", "degree_persistent_perdeg_eager", "tile_persistent_xcd", "persistent2",
        "persistent_perdeg_xcd", "elemtile2_perdeg", "elemtile4_interleaved_xcd_eager",
        "elemtile4_persistent"])
def test_schedules_do_not_change_results(dev, mem, ensemble_default, schedule):
    res = run_ensemble(dev, mem, STEPS, schedule=schedule)
    for i in range(len(mem)):
        assert res[i][1:] == ensemble_default[i][1:], i
        assert np.array_equal(res[i][0], ensemble_default[i][0]), i


def test_autotune_keeps_state(dev, mem, ensemble_default):
    res = run_ensemble(dev, mem, STEPS, autotune=True)
    for i in range(len(mem)):
        assert res[i][1:] == ensemble_default[i][1:], i
        assert np.array_equal(res[i][0], ensemble_default[i][0]), i
