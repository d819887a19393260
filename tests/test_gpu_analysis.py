"""GPU parity of the analysis kernels and solution integrals against the oracle.

- ArrivalTimes / MaximumElevation (okl/Analysis): element-wise, bit-exact against the oracle on
  the same inputs, and inside the time loop (calcArrivalTimes / calcMaximumElevation,
  src/deviceclass.cpp:1132-1141) equal to the oracle applied after every step to the states of a
  step-by-step run.
- Total mass / entropy (src/basis.cpp:1004-1111): the device reduction sums in another order than
  the reference's serial loop, so they agree to rounding: relative 1e-13.
"""
import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, RunParams
from esdgsem_mpiocca_amd import setup_api as S
from oracle import binding as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


def flood_case():
    return Case.flood(N=4, nel=12, seed=3)


def test_analysis_kernels_match_oracle(dev):
    case = flood_case()
    # h_zero = the reservoir level: the still reservoir has not arrived, disturbed nodes have
    rp = RunParams(h_zero=0.8, positivity=1, pos_pres_tol=1e-2)
    p = O.params(case, rp)
    E, n2 = case.Nelem, case.ngl ** 2
    rng = np.random.default_rng(0)
    q = case["q"].copy().reshape(E, 3, n2)
    q[:, 0] += np.where(rng.uniform(size=(E, n2)) < 0.2, rng.uniform(-1e-3, 1e-3, size=(E, n2)), 0.0)
    q = q.ravel()
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        dq, dB = ctx.array(q), ctx.array(case["b"])
        arr = np.full(E * n2, 1000.0)
        arr[::7] = 0.125                                  # already arrived: kept
        darr, dmx = ctx.array(arr), ctx.array(np.zeros(E * n2))
        ctx.arrival_times(E, 0.5, dq, dB, darr)
        ctx.maximum_elevation(E, 0.5, dq, dB, dmx)
        O.call("arrival_times", p, E, 0.5, q, case["b"], arr)
        mx = np.zeros(E * n2)
        O.call("maximum_elevation", p, E, q, case["b"], mx)
        assert np.array_equal(darr.download(), arr)
        assert np.array_equal(dmx.download(), mx)
        assert (arr == 0.5).any() and (arr == 1000.0).any()
        dJ, dw = ctx.array(case["Jac"]), ctx.array(case["w_GL"])
        m, s = np.zeros(1), np.zeros(1)
        ctx.total_mass_dev(E, dq, dJ, dw, m)
        ctx.total_entropy_dev(E, dq, dB, dJ, dw, s)
    finally:
        ctx.close()
    mo, so = np.zeros(1), np.zeros(1)
    O.call("total_mass", p, E, q, case["Jac"], case["w_GL"], mo)
    O.call("total_entropy", p, E, q, case["b"], case["Jac"], case["w_GL"], so)
    assert abs(m[0] - mo[0]) <= 1e-13 * abs(mo[0]), (m, mo)
    assert abs(s[0] - so[0]) <= 1e-13 * abs(so[0]), (s, so)
    assert mo[0] == S.total_mass(case, q) and so[0] == S.total_entropy(case, q, rp.g)


@pytest.mark.parametrize("fused", [0, 1])
def test_analysis_inside_the_time_loop(dev, fused):
    """Arrival times and maximum elevations after 12 steps equal the oracle applied after every
    step to the states of the same run taken one step at a time."""
    case = flood_case()
    base = dict(h_zero=0.8, positivity=1, pos_pres_tol=1e-2, CFL=0.1, num_flux=3,
                artificial_viscosity=1, eps0=0.005, sigma_min=-4.5, sigma_max=-2.5)
    rp_on = RunParams(**base, arrival_times=1, maximum_elevation=1)
    rp_off = RunParams(**base)
    steps = 12
    ctx = dev.Context(dev.make_params(case, rp_on))
    try:
        s = dev.Solver(ctx, case, rp_on, fused=fused)
        n, _ = s.run(steps)
        assert n == steps
        E, n2 = case.Nelem, case.ngl ** 2
        arr = s.view("ArrivalTimings", E * n2).download()
        mx = s.view("MaximumElevation", E * n2).download()
        s.close()
        s = dev.Solver(ctx, case, rp_off, fused=fused)
        p = O.params(case, rp_on)
        ref_arr, ref_mx = np.full(E * n2, 1000.0), np.zeros(E * n2)
        for _ in range(steps):
            s.run(1)
            q = s.get_q()
            O.call("arrival_times", p, E, s.time(), q, case["b"], ref_arr)
            O.call("maximum_elevation", p, E, q, case["b"], ref_mx)
        s.close()
    finally:
        ctx.close()
    assert np.array_equal(arr, ref_arr)
    assert np.array_equal(mx, ref_mx)
    assert (arr < 1000.0).any() and (mx > 0.0).any()


def test_single_kernel_stage_rejects_analysis(dev):
    case = Case.gaussian_bump(N=4, nel=8)
    rp = RunParams(num_flux=0, maximum_elevation=1)
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        with pytest.raises(dev.EsdgDeviceError):
            dev.Solver(ctx, case, rp, fused=2)
    finally:
        ctx.close()


def test_gauge_samples(dev):
    """esdg_gauge_sample: h + b - h_zero at global node ids (the index rule of
    src/deviceclass.cpp:850), exact; the last node of the mesh included."""
    case = Case.ocean_uplift(N=5, nel=8, amp=3.0)
    rp = RunParams(h_zero=4000.0)
    E, n2 = case.Nelem, case.ngl ** 2
    rng = np.random.default_rng(2)
    ids = np.concatenate([rng.integers(0, E * n2, size=40), [0, E * n2 - 1]]).astype(np.int32)
    Q = case["q"].reshape(E, 3, n2)
    expect = Q[ids // n2, 0, ids % n2] + case["b"][ids] - 4000.0
    ctx = dev.Context(dev.make_params(case, rp))
    try:
        out = ctx.empty(len(ids))
        ctx.gauge_sample(len(ids), ctx.array(ids), ctx.array(case["q"]), ctx.array(case["b"]), out)
        assert np.array_equal(out.download(), expect)
        ctx.gauge_sample(0, None, None, None, None)
    finally:
        ctx.close()
