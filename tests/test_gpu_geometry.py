"""The geometry built on the device (include/esdg_device_setup.h) against the host setup
(esdg_setup.h, itself bit-exact against the reference's own Mesh/SW2D/basis sources).

- Each geometry kernel on the host case's own inputs gives the host's arrays bit for bit:
  metrics and 1/J (Mesh.cpp:1722-1742, main.cpp:700), face normals (Mesh.cpp:1744-1806, 282-307),
  element sizes (basis.cpp:939-966), Bx/By (SW2D.cpp:1048-1105) -- on the curved mesh family and
  on an ISM-V2 mesh with curved boundaries and reversed faces.
- A whole structured case built on the device equals the host's: connectivity, nodes, metrics,
  normals, sizes bit for bit; bathymetry and state (per-node sin/cos/exp, two math libraries)
  within a few ulps.
- A solver fed from the device case (device-to-device inputs) matches the oracle run on the same
  arrays, and the 2048^2 N=7 case of the largest run builds on the device in seconds.
"""
import os
import time

import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, RunParams, indicator
from esdgsem_mpiocca_amd import setup_api as S
from helpers import rel_l2
from oracle import binding as O

pytestmark = pytest.mark.gpu

ISM_MESH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes", "CircleInCircle.mesh")

EXACT = ("x", "y", "xXi", "xEta", "yXi", "yEta", "Jac", "nx", "ny", "scal", "EleSizes", "x_GL", "w_GL", "D",
         "D0", "Dstrong", "Dhat", "VdmInv", "EdgeData", "EdgeReversed", "ElemToEdge", "ElemEdgeMasterSlave",
         "ElemEdgeOrientation")


@pytest.fixture(scope="module")
def dev():
    import esdgsem_mpiocca_amd as E
    return E.device()


def specs():
    return {
        "dam_break_curved_N7": lambda b: Case.dam_break(N=7, nel=12, noise=1e-3, seed=3, build=b),
        "flood_curved_N7": lambda b: Case.flood(N=7, nel=10, seed=1, build=b),
        "bump_cartesian_periodic_N4": lambda b: Case.gaussian_bump(N=4, nel=16, build=b),
        "ocean_cartesian_N5": lambda b: Case.ocean_uplift(N=5, nel=8, amp=3.0, build=b),
        "rectangular_curved_N3": lambda b: Case(S.CASE_DAM_BREAK, S.MESH_CURVED_SINE, 3, 7, 4, 0.0, 7.0, -1.0, 3.0,
                                                perturb_amp=0.1, perturb_len=2.0, amp=1e-3, seed=9, build=b),
        "periodic_x_only_N2": lambda b: Case(S.CASE_CONSTANT, S.MESH_CARTESIAN, 2, 3, 5, 0.0, 3.0, 0.0, 5.0, 1, 0,
                                             p=(1.0, 0.2, -0.1, 0), build=b),
    }


def ulp_close(a, b, ulps=8):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= ulps * np.finfo(np.float64).eps * scale


@pytest.mark.parametrize("name", list(specs()))
def test_device_case_equals_host_case(dev, name):
    host = specs()[name](True)
    ctx = dev.Context(dev.make_params(host, RunParams()))
    try:
        dc = dev.DeviceCase(ctx, specs()[name](False))
        assert (dc.Nelem, dc.Nfaces, dc.ngl) == (host.Nelem, host.Nfaces, host.ngl)
        assert dc.min_ele_size == host.min_ele_size
        for n in EXACT:
            d = dc.download(n)
            assert d.dtype == host[n].dtype and np.array_equal(d, host[n]), n
        b, q = dc.download("b"), dc.download("q")
        assert ulp_close(b, host["b"]) and ulp_close(q, host["q"])
        # the gradient kernel on the host's b gives the host's Bx, By bit for bit; the case's own
        # Bx, By come from its own b
        if np.array_equal(b, host["b"]):
            assert np.array_equal(dc.download("Bx"), host["Bx"]) and np.array_equal(dc.download("By"), host["By"])
        dB, dBx, dBy = ctx.array(host["b"]), ctx.empty(host.Nelem * host.ngl ** 2), ctx.empty(host.Nelem * host.ngl ** 2)
        ctx.bathymetry_gradient(host.Nelem, dc.view("D0"), dB, dc.view("yEta"), dc.view("yXi"), dc.view("xEta"),
                                dc.view("xXi"), dc.view("Jac"), dBx, dBy)
        assert np.array_equal(dBx.download(), host["Bx"]) and np.array_equal(dBy.download(), host["By"])
        dc.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("which", ["curved", "ism"])
def test_geometry_kernels_on_host_inputs(dev, which):
    """Metrics, 1/J, face normals, sizes and Bx/By from the host case's nodes and bathymetry."""
    if which == "curved":
        host = Case.dam_break(N=7, nel=9, noise=1e-3, seed=2)
    else:
        host = Case.from_ism(ISM_MESH, N=4, case_kind=S.CASE_GAUSSIAN_BUMP, amp=0.05)
        assert (host["EdgeReversed"] == 0).any()
    E, F, n2 = host.Nelem, host.Nfaces, host.ngl ** 2
    ctx = dev.Context(dev.make_params(host, RunParams()))
    try:
        D0, x, y = ctx.array(host["D0"]), ctx.array(host["x"]), ctx.array(host["y"])
        out = {k: ctx.empty(E * n2) for k in ("xXi", "xEta", "yXi", "yEta", "Jac")}
        ctx.mapped_metrics(E, D0, x, y, out["xXi"], out["xEta"], out["yXi"], out["yEta"], out["Jac"])
        for k, v in out.items():
            assert np.array_equal(v.download(), host[k]), k
        ED = ctx.array(host["EdgeData"])
        nx, ny, sc = (ctx.empty(F * host.ngl) for _ in range(3))
        ctx.face_normals(F, ED, out["xXi"], out["xEta"], out["yXi"], out["yEta"], out["Jac"], nx, ny, sc)
        assert np.array_equal(nx.download(), host["nx"]) and np.array_equal(ny.download(), host["ny"])
        assert np.array_equal(sc.download(), host["scal"])
        sizes, mn = ctx.empty(E), np.zeros(1)
        ctx.element_sizes(E, out["Jac"], ctx.array(host["w_GL"]), sizes, mn)
        assert np.array_equal(sizes.download(), host["EleSizes"]) and mn[0] == host.min_ele_size
        Bx, By = ctx.empty(E * n2), ctx.empty(E * n2)
        ctx.bathymetry_gradient(E, D0, ctx.array(host["b"]), out["yEta"], out["yXi"], out["xEta"], out["xXi"],
                                out["Jac"], Bx, By)
        assert np.array_equal(Bx.download(), host["Bx"]) and np.array_equal(By.download(), host["By"])
    finally:
        ctx.close()


def test_solver_from_device_case_matches_oracle(dev):
    """The 2048^2 run's physics (N=7 flood, limiter, AV) on a 12^2 case built on the device: the
    solver takes its inputs device to device; the oracle runs on the downloaded arrays."""
    spec = Case.flood(N=7, nel=12, seed=0, build=False)
    rp = indicator.av_run_params(positivity=1, pos_pres_tol=0.1, CFL=0.1)
    ctx = dev.Context(dev.make_params(spec, rp))
    try:
        dc = dev.DeviceCase(ctx, spec)
        host = dc.to_host_case()
        s = dev.Solver(ctx, dc, rp, fused=1)
        n, t = s.run(10)
        q = s.get_q()
        s.close()
        dc.close()
    finally:
        ctx.close()
    qo, no, to, _ = O.timeloop(host, rp, 10)
    assert n == no == 10 and abs(t - to) <= 1e-12 * to
    A, B = q.reshape(host.Nelem, 3, -1), qo.reshape(host.Nelem, 3, -1)
    assert max(rel_l2(A[:, v], B[:, v]) for v in range(3)) <= 1e-12


def test_device_case_2048_builds_in_seconds(dev):
    """The largest run's case (2048^2 curved quads, N=7: 268 M nodes) on the device: built in a
    few seconds, with the mesh's invariants -- the element areas sum to the domain's 100, every
    face normal is a unit vector, every interior face is shared by two elements."""
    spec = Case.flood(N=7, nel=2048, seed=0, build=False)
    rp = indicator.av_run_params(positivity=1, pos_pres_tol=0.1, CFL=0.1)
    ctx = dev.Context(dev.make_params(spec, rp))
    try:
        t0 = time.perf_counter()
        dc = dev.DeviceCase(ctx, spec)
        ctx.finish()
        dt = time.perf_counter() - t0
        print(f"2048^2 N=7 device case: {dt:.2f} s, {dc.Nelem} elements, {dc.Nfaces} faces", flush=True)
        assert dc.Nelem == 2048 * 2048 and dc.Nfaces == 2 * 2048 * 2049
        assert dt < 20.0
        sizes = dc.download("EleSizes")
        assert abs(sizes.sum() - 100.0) <= 1e-9 and (sizes > 0).all()
        assert dc.min_ele_size == np.sqrt(sizes.min())
        nx, ny = dc.download("nx"), dc.download("ny")
        assert np.abs(np.hypot(nx, ny) - 1.0).max() <= 1e-15
        del nx, ny
        ED = dc.download("EdgeData").reshape(-1, 8)
        assert (ED[:, 1] >= 0).sum() == 2 * 2047 * 2048 and (ED[:, 7] == np.arange(dc.Nfaces)).all()
        dc.close()
    finally:
        ctx.close()


def test_device_case_rejects_what_it_does_not_build(dev):
    """The context's degree must be the spec's; ISM meshes are read on the host."""
    spec7 = Case.dam_break(N=7, nel=4, build=False)
    ctx = dev.Context(dev.make_params(Case.dam_break(N=4, nel=4, build=False), RunParams()))
    try:
        with pytest.raises(dev.EsdgDeviceError, match="ngl"):
            dev.DeviceCase(ctx, spec7)
    finally:
        ctx.close()
    with pytest.raises(ValueError):
        Case.from_ism(ISM_MESH, N=4).__class__(S.CASE_CONSTANT, S.MESH_ISM, 4, 0, 0, 0.0, 1.0, 0.0, 1.0,
                                               mesh_file=ISM_MESH, build=False)
