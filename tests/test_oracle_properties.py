"""Pins of the oracle's kernels by what the mathematics fixes.

The reference ships no tests or golden vectors for these kernels and its okl path cannot be built
here, so each oracle function is checked against properties a plausible restatement mistake
(transposed operator, swapped metric, wrong sign or side index, dropped term) would break:
independent vectorised formulations, consistency and entropy identities of the two-point and
surface fluxes, free-stream preservation and well-balancing on curved meshes, discrete
conservation and semi-discrete entropy conservation of the whole residual, exactness of the BR1
gradient, the modal sensor's closed form, the limiter's invariants and the RK order.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as L

from esdgsem_mpiocca_amd import Case, RunParams
from esdgsem_mpiocca_amd import setup_api as S
from helpers import oracle_residual, perturbed_state, quad_weights
from oracle import binding as O

G = 9.81


# --------------------------------------------------------------------------------------------
# volume term: an independent vectorised restatement with p# = g/2 hL hR
# --------------------------------------------------------------------------------------------
def numpy_volume(case, q):
    ngl, E = case.ngl, case.Nelem
    D = case["D"].reshape(ngl, ngl)
    Q = q.reshape(E, 3, ngl, ngl)                  # [e, v, row(eta), col(xi)]
    h, hu, hv = Q[:, 0], Q[:, 1], Q[:, 2]
    inv = np.where(h > 0, 1.0 / np.where(h > 0, h, 1.0), 0.0)
    u, v = hu * inv, hv * inv
    met = {k: case[k].reshape(E, ngl, ngl) for k in ("yXi", "yEta", "xXi", "xEta", "Jac", "Bx", "By")}

    def two_point(axis, Y, X):
        # pairs along `axis` (2 = xi within a row, 1 = eta within a column); returns [e, a, b, comp]
        def pair(f):
            a = np.expand_dims(f, axis + 1)         # node a
            b = np.expand_dims(f, axis)             # node b
            return a, b
        hA, hB = pair(h); huA, huB = pair(hu); hvA, hvB = pair(hv)
        uA, uB = pair(u); vA, vB = pair(v); YA, YB = pair(Y); XA, XB = pair(X)
        p = 0.5 * G * hA * hB
        f = np.stack([0.5 * (huA + huB), 0.25 * (huA + huB) * (uA + uB) + p, 0.25 * (huA + huB) * (vA + vB)])
        g = np.stack([0.5 * (hvA + hvB), 0.25 * (hvA + hvB) * (uA + uB), 0.25 * (hvA + hvB) * (vA + vB) + p])
        return 0.5 * (YA + YB) * f - 0.5 * (XA + XB) * g

    Fx = two_point(2, met["yEta"], met["xEta"])        # [3, e, i, j, l]
    Fy = two_point(1, met["yXi"], met["xXi"])          # [3, e, i, l, j]
    div = np.einsum("jl,ceijl->ceij", D, Fx) - np.einsum("il,ceilj->ceij", D, Fy)
    out = -met["Jac"][None] * div
    out[1] -= G * h * met["Bx"]
    out[2] -= G * h * met["By"]
    return np.moveaxis(out, 0, 1).reshape(-1)


@pytest.mark.parametrize("mk", ["cart", "curved"])
def test_volume_matches_independent_form(mk):
    c = Case.gaussian_bump(N=4, nel=4) if mk == "cart" else Case.dam_break(N=6, nel=3, seed=1)
    rp = RunParams()
    q = perturbed_state(c, seed=2)
    Qt = np.zeros_like(q)
    a = c.arrays
    O.call("volume_fluxdiff", O.params(c, rp), c.Nelem, a["Jac"], a["yXi"], a["yEta"], a["xXi"], a["xEta"], q,
           a["D"], a["Bx"], a["By"], Qt)
    ref = numpy_volume(c, q)
    assert np.abs(Qt - ref).max() <= 1e-11 * np.abs(ref).max()


# --------------------------------------------------------------------------------------------
# surface fluxes on a synthetic set of face nodes (ngl = 1: one node per "face")
# --------------------------------------------------------------------------------------------
def flux(name, qL, qR, n, bL=None, bR=None, ES=1, tol=1e-12):
    M = len(qL)
    p = O.OracleParams(1, ES, G, 1.0, tol, 1e-5, 0.0, 0.0, 0.0)
    QL = np.ascontiguousarray(np.stack([qL[:, 0], qL[:, 1], qL[:, 2]], axis=1).reshape(-1))
    QR = np.ascontiguousarray(np.stack([qR[:, 0], qR[:, 1], qR[:, 2]], axis=1).reshape(-1))
    bL = np.zeros(M) if bL is None else np.ascontiguousarray(bL)
    bR = np.zeros(M) if bR is None else np.ascontiguousarray(bR)
    SP = np.zeros(3 * M)
    O.call(name, p, M, np.ones(M, np.int32), np.ascontiguousarray(n[:, 0]), np.ascontiguousarray(n[:, 1]),
           np.ones(M), QL, QR, bL, bR, SP)
    return SP.reshape(M, 3)


def random_faces(M, seed):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.5, 2.0, (M, 2))
    u = rng.uniform(-1, 1, (M, 2, 2))
    th = rng.uniform(0, 2 * np.pi, M)
    n = np.stack([np.cos(th), np.sin(th)], 1)
    qL = np.stack([h[:, 0], h[:, 0] * u[:, 0, 0], h[:, 0] * u[:, 0, 1]], 1)
    qR = np.stack([h[:, 1], h[:, 1] * u[:, 1, 0], h[:, 1] * u[:, 1, 1]], 1)
    return qL, qR, n


def phys_flux(q, n):
    h, hu, hv = q[:, 0], q[:, 1], q[:, 2]
    u, v = hu / h, hv / h
    un = u * n[:, 0] + v * n[:, 1]
    return np.stack([h * un, hu * un + 0.5 * G * h * h * n[:, 0], hv * un + 0.5 * G * h * h * n[:, 1]], 1)


def ent_vars(q, b):
    h, u, v = q[:, 0], q[:, 1] / q[:, 0], q[:, 2] / q[:, 0]
    return np.stack([G * (h + b) - 0.5 * (u * u + v * v), u, v], 1)


@pytest.mark.parametrize("name", ["num_flux_es", "num_flux_es_rotated"])
def test_surface_flux_consistency_and_entropy(name):
    qL, qR, n = random_faces(400, 3)
    # consistency: F*(u, u) = n . f(u) (with and without dissipation)
    for ES in (0, 1):
        F = flux(name, qL, qL, n, ES=ES)
        assert np.abs(F - phys_flux(qL, n)).max() < 1e-13
    # entropy conservation of the EC part (flat bottom): [v] . F_EC = [psi], psi = g/2 h^2 u_n
    Fec = flux(name, qL, qR, n, ES=0)
    b0 = np.zeros(len(qL))
    jv = ent_vars(qR, b0) - ent_vars(qL, b0)
    psi = lambda q: 0.5 * G * q[:, 0] ** 2 * (q[:, 1] * n[:, 0] + q[:, 2] * n[:, 1]) / q[:, 0]
    assert np.abs((jv * Fec).sum(1) - (psi(qR) - psi(qL))).max() < 1e-12
    # entropy dissipation of the ES flux, with bathymetry jumps in v1
    rng = np.random.default_rng(9)
    bL, bR = rng.uniform(-0.2, 0.2, len(qL)), rng.uniform(-0.2, 0.2, len(qL))
    Fes = flux(name, qL, qR, n, bL, bR, ES=1)
    Fec = flux(name, qL, qR, n, bL, bR, ES=0)
    jv = ent_vars(qR, bR) - ent_vars(qL, bL)
    assert ((jv * (Fes - Fec)).sum(1) <= 1e-13).all()
    assert ((jv * (Fes - Fec)).sum(1) < -1e-6).mean() > 0.9


def supersonic_faces(M, seed, sign, tangential=True):
    """Faces with both states moving along sign*n faster than the largest wave speed."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, M)
    n = np.stack([np.cos(th), np.sin(th)], 1)
    t = np.stack([-n[:, 1], n[:, 0]], 1)
    cmax = np.sqrt(G * 1.0)
    out = []
    for _ in range(2):
        h = rng.uniform(0.5, 1.0, M)
        un = sign * rng.uniform(2.0, 3.0, M) * cmax
        ut = rng.uniform(-1, 1, M) if tangential else np.zeros(M)
        vel = un[:, None] * n + ut[:, None] * t
        out.append(np.stack([h, h * vel[:, 0], h * vel[:, 1]], 1))
    return out[0], out[1], n


def test_roe_flux_consistency_and_upwinding():
    """calcRoeFlux.okl: F(u, u) = n.f(u); with every Roe eigenvalue positive (negative) and no
    tangential velocity the flux is the left (right) physical flux -- Roe's property of the
    sqrt(h)-weighted averages.  With tangential velocity it is not: the okl's second left
    eigenvector carries +u_t where -u_t belongs (calcRoeFlux.okl:161, calcCombinedFluxOcean.okl:120),
    and the oracle keeps that sign, so the defect shows exactly there."""
    qL, qR, n = random_faces(400, 4)
    scale = G * 4.0
    assert np.abs(flux("num_flux_roe", qL, qL, n) - phys_flux(qL, n)).max() < 1e-14 * scale
    for sign in (1.0, -1.0):
        qL, qR, n = supersonic_faces(300, 5, sign, tangential=False)
        F = flux("num_flux_roe", qL, qR, n)
        ref = phys_flux(qL, n) if sign > 0 else phys_flux(qR, n)
        assert np.abs(F - ref).max() < 1e-12 * scale, np.abs(F - ref).max()
        qL, qR, n = supersonic_faces(300, 5, sign, tangential=True)
        F = flux("num_flux_roe", qL, qR, n)
        ref = phys_flux(qL, n) if sign > 0 else phys_flux(qR, n)
        assert np.abs(F[:, 0] - ref[:, 0]).max() < 1e-12 * scale      # mass: the sign does not enter
        assert np.abs(F[:, 1:] - ref[:, 1:]).max() > 1e-3
    # subsonic: the Roe dissipation is not zero
    qL, qR, n = random_faces(400, 6)
    F = flux("num_flux_roe", qL, qR, n)
    assert np.abs(F - 0.5 * (phys_flux(qL, n) + phys_flux(qR, n))).max() > 1e-3


def test_ocean_flux_selects_roe_on_open_boundary_faces():
    """calcCombinedFluxOcean.okl: Roe on boundary faces with bL < 6.7, rotated ES everywhere else."""
    M = 500
    qL, qR, n = random_faces(M, 7)
    rng = np.random.default_rng(8)
    bL = rng.uniform(0.0, 20.0, M)
    bR = rng.uniform(0.0, 20.0, M)
    ED = np.zeros((M, 8), np.int32)
    bnd = rng.uniform(size=M) < 0.5
    ED[bnd, 6] = -1
    p = O.OracleParams(1, 1, G, 1.0, 1e-12, 1e-5, 0.0, 0.0, 0.0)
    QL = np.ascontiguousarray(qL.reshape(-1))
    QR = np.ascontiguousarray(qR.reshape(-1))
    ones = np.ones(M, np.int32)
    nx, ny = np.ascontiguousarray(n[:, 0]), np.ascontiguousarray(n[:, 1])
    SP = np.zeros(3 * M)
    O.call("num_flux_ocean", p, M, np.ascontiguousarray(ED.reshape(-1)), ones, nx, ny, np.ones(M), QL, QR, bL, bR, SP)
    roe = flux("num_flux_roe", qL, qR, n, bL, bR)
    rot = flux("num_flux_es_rotated", qL, qR, n, bL, bR)
    sel = bnd & (bL < 6.7)
    assert sel.any() and (~sel).any()
    expect = np.where(sel[:, None], roe, rot)
    assert np.array_equal(SP.reshape(M, 3), expect)


def test_es_dissipation_matrix_independent():
    """F_ES - F_EC = -1/2 R |Lambda| R^T [v] with the eigen-structure of calcESFlux.okl."""
    qL, qR, n = random_faces(200, 5)
    Fes = flux("num_flux_es", qL, qR, n, ES=1)
    Fec = flux("num_flux_es", qL, qR, n, ES=0)
    jv = ent_vars(qR, 0 * qR[:, 0]) - ent_vars(qL, 0 * qL[:, 0])
    for k in range(len(qL)):
        hL, hR = qL[k, 0], qR[k, 0]
        uA = 0.5 * (qL[k, 1] / hL + qR[k, 1] / hR)
        vA = 0.5 * (qL[k, 2] / hL + qR[k, 2] / hR)
        c = np.sqrt(0.5 * G * (hL + hR))
        nx, ny = n[k]
        un = uA * nx + vA * ny
        R = np.array([[1, 0, 1], [uA + nx * c, ny, uA - nx * c], [vA + ny * c, nx, vA - ny * c]])
        Lam = np.diag([abs(un + c) / (2 * G), 0.5 * (hL + hR) * abs(un), abs(un - c) / (2 * G)])
        d = -0.5 * R @ Lam @ R.T @ jv[k]
        assert np.allclose(Fes[k] - Fec[k], d, rtol=1e-12, atol=1e-14)


def test_rotated_equals_cartesian_on_axis_aligned_faces():
    qL, qR, n = random_faces(100, 7)
    for nn in ((1, 0), (0, 1), (-1, 0), (0, -1)):
        nrm = np.tile(np.array(nn, float), (len(qL), 1))
        a = flux("num_flux_es", qL, qR, nrm)
        b = flux("num_flux_es_rotated", qL, qR, nrm)
        assert np.abs(a - b).max() < 1e-12


def test_wall_gather_mirror_is_involution():
    c = Case.dam_break(N=3, nel=3, seed=5)
    q = perturbed_state(c, seed=5)
    p = O.params(c, RunParams())
    fn = c.Nfaces * c.ngl
    qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
    O.call("collect_edge_data_walls", p, c.Nfaces, c["EdgeData"], q, c["nx"], c["ny"], qL, qR)
    ED = c["EdgeData"].reshape(-1, 8)
    nx, ny = c["nx"].reshape(-1, c.ngl), c["ny"].reshape(-1, c.ngl)
    L3, R3 = qL.reshape(-1, 3, c.ngl), qR.reshape(-1, 3, c.ngl)
    bnd = ED[:, 1] == -1
    # the mirror keeps h, flips the normal momentum, keeps the tangential one
    assert np.array_equal(L3[bnd, 0], R3[bnd, 0])
    mnL = L3[bnd, 1] * nx[bnd] + L3[bnd, 2] * ny[bnd]
    mnR = R3[bnd, 1] * nx[bnd] + R3[bnd, 2] * ny[bnd]
    mtL = -L3[bnd, 1] * ny[bnd] + L3[bnd, 2] * nx[bnd]
    mtR = -R3[bnd, 1] * ny[bnd] + R3[bnd, 2] * nx[bnd]
    assert np.abs(mnL + mnR).max() < 1e-15 and np.abs(mtL - mtR).max() < 1e-15
    # interior faces: both traces sample the same physical points (x, y agree)
    x, y = c["x"].reshape(c.Nelem, -1), c["y"].reshape(c.Nelem, -1)
    side = {0: lambda k: k, 2: lambda k: c.ngl ** 2 - c.ngl + k, 3: lambda k: k * c.ngl,
            1: lambda k: (k + 1) * c.ngl - 1}
    for f in np.nonzero(~bnd)[0]:
        eL, eR, sL, sR = ED[f, :4]
        for k in range(c.ngl):
            assert abs(x[eL, side[sL](k)] - x[eR, side[sR](k)]) < 1e-13
            assert abs(y[eL, side[sL](k)] - y[eR, side[sR](k)]) < 1e-13


def test_boundary_variants_exterior_states():
    """CollectEdgeData variants (SolidWalls / InnerOuter / OceanBoundary): interior traces are the
    same in all three; boundary exteriors follow each okl's rule, and the ocean rule takes both of
    its branches on the shelf mesh."""
    c = Case.ocean_uplift(N=3, nel=5, amp=2.0)
    q = perturbed_state(c, seed=7)
    fn = c.Nfaces * c.ngl
    bL, bR = np.zeros(fn), np.zeros(fn)
    p0 = O.params(c, RunParams())
    O.call("collect_edge_data_bottom", p0, c.Nfaces, c["EdgeData"], c["b"], bL, bR)
    out = {}
    for bc in (0, 1, 2):
        p = O.params(c, RunParams(bc_kind=bc, h_zero=4000.0))
        qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
        O.call("collect_edge_data", p, c.Nfaces, c["EdgeData"], q, c["nx"], c["ny"], bL, bR, qL, qR)
        out[bc] = (qL.reshape(-1, 3, c.ngl), qR.reshape(-1, 3, c.ngl))
    wL, wR = np.zeros(fn * 3), np.zeros(fn * 3)
    O.call("collect_edge_data_walls", p0, c.Nfaces, c["EdgeData"], q, c["nx"], c["ny"], wL, wR)
    assert np.array_equal(out[0][0].ravel(), wL) and np.array_equal(out[0][1].ravel(), wR)
    ED = c["EdgeData"].reshape(-1, 8)
    bnd = ED[:, 6] == -1
    assert bnd.any() and not (ED[:, 5] == -1).any()
    for bc in (1, 2):
        assert np.array_equal(out[bc][0], out[0][0])              # left traces: always interior
        assert np.array_equal(out[bc][1][~bnd], out[0][1][~bnd])  # interior faces unchanged
    L, R1 = out[0][0][bnd], out[1][1][bnd]
    assert np.array_equal(R1[:, 0], L[:, 0]) and not R1[:, 1:].any()
    R2, Rw = out[2][1][bnd], out[0][1][bnd]
    bi = bL.reshape(-1, c.ngl)[bnd]
    wall = bi > 6.7
    assert wall.any() and (~wall).any()
    for v in range(3):
        assert np.array_equal(R2[:, v][wall], Rw[:, v][wall])
    assert np.array_equal(R2[:, 0][~wall], 4000.0 - bi[~wall])
    assert not R2[:, 1][~wall].any() and not R2[:, 2][~wall].any()


@pytest.mark.parametrize("t", [100.0, 400.0])
def test_three_mound_boundary_states(t):
    """3MoundInflow.okl exterior states on a channel: reflecting walls where n_x = 0, the inflow
    state (1/2, 1/2, 0) -- (1, 1, 0) after t = 300 -- at x = 0, a copy of the interior elsewhere."""
    c = Case(S.CASE_CONSTANT, S.MESH_CARTESIAN, 3, 6, 3, 0.0, 4.0, 0.0, 2.0, p=(1.0, 0.2, 0.1, 0))
    q = perturbed_state(c, seed=11)
    p = O.params(c, RunParams(bc_kind=4))
    fn = c.Nfaces * c.ngl
    bL, bR, qL, qR = np.zeros(fn), np.zeros(fn), np.zeros(fn * 3), np.zeros(fn * 3)
    O.call("collect_edge_data_xt", p, c.Nfaces, c["EdgeData"], q, c["x"], c["y"], c["nx"], c["ny"], bL, bR, t,
           qL, qR)
    ED = c["EdgeData"].reshape(-1, 8)
    L3, R3 = qL.reshape(-1, 3, c.ngl), qR.reshape(-1, 3, c.ngl)
    nx, ny = c["nx"].reshape(-1, c.ngl), c["ny"].reshape(-1, c.ngl)
    x = c["x"].reshape(c.Nelem, -1)
    side = {0: lambda k: k, 2: lambda k: c.ngl ** 2 - c.ngl + k, 3: lambda k: k * c.ngl,
            1: lambda k: (k + 1) * c.ngl - 1}
    inflow = (0.5, 0.5, 0.0) if t <= 300 else (1.0, 1.0, 0.0)
    seen = set()
    for f in np.nonzero(ED[:, 6] == -1)[0]:
        for k in range(c.ngl):
            L, R = L3[f, :, k], R3[f, :, k]
            if abs(nx[f, k]) < 1e-5:
                seen.add("wall")
                assert R[0] == L[0]
                mn = lambda v: v[1] * nx[f, k] + v[2] * ny[f, k]
                mt = lambda v: -v[1] * ny[f, k] + v[2] * nx[f, k]
                assert abs(mn(R) + mn(L)) < 1e-14 and abs(mt(R) - mt(L)) < 1e-14
            elif abs(x[ED[f, 0], side[ED[f, 2]](k)]) < 1e-5:
                seen.add("inflow")
                assert tuple(R) == inflow
            else:
                seen.add("outflow")
                assert np.array_equal(R, L)
    assert seen == {"wall", "inflow", "outflow"}


def test_open_boundaries_keep_still_water():
    """Still water h + b = h_zero at rest: the ocean boundary's exterior state equals the interior
    one on the open part and the residual vanishes to roundoff; a wrong still-water level does not.
    The inner/outer variant keeps a lake at rest likewise."""
    c = Case.ocean_uplift(N=4, nel=6, amp=0.0)
    sc = term_scale(c, c["q"])
    for nf in (0, 3):
        Qt = oracle_residual(c, RunParams(num_flux=nf, bc_kind=2, h_zero=4000.0), c["q"])
        assert np.abs(Qt).max() < 1e-14 * sc, np.abs(Qt).max() / sc
        Qt = oracle_residual(c, RunParams(num_flux=nf, bc_kind=1), c["q"])
        assert np.abs(Qt).max() < 1e-14 * sc
    Qt = oracle_residual(c, RunParams(num_flux=0, bc_kind=2, h_zero=4001.0), c["q"])
    assert np.abs(Qt).max() > 1e-6 * sc


# --------------------------------------------------------------------------------------------
# whole residual invariants
# --------------------------------------------------------------------------------------------
def term_scale(c, q):
    """Magnitude of the individual volume terms (g h^2 |D| |J^-1| |metric|): roundoff reference."""
    h = np.abs(q.reshape(c.Nelem, 3, -1)[:, 0]).max()
    met = max(np.abs(c[k]).max() for k in ("xXi", "xEta", "yXi", "yEta"))
    return G * h * h * np.abs(c["D"]).max() * np.abs(c["Jac"]).max() * met


def test_free_stream_and_well_balanced_curved():
    c = Case.dam_break(N=5, nel=5, seed=1)
    E = c.Nelem
    rp = RunParams(num_flux=3)
    # lake at rest over non-flat bathymetry: h + b = 2, u = v = 0
    q = np.zeros((E, 3, c.ngl ** 2))
    q[:, 0] = 2.0 - c["b"].reshape(E, -1)
    sc = term_scale(c, q.ravel())
    for nf in (0, 3):
        Qt = oracle_residual(c, RunParams(num_flux=nf), q.ravel())
        assert np.abs(Qt).max() < 1e-14 * sc
    # the same state with a dropped bathymetry source term is far from balanced
    c2 = Case.dam_break(N=5, nel=5, seed=1)
    c2.arrays["Bx"] = np.zeros_like(c2["Bx"])
    assert np.abs(oracle_residual(c2, rp, q.ravel())).max() > 1e-6 * sc


def test_free_stream_with_velocity_periodic():
    c = Case.gaussian_bump(N=4, nel=4)
    E = c.Nelem
    # constant state on flat bottom: every volume/surface term must cancel
    flat = Case(S.CASE_CONSTANT, S.MESH_CARTESIAN, 4, 4, 4, 0.0, 2.0, 0.0, 2.0, 1, 1, p=(1.3, 0.4, -0.2, 0))
    Qt = oracle_residual(flat, RunParams(num_flux=0), flat["q"])
    assert np.abs(Qt).max() < 1e-14 * term_scale(flat, flat["q"])
    # curved walls, flat bottom, fluid at rest: pressure terms cancel through the metric identities
    cv = Case.constant(N=6, nel=4, h=1.7)
    Qt = oracle_residual(cv, RunParams(num_flux=3), cv["q"])
    assert np.abs(Qt).max() < 1e-14 * term_scale(cv, cv["q"])


@pytest.mark.parametrize("mk", ["periodic", "walls_curved"])
def test_mass_conservation(mk):
    c = Case.gaussian_bump(N=4, nel=6) if mk == "periodic" else Case.dam_break(N=5, nel=4, seed=3)
    q = perturbed_state(c, seed=4)
    Qt = oracle_residual(c, RunParams(num_flux=0 if mk == "periodic" else 3), q).reshape(c.Nelem, 3, -1)
    wJ = quad_weights(c)
    assert abs((wJ * Qt[:, 0]).sum()) < 1e-12 * np.abs(wJ * Qt[:, 0]).sum()


def test_semi_discrete_entropy_conservation_periodic():
    """Periodic, EC volume + EC surface flux: sum w J v . Qt = 0; with ES dissipation it is < 0."""
    c = Case.gaussian_bump(N=4, nel=6)
    q = perturbed_state(c, seed=6, amp_u=0.5)
    E = c.Nelem
    Q = q.reshape(E, 3, -1)
    b = c["b"].reshape(E, -1)
    v = np.stack([G * (Q[:, 0] + b) - 0.5 * ((Q[:, 1] / Q[:, 0]) ** 2 + (Q[:, 2] / Q[:, 0]) ** 2),
                  Q[:, 1] / Q[:, 0], Q[:, 2] / Q[:, 0]], 1)
    wJ = quad_weights(c)
    for ES, sign in ((0, 0), (1, -1)):
        Qt = oracle_residual(c, RunParams(num_flux=0, ES=ES), q).reshape(E, 3, -1)
        dS = (wJ[:, None] * v * Qt).sum()
        scale = np.abs(wJ[:, None] * v * Qt).sum()
        if sign == 0:
            assert abs(dS) < 1e-12 * scale
        else:
            assert dS < -1e-8 * scale


# --------------------------------------------------------------------------------------------
# artificial viscosity chain
# --------------------------------------------------------------------------------------------
def gradient_of(c, q, rp):
    p = O.params(c, rp)
    a = c.arrays
    fn, nn = c.Nfaces * c.ngl, c.Nelem * c.ngl ** 2
    qL, qR = np.zeros(fn * 3), np.zeros(fn * 3)
    O.call("collect_edge_data_walls", p, c.Nfaces, a["EdgeData"], q, a["nx"], a["ny"], qL, qR)
    SGX, SGY = np.zeros(fn * 2), np.zeros(fn * 2)
    O.call("num_fluxes_gradient", p, c.Nfaces, a["EdgeData"], a["nx"], a["ny"], a["scal"], qL, qR, SGX, SGY)
    gX, gY = np.zeros(nn * 2), np.zeros(nn * 2)
    O.call("calc_gradient", p, c.Nelem, a["Jac"], a["yXi"], a["yEta"], a["xXi"], a["xEta"], q, a["Dhat"], gX, gY)
    O.call("surface_gradient", p, c.Nelem, a["Jac"], a["ElemEdgeMasterSlave"], a["ElemEdgeOrientation"],
           a["ElemToEdge"], SGX, SGY, gX, gY)
    O.call("scale_gradient", p, c.Nelem, q, gX, gY)
    return gX.reshape(c.Nelem, 2, -1), gY.reshape(c.Nelem, 2, -1)


@pytest.mark.parametrize("mk", ["cart", "curved"])
def test_br1_gradient_exact_for_linear_velocity(mk):
    c = (Case(S.CASE_CONSTANT, S.MESH_CARTESIAN, 4, 4, 3, 0.0, 10.0, 0.0, 10.0, p=(1.5, 0, 0, 0))
         if mk == "cart" else Case.constant(N=6, nel=4, h=1.5))
    E = c.Nelem
    x, y = c["x"].reshape(E, -1), c["y"].reshape(E, -1)
    h = 1.5
    u, v = 0.3 + 0.2 * x - 0.1 * y, -0.4 + 0.05 * x + 0.15 * y
    q = np.stack([np.full_like(x, h), h * u, h * v], 1).ravel()
    gX, gY = gradient_of(c, q, RunParams())
    tol = 1e-12 if mk == "cart" else 1e-10
    assert np.abs(gX[:, 0] - h * 0.2).max() < tol and np.abs(gY[:, 0] + h * 0.1).max() < tol
    assert np.abs(gX[:, 1] - h * 0.05).max() < tol and np.abs(gY[:, 1] - h * 0.15).max() < tol


def test_viscous_terms_vanish_with_zero_viscosity_and_are_dissipative():
    c = Case.dam_break(N=4, nel=4, seed=2)
    q = perturbed_state(c, seed=8)
    rp = RunParams(num_flux=3)
    base = oracle_residual(c, rp, q)
    zero = oracle_residual(c, rp, q, av_visc=np.zeros(c.Nelem))
    assert np.array_equal(base, zero)
    # constant viscosity: the added term dissipates kinetic energy  sum w J (u, v) . dQt <= 0
    eps = np.full(c.Nelem, 0.01)
    visc = oracle_residual(c, rp, q, av_visc=eps) - base
    E = c.Nelem
    Q = q.reshape(E, 3, -1)
    dV = visc.reshape(E, 3, -1)
    wJ = quad_weights(c)
    rate = (wJ * (Q[:, 1] / Q[:, 0] * dV[:, 1] + Q[:, 2] / Q[:, 0] * dV[:, 2])).sum()
    assert rate < 0


def test_shock_sensor_closed_form():
    """h + b = c0 + a * phi_k(xi) on one element: only one non-constant mode, so the energy ratios
    and the viscosity ramp follow in closed form."""
    N = 6
    c = Case(S.CASE_CONSTANT, S.MESH_CARTESIAN, N, 1, 1, -1.0, 1.0, -1.0, 1.0, p=(2.0, 0, 0, 0))
    xi = c["x_GL"]
    ngl = N + 1
    phi = lambda k, x: L.legval(x, np.eye(ngl)[k]) * np.sqrt(k + 0.5)
    rp = RunParams(eps0=0.02, sigma_min=-4.0, sigma_max=-2.0)
    p = O.params(c, rp)
    for k, a in ((N, 0.05), (N, 0.002), (N - 1, 0.01), (N - 2, 0.3), (N, 1e-4)):
        h = (2.0 + a * phi(k, xi))[None, :].repeat(ngl, 0)     # [eta, xi]
        q = np.stack([h.ravel(), 0 * h.ravel(), 0 * h.ravel()])[None].ravel()
        visc = np.zeros(1)
        O.call("shock_capturing", p, 1, q, np.zeros(ngl * ngl), c["VdmInv"], visc, None)
        m00 = 2.0 * np.sqrt(2.0) * np.sqrt(2.0)               # constant 2 -> mode (0,0) coefficient
        e_all = m00 ** 2 + a ** 2 * 2.0
        if k == N:
            sigma = np.log10(a * a * 2.0 / e_all)
        elif k == N - 1:
            sigma = np.log10(a * a * 2.0 / (e_all))           # term 2 over the modes below N
        else:
            sigma = -np.inf
        if sigma <= rp.sigma_min:
            expect = 0.0
        elif sigma <= rp.sigma_max:
            expect = 0.5 * rp.eps0 * (1 + np.sin(np.pi * (sigma - 0.5 * (rp.sigma_max + rp.sigma_min)) /
                                                 (rp.sigma_max - rp.sigma_min)))
        else:
            expect = rp.eps0
        assert abs(visc[0] - expect) < 1e-10 * rp.eps0, (k, a, visc[0], expect)


# --------------------------------------------------------------------------------------------
# limiter and RK
# --------------------------------------------------------------------------------------------
def test_positivity_limiter_invariants():
    c = Case.flood(N=4, nel=6, seed=2)
    E = c.Nelem
    rng = np.random.default_rng(3)
    Q = c["q"].reshape(E, 3, -1).copy() + 0.0
    Q[:, 0] += rng.uniform(-0.05, 0.05, Q[:, 0].shape)
    Q[:, 1] = rng.uniform(-0.1, 0.1, Q[:, 0].shape)
    q = Q.ravel().copy()
    rp = RunParams(pos_pres_tol=1e-3)
    O.call("pos_pres", O.params(c, rp), E, c["EleSizes"], c["w_GL"], c["Jac"], q)
    Q2 = q.reshape(E, 3, -1)
    wJ = quad_weights(c)
    mean0 = (wJ * Q[:, 0]).sum(1) / c["EleSizes"]
    mean1 = (wJ * Q2[:, 0]).sum(1) / c["EleSizes"]
    limited = (Q[:, 0] < 0).any(1) & (mean0 > 1e-5)
    assert limited.any()
    assert np.abs(mean1[limited] - mean0[limited]).max() < 1e-13
    assert (Q2[limited, 0] >= 0).all()
    untouched = (Q[:, 0] >= 1e-3).all(1)
    assert np.array_equal(Q2[untouched], Q[untouched])
    neg = mean0 < -1e-5
    assert (np.isneginf(Q2[neg, 0])).all()


def test_lsrk4_is_fourth_order():
    lam = -1.3 + 0.7j
    A, B, Cc = (np.array(x) for x in zip(*[(a, b, c) for a, b, c in zip(*__import__("esdgsem_mpiocca_amd").rk_coefficients(4))]))
    p = O.OracleParams(1, 1, G, 1.0, 0.0, 1e-5, 0.0, 0.0, 0.0)
    errs = []
    for nsteps in (20, 40, 80):
        dt = 1.0 / nsteps
        y = np.array([1.0, 0.0, 0.0])      # (Re, Im, pad): one "element" of ngl=1 holds 3 values
        g = np.zeros(3)
        for _ in range(nsteps):
            for s in range(5):
                z = y[0] + 1j * y[1]
                f = lam * z
                Qt = np.array([f.real, f.imag, 0.0])
                O.call("rk_ls_calc", p, 1, Qt, float(A[s]), g)
                O.call("rk_ls_update", p, 1, float(Cc[s]), dt, g, y)
        errs.append(abs(y[0] + 1j * y[1] - np.exp(lam)))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert (orders > 3.8).all(), orders


# --------------------------------------------------------------------------------------------
# time loop
# --------------------------------------------------------------------------------------------
def test_timeloop_lake_at_rest_and_mass():
    c = Case.gaussian_bump(N=4, nel=16, lake_at_rest=True)
    rp = RunParams(num_flux=0)
    q, n, t, dts = O.timeloop(c, rp, 200)
    Q = q.reshape(c.Nelem, 3, -1)
    assert n == 200
    assert np.abs(Q[:, 1:]).max() <= 1e-13
    # dt rule of DGtimeloop: sqrt(min area)/ngl * CFL / lambda_max
    lam0 = np.sqrt(G * c["q"].reshape(c.Nelem, 3, -1)[:, 0].max())
    assert abs(dts[0] - c.min_ele_size / c.ngl * rp.CFL / lam0) < 1e-15 * dts[0]


def test_timeloop_mass_conservation_bump():
    c = Case.gaussian_bump(N=4, nel=8)
    q, n, t, _ = O.timeloop(c, RunParams(num_flux=0), 50)
    wJ = quad_weights(c)
    m0 = (wJ * c["q"].reshape(c.Nelem, 3, -1)[:, 0]).sum()
    m1 = (wJ * q.reshape(c.Nelem, 3, -1)[:, 0]).sum()
    assert abs(m1 - m0) < 1e-13 * m0


# mean depth H0 and wave number k of each convergence Testcase (SW2D.cpp:312-318, 388-393, 760-775)
MANUFACTURED = {1: (8.0, 1.0), 8: (20.0, 1.0), 88: (20.0, 1e-3), 89: (20000.0, 1e-6)}


def manufactured(case, t, testcase=1, dt=False):
    """A convergence test's exact solution: h = H0 + cos(kx) sin(ky) cos t - b, hu = h/2, hv = 3h/2
    (with dt=True its time derivative)."""
    H0, k = MANUFACTURED[testcase]
    x, y, b = case["x"], case["y"], case["b"]
    if dt:
        h = -np.cos(k * x) * np.sin(k * y) * np.sin(t)
    else:
        h = H0 + np.cos(k * x) * np.sin(k * y) * np.cos(t) - b
    q = np.zeros((case.Nelem, 3, case.ngl ** 2))
    q[:, 0] = h.reshape(case.Nelem, -1)
    q[:, 1] = 0.5 * q[:, 0]
    q[:, 2] = 1.5 * q[:, 0]
    return q.ravel()


def manufactured_residual(case, rp, q, t):
    """Qt of one stage of a convergence test at time t: the Dirichlet gather, volume, ES flux,
    surface lift and the manufactured source (the reference's kernel order, deviceclass.cpp:957-1102)."""
    p = O.params(case, rp)
    a = case.arrays
    Ne, Nf, ngl = case.Nelem, case.Nfaces, case.ngl
    fn = Nf * ngl
    qL, qR, bL, bR = np.zeros(fn * 3), np.zeros(fn * 3), np.zeros(fn), np.zeros(fn)
    O.call("collect_edge_data_bottom", p, Nf, a["EdgeData"], a["b"], bL, bR)
    O.call("collect_edge_data_xt", p, Nf, a["EdgeData"], q, a["x"], a["y"], a["nx"], a["ny"], bL, bR, t, qL, qR)
    Qt = np.zeros(Ne * ngl * ngl * 3)
    O.call("volume_fluxdiff", p, Ne, a["Jac"], a["yXi"], a["yEta"], a["xXi"], a["xEta"], q, a["D"], a["Bx"],
           a["By"], Qt)
    SP, DB = np.zeros(fn * 3), np.zeros(fn)
    O.call("num_flux_es", p, Nf, a["EdgeReversed"], a["nx"], a["ny"], a["scal"], qL, qR, bL, bR, SP)
    O.call("surface", p, Ne, a["Jac"], a["ElemEdgeMasterSlave"], a["ElemEdgeOrientation"], a["ElemToEdge"], SP,
           DB, DB, Qt)
    O.call("add_source", p, Ne, a["Bx"], a["By"], a["b"], a["x"], a["y"], t, Qt)
    return Qt


@pytest.mark.parametrize("testcase", [1, 8, 88, 89])
def test_manufactured_source_balances_the_residual(testcase):
    """Dirichlet_ConvTest{,3,4,5}.okl + S_ConvTest{,3,4,5}.okl: on the exact solution the stage
    residual plus the manufactured source is the exact time derivative up to the spatial truncation
    error, which falls at order about N under refinement.  A wrong mean depth H0, a wrong or missing
    wave-number factor k in the source's derivatives, or a sign error leaves an O(1) remainder."""
    N, t = 3, 0.7
    rp = RunParams(num_flux=0, bc_kind=S.CONVERGENCE_BC[testcase])
    errs = []
    for nel in (4, 8):
        c = Case.convergence(N=N, nel=nel, testcase=testcase)
        q_t = manufactured(c, t, testcase, dt=True)
        Qt = manufactured_residual(c, rp, manufactured(c, t, testcase), t)
        errs.append(np.abs(Qt - q_t).max() / np.abs(q_t).max())
    eoc = np.log2(errs[0] / errs[1])
    assert errs[1] < 5e-3 and eoc >= N - 0.6, (errs, eoc)
    # the pin is sharp: the source of another Testcase leaves an O(1) remainder
    other = {1: 8, 8: 1, 88: 89, 89: 88}[testcase]
    c = Case.convergence(N=N, nel=8, testcase=testcase)
    Qt = manufactured_residual(c, RunParams(num_flux=0, bc_kind=S.CONVERGENCE_BC[other]),
                               manufactured(c, t, testcase), t)
    q_t = manufactured(c, t, testcase, dt=True)
    assert np.abs(Qt - q_t).max() / np.abs(q_t).max() > 0.1


@pytest.mark.parametrize("N,eoc_min", [(2, 2.7), (3, 3.7)])
def test_manufactured_solution_converges_at_order_N_plus_1(N, eoc_min):
    """Whole stage + time loop (volume, ES flux, surface lift, Dirichlet exterior states, manufactured
    source, LSRK4) against the exact solution of the convergence test: max-norm error falls as
    h^(N+1).  A wrong sign or index anywhere in the residual stalls the convergence."""
    errs = []
    for nel in (4, 8):
        c = Case.convergence(N=N, nel=nel)
        rp = RunParams(num_flux=0, bc_kind=3, CFL=0.2, T=0.1)
        q, n, t, _ = O.timeloop(c, rp, 100000)
        assert t == 0.1
        errs.append(np.abs(q - manufactured(c, t)).max())
    eoc = np.log2(errs[0] / errs[1])
    assert eoc >= eoc_min, (errs, eoc)


def test_testcase_8_converges_at_order_N_plus_1():
    """Testcase 8 (Dirichlet_ConvTest3.okl + S_ConvTest3.okl, b = 2 + sin y + cos x, H0 = 20) through
    the oracle's time loop: the error against the exact solution falls as h^(N+1)."""
    errs = []
    for nel in (4, 8):
        c = Case.convergence(N=3, nel=nel, testcase=8)
        q, n, t, _ = O.timeloop(c, RunParams(num_flux=0, bc_kind=5, CFL=0.2, T=0.1), 100000)
        assert t == 0.1
        errs.append(np.abs(q - manufactured(c, t, 8)).max())
    eoc = np.log2(errs[0] / errs[1])
    assert eoc >= 3.7, (errs, eoc)


# --------------------------------------------------------------------------------------------
# analysis kernels and solution integrals (okl/Analysis, src/basis.cpp:1004-1111)
# --------------------------------------------------------------------------------------------
def test_integrals_closed_forms():
    """Mass of h = 1 is the domain area; the lake-at-rest entropy integral matches an independent
    tensor-product quadrature; the setup library's host diagnostics equal the oracle bit for bit."""
    case = Case.dam_break(N=4, nel=6, noise=0.0)
    p = O.params(case, RunParams())
    ngl, E = case.ngl, case.Nelem
    q = np.zeros((E, 3, ngl * ngl))
    q[:, 0] = 1.0
    out = np.zeros(1)
    O.call("total_mass", p, E, q.ravel(), case["Jac"], case["w_GL"], out)
    assert abs(out[0] - 100.0) <= 1e-12 * 100.0
    b = case["b"].reshape(E, ngl * ngl)
    q[:, 0] = 2.0 - b
    O.call("total_entropy", p, E, q.ravel(), case["b"], case["Jac"], case["w_GL"], out)
    w = case["w_GL"]
    W = np.outer(w, w).ravel()                      # node j*ngl + i -> w_j w_i
    J = 1.0 / case["Jac"].reshape(E, ngl * ngl)
    h = q[:, 0]
    ref = np.sum((0.5 * G * h * h + G * h * b) * J * W)
    assert abs(out[0] - ref) <= 1e-12 * abs(ref)
    assert out[0] == S.total_entropy(case, q.ravel(), G)
    O.call("total_mass", p, E, q.ravel(), case["Jac"], case["w_GL"], out)
    assert out[0] == S.total_mass(case, q.ravel())


def test_arrival_times_and_maximum_elevation():
    """Arrival time: set once, at the first call where |h + b - h_zero| > 5e-5; maximum
    elevation: the running max of h + b - h_zero, starting from 0."""
    case = Case.dam_break(N=3, nel=4, noise=0.0)
    rp = RunParams(h_zero=1.5)
    p = O.params(case, rp)
    E, n2 = case.Nelem, case.ngl ** 2
    b = case["b"].reshape(E, n2)
    rng = np.random.default_rng(3)
    arr = np.full(E * n2, 1000.0)
    mx = np.zeros(E * n2)
    q = np.zeros((E, 3, n2))
    q[:, 0] = 1.5 - b                                 # at rest: nothing arrives
    O.call("arrival_times", p, E, 0.25, q.ravel(), case["b"], arr)
    assert (arr == 1000.0).all()
    wet = rng.uniform(size=(E, n2)) < 0.3
    q[:, 0] = 1.5 - b + np.where(wet, 1e-3, 2e-5)     # 2e-5 stays below the 5e-5 threshold
    O.call("arrival_times", p, E, 0.5, q.ravel(), case["b"], arr)
    assert np.array_equal(arr.reshape(E, n2) == 0.5, wet) and (arr[~wet.ravel()] == 1000.0).all()
    q[:, 0] = 1.5 - b + 1e-2
    O.call("arrival_times", p, E, 0.75, q.ravel(), case["b"], arr)
    assert np.array_equal(arr.reshape(E, n2), np.where(wet, 0.5, 0.75))
    for level in (-0.2, 0.3, 0.1):
        q[:, 0] = 1.5 - b + level * rng.uniform(size=(E, n2))
        O.call("maximum_elevation", p, E, q.ravel(), case["b"], mx)
    # the running max from 0: recompute independently
    rng = np.random.default_rng(3)
    rng.uniform(size=(E, n2))
    ref = np.zeros((E, n2))
    for level in (-0.2, 0.3, 0.1):
        ref = np.maximum(ref, (1.5 - b + level * rng.uniform(size=(E, n2))) + b - 1.5)
    assert np.allclose(mx.reshape(E, n2), ref, rtol=0, atol=1e-15)


# --------------------------------------------------------------------------------------------
# wave speed, Lax-Friedrichs flux, standard volume term, SSP-RK3 update
# --------------------------------------------------------------------------------------------
def spectral_radii(q):
    """Spectral radii of the x- and y-flux Jacobians at states q[M, 3], from complex-step
    Jacobians of the physical flux and their numerical eigenvalues (no wave-speed formula used)."""
    def fx(s):
        h, hu, hv = s
        return np.array([hu, hu * hu / h + 0.5 * G * h * h, hu * hv / h])

    def fy(s):
        h, hu, hv = s
        return np.array([hv, hu * hv / h, hv * hv / h + 0.5 * G * h * h])

    out = np.zeros((len(q), 2))
    for m, s in enumerate(q):
        for k, f in enumerate((fx, fy)):
            Jm = np.zeros((3, 3))
            for c in range(3):
                e = np.zeros(3, complex)
                e[c] = 1e-30j
                Jm[:, c] = f(s + e).imag / 1e-30
            out[m, k] = np.abs(np.linalg.eigvals(Jm)).max()
    return out


def test_find_lambda_max_is_the_largest_flux_jacobian_eigenvalue():
    """FindLambdaMax.okl:33-50: per element, the largest eigenvalue magnitude of the x- and
    y-flux Jacobians over its nodes; a node with h <= PosPresTOL counts with zero velocity.
    |u|+|v| or sqrt(u^2+v^2) in place of max(|u|,|v|), or a missing c, fails this."""
    ngl, E, tol = 3, 40, 1e-3
    n2 = ngl * ngl
    rng = np.random.default_rng(11)
    h = rng.uniform(0.3, 2.0, (E, n2))
    u = rng.uniform(-2.0, 2.0, (E, n2)) * rng.choice([0.2, 1.0, 3.0], (E, 1))
    v = rng.uniform(-2.0, 2.0, (E, n2)) * rng.choice([0.2, 1.0, 3.0], (E, 1))
    q = np.stack([h, h * u, h * v], 1)                          # [e][var][node]
    lam = np.zeros(E)
    p = O.OracleParams(ngl, 1, G, 1.0, tol, 1e-5, 0.0, 0.0, 0.0)
    O.call("find_lambda_max", p, E, np.ascontiguousarray(q.ravel()), lam)
    states = np.moveaxis(q, 1, 2).reshape(-1, 3)
    ref = spectral_radii(states).max(1).reshape(E, n2).max(1)
    assert np.abs(lam - ref).max() <= 1e-13 * ref.max()
    # dry nodes: velocity ignored at h <= tol (h == tol included), counted above it
    qd = q.copy()
    qd[0, 0, :] = 0.5 * tol
    qd[0, 1:, :] = 50.0                                         # dry element, large momentum
    qd[1, 0, :] = tol
    qd[1, 1:, :] = 50.0
    qd[2, 0, :] = 2.0 * tol
    qd[2, 1, :] = 3.0 * tol
    qd[2, 2, :] = -1.0 * tol                                    # u = 1.5, v = -0.5: wet, counted
    O.call("find_lambda_max", p, E, np.ascontiguousarray(qd.ravel()), lam)
    assert lam[0] == np.sqrt(G * 0.5 * tol)
    assert lam[1] == np.sqrt(G * tol)
    ref2 = spectral_radii(np.moveaxis(qd[2:3], 1, 2).reshape(-1, 3)).max()
    assert abs(lam[2] - ref2) <= 1e-13 * ref2 and lam[2] > 1.5


def lf_call(qL, qR, n, scal=1.0, geom=1.0, tol=1e-12):
    M = len(qL)
    p = O.OracleParams(1, 1, G, geom, tol, 1e-5, 0.0, 0.0, 0.0)
    SP = np.zeros(3 * M)
    O.call("num_flux_lf", p, M, np.ones(M, np.int32), np.ascontiguousarray(n[:, 0]),
           np.ascontiguousarray(n[:, 1]), np.full(M, scal), np.ascontiguousarray(qL.reshape(-1)),
           np.ascontiguousarray(qR.reshape(-1)), np.zeros(M), np.zeros(M), SP)
    return SP.reshape(M, 3)


def test_lax_friedrichs_flux_consistency_conservation_and_dissipation():
    """calcLaxFriedrich.okl:17-170: F = n.(f(qL)+f(qR))/2 - lam/2 (qR - qL), with lam the largest
    flux-Jacobian eigenvalue magnitude of both states in both coordinate directions; F(q, q) =
    n.f(q); F(qL, qR, n) = -F(qR, qL, -n); the output is scaled by scal * geomFace."""
    qL, qR, n = random_faces(300, 21)
    F = lf_call(qL, qL, n)
    assert np.abs(F - phys_flux(qL, n)).max() < 1e-13
    F = lf_call(qL, qR, n)
    Fm = lf_call(qR, qL, -n)
    assert np.abs(F + Fm).max() < 1e-13
    lam = np.maximum(spectral_radii(qL).max(1), spectral_radii(qR).max(1))
    central = 0.5 * (phys_flux(qL, n) + phys_flux(qR, n))
    ref = central - 0.5 * lam[:, None] * (qR - qL)
    assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max()
    # the dissipation removes energy-norm of the jump: (qR - qL) . (F - central) < 0
    assert (((qR - qL) * (F - central)).sum(1) < 0).all()
    Fs = lf_call(qL, qR, n, scal=2.5, geom=3.0)
    assert np.abs(Fs - 7.5 * F).max() <= 1e-15 * np.abs(Fs).max()
    # a side with h <= PosPresTOL enters with zero momentum and zero wave speed
    tol = 1e-3
    qLd = qL.copy()
    qLd[:, 0] = tol
    qLd[:, 1:] = 7.0
    Fd = lf_call(qLd, qR, n, tol=tol)
    q0 = qLd.copy()
    q0[:, 1:] = 0.0
    lam_d = spectral_radii(qR).max(1)
    fL = np.stack([np.zeros(len(qL)), 0.5 * G * tol * tol * n[:, 0], 0.5 * G * tol * tol * n[:, 1]], 1)
    ref_d = 0.5 * (fL + phys_flux(qR, n)) - 0.5 * lam_d[:, None] * (qR - q0)
    assert np.abs(Fd - ref_d).max() <= 1e-13 * np.abs(ref_d).max()


def lgl_nodes(ngl):
    """LGL nodes: -1, 1 and the roots of P_N' (numpy Legendre)."""
    N = ngl - 1
    inner = np.sort(L.legroots(L.legder([0] * N + [1]))) if N > 1 else np.array([])
    return np.concatenate([[-1.0], inner, [1.0]])


def d0_matrix(case):
    """The plain nodal derivative matrix: Dstrong without its two corner terms."""
    ngl = case.ngl
    D0 = case["Dstrong"].reshape(ngl, ngl).copy()
    w = case["w_GL"]
    D0[0, 0] -= 1.0 / w[0]
    D0[-1, -1] += 1.0 / w[-1]
    return D0


def test_standard_volume_term_is_the_exact_divergence_on_affine_elements():
    """VolumeKernelV99.okl:17-167 with the plain derivative matrix returns -(df/dx + dg/dy) -
    g h (0, Bx, By) exactly for fluxes of degree <= N on skewed affine elements (the divergence is
    formed here by complex-step differentiation of the physical flux), and zero for a constant
    state on the curved mesh.  A transposed D, a swapped metric or a dropped term fails this."""
    case = Case.gaussian_bump(N=4, nel=2)
    ngl, n2 = case.ngl, case.ngl ** 2
    D0 = d0_matrix(case)
    assert np.abs(D0 @ lgl_nodes(ngl) - 1.0).max() < 1e-13           # D0 differentiates x exactly
    xi = lgl_nodes(ngl)
    XI, ETA = np.meshgrid(xi, xi)                                      # [eta row][xi col]
    rng = np.random.default_rng(5)
    maps = [(0.3, 0.7, 0.9, 0.2, -0.25, 1.1), (-0.2, 0.4, 1.3, -0.5, 0.6, 0.8), (1.0, -1.0, 0.5, 0.0, 0.0, 0.5)]
    E = len(maps)

    def state(x, y):
        h = 1.5 + 0.3 * x - 0.2 * y + 0.1 * x * y - 0.05 * y * y
        u = 0.4 - 0.3 * x + 0.2 * y
        v = -0.1 + 0.25 * x + 0.15 * y
        return h, h * u, h * v

    def fluxes(x, y):
        h, hu, hv = state(x, y)
        f = [hu, hu * hu / h + 0.5 * G * h * h, hu * hv / h]
        g = [hv, hu * hv / h, hv * hv / h + 0.5 * G * h * h]
        return f, g

    arr = {k: np.zeros((E, n2)) for k in ("xXi", "xEta", "yXi", "yEta", "Jac", "Bx", "By")}
    q = np.zeros((E, 3, n2))
    ref = np.zeros((E, 3, n2))
    for e, (x0, y0, a, b, c, d) in enumerate(maps):
        x = (x0 + a * XI + b * ETA).ravel()
        y = (y0 + c * XI + d * ETA).ravel()
        arr["xXi"][e], arr["xEta"][e], arr["yXi"][e], arr["yEta"][e] = a, b, c, d
        arr["Jac"][e] = 1.0 / (a * d - b * c)
        arr["Bx"][e] = rng.uniform(-1, 1, n2)
        arr["By"][e] = rng.uniform(-1, 1, n2)
        q[e] = np.array(state(x, y))
        hstep = 1e-30
        fx = np.array(fluxes(x + 1j * hstep, y)[0]).imag / hstep
        gy = np.array(fluxes(x, y + 1j * hstep)[1]).imag / hstep
        ref[e] = -(fx + gy)
        ref[e, 1] -= G * q[e, 0] * arr["Bx"][e]
        ref[e, 2] -= G * q[e, 0] * arr["By"][e]
    Qt = np.zeros(E * 3 * n2)
    p = O.OracleParams(ngl, 1, G, 1.0, 1e-12, 1e-5, 0.0, 0.0, 0.0)
    O.call("volume_std", p, E, *(np.ascontiguousarray(arr[k].ravel()) for k in ("Jac", "yXi", "yEta", "xXi", "xEta")),
           np.ascontiguousarray(q.ravel()), np.ascontiguousarray(D0.ravel()),
           np.ascontiguousarray(arr["Bx"].ravel()), np.ascontiguousarray(arr["By"].ravel()), Qt)
    assert np.abs(Qt - ref.ravel()).max() <= 1e-12 * np.abs(ref).max()
    # free stream on the curved mesh (flat bottom): the discrete metric identities cancel every term
    cv = Case.dam_break(N=5, nel=3, seed=2)
    qc = np.zeros((cv.Nelem, 3, cv.ngl ** 2))
    qc[:, 0], qc[:, 1], qc[:, 2] = 1.3, 0.4, -0.2
    Qc = np.zeros(qc.size)
    z = np.zeros(cv.Nelem * cv.ngl ** 2)
    a_ = cv.arrays
    O.call("volume_std", O.params(cv, RunParams()), cv.Nelem, a_["Jac"], a_["yXi"], a_["yEta"], a_["xXi"],
           a_["xEta"], np.ascontiguousarray(qc.ravel()), np.ascontiguousarray(d0_matrix(cv).ravel()), z, z, Qc)
    assert np.abs(Qc).max() < 1e-14 * term_scale(cv, qc.ravel())


def test_ssp_rk3_update_is_third_order():
    """UpdateKernelSSP_V1.okl:17-98 (q = a qTmp + b q + c dt Qt, qTmp the step's start state)
    driven with the reference's SSP table (RungeKutta.cpp:25-38) integrates y' = lam y at third
    order; swapping a and b, or dropping dt, breaks the order."""
    from esdgsem_mpiocca_amd import rk_coefficients
    A, B, Cc = rk_coefficients(3, 1)
    assert len(A) == 3
    lam = -1.1 + 0.9j
    p = O.OracleParams(1, 1, G, 1.0, 0.0, 1e-5, 0.0, 0.0, 0.0)
    errs = []
    for nsteps in (20, 40, 80):
        dt = 1.0 / nsteps
        y = np.array([1.0, 0.0, 0.0])
        for _ in range(nsteps):
            y0 = y.copy()
            for s in range(3):
                f = lam * (y[0] + 1j * y[1])
                Qt = np.array([f.real, f.imag, 0.0])
                O.call("rk_ssp_update", p, 1, float(A[s]), float(B[s]), float(Cc[s]), dt, Qt, y0, y)
        errs.append(abs(y[0] + 1j * y[1] - np.exp(lam)))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert (orders > 2.85).all() and (orders < 3.3).all(), orders
