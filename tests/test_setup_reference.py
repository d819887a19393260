"""Product host setup vs the reference's own setup code (oracle/_ref), bit for bit.

oracle/_ref/libesdg_ref_setup.so is built from the reference's src/basis.cpp, src/RungeKutta.cpp,
src/Mesh.cpp, src/SW2D.cpp and src/ReadInput.cpp (compiled where they lie) plus a small driver.  It
exists only where the reference sources are present; elsewhere the same arrays are checked
against tests/golden/setup_golden.npz, written by tests/golden/make_setup_golden.py from that build.
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case, basis, rk_coefficients, total_entropy, total_mass
from esdgsem_mpiocca_amd import setup_api as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libesdg_ref_setup.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "setup_golden.npz")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def same_bits(a, b):
    return np.array_equal(bits(a), bits(b))


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return C.CDLL(REF)


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def ref_basis(ref, N):
    ngl = N + 1
    r = {k: np.zeros(ngl if k in ("x_GL", "w_GL") else ngl * ngl)
         for k in ("x_GL", "w_GL", "D", "D0", "Dstrong", "Dhat", "VdmInv")}
    ref.ref_basis(N, 1, *[P(r[k]) for k in ("x_GL", "w_GL", "D", "D0", "Dstrong", "Dhat", "VdmInv")])
    return r


@pytest.mark.parametrize("N", range(1, 11))
def test_basis_bit_exact(ref, N):
    r = ref_basis(ref, N)
    m = basis(N)
    for k in r:
        assert same_bits(r[k], m[k]), k


@pytest.mark.parametrize("order,ssp", [(3, 0), (3, 1), (4, 0)])
def test_rk_bit_exact(ref, order, ssp):
    A, B, Cc = (np.zeros(5) for _ in range(3))
    st = C.c_int()
    ref.ref_rk(order, ssp, C.byref(st), P(A), P(B), P(Cc))
    a, b, c = rk_coefficients(order, ssp)
    n = st.value
    assert same_bits(A[:n], a) and same_bits(B[:n], b) and same_bits(Cc[:n], c)


def ref_cartesian(ref, N, nx, ny, xL, xR, yL, yR, px, py):
    ngl = N + 1
    Ne = nx * ny
    nn = Ne * ngl * ngl
    Nf = (nx + 1) * ny + (ny + 1) * nx - (nx if px else 0) - (ny if py else 0)
    arr = {k: np.zeros(nn) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J")}
    fa = {k: np.zeros(Nf * ngl) for k in ("nx", "ny", "scal")}
    EI = np.zeros(Nf * 7, np.int32)
    E2E, MS, OR = (np.zeros(Ne * 4, np.int32) for _ in range(3))
    nf = C.c_int()
    with tempfile.TemporaryDirectory() as d:
        rc = ref.ref_cartesian_mesh(d.encode(), N, nx, ny, C.c_double(xL), C.c_double(xR), C.c_double(yL),
                                    C.c_double(yR), px, py,
                                    *[P(arr[k]) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J")],
                                    *[P(fa[k]) for k in ("nx", "ny", "scal")], P(EI), P(E2E), P(MS), P(OR),
                                    C.byref(nf))
    assert rc == 0 and nf.value == Nf
    return arr, fa, EI.reshape(-1, 7), E2E, MS, OR


CART = [(4, 16, 16, 0.0, 2.0, 0.0, 2.0, 1, 1), (7, 5, 3, 0.0, 10.0, 0.0, 10.0, 0, 0),
        (3, 4, 6, -5.0, 5.0, -1.0, 3.0, 1, 0), (5, 3, 4, -5.0, 5.0, -1.0, 3.0, 0, 1)]


@pytest.mark.parametrize("spec", CART, ids=[f"N{c[0]}_{c[1]}x{c[2]}_p{c[7]}{c[8]}" for c in CART])
def test_cartesian_mesh_bit_exact(ref, spec):
    N, nx, ny, xL, xR, yL, yR, px, py = spec
    arr, fa, EI, E2E, MS, OR = ref_cartesian(ref, *spec)
    c = Case(S.CASE_GAUSSIAN_BUMP, S.MESH_CARTESIAN, N, nx, ny, xL, xR, yL, yR, px, py)
    for k in ("x", "y", "xXi", "xEta", "yXi", "yEta"):
        assert same_bits(arr[k], c[k]), k
    assert same_bits(1.0 / arr["J"], c["Jac"])          # main.cpp:700
    for k in ("nx", "ny", "scal"):
        assert same_bits(fa[k], c[k]), k
    ED = c["EdgeData"].reshape(-1, 8)
    # EdgeData = MyEdgeInfo[2..9] (MeshPartitioning.cpp:104-131, main.cpp:653-663)
    assert np.array_equal(ED[:, 0], EI[:, 2]) and np.array_equal(ED[:, 1], EI[:, 3])
    assert np.array_equal(ED[:, 2], EI[:, 4]) and np.array_equal(ED[:, 3], EI[:, 5])
    assert np.array_equal(ED[:, 4], EI[:, 6]) and np.array_equal(ED[:, 5], np.zeros(len(ED)))
    assert np.array_equal(ED[:, 6], np.where(EI[:, 3] == -1, -1, 0))
    assert np.array_equal(ED[:, 7], np.arange(len(ED)))
    assert np.array_equal(c["EdgeReversed"], EI[:, 6])
    assert np.array_equal(c["ElemToEdge"], E2E)
    assert np.array_equal(c["ElemEdgeMasterSlave"], MS)
    assert np.array_equal(c["ElemEdgeOrientation"], OR)


CASES = {
    "bump_cart": lambda: Case.gaussian_bump(N=4, nel=16),
    "dam_curved": lambda: Case.dam_break(N=7, nel=6, seed=2),
    "flood_curved": lambda: Case.flood(N=5, nel=5, seed=3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_bathymetry_gradient_bit_exact(ref, name):
    c = CASES[name]()
    n = c.Nelem * c.ngl ** 2
    Bx, By = np.zeros(n), np.zeros(n)
    ref.ref_bathymetry_gradient(c.Nelem, c.ngl, C.c_double(9.81), P(c["x"]), P(c["y"]), P(c["b"]), P(c["D0"]),
                                P(c["yEta"]), P(c["yXi"]), P(c["xEta"]), P(c["xXi"]), P(Bx), P(By), P(c["Jac"]))
    assert same_bits(Bx, c["Bx"]) and same_bits(By, c["By"])


@pytest.mark.parametrize("name", list(CASES))
def test_element_sizes_and_diagnostics_bit_exact(ref, name):
    c = CASES[name]()
    sizes = np.zeros(c.Nelem)
    m = C.c_double()
    ref.ref_element_sizes(c.N, c.Nelem, P(c["Jac"]), P(sizes), C.byref(m))
    assert same_bits(sizes, c["EleSizes"])
    assert same_bits(np.array([m.value]), np.array([c.min_ele_size]))
    ent, mass = C.c_double(), C.c_double()
    ref.ref_total_entropy_mass(c.N, c.Nelem, C.c_double(9.81), P(c["q"]), P(c["b"]), P(c["Jac"]),
                               C.byref(ent), C.byref(mass))
    assert ent.value == total_entropy(c, c["q"], 9.81)
    assert mass.value == total_mass(c, c["q"])


def test_golden_setup_fixture():
    """The same pins without the reference build: committed fixture written from oracle/_ref."""
    if not os.path.exists(GOLDEN):
        pytest.skip("golden fixture not generated")
    g = np.load(GOLDEN, allow_pickle=False)
    for N in range(1, 11):
        m = basis(N)
        for k in ("x_GL", "w_GL", "D", "D0", "Dstrong", "Dhat", "VdmInv"):
            assert same_bits(g[f"basis{N}_{k}"], m[k]), (N, k)
    for order, ssp in ((3, 0), (3, 1), (4, 0)):
        a, b, c = rk_coefficients(order, ssp)
        assert same_bits(g[f"rk{order}{ssp}"], np.concatenate([a, b, c]))
    c = Case.gaussian_bump(N=4, nel=16)
    for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "Jac", "nx", "ny", "scal", "Bx", "By", "EleSizes"):
        assert same_bits(g[f"bump_{k}"], c[k]), k
    for k in ("EdgeData", "ElemToEdge", "ElemEdgeMasterSlave", "ElemEdgeOrientation"):
        assert np.array_equal(g[f"bump_{k}"], c[k]), k
    for name in ("PeriodicXandY3", "CircleInCircle"):
        c = Case.from_ism(os.path.join(ROOT, "tests", "golden", "meshes", name + ".mesh"), 4)
        for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "Jac", "nx", "ny", "scal"):
            assert same_bits(g[f"ism_{name}_{k}"], c[k]), (name, k)
        EI = g[f"ism_{name}_EdgeInfo"].reshape(-1, 7)
        assert np.array_equal(c["EdgeData"].reshape(-1, 8)[:, :5], EI[:, 2:7]), name
        for k in ("ElemToEdge", "ElemEdgeMasterSlave", "ElemEdgeOrientation"):
            assert np.array_equal(g[f"ism_{name}_{k}"], c[k]), (name, k)


# --------------------------------------------------------------------------------------------
# ISM-V2 meshes (Mesh::ReadMesh, src/Mesh.cpp:698-1808) against the reference's own reader
# --------------------------------------------------------------------------------------------
MESH_DIR = "/root/reference/meshFiles"          # the reference's shipped mesh files (read here only)
FIXTURE_MESHES = os.path.join(ROOT, "tests", "golden", "meshes")


def ref_ism(ref, path, N, read_bottom=0, to_km=0):
    ne, nf = C.c_int(), C.c_int()
    null = [None] * 15
    assert ref.ref_ism_mesh(path.encode(), N, read_bottom, to_km, C.byref(ne), C.byref(nf), *null) == 0
    Ne, Nf, ngl = ne.value, nf.value, N + 1
    nn = Ne * ngl * ngl
    arr = {k: np.zeros(nn) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J", "b")}
    fa = {k: np.zeros(Nf * ngl) for k in ("nx", "ny", "scal")}
    EI = np.zeros(Nf * 7, np.int32)
    E2E, MS, OR = (np.zeros(Ne * 4, np.int32) for _ in range(3))
    assert ref.ref_ism_mesh(path.encode(), N, read_bottom, to_km, C.byref(ne), C.byref(nf),
                            *[P(arr[k]) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J", "b")],
                            *[P(fa[k]) for k in ("nx", "ny", "scal")], P(EI), P(E2E), P(MS), P(OR)) == 0
    return arr, fa, EI.reshape(-1, 7), E2E, MS, OR


def check_ism(ref, path, N, read_bottom=0, to_km=0):
    arr, fa, EI, E2E, MS, OR = ref_ism(ref, path, N, read_bottom, to_km)
    c = Case.from_ism(path, N, read_bottom=read_bottom, convert_to_km=to_km)
    assert c.Nelem * c.ngl ** 2 == len(arr["x"]) and c.Nfaces == len(EI)
    for k in ("x", "y", "xXi", "xEta", "yXi", "yEta"):
        assert same_bits(arr[k], c[k]), k
    assert same_bits(1.0 / arr["J"], c["Jac"])
    for k in ("nx", "ny", "scal"):
        assert same_bits(fa[k], c[k]), k
    if read_bottom:
        assert same_bits(arr["b"], c["b"])
    ED = c["EdgeData"].reshape(-1, 8)
    assert np.array_equal(ED[:, :5], EI[:, 2:7])
    assert np.array_equal(ED[:, 6], np.where(EI[:, 3] == -1, -1, 0))
    assert np.array_equal(c["ElemToEdge"], E2E)
    assert np.array_equal(c["ElemEdgeMasterSlave"], MS)
    assert np.array_equal(c["ElemEdgeOrientation"], OR)
    return c


def shipped_meshes():
    return sorted(os.listdir(MESH_DIR)) if os.path.isdir(MESH_DIR) else []


@pytest.mark.parametrize("N", [3, 7])
@pytest.mark.parametrize("name", shipped_meshes())
def test_ism_mesh_bit_exact(ref, name, N):
    """Every mesh file the reference ships, read by the product and by the reference."""
    c = check_ism(ref, os.path.join(MESH_DIR, name), N)
    # the geometry is sound: positive Jacobians, unit normals, sum of scal w = side lengths > 0
    assert (c["Jac"] > 0).all()
    assert np.allclose(np.hypot(c["nx"], c["ny"]), 1.0, rtol=0, atol=1e-14)


def with_bottom_column(src, dst):
    """An ISM-V2 file with a third (bathymetry) column on node and curve lines:
    b = 0.1 x - 0.05 y + 0.02 x y, evaluated from the file's own coordinates."""
    lines = open(src).read().splitlines()
    nn, ne, nel, order = (int(v) for v in lines[1].split())
    bfun = lambda x, y: 0.1 * x - 0.05 * y + 0.02 * x * y      # noqa: E731
    out = lines[:2]
    i = 2
    for _ in range(nn):
        x, y = (float(v) for v in lines[i].split()[:2])
        out.append(f"{lines[i]} {bfun(x, y)!r}")
        i += 1
    out += lines[i:i + ne]
    i += ne
    for _ in range(nel):
        out.append(lines[i]); out.append(lines[i + 1])
        flags = [int(v) for v in lines[i + 1].split()]
        i += 2
        for f in flags:
            if f:
                for _ in range(order + 1):
                    x, y = (float(v) for v in lines[i].split()[:2])
                    out.append(f"{lines[i]} {bfun(x, y)!r}")
                    i += 1
        out.append(lines[i])
        i += 1
    open(dst, "w").write("\n".join(out) + "\n")


@pytest.mark.parametrize("name", ["PeriodicXandY3.mesh", "CircleInCircle.mesh"])
def test_ism_mesh_bottom_column_bit_exact(ref, name, tmp_path):
    """ReadBottom: b from the file's third column, mapped like the coordinates."""
    dst = str(tmp_path / name)
    with_bottom_column(os.path.join(FIXTURE_MESHES, name), dst)
    c = check_ism(ref, dst, 4, read_bottom=1)
    assert np.ptp(c["b"]) > 0.1


def test_ism_fixture_meshes_match_the_shipped_files():
    """The fixture copies under tests/golden/meshes are the reference's files, unchanged."""
    if not shipped_meshes():
        pytest.skip("reference mesh files absent")
    for name in os.listdir(FIXTURE_MESHES):
        assert open(os.path.join(FIXTURE_MESHES, name), "rb").read() == \
            open(os.path.join(MESH_DIR, name), "rb").read()


def test_ism_mesh_degrees_to_km_bit_exact(ref):
    """ConvertToKM: the fixture mesh read as degrees (it spans +-1 degree) and projected to km."""
    c = check_ism(ref, os.path.join(FIXTURE_MESHES, "PeriodicXandY3.mesh"), 4, to_km=1)
    assert 200.0 < np.ptp(c["x"]) < 250.0        # 2 degrees of longitude at the equator: 222.6 km
