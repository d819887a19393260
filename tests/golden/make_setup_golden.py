"""Write tests/golden/setup_golden.npz from the reference's own setup code (oracle/_ref).

Run in a checkout where oracle/_ref/libesdg_ref_setup.so is built (`make -C oracle ref`, which
needs the reference sources).  The fixture holds only data: the LGL basis operators for N = 1..10,
the RK tables, the 16x16 periodic Cartesian N = 4 mesh arrays with the Gaussian-bump
bathymetry gradient, and the N = 4 geometry of the two ISM-V2 fixture meshes (tests/golden/meshes)
-- every array as the reference computes it.
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from esdgsem_mpiocca_amd import Case  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libesdg_ref_setup.so")


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def main():
    ref = C.CDLL(REF)
    out = {}
    for N in range(1, 11):
        ngl = N + 1
        r = {k: np.zeros(ngl if k in ("x_GL", "w_GL") else ngl * ngl)
             for k in ("x_GL", "w_GL", "D", "D0", "Dstrong", "Dhat", "VdmInv")}
        ref.ref_basis(N, 1, *[P(r[k]) for k in ("x_GL", "w_GL", "D", "D0", "Dstrong", "Dhat", "VdmInv")])
        for k, v in r.items():
            out[f"basis{N}_{k}"] = v
    for order, ssp in ((3, 0), (3, 1), (4, 0)):
        A, B, Cc = (np.zeros(5) for _ in range(3))
        st = C.c_int()
        ref.ref_rk(order, ssp, C.byref(st), P(A), P(B), P(Cc))
        n = st.value
        out[f"rk{order}{ssp}"] = np.concatenate([A[:n], B[:n], Cc[:n]])
    N, nel = 4, 16
    ngl = N + 1
    nn, Nf = nel * nel * ngl * ngl, 2 * nel * nel
    arr = {k: np.zeros(nn) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J")}
    fa = {k: np.zeros(Nf * ngl) for k in ("nx", "ny", "scal")}
    EI = np.zeros(Nf * 7, np.int32)
    E2E, MS, OR = (np.zeros(nel * nel * 4, np.int32) for _ in range(3))
    nf = C.c_int()
    with tempfile.TemporaryDirectory() as d:
        rc = ref.ref_cartesian_mesh(d.encode(), N, nel, nel, C.c_double(0.0), C.c_double(2.0), C.c_double(0.0),
                                    C.c_double(2.0), 1, 1,
                                    *[P(arr[k]) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J")],
                                    *[P(fa[k]) for k in ("nx", "ny", "scal")], P(EI), P(E2E), P(MS), P(OR),
                                    C.byref(nf))
    assert rc == 0
    EI = EI.reshape(-1, 7)
    ED = np.zeros((Nf, 8), np.int32)
    ED[:, 0:5] = EI[:, 2:7]
    ED[:, 6] = np.where(EI[:, 3] == -1, -1, 0)
    ED[:, 7] = np.arange(Nf)
    for k in ("x", "y", "xXi", "xEta", "yXi", "yEta"):
        out[f"bump_{k}"] = arr[k]
    out["bump_Jac"] = 1.0 / arr["J"]
    for k in ("nx", "ny", "scal"):
        out[f"bump_{k}"] = fa[k]
    out["bump_EdgeData"] = ED.ravel()
    out["bump_ElemToEdge"] = E2E
    out["bump_ElemEdgeMasterSlave"] = MS
    out["bump_ElemEdgeOrientation"] = OR
    # bathymetry gradient and element sizes of the reference on the same mesh + bump bathymetry
    c = Case.gaussian_bump(N=N, nel=nel)
    b = c["b"]
    Bx, By = np.zeros(nn), np.zeros(nn)
    ref.ref_bathymetry_gradient(nel * nel, ngl, C.c_double(9.81), P(arr["x"]), P(arr["y"]), P(b), P(out["basis4_D0"]),
                                P(arr["yEta"]), P(arr["yXi"]), P(arr["xEta"]), P(arr["xXi"]), P(Bx), P(By),
                                P(out["bump_Jac"]))
    out["bump_Bx"], out["bump_By"] = Bx, By
    sizes = np.zeros(nel * nel)
    m = C.c_double()
    ref.ref_element_sizes(N, nel * nel, P(out["bump_Jac"]), P(sizes), C.byref(m))
    out["bump_EleSizes"] = sizes
    # the reference's reader on the two fixture meshes (tests/golden/meshes), N = 4
    for name in ("PeriodicXandY3", "CircleInCircle"):
        mpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes", name + ".mesh")
        ne, nf = C.c_int(), C.c_int()
        ref.ref_ism_mesh(mpath.encode(), 4, 0, 0, C.byref(ne), C.byref(nf), *([None] * 15))
        mn = ne.value * 25
        a = {k: np.zeros(mn) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J", "b")}
        f = {k: np.zeros(nf.value * 5) for k in ("nx", "ny", "scal")}
        ei = np.zeros(nf.value * 7, np.int32)
        e2e, ms, orr = (np.zeros(ne.value * 4, np.int32) for _ in range(3))
        ref.ref_ism_mesh(mpath.encode(), 4, 0, 0, C.byref(ne), C.byref(nf),
                         *[P(a[k]) for k in ("x", "y", "xXi", "xEta", "yXi", "yEta", "J", "b")],
                         *[P(f[k]) for k in ("nx", "ny", "scal")], P(ei), P(e2e), P(ms), P(orr))
        for k in ("x", "y", "xXi", "xEta", "yXi", "yEta"):
            out[f"ism_{name}_{k}"] = a[k]
        out[f"ism_{name}_Jac"] = 1.0 / a["J"]
        for k in ("nx", "ny", "scal"):
            out[f"ism_{name}_{k}"] = f[k]
        out[f"ism_{name}_EdgeInfo"] = ei
        out[f"ism_{name}_ElemToEdge"] = e2e
        out[f"ism_{name}_ElemEdgeMasterSlave"] = ms
        out[f"ism_{name}_ElemEdgeOrientation"] = orr
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "setup_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
