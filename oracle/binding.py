"""ctypes binding of the TEST ORACLE (oracle/liboracle.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OracleParams(C.Structure):
    _fields_ = [("ngl", C.c_int), ("ES", C.c_int), ("g", C.c_double), ("geomFace", C.c_double),
                ("PosPresTOL", C.c_double), ("ZeroTOL", C.c_double), ("eps0", C.c_double),
                ("sigmaMin", C.c_double), ("sigmaMax", C.c_double), ("bc_kind", C.c_int),
                ("h_zero", C.c_double)]


_RUN_DOUBLES = ("Jac", "Yxi", "Yeta", "Xxi", "Xeta", "Bx", "By", "B", "D", "Dhat", "Dstrong",
                "VdmInv", "nx", "ny", "scal", "EleSizes", "GLw")
_RUN_INTS = ("EdgeData", "EdgeReversed", "ElemToEdge", "EMS", "EOR")
_RUN_COORDS = ("x", "y")


class OracleRun(C.Structure):
    _fields_ = ([("Nelem", C.c_int), ("Nfaces", C.c_int), ("CFL", C.c_double), ("DFL", C.c_double),
                 ("T", C.c_double), ("min_ele_size", C.c_double), ("num_flux", C.c_int),
                 ("artificial_viscosity", C.c_int), ("positivity", C.c_int), ("rk_order", C.c_int),
                 ("rk_ssp", C.c_int), ("max_steps", C.c_int)]
                + [(n, C.POINTER(C.c_double)) for n in _RUN_DOUBLES]
                + [(n, C.POINTER(C.c_int)) for n in _RUN_INTS]
                + [(n, C.POINTER(C.c_double)) for n in _RUN_COORDS])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build with `make -C oracle`")
        _LIB = C.CDLL(path)
    return _LIB


def params(case, rp) -> OracleParams:
    return OracleParams(case.ngl, rp.ES, rp.g, case.geom_face, rp.pos_pres_tol, rp.zero_tol,
                        rp.eps0, rp.sigma_min, rp.sigma_max, rp.bc_kind, rp.h_zero)


def _arg(a):
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, np.ndarray):
        if a.dtype == np.float64:
            assert a.flags.c_contiguous
            return a.ctypes.data_as(C.c_void_p)
        if a.dtype == np.int32:
            assert a.flags.c_contiguous
            return a.ctypes.data_as(C.c_void_p)
        raise TypeError(a.dtype)
    if isinstance(a, (bool, int, np.integer)):
        return C.c_int(int(a))
    if isinstance(a, (float, np.floating)):
        return C.c_double(float(a))
    raise TypeError(type(a))


def call(name: str, p: OracleParams, *args):
    """Call oracle_<name>(&p, args...) with numpy arrays (modified in place)."""
    fn = getattr(lib(), "oracle_" + name)
    fn.restype = None
    fn(C.byref(p), *[_arg(a) for a in args])


def case_run_inputs(case, rp, steps, T=None):
    a = case.arrays
    m = {"Jac": a["Jac"], "Yxi": a["yXi"], "Yeta": a["yEta"], "Xxi": a["xXi"], "Xeta": a["xEta"],
         "Bx": a["Bx"], "By": a["By"], "B": a["b"], "D": a["D"], "Dhat": a["Dhat"],
         "Dstrong": a["Dstrong"], "VdmInv": a["VdmInv"], "nx": a["nx"], "ny": a["ny"],
         "scal": a["scal"], "EleSizes": a["EleSizes"], "GLw": a["w_GL"],
         "EdgeData": a["EdgeData"], "EdgeReversed": a["EdgeReversed"],
         "ElemToEdge": a["ElemToEdge"], "EMS": a["ElemEdgeMasterSlave"],
         "EOR": a["ElemEdgeOrientation"]}
    r = OracleRun()
    r.Nelem, r.Nfaces = case.Nelem, case.Nfaces
    r.CFL, r.DFL = rp.CFL, rp.DFL
    r.T = rp.T if T is None else T
    r.min_ele_size = case.min_ele_size
    r.num_flux, r.artificial_viscosity, r.positivity = rp.num_flux, rp.artificial_viscosity, rp.positivity
    r.rk_order, r.rk_ssp, r.max_steps = rp.rk_order, rp.rk_ssp, steps
    keep = {}
    for n in _RUN_DOUBLES:
        keep[n] = np.ascontiguousarray(m[n], dtype=np.float64)
        setattr(r, n, keep[n].ctypes.data_as(C.POINTER(C.c_double)))
    for n in _RUN_INTS:
        keep[n] = np.ascontiguousarray(m[n], dtype=np.int32)
        setattr(r, n, keep[n].ctypes.data_as(C.POINTER(C.c_int)))
    for n in _RUN_COORDS:
        keep[n] = np.ascontiguousarray(a[n], dtype=np.float64)
        setattr(r, n, keep[n].ctypes.data_as(C.POINTER(C.c_double)))
    return r, keep


def timeloop(case, rp, steps, q=None, T=None):
    """Advance `steps` steps (or to T) from q (default: the case's initial state).
    Returns (q, steps_taken, t, dt_history)."""
    if not rp.flux_differencing:
        raise ValueError("the oracle time loop follows the flux-differencing configuration")
    r, keep = case_run_inputs(case, rp, steps, T)
    q = np.array(case["q"] if q is None else q, dtype=np.float64, copy=True)
    t = C.c_double()
    dts = np.zeros(max(steps, 1))
    fn = lib().oracle_timeloop
    fn.restype = C.c_int
    p = params(case, rp)
    n = fn(C.byref(p), C.byref(r), q.ctypes.data_as(C.c_void_p), C.byref(t),
           dts.ctypes.data_as(C.c_void_p))
    return q, n, t.value, dts[:n]
