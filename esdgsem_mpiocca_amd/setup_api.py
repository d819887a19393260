"""ctypes binding of libesdg_setup.so (include/esdg_setup.h): basis, RK tables and test cases.

A :class:`Case` owns the host arrays that the reference builds in main.cpp before
``deviceclass::copyDeviceVariables`` (geometry, connectivity, bathymetry, initial state), and a
:class:`RunParams` holds the run flags of input.dat that the hot path reads.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class EsdgError(RuntimeError):
    pass


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_PKG, "libesdg_setup.so")
        if not os.path.exists(path):
            raise EsdgError(f"{path} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(path)
        lib.esdg_setup_last_error.restype = C.c_char_p
        lib.esdg_case_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        lib.esdg_case_create.argtypes = [C.POINTER(CaseSpec), C.POINTER(C.c_void_p)]
        lib.esdg_case_get_info.argtypes = [C.c_void_p, C.POINTER(CaseInfo)]
        lib.esdg_case_destroy.argtypes = [C.c_void_p]
        lib.esdg_case_destroy.restype = None
        _LIB = lib
    return _LIB


def _check(rc):
    if rc != 0:
        raise EsdgError(_lib().esdg_setup_last_error().decode())


class CaseSpec(C.Structure):
    _fields_ = [
        ("case_kind", C.c_int), ("mesh_kind", C.c_int), ("N", C.c_int),
        ("nelx", C.c_int), ("nely", C.c_int),
        ("xL", C.c_double), ("xR", C.c_double), ("yL", C.c_double), ("yR", C.c_double),
        ("periodic_x", C.c_int), ("periodic_y", C.c_int),
        ("perturb_amp", C.c_double), ("perturb_len", C.c_double), ("g", C.c_double),
        ("amp", C.c_double), ("p0", C.c_double), ("p1", C.c_double), ("p2", C.c_double),
        ("p3", C.c_double), ("seed", C.c_uint64), ("mesh_file", C.c_char_p), ("read_bottom", C.c_int),
        ("convert_to_km", C.c_int),
    ]


class CaseInfo(C.Structure):
    _fields_ = [("N", C.c_int), ("ngl", C.c_int), ("Nelem", C.c_int), ("Nfaces", C.c_int),
                ("min_ele_size", C.c_double)]


CASE_GAUSSIAN_BUMP, CASE_DAM_BREAK, CASE_FLOOD, CASE_OCEAN_UPLIFT, CASE_CONSTANT, CASE_CONVERGENCE = range(6)
MESH_CARTESIAN, MESH_CURVED_SINE, MESH_ISM = 0, 1, 2

# boundary kind (Dirichlet exterior state + manufactured source) of each convergence Testcase
# (src/deviceclass.cpp:363-387): Dirichlet_ConvTest.okl, ...ConvTest3, ...ConvTest4, ...ConvTest5
CONVERGENCE_BC = {1: 3, 8: 5, 88: 6, 89: 7}

DOUBLE_ARRAYS = ("x", "y", "xXi", "xEta", "yXi", "yEta", "Jac", "b", "Bx", "By", "q", "nx", "ny",
                 "scal", "EleSizes", "x_GL", "w_GL", "D", "D0", "Dstrong", "Dhat", "VdmInv")
INT_ARRAYS = ("EdgeData", "EdgeReversed", "ElemToEdge", "ElemEdgeMasterSlave",
              "ElemEdgeOrientation")


@dataclasses.dataclass
class RunParams:
    """Run flags and constants of the hot path (input.dat lines 2-30; src/deviceclass.cpp:131-218)."""
    g: float = 9.81
    CFL: float = 0.5
    DFL: float = 0.1
    T: float = 1.0e30
    ES: int = 1
    num_flux: int = 0           # 0 ES (calcESFlux), 3 rotated ES (calcESFluxRotated)
    flux_differencing: int = 1
    artificial_viscosity: int = 0
    positivity: int = 0
    pos_pres_tol: float = 1.0e-8
    eps0: float = 0.0
    sigma_min: float = -4.0
    sigma_max: float = -2.0
    rk_order: int = 4
    rk_ssp: int = 0
    zero_tol: float = 1.0e-5    # hard-coded in the reference (src/deviceclass.cpp:148)
    bc_kind: int = 0            # 0 SolidWalls, 1 InnerOuter, 2 OceanBoundary, 3 and 5-7 manufactured
                                # solutions, 4 3MoundInflow (deviceclass.cpp:363-415)
    h_zero: float = 0.0         # still-water level of the ocean boundary (define h_zero)
    arrival_times: int = 0      # calcArrivalTimes (deviceclass.cpp:1132-1135)
    maximum_elevation: int = 0  # calcMaximumElevation (deviceclass.cpp:1138-1141)


class Case:
    """A mesh + state built by libesdg_setup.  Arrays are numpy copies (safe to keep)."""

    def __init__(self, case_kind, mesh_kind, N, nelx, nely, xL, xR, yL, yR, periodic_x=0,
                 periodic_y=0, perturb_amp=0.0, perturb_len=1.0, g=9.81, amp=0.0, p=(0, 0, 0, 0),
                 seed=0, mesh_file=None, read_bottom=0, convert_to_km=0, build=True):
        """build=False keeps only the spec and the sizes (for esdg_device_case_create, which
        builds the arrays on the device); `arrays` is then None."""
        self._mesh_file = None if mesh_file is None else os.fsencode(mesh_file)
        spec = CaseSpec(case_kind, mesh_kind, N, nelx, nely, xL, xR, yL, yR, periodic_x,
                        periodic_y, perturb_amp, perturb_len, g, amp, p[0], p[1], p[2], p[3], seed,
                        self._mesh_file, int(read_bottom), int(convert_to_km))
        self.spec = spec
        if not build:
            if mesh_kind == MESH_ISM:
                raise ValueError("an ISM mesh is read on the host: build=True")
            self.N, self.ngl, self.Nelem = N, N + 1, nelx * nely
            self.Nfaces = ((nelx + 1) * nely + (nely + 1) * nelx - (nelx if periodic_x else 0)
                           - (nely if periodic_y else 0))
            self.min_ele_size = None
            self.arrays = None
            return
        h = C.c_void_p()
        _check(_lib().esdg_case_create(C.byref(spec), C.byref(h)))
        try:
            info = CaseInfo()
            _check(_lib().esdg_case_get_info(h, C.byref(info)))
            self.N, self.ngl, self.Nelem, self.Nfaces = info.N, info.ngl, info.Nelem, info.Nfaces
            self.min_ele_size = info.min_ele_size
            self.arrays = {}
            for name in DOUBLE_ARRAYS + INT_ARRAYS:
                ptr, cnt, isint = C.c_void_p(), C.c_int64(), C.c_int()
                _check(_lib().esdg_case_array(h, name.encode(), C.byref(ptr), C.byref(cnt),
                                              C.byref(isint)))
                out = np.empty(cnt.value, dtype=np.int32 if isint.value else np.float64)
                if cnt.value:
                    C.memmove(out.ctypes.data, ptr.value, out.nbytes)
                self.arrays[name] = out
        finally:
            _lib().esdg_case_destroy(h)

    def __getitem__(self, name):
        return self.arrays[name]

    @property
    def geom_face(self) -> float:
        w = self.arrays["w_GL"] if self.arrays is not None else basis(self.N)["w_GL"]
        return 1.0 / float(w[0])

    @classmethod
    def from_ism(cls, path, N, case_kind=CASE_CONSTANT, p=(1.0, 0.0, 0.0, 0.0), amp=0.0, seed=0,
                 read_bottom=0, convert_to_km=0, bounds=(0.0, 1.0, 0.0, 1.0)):
        """An ISM-V2 mesh file (Mesh::ReadMesh) with a case's state; `bounds` only feeds the case
        formulas that use the domain extent."""
        xL, xR, yL, yR = bounds
        return cls(case_kind, MESH_ISM, N, 0, 0, xL, xR, yL, yR, amp=amp, p=p, seed=seed,
                   mesh_file=path, read_bottom=read_bottom, convert_to_km=convert_to_km)

    # --- the benchmark configurations -------------------------------------------------
    @classmethod
    def gaussian_bump(cls, N=4, nel=16, lake_at_rest=False, build=True):
        """16x16 periodic Cartesian mesh on [0,2]^2, b = 0.8 exp(-5 r^2), h = 1-b (+0.01 bump)."""
        return cls(CASE_GAUSSIAN_BUMP, MESH_CARTESIAN, N, nel, nel, 0.0, 2.0, 0.0, 2.0, 1, 1,
                   amp=0.0 if lake_at_rest else 0.01, build=build)

    @classmethod
    def dam_break(cls, N=7, nel=512, noise=1e-3, seed=0, perturb=0.05, ratio=0.0, angle=0.0, build=True):
        """nel^2 curved quads on [0,10]^2 (x,y += 0.05 sin(pi x/5) sin(pi y/5)), walls, dam break:
        h = 2 | 1 across x = 5, or with ratio/angle set h = ratio | 1 across the line through (5, 5)
        with normal angle `angle`."""
        return cls(CASE_DAM_BREAK, MESH_CURVED_SINE, N, nel, nel, 0.0, 10.0, 0.0, 10.0, 0, 0,
                   perturb_amp=perturb, perturb_len=5.0, amp=noise, p=(5.0, ratio, angle, 0),
                   seed=seed, build=build)

    @classmethod
    def flood(cls, N=7, nel=2048, seed=0, level_left=0.8, x_dam=3.0, level_right=-0.2,
              perturb=0.05, build=True):
        """nel^2 curved quads on [0,10]^2, random Fourier bathymetry, reservoir over dry land."""
        return cls(CASE_FLOOD, MESH_CURVED_SINE, N, nel, nel, 0.0, 10.0, 0.0, 10.0, 0, 0,
                   perturb_amp=perturb, perturb_len=5.0, p=(level_left, x_dam, level_right, 0),
                   seed=seed, build=build)

    def with_state(self, q):
        """A copy of this case with another initial state q (mesh arrays shared, not copied)."""
        import copy
        c = copy.copy(self)
        c.arrays = dict(self.arrays)
        c.arrays["q"] = np.ascontiguousarray(q, dtype=np.float64).reshape(self.arrays["q"].shape)
        return c

    @classmethod
    def ocean_uplift(cls, N=4, nel=64, amp=1.0, axis_a=50e3, axis_c=30e3, angle=0.0, length=800e3, build=True):
        """nel^2 Cartesian quads on [0, length]^2: 4000 m ocean with a linear shelf to 50 m over the
        last 20% of x, still-water level 4000 (b = 4000 - depth), elliptic Gaussian sea-surface
        uplift of amplitude amp, axes axis_a, axis_c (m) rotated by angle, centred at (40%, 50%).
        Run it with RunParams(bc_kind=2, h_zero=4000): open boundary over the deep ocean (b <= 6.7),
        walls over the shelf."""
        return cls(CASE_OCEAN_UPLIFT, MESH_CARTESIAN, N, nel, nel, 0.0, length, 0.0, length, 0, 0,
                   amp=amp, p=(axis_a, axis_c, angle, 0), build=build)

    @classmethod
    def convergence(cls, N=3, nel=8, curved=False, perturb=0.02, testcase=1, length=None):
        """Manufactured solution of convergence Testcase 1, 8, 88 or 89 on nel^2 quads of
        [0, length]^2 (optionally the curved mesh family).  The default length puts the solution's
        wave number k times length at 1 (1, 1, 1e3, 1e6).  Run it with
        RunParams(bc_kind=CONVERGENCE_BC[testcase]): Dirichlet exterior states and the manufactured
        source."""
        if testcase not in CONVERGENCE_BC:
            raise ValueError("convergence Testcase must be 1, 8, 88 or 89")
        if length is None:
            length = {1: 1.0, 8: 1.0, 88: 1.0e3, 89: 1.0e6}[testcase]
        return cls(CASE_CONVERGENCE, MESH_CURVED_SINE if curved else MESH_CARTESIAN, N, nel, nel, 0.0, length,
                   0.0, length, 0, 0, perturb_amp=perturb * length if curved else 0.0, perturb_len=0.5 * length,
                   p=(testcase, 0, 0, 0))

    @classmethod
    def constant(cls, N, nel, h=1.0, hu=0.0, hv=0.0, curved=True, perturb=0.05):
        """Constant state on flat bottom (free-stream preservation)."""
        return cls(CASE_CONSTANT, MESH_CURVED_SINE if curved else MESH_CARTESIAN, N, nel, nel,
                   0.0, 10.0, 0.0, 10.0, 0, 0, perturb_amp=perturb if curved else 0.0,
                   perturb_len=5.0, p=(h, hu, hv, 0))


def basis(N: int, flux_differencing: int = 1):
    ngl = N + 1
    out = {k: np.zeros(ngl if k in ("x_GL", "w_GL") else ngl * ngl)
           for k in ("x_GL", "w_GL", "D", "D0", "Dstrong", "Dhat", "VdmInv")}
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    _check(_lib().esdg_basis_build(N, flux_differencing, ptr(out["x_GL"]), ptr(out["w_GL"]),
                                   ptr(out["D"]), ptr(out["D0"]), ptr(out["Dstrong"]),
                                   ptr(out["Dhat"]), ptr(out["VdmInv"])))
    return out


def rk_coefficients(order: int, ssp: int = 0):
    A, B, Cc = (np.zeros(5) for _ in range(3))
    st = C.c_int()
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    _check(_lib().esdg_rk_coefficients(order, ssp, C.byref(st), ptr(A), ptr(B), ptr(Cc)))
    n = st.value
    return A[:n].copy(), B[:n].copy(), Cc[:n].copy()


def _dptr(keep: list, a):
    """Pointer to a contiguous float64 view of a; the array is kept alive in `keep`."""
    arr = np.ascontiguousarray(a, dtype=np.float64)
    keep.append(arr)
    return arr.ctypes.data_as(C.POINTER(C.c_double))


def total_entropy(case: Case, q: np.ndarray, g: float) -> float:
    out, keep = C.c_double(), []
    _check(_lib().esdg_total_entropy(case.ngl, case.Nelem, C.c_double(g), _dptr(keep, case["w_GL"]),
                                     _dptr(keep, q), _dptr(keep, case["b"]), _dptr(keep, case["Jac"]),
                                     C.byref(out)))
    return out.value


def total_mass(case: Case, q: np.ndarray) -> float:
    out, keep = C.c_double(), []
    _check(_lib().esdg_total_mass(case.ngl, case.Nelem, _dptr(keep, case["w_GL"]), _dptr(keep, q),
                                  _dptr(keep, case["Jac"]), C.byref(out)))
    return out.value


# ---- Tecplot output (include/esdg_output.h; the reference's src/Plots.cpp files) ----------------
def plot_solution(path, case: Case, q: np.ndarray, plot_var: int = 1, h_0: float = 0.0):
    """PlotSolution: plot_var 1 writes H = h + b, hu, hv, b; 2 the free surface h + b - h_0."""
    keep = []
    _check(_lib().esdg_plot_solution(os.fsencode(path), case.Nelem, case.ngl, int(plot_var), _dptr(keep, case["x"]),
                                     _dptr(keep, case["y"]), _dptr(keep, q), _dptr(keep, case["b"]),
                                     C.c_double(h_0)))


def plot_nodal(path, case: Case, values: np.ndarray, kind: str):
    """PlotArrivalTimings (kind "arrival") or PlotMaximumElevation (kind "maximum") of a per-node buffer."""
    fn = {"arrival": _lib().esdg_plot_arrival_times, "maximum": _lib().esdg_plot_maximum_elevation}[kind]
    keep = []
    _check(fn(os.fsencode(path), case.Nelem, case.ngl, _dptr(keep, case["x"]), _dptr(keep, case["y"]),
              _dptr(keep, values)))


def plot_viscosity(path, case: Case, eps: np.ndarray):
    """PlotViscoseParameter: the element viscosity at every node of its element."""
    keep = []
    _check(_lib().esdg_plot_viscosity(os.fsencode(path), case.Nelem, case.ngl, _dptr(keep, case["x"]),
                                      _dptr(keep, case["y"]), _dptr(keep, eps)))


def plot_series(path, times: np.ndarray, totals: np.ndarray, kind: str):
    """PlotEntropy (kind "entropy") or PlotMass (kind "mass", written as (m_i - m_0) / m_0)."""
    fn = {"entropy": _lib().esdg_plot_entropy, "mass": _lib().esdg_plot_mass}[kind]
    keep = []
    _check(fn(os.fsencode(path), len(times), _dptr(keep, times), _dptr(keep, totals)))


REFERENCE_GAUGES = ("Chennai", "Tuticorin", "Visakhapatnam", "Paradip", "Kochi", "Mormugao", "Okha")


def plot_time_series(path, times: np.ndarray, series: np.ndarray, names=REFERENCE_GAUGES):
    """PlotTimeSeries: gauge records; series is [ngauges][len(times)], one column per name."""
    series = np.ascontiguousarray(series, dtype=np.float64).reshape(len(names), len(times))
    keep = []
    cnames = (C.c_char_p * len(names))(*[n.encode() for n in names])
    _check(_lib().esdg_plot_time_series(os.fsencode(path), len(times), _dptr(keep, times), len(names), cnames,
                                        _dptr(keep, series)))
