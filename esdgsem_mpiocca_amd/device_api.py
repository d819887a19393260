"""ctypes binding of libesdg.so (include/esdg.h): the gfx950 hot path.

There is no CPU fallback: every call goes to the HIP library, and a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .setup_api import Case, RunParams

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ESDG_H2D, ESDG_D2H, ESDG_D2D = 0, 1, 2


class EsdgDeviceError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("ngl", C.c_int), ("ES", C.c_int), ("num_flux", C.c_int), ("bc_kind", C.c_int),
                ("g", C.c_double), ("geomFace", C.c_double), ("PosPresTOL", C.c_double),
                ("ZeroTOL", C.c_double), ("eps0", C.c_double), ("sigmaMin", C.c_double),
                ("sigmaMax", C.c_double), ("h_zero", C.c_double)]


class RunParamsC(C.Structure):
    _fields_ = [("CFL", C.c_double), ("DFL", C.c_double), ("T", C.c_double),
                ("min_ele_size", C.c_double), ("artificial_viscosity", C.c_int),
                ("positivity", C.c_int), ("rk_order", C.c_int), ("rk_ssp", C.c_int),
                ("flux_differencing", C.c_int), ("fused", C.c_int), ("arrival_times", C.c_int),
                ("maximum_elevation", C.c_int)]


_INPUT_DOUBLES = ("GLw", "nx", "ny", "scal", "Yxi", "Yeta", "Xxi", "Xeta", "B", "Bx", "By", "D",
                  "Dstrong", "Dhat", "Jac", "x", "y", "q", "EleSizes", "VdmInv")
_INPUT_INTS = ("ElemEdgeMasterSlave", "ElemEdgeOrientation", "ElemToEdge", "EdgeData",
               "EdgeReversed")


class Inputs(C.Structure):
    _fields_ = ([("Nelem", C.c_int), ("Nfaces", C.c_int)]
                + [(n, C.POINTER(C.c_double)) for n in _INPUT_DOUBLES]
                + [(n, C.POINTER(C.c_int)) for n in _INPUT_INTS])


def library_path() -> str:
    return os.path.join(_PKG, "libesdg.so")


def lib():
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise EsdgDeviceError(f"{path} is missing: the HIP library must be built "
                                  "(__graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(path)
        L.esdg_last_error.restype = C.c_char_p
        L.esdg_last_error.argtypes = [C.c_void_p]
        L.esdg_bytes_allocated.restype = C.c_size_t
        L.esdg_bytes_allocated.argtypes = [C.c_void_p]
        L.esdg_stream.restype = C.c_void_p
        L.esdg_stream.argtypes = [C.c_void_p]
        L.esdg_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(Params)]
        L.esdg_malloc.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_size_t]
        L.esdg_free.argtypes = [C.c_void_p, C.c_void_p]
        L.esdg_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.esdg_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_size_t]
        L.esdg_profile_query.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_double)]
        L.esdg_solver_create.argtypes = [C.c_void_p, C.POINTER(Inputs), C.POINTER(RunParamsC),
                                         C.POINTER(C.c_void_p)]
        L.esdg_solver_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.esdg_solver_get_q.argtypes = [C.c_void_p, C.c_void_p]
        L.esdg_solver_set_q.argtypes = [C.c_void_p, C.c_void_p]
        L.esdg_solver_time.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.esdg_solver_last_dt.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.esdg_solver_buffer.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.esdg_solver_destroy.argtypes = [C.c_void_p]
        L.esdg_solver_create_dev.argtypes = [C.c_void_p, C.POINTER(Inputs), C.POINTER(RunParamsC),
                                             C.POINTER(C.c_void_p)]
        L.esdg_device_case_inputs.argtypes = [C.c_void_p, C.POINTER(Inputs)]
        L.esdg_device_case_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p),
                                             C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.esdg_device_case_get_info.argtypes = [C.c_void_p, C.c_void_p]
        L.esdg_device_case_destroy.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def make_params(case: Case, rp: RunParams) -> Params:
    return Params(case.ngl, rp.ES, rp.num_flux, rp.bc_kind, rp.g, case.geom_face, rp.pos_pres_tol,
                  rp.zero_tol, rp.eps0, rp.sigma_min, rp.sigma_max, rp.h_zero)


class DeviceArray:
    """A device allocation of the context, shaped like a numpy array."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(lib().esdg_malloc(ctx.handle, C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @property
    def _as_parameter_(self):
        return C.c_void_p(self.ptr)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes, (host.nbytes, self.nbytes)
        self.ctx.check(lib().esdg_memcpy(self.ctx.handle, C.c_void_p(self.ptr),
                                         host.ctypes.data_as(C.c_void_p), self.nbytes, ESDG_H2D))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(lib().esdg_memcpy(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                         C.c_void_p(self.ptr), self.nbytes, ESDG_D2H))
        return out

    def free(self):
        if self.ptr:
            lib().esdg_free(self.ctx.handle, C.c_void_p(self.ptr))
            self.ptr = None


class DeviceView:
    """Non-owning view of a device buffer owned by the library (e.g. a solver's state)."""

    def __init__(self, ctx: "Context", ptr: int, shape, dtype):
        self.ctx, self.ptr = ctx, ptr
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(lib().esdg_memcpy(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                         C.c_void_p(self.ptr), self.nbytes, ESDG_D2H))
        return out


class Context:
    """One HIP device + stream (esdg_ctx).  Kernel entries are called as ``ctx.<name>(...)``."""

    def __init__(self, params: Params, device: int = 0):
        h = C.c_void_p()
        rc = lib().esdg_create(C.byref(h), device, C.byref(params))
        if rc != 0:
            raise EsdgDeviceError(lib().esdg_last_error(None).decode())
        self.handle = h.value
        self.params = params

    def check(self, rc):
        if rc != 0:
            raise EsdgDeviceError(lib().esdg_last_error(C.c_void_p(self.handle)).decode())

    def array(self, host: np.ndarray) -> DeviceArray:
        host = np.ascontiguousarray(host)
        return DeviceArray(self, host.shape, host.dtype).upload(host)

    def empty(self, shape, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def call(self, name: str, *args):
        fn = getattr(lib(), "esdg_" + name)
        cargs = [C.c_void_p(self.handle)]
        for a in args:
            if isinstance(a, (DeviceArray, DeviceView)):
                cargs.append(C.c_void_p(a.ptr))
            elif a is None:
                cargs.append(C.c_void_p(0))
            elif isinstance(a, (bool, int, np.integer)):
                cargs.append(C.c_int(int(a)))
            elif isinstance(a, (float, np.floating)):
                cargs.append(C.c_double(float(a)))
            elif isinstance(a, np.ndarray):   # host output (e.g. an integral), C-contiguous
                if not a.flags.c_contiguous:
                    raise TypeError("host arrays must be C-contiguous")
                cargs.append(C.c_void_p(a.ctypes.data))
            else:
                raise TypeError(f"unsupported argument {a!r}")
        self.check(fn(*cargs))

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return lambda *args: self.call(name, *args)

    def finish(self):
        self.check(lib().esdg_finish(C.c_void_p(self.handle)))

    def profile(self, on=True):
        self.check(lib().esdg_profile_enable(C.c_void_p(self.handle), int(on)))

    def profile_reset(self):
        self.check(lib().esdg_profile_reset(C.c_void_p(self.handle)))

    def profile_query(self, name):
        n, ms = C.c_int64(), C.c_double()
        self.check(lib().esdg_profile_query(C.c_void_p(self.handle), name.encode(), C.byref(n),
                                            C.byref(ms)))
        return n.value, ms.value

    def bytes_allocated(self):
        return lib().esdg_bytes_allocated(C.c_void_p(self.handle))

    def close(self):
        if self.handle:
            lib().esdg_destroy(C.c_void_p(self.handle))
            self.handle = None


def make_inputs(case: Case, rp: RunParams, q=None):
    """esdg_inputs of a Case (host pointers into the Case's own arrays, so two calls on the same
    Case hand over the same pointers) and the list of arrays that must outlive the call."""
    a = case.arrays
    src = {
        "GLw": a["w_GL"], "nx": a["nx"], "ny": a["ny"], "scal": a["scal"], "Yxi": a["yXi"],
        "Yeta": a["yEta"], "Xxi": a["xXi"], "Xeta": a["xEta"], "B": a["b"], "Bx": a["Bx"],
        "By": a["By"], "D": a["D"], "Dstrong": a["Dstrong"], "Dhat": a["Dhat"],
        "Jac": a["Jac"], "x": a["x"], "y": a["y"], "q": a["q"] if q is None else q,
        "EleSizes": a["EleSizes"], "VdmInv": a["VdmInv"],
        "ElemEdgeMasterSlave": a["ElemEdgeMasterSlave"],
        "ElemEdgeOrientation": a["ElemEdgeOrientation"], "ElemToEdge": a["ElemToEdge"],
        "EdgeData": a["EdgeData"], "EdgeReversed": a["EdgeReversed"]}
    if not rp.flux_differencing:
        # FluxDifferencing = 0 builds D without the factor 2 (src/basis.cpp:108-119)
        src["D"] = a["Dstrong"]
    keep = []
    inp = Inputs()
    inp.Nelem, inp.Nfaces = case.Nelem, case.Nfaces
    for n in _INPUT_DOUBLES:
        arr = np.ascontiguousarray(src[n], dtype=np.float64)
        keep.append(arr)
        setattr(inp, n, arr.ctypes.data_as(C.POINTER(C.c_double)))
    for n in _INPUT_INTS:
        arr = np.ascontiguousarray(src[n], dtype=np.int32)
        keep.append(arr)
        setattr(inp, n, arr.ctypes.data_as(C.POINTER(C.c_int)))
    return inp, keep


def make_run_params(case: Case, rp: RunParams, fused: int = 0) -> RunParamsC:
    return RunParamsC(rp.CFL, rp.DFL, rp.T, case.min_ele_size, rp.artificial_viscosity,
                      rp.positivity, rp.rk_order, rp.rk_ssp, rp.flux_differencing, int(fused),
                      int(rp.arrival_times), int(rp.maximum_elevation))


class DeviceCase:
    """esdg_device_case (include/esdg_device_setup.h): a structured-mesh case (Case.<config>(...,
    build=False) gives its spec) built on the device of ctx.  Has the Case attributes the solver
    uses (N, ngl, Nelem, Nfaces, min_ele_size); arrays stay on the device."""

    def __init__(self, ctx: Context, case: Case):
        from .setup_api import CaseInfo
        self.ctx = ctx
        h = C.c_void_p()
        L = lib()
        L.esdg_device_case_create.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        ctx.check(L.esdg_device_case_create(C.c_void_p(ctx.handle), C.byref(case.spec), C.byref(h)))
        self.handle = h.value
        info = CaseInfo()
        ctx.check(L.esdg_device_case_get_info(C.c_void_p(self.handle), C.byref(info)))
        self.N, self.ngl, self.Nelem, self.Nfaces = info.N, info.ngl, info.Nelem, info.Nfaces
        self.min_ele_size = info.min_ele_size
        self.spec = case.spec

    @property
    def geom_face(self) -> float:
        from .setup_api import basis
        return 1.0 / float(basis(self.N)["w_GL"][0])

    def array_ptr(self, name: str):
        p, n, isint = C.c_void_p(), C.c_int64(), C.c_int()
        self.ctx.check(lib().esdg_device_case_array(C.c_void_p(self.handle), name.encode(), C.byref(p),
                                                    C.byref(n), C.byref(isint)))
        return p.value, n.value, bool(isint.value)

    def view(self, name: str) -> DeviceView:
        p, n, isint = self.array_ptr(name)
        return DeviceView(self.ctx, p, n, np.int32 if isint else np.float64)

    def download(self, name: str) -> np.ndarray:
        return self.view(name).download()

    def to_host_case(self) -> Case:
        """A host Case holding copies of every device array (for the oracle and host checks)."""
        from .setup_api import DOUBLE_ARRAYS, INT_ARRAYS
        c = Case.__new__(Case)
        c.spec = self.spec
        c._mesh_file = None
        c.N, c.ngl, c.Nelem, c.Nfaces = self.N, self.ngl, self.Nelem, self.Nfaces
        c.min_ele_size = self.min_ele_size
        c.arrays = {n: self.download(n) for n in DOUBLE_ARRAYS + INT_ARRAYS}
        return c

    def close(self):
        if self.handle:
            lib().esdg_device_case_destroy(C.c_void_p(self.handle))
            self.handle = None


class Solver:
    """deviceclass mirror: buffers + DGtimeloop on the device for one Case (host arrays) or
    DeviceCase (arrays already on the device, copied device to device).
    fused: 0 the reference's kernel sequence, 1 fused stage kernels, 2 single-kernel stage."""

    def __init__(self, ctx: Context, case, rp: RunParams, fused: int = 0):
        self.ctx, self.case, self.rp = ctx, case, rp
        rpc = make_run_params(case, rp, fused)
        h = C.c_void_p()
        if isinstance(case, DeviceCase):
            if not rp.flux_differencing:
                raise ValueError("a DeviceCase carries the flux-differencing D only")
            inp = Inputs()
            ctx.check(lib().esdg_device_case_inputs(C.c_void_p(case.handle), C.byref(inp)))
            self._keep = [case]
            ctx.check(lib().esdg_solver_create_dev(C.c_void_p(ctx.handle), C.byref(inp), C.byref(rpc),
                                                   C.byref(h)))
        else:
            inp, self._keep = make_inputs(case, rp)
            ctx.check(lib().esdg_solver_create(C.c_void_p(ctx.handle), C.byref(inp), C.byref(rpc),
                                               C.byref(h)))
        self.handle = h.value

    def run(self, max_steps: int):
        n, t = C.c_int(), C.c_double()
        self.ctx.check(lib().esdg_solver_run(C.c_void_p(self.handle), max_steps, C.byref(n),
                                             C.byref(t)))
        return n.value, t.value

    def get_q(self) -> np.ndarray:
        out = np.empty(self.case.Nelem * self.case.ngl ** 2 * 3)
        self.ctx.check(lib().esdg_solver_get_q(C.c_void_p(self.handle),
                                               out.ctypes.data_as(C.c_void_p)))
        return out

    def set_q(self, q: np.ndarray):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self.ctx.check(lib().esdg_solver_set_q(C.c_void_p(self.handle),
                                               q.ctypes.data_as(C.c_void_p)))

    def last_dt(self) -> float:
        d = C.c_double()
        self.ctx.check(lib().esdg_solver_last_dt(C.c_void_p(self.handle), C.byref(d)))
        return d.value

    def time(self) -> float:
        t = C.c_double()
        self.ctx.check(lib().esdg_solver_time(C.c_void_p(self.handle), C.byref(t)))
        return t.value

    def checkpoint(self, path):
        """Binary checkpoint between steps (t, last dt, q, the RK register, analysis buffers)."""
        self.ctx.check(lib().esdg_solver_checkpoint(C.c_void_p(self.handle), os.fsencode(path)))

    def restart(self, path):
        """Restart from a checkpoint of a solver with the same degree, mesh size and buffers."""
        self.ctx.check(lib().esdg_solver_restart(C.c_void_p(self.handle), os.fsencode(path)))

    def set_indicator(self, mlp, lo: float = 0.5, hi: float = 0.5):
        """Learned troubled-cell indicator in the AV path (None: the modal sensor)."""
        self.ctx.check(lib().esdg_solver_set_indicator(C.c_void_p(self.handle),
                                                       C.c_void_p(mlp.handle if mlp else None),
                                                       C.c_float(lo), C.c_float(hi)))

    def indicator_audit(self, on: bool = True):
        self.ctx.check(lib().esdg_solver_indicator_audit(C.c_void_p(self.handle), int(on)))

    def indicator_counts(self):
        v = [C.c_int64() for _ in range(5)]
        self.ctx.check(lib().esdg_solver_indicator_counts(C.c_void_p(self.handle), *[C.byref(x) for x in v]))
        return dict(zip(("viscous", "false_negatives", "sensor_positives", "not_smooth", "element_stages"),
                        [x.value for x in v]))

    def view(self, name: str, shape, dtype=np.float64) -> DeviceView:
        return DeviceView(self.ctx, self.buffer_ptr(name), shape, dtype)

    def buffer_ptr(self, name: str) -> int:
        p = C.c_void_p()
        self.ctx.check(lib().esdg_solver_buffer(C.c_void_p(self.handle), name.encode(), C.byref(p)))
        return p.value

    def close(self):
        if self.handle:
            lib().esdg_solver_destroy(C.c_void_p(self.handle))
            self.handle = None


class EnsembleSchedule(C.Structure):
    _fields_ = [("order", C.c_int), ("blocks_per_cu", C.c_int), ("per_degree", C.c_int),
                ("xcd", C.c_int), ("graph", C.c_int), ("tile", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Ensemble:
    """esdg_ensemble (include/esdg_ensemble.h): independent scenarios advanced together.

    members: list of (case, rp, q) with q None for the case's own initial state.  Members that
    share a Case share its mesh arrays on the device."""

    def __init__(self, ctx: Context, members):
        self.ctx = ctx
        n = len(members)
        self.members = members
        P = (Params * n)()
        I = (Inputs * n)()
        R = (RunParamsC * n)()
        self._keep = []
        for i, (case, rp, q) in enumerate(members):
            P[i] = make_params(case, rp)
            inp, keep = make_inputs(case, rp, q)
            I[i] = inp
            self._keep.append(keep)
            R[i] = make_run_params(case, rp)
        h = C.c_void_p()
        ctx.check(lib().esdg_ensemble_create(C.c_void_p(ctx.handle), n, P, I, R, C.byref(h)))
        self.handle = h.value
        self._keep = None      # the library holds device copies now

    def set_schedule(self, **kw):
        sc = self.schedule()
        for k, v in kw.items():
            setattr(sc, k, int(v))
        self.ctx.check(lib().esdg_ensemble_set_schedule(C.c_void_p(self.handle), C.byref(sc)))

    def schedule(self) -> EnsembleSchedule:
        sc = EnsembleSchedule()
        self.ctx.check(lib().esdg_ensemble_get_schedule(C.c_void_p(self.handle), C.byref(sc)))
        return sc

    def autotune(self, trial_steps=2):
        sc, ms = EnsembleSchedule(), C.c_double()
        self.ctx.check(lib().esdg_ensemble_autotune(C.c_void_p(self.handle), int(trial_steps),
                                                    C.byref(sc), C.byref(ms)))
        return sc.as_dict(), ms.value

    def run(self, steps: int):
        self.ctx.check(lib().esdg_ensemble_run(C.c_void_p(self.handle), int(steps)))

    def run_async(self, steps: int):
        self.ctx.check(lib().esdg_ensemble_run_async(C.c_void_p(self.handle), int(steps)))

    def get_q(self, i: int) -> np.ndarray:
        case = self.members[i][0]
        out = np.empty(case.Nelem * case.ngl ** 2 * 3)
        self.ctx.check(lib().esdg_ensemble_get_q(C.c_void_p(self.handle), int(i),
                                                 out.ctypes.data_as(C.c_void_p)))
        return out

    def status(self, i: int):
        t, n, dt = C.c_double(), C.c_int(), C.c_double()
        self.ctx.check(lib().esdg_ensemble_status(C.c_void_p(self.handle), int(i), C.byref(t),
                                                  C.byref(n), C.byref(dt)))
        return t.value, n.value, dt.value

    def nbytes(self) -> int:
        b = C.c_size_t()
        self.ctx.check(lib().esdg_ensemble_bytes(C.c_void_p(self.handle), C.byref(b)))
        return b.value

    def set_indicator(self, mlp, lo: float = 0.5, hi: float = 0.5):
        """Learned indicator for the AV members with N = 7 (None: the modal sensor everywhere).
        The network object must stay open while the ensemble uses it."""
        self._mlp = mlp
        self.ctx.check(lib().esdg_ensemble_set_indicator(C.c_void_p(self.handle),
                                                         C.c_void_p(mlp.handle if mlp else None),
                                                         C.c_float(lo), C.c_float(hi)))

    def indicator_audit(self, on: bool = True):
        self.ctx.check(lib().esdg_ensemble_indicator_audit(C.c_void_p(self.handle), int(on)))

    def indicator_counts(self):
        v = [C.c_int64() for _ in range(5)]
        self.ctx.check(lib().esdg_ensemble_indicator_counts(C.c_void_p(self.handle), *[C.byref(x) for x in v]))
        return dict(zip(("viscous", "false_negatives", "sensor_positives", "not_smooth", "element_stages"),
                        [x.value for x in v]))

    def close(self):
        if self.handle:
            lib().esdg_ensemble_destroy(C.c_void_p(self.handle))
            self.handle = None


MLP_NIN, MLP_NHID = 72, 32
MLP_NPARAM = 72 * 32 + 32 + 32 * 32 + 32 + 32 + 1


class MLP:
    """esdg_mlp (include/esdg_mlp.h): the learned troubled-cell indicator on the device."""

    def __init__(self, ctx: Context, seed: int = 0, params=None):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().esdg_mlp_create(C.c_void_p(ctx.handle), C.c_uint64(seed), C.byref(h)))
        self.handle = h.value
        if params is not None:
            self.set_params(params)

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        assert a.size == MLP_NPARAM
        self.ctx.check(lib().esdg_mlp_set_params(C.c_void_p(self.handle), a.ctypes.data_as(C.c_void_p)))

    def get_params(self) -> np.ndarray:
        out = np.empty(MLP_NPARAM, dtype=np.float32)
        self.ctx.check(lib().esdg_mlp_get_params(C.c_void_p(self.handle), out.ctypes.data_as(C.c_void_p)))
        return out

    def infer(self, feat_dev: DeviceArray, n: int, out_dev: DeviceArray):
        self.ctx.check(lib().esdg_mlp_infer(C.c_void_p(self.handle), int(n), C.c_void_p(feat_dev.ptr),
                                            C.c_void_p(out_dev.ptr)))

    def indicator(self, Nelem, q_dev, B_dev, V_dev, out_dev):
        self.ctx.check(lib().esdg_mlp_indicator(C.c_void_p(self.handle), int(Nelem), C.c_void_p(q_dev.ptr),
                                                C.c_void_p(B_dev.ptr), C.c_void_p(V_dev.ptr),
                                                C.c_void_p(out_dev.ptr)))

    def loss_grad(self, n, feat_dev, label_dev, idx_dev, pos_weight, grad_dev, loss_dev):
        self.ctx.check(lib().esdg_mlp_loss_grad(C.c_void_p(self.handle), int(n), C.c_void_p(feat_dev.ptr),
                                                C.c_void_p(label_dev.ptr),
                                                C.c_void_p(idx_dev.ptr if idx_dev is not None else None),
                                                C.c_float(pos_weight), C.c_void_p(grad_dev.ptr),
                                                C.c_void_p(loss_dev.ptr)))

    def train_step(self, n, feat_dev, label_dev, idx_dev, pos_weight, lr, want_loss=False) -> float:
        loss = C.c_float(float("nan"))
        self.ctx.check(lib().esdg_mlp_train_step(C.c_void_p(self.handle), int(n), C.c_void_p(feat_dev.ptr),
                                                 C.c_void_p(label_dev.ptr),
                                                 C.c_void_p(idx_dev.ptr if idx_dev is not None else None),
                                                 C.c_float(pos_weight), C.c_float(lr),
                                                 C.byref(loss) if want_loss else None))
        return loss.value

    def close(self):
        if self.handle:
            lib().esdg_mlp_destroy(C.c_void_p(self.handle))
            self.handle = None


def mlp_features(ctx: Context, Nelem, q_dev, B_dev, V_dev, feat_dev):
    ctx.check(lib().esdg_mlp_features(C.c_void_p(ctx.handle), int(Nelem), C.c_void_p(q_dev.ptr),
                                      C.c_void_p(B_dev.ptr), C.c_void_p(V_dev.ptr), C.c_void_p(feat_dev.ptr)))
