"""Tecplot output (include/esdg_output.h) against the reference's own writers (src/Plots.cpp,
compiled into oracle/_ref): the files must be identical byte for byte.  The structural check
(zones, node order, values to the 6 significant digits of the stream's default format) runs
without the reference build too."""
import ctypes as C
import os

import numpy as np
import pytest

from esdgsem_mpiocca_amd import Case
from esdgsem_mpiocca_amd import setup_api as S
from helpers import perturbed_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libesdg_ref_setup.so")


@pytest.fixture(scope="module")
def data():
    case = Case.dam_break(N=4, nel=3, noise=1e-3, seed=1)
    q = perturbed_state(case, seed=2)
    n = case.Nelem * case.ngl ** 2
    rng = np.random.default_rng(3)
    arr = np.where(rng.uniform(size=n) < 0.5, 1000.0, rng.uniform(0, 3, size=n))
    mx = rng.normal(size=n) * 1e-3
    eps = np.where(rng.uniform(size=case.Nelem) < 0.5, 0.0, rng.uniform(0, 5e-3, size=case.Nelem))
    return case, q, arr, mx, eps


def ours(tmp_path, case, q, arr, mx, eps):
    out = {}
    for var, h0 in ((1, 0.0), (2, 1.5)):
        p = tmp_path / f"ours_plot{var}.tec"
        S.plot_solution(p, case, q, var, h0)
        out[f"plot{var}"] = p.read_bytes()
    S.plot_nodal(tmp_path / "ours_arr.tec", case, arr, "arrival")
    S.plot_nodal(tmp_path / "ours_max.tec", case, mx, "maximum")
    S.plot_viscosity(tmp_path / "ours_visc.tec", case, eps)
    out["arr"] = (tmp_path / "ours_arr.tec").read_bytes()
    out["max"] = (tmp_path / "ours_max.tec").read_bytes()
    out["visc"] = (tmp_path / "ours_visc.tec").read_bytes()
    return out


def test_plot_files_equal_the_reference_writers(tmp_path, data):
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    case, q, arr, mx, eps = data
    mine = ours(tmp_path, case, q, arr, mx, eps)
    ref = C.CDLL(REF)
    ref.ref_plot.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_int, C.c_double]
    P = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)
    x, y, b = case["x"], case["y"], case["b"]
    wd = os.fsencode(str(tmp_path))
    for var, h0 in ((1, 0.0), (2, 1.5)):
        assert ref.ref_plot(wd, 0, case.Nelem, case.ngl, var, P(x), P(y), P(q), P(b), 10 + var, h0) == 0
        assert (tmp_path / "movie" / f"plot_{10 + var}.tec").read_bytes() == mine[f"plot{var}"]
    assert ref.ref_plot(wd, 1, case.Nelem, case.ngl, 0, P(x), P(y), P(arr), P(b), 0, 0.0) == 0
    assert (tmp_path / "movie" / "Arrivals.tec").read_bytes() == mine["arr"]
    assert ref.ref_plot(wd, 2, case.Nelem, case.ngl, 0, P(x), P(y), P(mx), P(b), 0, 0.0) == 0
    assert (tmp_path / "movie" / "MaximumElevations.tec").read_bytes() == mine["max"]
    assert ref.ref_plot(wd, 3, case.Nelem, case.ngl, 0, P(x), P(y), P(eps), P(b), 4, 0.0) == 0
    assert (tmp_path / "movie" / "visc_4.tec").read_bytes() == mine["visc"]


def test_plot_file_structure(tmp_path, data):
    case, q, arr, mx, eps = data
    mine = ours(tmp_path, case, q, arr, mx, eps)
    ngl, E = case.ngl, case.Nelem
    lines = mine["plot2"].decode().split("\n")
    assert lines[0] == "TITLE = H_solution.tec"
    assert lines[1] == 'VARIABLES = "x","y","FreeSurface","hu","hv","bottom"'
    assert len(lines) == 2 + E * (ngl * ngl + 1) + 1 and lines[-1] == ""
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[2:-1] if not ln.startswith("ZONE")])
    assert sum(ln == f"ZONE I ={ngl},J={ngl},F=POINT" for ln in lines) == E
    Q = q.reshape(E, 3, ngl * ngl)
    want = np.stack([case["x"], case["y"], Q[:, 0].ravel() + case["b"] - 1.5, Q[:, 1].ravel(), Q[:, 2].ravel(),
                     case["b"] - 1.5], axis=1)
    # six significant digits
    assert np.all(np.abs(rows - want) <= 5e-6 * np.maximum(np.abs(want), 1e-300) + 1e-300)
    visc = mine["visc"].decode()
    assert visc.endswith(" \nTest")
    v = np.array([float(ln.split()[2]) for ln in visc.split("\n")[2:] if ln and not ln.startswith(("ZONE", "Test"))])
    assert np.allclose(v, np.repeat(eps, ngl * ngl), rtol=5e-6, atol=0)


def test_series_files_equal_the_reference_writers(tmp_path):
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(4)
    t = np.cumsum(rng.uniform(1e-4, 1e-2, size=17))
    ent = 123.456 + np.cumsum(rng.normal(size=17)) * 1e-7
    mass = 40.0 + rng.normal(size=17) * 1e-12
    S.plot_series(tmp_path / "ours_ent.tec", t, ent, "entropy")
    S.plot_series(tmp_path / "ours_mass.tec", t, mass, "mass")
    ref = C.CDLL(REF)
    ref.ref_plot.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_int, C.c_double]
    P = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)
    wd = os.fsencode(str(tmp_path))
    assert ref.ref_plot(wd, 4, 0, 0, 0, P(t), None, P(ent), None, len(t), 0.0) == 0
    assert ref.ref_plot(wd, 5, 0, 0, 0, P(t), None, P(mass), None, len(t), 0.0) == 0
    assert (tmp_path / "movie" / "Entropyplot.tec").read_bytes() == (tmp_path / "ours_ent.tec").read_bytes()
    assert (tmp_path / "movie" / "Massplot.tec").read_bytes() == (tmp_path / "ours_mass.tec").read_bytes()


def test_time_series_file_equals_the_reference_writer(tmp_path):
    """PlotTimeSeries (src/Plots.cpp:474-517) with the reference's seven gauges, byte for byte."""
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(8)
    t = np.cumsum(rng.uniform(1.0, 30.0, size=23))
    series = rng.normal(size=(7, 23)) * np.array([1e-3, 0.5, 2.0, 1e-9, 3.0, 0.25, 7.0])[:, None]
    series[2, 5] = 0.0
    S.plot_time_series(tmp_path / "ours_ts.tec", t, series)
    ref = C.CDLL(REF)
    ref.ref_plot.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_int, C.c_double]
    P = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)
    assert ref.ref_plot(os.fsencode(str(tmp_path)), 6, 0, 0, 0, P(t), None, P(series), None, len(t), 0.0) == 0
    assert (tmp_path / "movie" / "Timeseries.tec").read_bytes() == (tmp_path / "ours_ts.tec").read_bytes()
    # other gauge sets: one column per name
    S.plot_time_series(tmp_path / "two.tec", t[:3], series[:2, :3], names=("A", "B"))
    lines = (tmp_path / "two.tec").read_text().splitlines()
    assert lines[1] == 'VARIABLES = "t","A","B"' and len(lines) == 5 and len(lines[2].split()) == 3
