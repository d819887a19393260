"""The C-ABI libraries load and export every function their headers declare (no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "esdgsem_mpiocca_amd")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(esdg_\w+)\s*\(", text)
    return sorted(set(names))


HEADERS = [("esdg.h", "libesdg.so"), ("esdg_ensemble.h", "libesdg.so"), ("esdg_mlp.h", "libesdg.so"),
           ("esdg_device_setup.h", "libesdg.so"),
           ("esdg_setup.h", "libesdg_setup.so"), ("esdg_output.h", "libesdg_setup.so")]


def test_every_header_is_checked():
    assert sorted(h for h, _ in HEADERS) == sorted(f for f in os.listdir(os.path.join(ROOT, "include"))
                                                   if f.endswith(".h"))


@pytest.mark.parametrize("header,lib", HEADERS)
def test_exports(header, lib):
    path = os.path.join(PKG, lib)
    if not os.path.exists(path):
        pytest.fail(f"{lib} is not built: run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 4
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_device_layer_has_no_cpu_fallback(monkeypatch):
    """Without the HIP library the device layer raises instead of computing anything."""
    from esdgsem_mpiocca_amd import device_api
    monkeypatch.setattr(device_api, "_LIB", None)
    monkeypatch.setattr(device_api, "library_path", lambda: os.path.join(ROOT, "no_such_lib.so"))
    with pytest.raises(device_api.EsdgDeviceError):
        device_api.lib()


def test_create_without_gpu_reports_error():
    """esdg_create on a machine without a HIP device returns an error code and a message."""
    from esdgsem_mpiocca_amd import device_api as D
    import esdgsem_mpiocca_amd as E
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    c = E.Case.gaussian_bump(N=2, nel=2)
    with pytest.raises(D.EsdgDeviceError):
        D.Context(D.make_params(c, E.RunParams()))
