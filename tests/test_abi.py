"""The C-ABI libraries load and export every entry point their headers declare
(no device calls: runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dyn_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("dynasor_gpu.h", "libdynasor_gpu.so"),
                                        ("dynasor_host.h", "libdynasor_host.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = ROOT / "dynasor_plus_amd" / "lib" / lib
    assert path.exists(), f"{path} not built"
    if lib == "libdynasor_host.so":
        C.CDLL(str(ROOT / "dynasor_plus_amd" / "lib" / "libdynasor_gpu.so"), mode=C.RTLD_GLOBAL)
    handle = C.CDLL(str(path))
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, missing


def test_python_binding_covers_the_gpu_header():
    from dynasor_plus_amd import _lib
    names = set(declared("dynasor_gpu.h"))
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_abi_version_without_device():
    from dynasor_plus_amd import load_library
    lib = load_library()
    assert lib.dyn_gpu_abi_version() == 1


def test_create_without_device_reports_error():
    """No GPU in this container: context creation fails with a status code."""
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from dynasor_plus_amd import load_library
    lib = load_library()
    h = C.c_void_p()
    rc = lib.dyn_gpu_create(0, C.byref(h))
    assert rc != 0 and not h.value
