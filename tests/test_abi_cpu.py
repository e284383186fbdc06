"""CPU checks of the C ABI and the drop-in module's argument handling.

No compute call runs here (there is no GPU in the build container): the
library must load, export every function include/wenox.h declares, reject
the shapes the reference rejects before touching a device, and report a
missing device as an error rather than computing on the CPU.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wenox.h")
LIB = os.path.join(ROOT, "weno-cfd-solver_amd", "wenox", "libwenox.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(wenox_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_dropin_surface():
    names = declared_functions()
    for fn in ("flux_x", "flux_y", "con2primi", "weno_x", "weno_y", "hllc_x", "hllc_y", "hlle_x",
               "hlle_y", "hllc_hlle_x", "hllc_hlle_y", "l_local"):
        assert f"wenox_{fn}" in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libwenox.so first (python weno-cfd-solver_amd/build.py)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (wenox_[a-z0-9_]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    from wenox import _lib
    L = _lib.load()
    for name in declared_functions():
        assert hasattr(L, name)
    assert L.wenox_version().decode().startswith("wenox")


def test_library_has_gfx950_code_object(tmp_path):
    """The device code bundle (built here from source) targets gfx950."""
    blob = tmp_path / "fatbin.bin"
    r = subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, str(blob)],
                       capture_output=True)
    if r.returncode != 0:
        pytest.skip("objcopy unavailable")
    data = blob.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


@pytest.fixture
def wx():
    import weno_ext
    return weno_ext


def test_type_errors_like_pyo3(wx):
    q = np.zeros((10, 10, 8))
    with pytest.raises(TypeError):
        wx.l_local_py(q.tolist(), 0.1, 1.4, False, (0.5, 2.0), False)
    with pytest.raises(TypeError):
        wx.l_local_py(q.astype(np.float32), 0.1, 1.4, False, (0.5, 2.0), False)
    with pytest.raises(TypeError):
        wx.l_local_py(q[..., 0], 0.1, 1.4, False, (0.5, 2.0), False)
    with pytest.raises(TypeError):
        wx.l_local_py(q, 0.1, 1.4, 1, (0.5, 2.0), False)          # int for bool
    with pytest.raises(TypeError):
        wx.l_local_py(q, 0.1, 1.4, False, [0.5, 2.0], False)      # list for tuple
    with pytest.raises(ValueError):
        wx.l_local_py(q, 0.1, 1.4, False, (0.5, 2.0, 3.0), False)
    with pytest.raises(TypeError):
        wx.flux_x_py(q, "1.4")
    with pytest.raises(TypeError):
        wx.hllc_x_rs_py(q, q, q, None, 1.4)


def test_panics_before_any_device_use(wx):
    P = wx.PanicException
    assert issubclass(P, BaseException) and not issubclass(P, Exception)
    with pytest.raises(P):
        wx.l_local_py(np.zeros((5, 10, 8)), 0.1, 1.4, True, (0.5, 2.0))
    with pytest.raises(P):
        wx.l_local_py(np.zeros((10, 10, 3)), 0.1, 1.4, True, (0.5, 2.0))
    with pytest.raises(P):   # blend sensor needs 2 x 2 faces: 7 x 7 padded is too small
        wx.l_local_py(np.zeros((7, 7, 8)), 0.1, 1.4, False, (0.5, 2.0))
    with pytest.raises(P):
        wx.weno_x_py(np.zeros((4, 10, 8)), 1.4)
    with pytest.raises(P):
        wx.weno_y_py(np.zeros((10, 4, 8)), 1.4)
    with pytest.raises(P):
        a = np.zeros((4, 4, 4))
        wx.hllc_x_rs_py(a, a, a[:3], a, 1.4)
    with pytest.raises(P):
        a = np.zeros((1, 4, 4))
        wx.hllc_hlle_x_rs_py(a, a, a, a, 1.4, 0.5, 2.0, False)


def test_no_cpu_fallback_without_device(wx):
    import wenox
    if wenox.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(wenox.WenoxError):
        wx.l_local_py(np.ones((12, 12, 8)), 0.1, 1.4, True, (0.5, 2.0))
    with pytest.raises(wenox.WenoxError):
        wenox.Solver(16, 16, 1 / 16)


def test_empty_arrays_follow_the_reference(wx):
    """Zero-sized inputs never index, so the reference returns empty arrays."""
    import wenox
    if wenox.device_count() == 0:
        pytest.skip("needs the device for the (empty) call path")
    e = np.zeros((0, 5, 4))
    assert wx.flux_x_py(e, 1.4).shape == (0, 5, 4)


def test_solver_argument_checks_before_device_use():
    """Descriptor errors are reported before any device is touched (here and
    on the GPU alike): the standard double Mach reflection's domain height
    must be 1; the no-allocation l_local rejects null and undersized views."""
    import ctypes
    import wenox
    from wenox import _lib
    with pytest.raises(ValueError, match="domain height"):
        wenox.Solver(40, 34, 4.0 / 40, bc="dmr", engine="fused")
    L = _lib.load()
    assert L.wenox_l_local_noalloc(None, 0.1, 1.4, 0, 0.5, 2.0, 0, None, None) == _lib.E_ARG
    v = _lib.View3(8, (ctypes.c_int64 * 3)(5, 10, 4), (ctypes.c_int64 * 3)(40, 4, 1))
    assert L.wenox_l_local_noalloc(ctypes.byref(v), 0.1, 1.4, 1, 0.5, 2.0, 0, None, None) == _lib.E_SHAPE


def test_type_stubs_list_the_module_surface():
    """weno_ext/__init__.pyi declares every exported function with the
    module's own parameter names, order and defaults."""
    import ast
    import inspect

    import weno_ext
    path = os.path.join(os.path.dirname(weno_ext.__file__), "__init__.pyi")
    tree = ast.parse(open(path).read())
    stubs = {n.name: n for n in tree.body if isinstance(n, ast.FunctionDef)}
    names = [n for n in weno_ext.__all__ if n != "PanicException"] + ["set_variant"]
    assert set(names) <= set(stubs)
    for name in names:
        params = list(inspect.signature(getattr(weno_ext, name)).parameters.values())
        args = stubs[name].args
        assert [a.arg for a in args.args] == [p.name for p in params], name
        defaults = [p.name for p in params if p.default is not inspect.Parameter.empty]
        assert len(args.defaults) == len(defaults), name
