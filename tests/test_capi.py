"""The C-ABI library loads and exports every entry point declared in include/*.h (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(bemfmm_[a-z_0-9]+)\s*\(", src))
    return sorted(names)


def test_header_symbols_exported():
    from bem_for_nonlinear_waves_amd import _lib
    names = _declared()
    assert len(names) >= 30
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_binding_loads():
    from bem_for_nonlinear_waves_amd import _lib
    lib = _lib.load()
    o = _lib.Opts()
    lib.bemfmm_default_opts(ctypes.byref(o))
    assert (o.p, o.theta, o.leaf_max, o.max_level, o.q) == (10, 0.25, 50, 7, 6)


def test_server_default_opts():
    from bem_for_nonlinear_waves_amd import serve
    o = serve.default_opts()
    assert o.slots >= 1 and o.hbm_budget_bytes > 0


def test_server_failure_path_without_gpu(tiny_tank):
    """Without a GPU a request fails cleanly (context creation), is counted, and the server shuts down."""
    import numpy as np
    import pytest
    from bem_for_nonlinear_waves_amd import _lib
    from bem_for_nonlinear_waves_amd.bvp import BEMError, is_available
    from bem_for_nonlinear_waves_amd.serve import BVPServer
    if is_available():
        pytest.skip("a GPU is present: the request would run")
    m, d, phi, phin = tiny_tank
    srv = BVPServer(slots=2)
    t = srv.submit(m, d, phi, phin)
    with pytest.raises(BEMError) as e:
        srv.wait(t)
    assert e.value.code == _lib.E_DEVICE
    st = srv.stats()
    assert st["failed"] == 1 and st["completed"] == 0
    srv.close()
