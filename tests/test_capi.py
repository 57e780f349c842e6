"""The C-ABI library loads and exports every entry point declared in include/bemfmm.h (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "bemfmm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bemfmm_[a-z_]+)\s*\(", src)))


def test_header_symbols_exported():
    from bem_for_nonlinear_waves_amd import _lib
    names = _declared()
    assert len(names) >= 18
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
