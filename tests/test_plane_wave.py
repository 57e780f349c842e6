"""Plane-wave M2L (Greengard & Rokhlin 1997 section 7; the reference's lib_plane_wave_m2l.hpp:293-490, whose
rotation to the six directions is switched off at :306-313): the composed operator of every lattice offset
(rotation of the translation's dominant axis onto +z, multipole to exponentials, diagonal shift, exponentials to
local, rotation back) against the SimpleM2L operator of the same offset (GR97 Theorem 3.5, pinned in
tests/test_oracle.py). The quadrature rules are the published 10- and 6-digit ones, so the two must agree to
about 1e-11 and 1e-7 (relative to the operator's largest entry) for offsets in all six directions, from the
smallest the MAC admits (|o| = 7 box sides) to the largest of the tank plans.
"""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
OFFSETS = [(0, 0, 7), (0, 0, -7), (7, 0, 0), (-7, 1, 0), (1, 7, 2), (2, -9, 3), (4, 4, 4), (-4, -4, -4),
           (-5, 3, -6), (12, -3, 5), (1, -2, -15), (-3, 16, -2), (0, -8, 1), (9, 9, -1), (-11, 0, 11)]


@pytest.fixture(scope="module")
def emu():
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emu")], check=True)
    lib = ctypes.CDLL(os.path.join(HERE, "emu", "_build", "libplanemu.so"))
    lib.emu_operator.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p]
    return lib


def _op(lib, P, kind, o):
    C = (P + 1) ** 2
    A = np.empty(C * C)
    lib.emu_operator(P, kind, o[0], o[1], o[2], A.ctypes.data)
    return A.reshape(C, C)


def test_offsets_cover_the_six_directions():
    axes = set()
    for o in OFFSETS:
        a = int(np.argmax(np.abs(o)))
        axes.add((a, int(np.sign(-o[a]))))  # the translation (target - source) is -o
    assert len(axes) == 6


@pytest.mark.parametrize("P", [8, 10])
@pytest.mark.parametrize("digits,tol", [(10, 1e-10), (6, 1e-6)])
def test_plane_wave_operator_equals_simple_m2l(emu, P, digits, tol):
    worst = 0.0
    for o in OFFSETS:
        ref = _op(emu, P, 2, o)
        pw = _op(emu, P, 4 if digits == 10 else 5, o)
        err = np.abs(pw - ref).max() / np.abs(ref).max()
        worst = max(worst, err)
        assert err <= tol, (o, err)
    print(f"P={P} digits={digits}: worst relative difference {worst:.2e}")
