"""The oracle against routines compiled from the reference's own sources:
getP2 and getGamma (utilities/gamma_to_p2.f90) and the constants of
common/puffin_constants.f90, built into oracle/_ref/libpuffin_ref.so by
`make -C oracle ref` where the reference tree is present (build() does it).
Skipped where that library was not built.

getP2 sits under every stage of the path (the z2 rate and the (1 + eta p2)
factor of the sources and equations), so the oracle's restatement is held to
the compiled reference bit for bit, over the parameter ranges of the
workloads; the reference's default-real pi and the single-precision c and q_e
of the current diagnostic are read from its own constants module."""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import pytest

import oracle as orc
from puffin_amd import h5io
from puffin_amd.params import PI_REF

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libpuffin_ref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/_ref not built (no reference tree)")

_D = ctypes.POINTER(ctypes.c_double)


def _ref():
    lib = ctypes.CDLL(LIB)
    for name in ("ref_get_p2", "ref_get_gamma"):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [ctypes.c_int64, _D, _D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]
    lib.ref_constants.restype = None
    lib.ref_constants.argtypes = [_D]
    return lib


def _call(f, a, b, c, eta, g0, aw):
    a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c))
    out = np.empty_like(a)
    f(a.size, a.ctypes.data_as(_D), b.ctypes.data_as(_D), c.ctypes.data_as(_D), eta, g0, aw,
      out.ctypes.data_as(_D))
    return out


# (eta, gamma_r, aw): the 1-D test deck's scale, C0/C3 (eta 1e-4, gamma_r 100),
# C1/C4 (eta 1e-6, gamma_r 866) and a low-energy case
PARAMS = [(1.0e-4, 100.0, 1.0), (1.0376e-6, 866.0, 1.0), (3.0e-3, 30.0, 1.0 / math.sqrt(2.0)),
          (1.0e-2, 8.0, 1.2)]


@pytest.mark.parametrize("eta,g0,aw", PARAMS)
def test_oracle_p2_is_the_compiled_reference_bit_for_bit(eta, g0, aw):
    rng = np.random.default_rng(11)
    n = 4000
    gam = 1.0 + 1e-2 * rng.standard_normal(n)
    px, py = np.clip(rng.standard_normal((2, n)), -3.0, 3.0)
    want = _call(_ref().ref_get_p2, gam, px, py, eta, g0, aw)
    got = np.array([orc.get_p2(gam[i], px[i], py[i], eta, g0, aw) for i in range(n)])
    assert np.all(np.isfinite(want))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("eta,g0,aw", PARAMS)
def test_oracle_gamma_is_the_compiled_reference(eta, g0, aw):
    rng = np.random.default_rng(12)
    n = 4000
    gam = 1.0 + 1e-2 * rng.standard_normal(n)
    px, py = np.clip(rng.standard_normal((2, n)), -3.0, 3.0)
    p2 = _call(_ref().ref_get_p2, gam, px, py, eta, g0, aw)
    want = _call(_ref().ref_get_gamma, p2, px, py, eta, g0, aw)
    got = np.array([orc.get_gamma(p2[i], px[i], py[i], eta, g0, aw) for i in range(n)])
    assert np.array_equal(got, want)
    # and the reference's own round trip, to the conditioning of p2 (D = 1 - 1/u^2)
    assert np.max(np.abs(want - gam) / gam) < 1e-9


def test_constants_are_the_reference_modules():
    c = (ctypes.c_double * 3)()
    _ref().ref_constants(c)
    pi, c_light, q_e = c[0], c[1], c[2]
    # default-real literals widened to REAL(8): pi to 3.1415927410125732
    assert pi == PI_REF == float(np.float32(math.pi))
    assert c_light == h5io.C_LIGHT
    assert q_e == h5io.Q_E
