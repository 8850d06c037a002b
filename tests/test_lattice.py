"""Lattice elements on the CPU: the lattice-file parser (readLatt,
acc_lattice.f90:243-406) and pins of the oracle's element transforms by
invariants and closed forms (no GPU needed)."""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import pytest

import oracle as orc
from puffin_amd.lattice import Chicane, Drift, Modulation, Quad, Undulator, parse_lattice
from puffin_amd.params import (ModuleParams, PI_REF, UND_CURVED, UND_ELLIPTICAL, UND_HELICAL,
                               UND_PLANEPOLE, scaling)

# A lattice in the reference's format: a header line that is no element,
# '!' comments, blank lines, quoted and empty undulator types, comma
# separators, Fortran exponents.
LATTICE = """Module file for Puffin
! comment: QU in a comment is no quad
QU -17.46 17.46

UN  'planepole'  114   1.0   0.0   30   1.0   1.0   0.0    0.0
DR 14.633333
CH  0.5  0.136  1.5d-3
UN  ''  25   1.945667391924941   -0.01   20   0.3   0.7   0.1    0.2
MO 2.5 1.0e-3
UN 'helical', 10, 1.0, 0.0, 16, 0.0, 0.5, 0.0, 0.0
UN 'curved' 5 1.0 0.0 10 1.0 0.0 0.0 0.0
UN 'Helical' 5 1.0 0.0 10 0.2 0.4 0.0 0.0
"""


def test_parse_lattice_elements_and_scalings():
    rho = 0.005
    els = parse_lattice(LATTICE, rho)
    kinds = [type(e).__name__ for e in els]
    assert kinds == ["Quad", "Undulator", "Drift", "Chicane", "Undulator", "Modulation", "Undulator",
                     "Undulator", "Undulator"]
    q, u1, d, c, u2, m, u3, u4, u5 = els
    assert (q.fx, q.fy) == (-17.46, 17.46)
    lam = 4.0 * PI_REF * rho                      # one undulator period in zbar
    assert u1.und_type == UND_PLANEPOLE and (u1.ux, u1.uy) == (0.0, 1.0)   # fixed polarisation
    assert u1.nsteps == 114 * 30 and u1.h == lam / 30
    assert d.zbar == pytest.approx(14.633333 * lam, rel=1e-15)
    assert c.zbar == pytest.approx(0.5 * lam, rel=1e-15) and c.slip == pytest.approx(0.136 * lam, rel=1e-15)
    assert c.disp == 1.5e-3
    assert u2.und_type == UND_ELLIPTICAL and (u2.ux, u2.uy) == (0.3, 0.7)
    assert (u2.alpha, u2.taper, u2.kbeta_x, u2.kbeta_y) == (1.945667391924941, -0.01, 0.1, 0.2)
    assert u2.nsteps == 25 * 20 and u2.h == lam / 20
    assert (m.wavenum, m.mag) == (2.5, 1.0e-3)
    assert u3.und_type == UND_HELICAL and (u3.ux, u3.uy) == (1.0, 1.0)
    assert u4.und_type == UND_CURVED and (u4.ux, u4.uy) == (0.0, 1.0)
    # type names are compared as written: 'Helical' is an elliptical module
    assert u5.und_type == UND_ELLIPTICAL and (u5.ux, u5.uy) == (0.2, 0.4)


def test_parse_lattice_pi_is_the_reference_constant():
    """Drift lengths use pi as the reference's default-real literal."""
    (d,) = parse_lattice("DR 1.0\n", 0.01)
    assert d.zbar == pytest.approx(4.0 * PI_REF * 0.01, rel=1e-15)
    assert d.zbar != pytest.approx(4.0 * math.pi * 0.01, rel=1e-9)


# ---------------------------------------------------------------------------
# oracle element pins
# ---------------------------------------------------------------------------

def _mp(und_type=UND_ELLIPTICAL, one_d=0, periodic=0):
    rho = 0.01
    eta, kappa, fx, fy = scaling(rho, 1.0, 100.0, 1 / math.sqrt(2), 1 / math.sqrt(2))
    mp = ModuleParams(nx=9, ny=9, nz2=40, dx=0.5, dy=0.5, dz2=4 * PI_REF * rho / 16, rho=rho, eta=eta,
                      kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy, periodic=periodic, one_d=one_d,
                      und_type=und_type)
    if und_type == UND_CURVED:
        k = math.sqrt(eta / (8.0 * rho ** 2))
        mp = dataclasses.replace(mp, kx_und=k, ky_und=k, fx=0.0, fy=1.0)
    elif und_type == UND_PLANEPOLE:
        mp = dataclasses.replace(mp, fx=0.0, fy=1.0)
    return mp


def _parts(n=500, seed=3):
    rng = np.random.default_rng(seed)
    return dict(x=rng.standard_normal(n), y=rng.standard_normal(n), z2=rng.uniform(0.0, 3.0, n),
                px=0.3 * rng.standard_normal(n), py=0.3 * rng.standard_normal(n),
                gam=1.0 + 5e-3 * rng.standard_normal(n), chi=rng.uniform(0.5, 1.5, n))


def _copy(p):
    return {k: v.copy() for k, v in p.items()}


def test_oracle_drift_additive_and_consistent_with_gamma():
    mp = _mp()
    p0 = _parts()
    a, b = _copy(p0), _copy(p0)
    orc.drift(mp, a, 0.3)
    orc.drift(mp, a, 0.2)
    orc.drift(mp, b, 0.5)
    for k in ("x", "y", "z2"):
        np.testing.assert_allclose(a[k], b[k], rtol=0, atol=1e-13)
    for k in ("px", "py", "gam"):
        assert np.array_equal(b[k], p0[k])
    # the z2 rate is the p2 for which getGamma returns each particle's Gamma
    p2 = (b["z2"] - p0["z2"]) / 0.5
    g = np.array([orc.get_gamma(float(a2), float(a3), float(a4), mp.eta, mp.gamma_r, mp.aw)
                  for a2, a3, a4 in zip(p2, p0["px"], p0["py"])])
    np.testing.assert_allclose(g, p0["gam"], rtol=1e-9)
    # x moves with +px, y against +py, at the same rate per unit momentum
    dx, dy = b["x"] - p0["x"], b["y"] - p0["y"]
    np.testing.assert_allclose(dx * p0["py"], -dy * p0["px"], rtol=1e-12, atol=1e-15)
    rate = dx / (p0["px"] * 0.5) * p0["gam"] / (1.0 + mp.eta * p2)
    np.testing.assert_allclose(rate, 2.0 * mp.rho * mp.kappa / math.sqrt(mp.eta), rtol=1e-9)
    # 1-D: only z2 moves
    c = _copy(p0)
    orc.drift(_mp(one_d=1), c, 0.5)
    assert np.array_equal(c["x"], p0["x"]) and np.array_equal(c["y"], p0["y"])
    np.testing.assert_allclose(c["z2"], b["z2"], rtol=0, atol=1e-15)


def test_oracle_chicane_slip_dispersion_and_wrap():
    mp = _mp()
    p0 = _parts()
    a = _copy(p0)
    a["gam"][:] = 1.0
    orc.chicane(mp, a, 0.7, 0.25)                 # Gamma = 1: a pure slip
    np.testing.assert_allclose(a["z2"], p0["z2"] + 0.25, rtol=0, atol=1e-15)
    b = _copy(p0)
    orc.chicane(mp, b, 0.7, 0.0)                  # energy-dependent delay, -2 D (Gamma - 1)
    np.testing.assert_allclose(b["z2"], p0["z2"] - 1.4 * (p0["gam"] - 1.0), rtol=0, atol=1e-15)
    mpp = _mp(periodic=1)
    lenz2 = mpp.dz2 * (mpp.nz2 - 1)
    c = _copy(p0)
    orc.chicane(mpp, c, 0.0, 2.5 * lenz2)
    assert (c["z2"] >= 0).all() and (c["z2"] < lenz2).all()
    r = np.mod(p0["z2"] + 2.5 * lenz2, lenz2)
    np.testing.assert_allclose(c["z2"], r, rtol=0, atol=1e-12)


def test_oracle_quad_linear_kick_and_inverse():
    mp = _mp()
    p0 = _parts()
    a = _copy(p0)
    orc.quad(mp, a, 2.0, -3.0)
    dpx, dpy = a["px"] - p0["px"], a["py"] - p0["py"]
    kx = dpx * 2.0 / p0["x"]
    ky = dpy * -3.0 / p0["y"]
    # one strength for every particle, sqrt(eta)/(2 rho kappa), focusing in x for fx > 0
    c = math.sqrt(mp.eta) / (2.0 * mp.rho * mp.kappa)
    np.testing.assert_allclose(kx, c, rtol=1e-9)
    np.testing.assert_allclose(ky, -c, rtol=1e-9)
    orc.quad(mp, a, -2.0, 3.0)                    # the opposite quad undoes it
    for k in ("px", "py"):
        np.testing.assert_allclose(a[k], p0[k], rtol=0, atol=1e-14)
    for k in ("x", "y", "z2", "gam"):
        assert np.array_equal(a[k], p0[k])


def test_oracle_modulation_closed_form():
    p0 = _parts()
    a = _copy(p0)
    orc.modulation(a, 3.0, 0.0)
    assert np.array_equal(a["gam"], p0["gam"])
    a["z2"][:4] = [0.0, math.pi / 3.0, math.pi / 6.0, 2 * math.pi / 3.0]
    g0 = a["gam"].copy()
    orc.modulation(a, 3.0, 1e-3)
    np.testing.assert_allclose(a["gam"][:4] - g0[:4], [1e-3, -1e-3, 0.0, 1e-3], rtol=0, atol=1e-15)


def test_oracle_correct_trans_centres_the_beam():
    p0 = _parts()
    a = _copy(p0)
    orc.correct_trans(a)
    w = a["chi"]
    for k in ("x", "y", "px", "py"):
        assert abs(np.sum(w * a[k]) / np.sum(w)) < 1e-15
        np.testing.assert_allclose(np.diff(a[k]), np.diff(p0[k]), rtol=0, atol=1e-13)   # a shift
    b = _copy(a)
    orc.correct_trans(b)
    for k in ("x", "y", "px", "py"):
        np.testing.assert_allclose(b[k], a[k], rtol=0, atol=1e-15)


@pytest.mark.parametrize("und_type", [UND_ELLIPTICAL, UND_HELICAL])
def test_oracle_match_in_out_round_trip(und_type):
    """matchOut undoes matchIn at the same position and strength when the
    offsets do not depend on x, y."""
    mp = _mp(und_type)
    p0 = _parts()
    a = _copy(p0)
    for sz in (0.0, 0.37):
        orc.match(mp, 1.3, a, sz, +1)
        orc.match(mp, 1.3, a, sz, -1)
    for k in ("x", "y", "px", "py"):
        np.testing.assert_allclose(a[k], p0[k], rtol=0, atol=1e-13)


def test_oracle_match_planepole_and_curved_entry_at_zero():
    """At sz = 0 the entry puts px at -uy n2col less the pole-face term and
    moves only y (cos = 1, sin = 0); with ux = 0 nothing moves in y."""
    n2col = 1.2
    mp = _mp(UND_PLANEPOLE)
    p0 = _parts()
    a = _copy(p0)
    orc.match(mp, n2col, a, 0.0, +1)
    np.testing.assert_allclose(a["px"] - p0["px"], -n2col - mp.eta / (8 * mp.rho ** 2) * p0["x"] ** 2,
                               rtol=1e-13)
    for k in ("x", "y", "py"):
        np.testing.assert_allclose(a[k], p0[k], rtol=0, atol=1e-15)
    mc = _mp(UND_CURVED)
    b = _copy(p0)
    orc.match(mc, n2col, b, 0.0, +1)
    k2 = mc.kx_und ** 2
    np.testing.assert_allclose(b["px"] - p0["px"], -n2col - 0.5 * k2 * (p0["x"] ** 2 + p0["y"] ** 2),
                               rtol=1e-13)
    np.testing.assert_allclose(b["py"] - p0["py"], k2 * p0["x"] * p0["y"], rtol=1e-13, atol=1e-15)
    for k in ("x", "y"):
        np.testing.assert_allclose(b[k], p0[k], rtol=0, atol=1e-15)
