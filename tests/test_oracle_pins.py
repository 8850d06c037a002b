"""Pins of the oracle's 3-D undulator fields and transverse equations that the
reference's own 1-D golden vectors cannot reach (its test deck has bz = 0):
the bz coupling terms of dp_r / dp_i (puffin_equations.f90:31-35, 61-65),
the main-section bz of every undulator type (bfields.f90:528-881), and the
entry and exit ramps (bfields.f90:84-109, 129-148, 239-278, 380-415,
481-500, 568-635, 763-867).  Each check comes from physics or calculus, not
from the formulas the oracle restates:

* Maxwell: a static undulator field is curl-free.  In the scaled
  coordinates (x, y in units of sqrt(l_g l_c), zbar in units of l_g, so
  that the undulator wavenumber in x units is sqrt(eta) / (2 rho)) that
  reads  d bz / d x = sqrt(eta) d bx / d zbar,  d bz / d y = sqrt(eta)
  d by / d zbar,  d bx / d y = d by / d x.  Finite differences of the
  oracle's fields must satisfy them, up to the reference's own stated
  departures (read from bfields.f90 and written out below).
* Natural focusing: averaged over the wiggle, the uncoupled equations give
  betatron motion at k_x = kappa (1 + eta p2) n2col fx / (sqrt 2 Gamma) (and
  fy for y).  The uncoupled 3-D orbit must follow x0 cos(k_x zbar).
* Ramps: with bz = 0 (1-D) and no field, d p_r / d u = n2col by and
  d p_i / d u = n2col bx (u = zbar / 2 rho), so the momenta are the
  integrals of the ramp envelopes times sin / cos, written here from their
  antiderivatives.  A beam entering at rest leaves the entry ramp on the
  matched orbit and the exit ramp with p_perp = 0.
A flipped sign, a wrong factor (kappa, n2col, 1 / Gamma, sqrt(eta) / 2 rho)
or a transposed x / y term fails these by orders of magnitude more than
their tolerances.  Not pinned: the 1 + eta p2 factor of the bz terms, a
1e-4 effect at these parameters.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as orc
from puffin_amd.configs import match_in
from puffin_amd.params import (PI_REF, UND_CURVED, UND_ELLIPTICAL, UND_HELICAL, UND_PLANEPOLE,
                               ModuleParams, gamma_r_for_eta, scaling)

P = PI_REF   # the reference's pi (a default-real literal) sets the ramp lengths
RHO = 0.01


def _mp(und=UND_ELLIPTICAL, ux=0.8, uy=0.6, **kw):
    gr = gamma_r_for_eta(1e-4, 1.0, max(ux, 1e-9), max(uy, 1e-9), und)
    eta, kappa, fx, fy = scaling(RHO, 1.0, gr, ux, uy, und)
    base = dict(nx=3, ny=3, nz2=3, dx=1e3, dy=1e3, dz2=0.01, rho=RHO, eta=eta, kappa=kappa, gamma_r=gr,
                aw=1.0, fx=fx, fy=fy, periodic=1, und_type=und, coupling=0, field_evolve=0)
    base.update(kw)
    return ModuleParams(**base)


def _grad(mp, place, z, x, y, e=1e-5):
    """Central differences of (bx, by, bz) in x, y and zbar."""
    f = lambda zz, xx, yy: np.array(orc.bfields(mp, place, zz, xx, yy))
    dx = (f(z, x + e, y) - f(z, x - e, y)) / (2 * e)
    dy = (f(z, x, y + e) - f(z, x, y - e)) / (2 * e)
    ez = e * mp.rho
    dz = (f(z + ez, x, y) - f(z - ez, x, y)) / (2 * ez)
    return dx, dy, dz, f(z, x, y)


POINTS = [(0.37, 0.3, -0.45), (1.13, -0.8, 0.25), (2.9, 0.6, 0.7)]   # (u = zbar / 2 rho, x, y)


# ---------------------------------------------------------------------------
# main section: curl-free fields
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("und", [UND_ELLIPTICAL, UND_HELICAL])
def test_main_section_elliptical_field_is_curl_free(und):
    mp = _mp(und) if und == UND_ELLIPTICAL else _mp(und, ux=1.0, uy=1.0)
    se = math.sqrt(mp.eta)
    for u, x, y in POINTS:
        dx, dy, dz, b = _grad(mp, 1, 2 * RHO * u, x, y)
        scale = se / (2 * RHO)
        assert dx[2] == pytest.approx(se * dz[0], abs=1e-7 * scale)
        assert dy[2] == pytest.approx(se * dz[1], abs=1e-7 * scale)
        assert dy[0] == pytest.approx(dx[1], abs=1e-9)
        # and the field is not trivially zero there
        assert abs(dx[2]) + abs(dy[2]) > 0.1 * scale


def test_main_section_planepole_bz_is_curl_free_but_by_drops_cosh():
    """bfields.f90:409-415 writes by = sin(u) without the cosh(k y) of
    doc/manual.tex:818-822, while bz = sinh(k y) cos(u) (:630-635) keeps
    its y-dependence: the curl then holds up to that factor."""
    mp = _mp(UND_PLANEPOLE, ux=0.0, uy=1.0)
    se = math.sqrt(mp.eta)
    k = se / (2 * RHO)
    for u, x, y in POINTS:
        dx, dy, dz, b = _grad(mp, 1, 2 * RHO * u, x, y)
        assert dy[2] == pytest.approx(math.cosh(k * y) * se * dz[1], rel=1e-7)
        assert dx[2] == pytest.approx(se * dz[0], abs=1e-9)
        assert b[0] == 0.0


def test_main_section_curved_pole_bz_carries_ky_over_kx():
    """The curved-pole bz of the entry ramp divides by ky (bfields.f90:568-575),
    which makes the field curl-free; the main section and the exit ramp divide
    by kx instead (:577-595).  So in the main section the curl holds up to
    the factor ky / kx, for both transverse components, and the x-y curl
    component holds exactly."""
    mp = _mp(UND_CURVED, ux=0.0, uy=1.0, kx_und=0.3, ky_und=0.5)
    se = math.sqrt(mp.eta)
    q = mp.ky_und / mp.kx_und
    for u, x, y in POINTS:
        dx, dy, dz, b = _grad(mp, 1, 2 * RHO * u, x, y)
        assert dx[2] == pytest.approx(q * se * dz[0], rel=1e-7)
        assert dy[2] == pytest.approx(q * se * dz[1], rel=1e-7)
        assert dy[0] == pytest.approx(dx[1], rel=1e-7)


# ---------------------------------------------------------------------------
# ramps: bz follows the envelopes of bx and by
# ---------------------------------------------------------------------------

def _ramp_mp(und, **kw):
    ux, uy = (0.8, 0.6) if und == UND_ELLIPTICAL else (0.0, 1.0)
    szfs = 8 * P * RHO
    return _mp(und, ux=ux, uy=uy, und_ends=1, szfs=szfs, szfe=szfs + 12 * P * RHO, **kw)


@pytest.mark.parametrize("u", [0.3, 1.9, 4.4, 6.1, 10.7, 12.3])
def test_elliptical_ramp_bz_follows_bx_and_by_envelopes(u):
    """Where a ramp scales bx by E_x(zbar) and by by E_y(zbar), a field that
    stays curl-free to leading order scales the x- and y-parts of bz by the
    same envelopes: bz_x cos(u) = -k bx sin(u) / fx * fx and
    bz_y sin(u) = k by cos(u), with k = sqrt(eta) / (2 rho).  The entry ramp
    and the y-part of the exit ramp do; the x-part of the exit ramp on
    pi rho < zbar - sZFE < 7 pi rho has the opposite sign (bfields.f90:803-809
    lacks the minus of bx's envelope at :259-260), a reference quirk the
    oracle keeps."""
    mp = _ramp_mp(UND_ELLIPTICAL)
    k = math.sqrt(mp.eta) / (2 * RHO)
    for place, z in ((0, 2 * RHO * u), (2, mp.szfe + 2 * RHO * u)):
        if place == 0 and z > mp.szfs:
            continue
        uu = (z if place == 0 else z - mp.szfe) / (2 * RHO)
        x, y = 0.7, -0.4
        dx, dy, dz, b = _grad(mp, place, z, x, y)
        sign = 1.0
        if place == 2 and P * RHO < z - mp.szfe < 7 * P * RHO:
            sign = -1.0
        assert dx[2] * math.cos(uu) == pytest.approx(-sign * k * b[0] * math.sin(uu), abs=1e-7 * k)
        assert dy[2] * math.sin(uu) == pytest.approx(k * b[1] * math.cos(uu), abs=1e-7 * k)


@pytest.mark.parametrize("und", [UND_PLANEPOLE, UND_CURVED])
@pytest.mark.parametrize("u", [0.4, 3.3, 7.9])
def test_planar_ramps_scale_bz_with_the_by_envelope(und, u):
    """Plane-pole and curved-pole ramps scale bx, by and bz by one envelope
    (u / 4 pi in, 1 - u / 4 pi out) of the main section's fields at the same
    (x, y) and phase.  The curved pole's entry bz divides by ky where the
    main section divides by kx (bfields.f90:568-575 against :588-595), so
    there the ratio carries kx / ky as well."""
    mp = _ramp_mp(und, kx_und=0.3, ky_und=0.5)
    x, y = 0.4, 0.6
    for place, z in ((0, 2 * RHO * u), (2, mp.szfe + 2 * RHO * u)):
        uu = (z if place == 0 else z - mp.szfe) / (2 * RHO)
        ramp = np.array(orc.bfields(mp, place, z, x, y))
        main = np.array(orc.bfields(mp, 1, 2 * RHO * uu, x, y))
        env = ramp[1] / main[1]
        assert env == pytest.approx(uu / (4 * P) if place == 0 else 1 - uu / (4 * P), rel=1e-9)
        q = mp.kx_und / mp.ky_und if (und == UND_CURVED and place == 0) else 1.0
        assert ramp[2] == pytest.approx(q * env * main[2], rel=1e-9, abs=1e-15)
        assert ramp[0] == pytest.approx(env * main[0], rel=1e-9, abs=1e-15)


# ---------------------------------------------------------------------------
# natural focusing of the uncoupled 3-D orbit (bz terms of dp_r, dp_i)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("gam", [1.0, 1.02])
def test_uncoupled_orbit_follows_natural_focusing_closed_form(gam):
    mp = _mp(UND_ELLIPTICAL)
    n2col = 0.9
    x0 = np.array([1.0, 0.0])
    y0 = np.array([0.0, 1.0])
    p = dict(x=x0.copy(), y=y0.copy(), z2=np.full(2, 0.1), px=np.zeros(2), py=np.zeros(2),
             gam=np.full(2, gam), chi=np.full(2, 1e-9))
    match_in(mp, n2col, p)
    xo, yo = p["x"] - x0, p["y"] - y0      # wiggle offsets at whole periods
    o = orc.Oracle(mp, p, (np.zeros(mp.nmesh), np.zeros(mp.nmesh)), taper=(n2col, n2col, 0.0))
    p2 = 0.5 * (orc.get_p2(gam, mp.fy * n2col, 0.0, mp.eta, mp.gamma_r, mp.aw) +
                orc.get_p2(gam, 0.0, mp.fx * n2col, mp.eta, mp.gamma_r, mp.aw))
    u = 1.0 + mp.eta * p2
    kx = mp.kappa * u * n2col * mp.fx / (math.sqrt(2.0) * gam)
    ky = mp.kappa * u * n2col * mp.fy / (math.sqrt(2.0) * gam)
    spp = 30
    h = 4 * P * RHO / spp
    szl = 0.0
    worst = 0.0
    for period in range(1, 41):
        for _ in range(spp):
            assert o.rk4par(szl, h) == 0
            szl += h
        z = period * 4 * P * RHO
        worst = max(worst, abs(o.parts["x"][0] - xo[0] - x0[0] * math.cos(kx * z)),
                    abs(o.parts["y"][1] - yo[1] - y0[1] * math.cos(ky * z)))
    # 40 periods turn x through 1.45 rad; the averaged closed form holds to
    # ~1e-4 of the amplitude (wiggle ripple, (k_beta / k_u)^2 terms)
    assert kx * z > 1.0
    assert worst < 5e-4, worst


# ---------------------------------------------------------------------------
# entry and exit ramps of the momenta (1-D: bz = 0)
# ---------------------------------------------------------------------------

def _anti_bx_in(u):   # antiderivative of (2u - P) / (6P) cos u
    return ((2 * u - P) * math.sin(u) + 2 * math.cos(u)) / (6 * P)


def _anti_bx_out(u):  # antiderivative of (7P - 2u) / (6P) cos u
    return ((7 * P - 2 * u) * math.sin(u) - 2 * math.cos(u)) / (6 * P)


def _int_bx_entry(u):
    """int_0^u of the entry bx envelope times cos: 0 below pi rho, a linear
    ramp to pi rho ... 7 pi rho, then cos."""
    if u <= P / 2:
        return 0.0
    if u <= 7 * P / 2:
        return _anti_bx_in(u) - _anti_bx_in(P / 2)
    return _anti_bx_in(7 * P / 2) - _anti_bx_in(P / 2) + math.sin(u) - math.sin(7 * P / 2)


def _int_by_entry(u):  # int_0^u (v / 4P) sin v
    return (math.sin(u) - u * math.cos(u)) / (4 * P)


def _int_bx_exit(u):
    if u <= P / 2:
        return math.sin(u)
    if u <= 7 * P / 2:
        return math.sin(P / 2) + _anti_bx_out(u) - _anti_bx_out(P / 2)
    return math.sin(P / 2) + _anti_bx_out(7 * P / 2) - _anti_bx_out(P / 2)


def _int_by_exit(u):   # int_0^u (1 - v / 4P) sin v
    return (1 - math.cos(u)) - (math.sin(u) - u * math.cos(u)) / (4 * P)


@pytest.mark.parametrize("und", [UND_ELLIPTICAL, UND_PLANEPOLE])
def test_ramps_match_the_orbit_in_and_out(und):
    """A macroparticle entering at rest (no matchIn) is put on the matched
    wiggle by the 2-period entry ramp and brought back to p_perp = 0 by the
    exit ramp; p_r and p_i follow the integrals of the envelopes at every
    step (120 steps per period)."""
    mp = _ramp_mp(und, nx=1, ny=1, nz2=5, dx=1.0, dy=1.0, one_d=1)
    n2col = 0.9
    gx = mp.fx if und == UND_ELLIPTICAL else 0.0
    gy = mp.fy if und == UND_ELLIPTICAL else 1.0
    p = dict(x=np.zeros(1), y=np.zeros(1), z2=np.full(1, 0.01), px=np.zeros(1), py=np.zeros(1),
             gam=np.ones(1), chi=np.full(1, 1e-9))
    o = orc.Oracle(mp, p, (np.zeros(mp.nmesh), np.zeros(mp.nmesh)), taper=(n2col, n2col, 0.0))
    spp = 120
    h = 4 * P * RHO / spp
    us, ue = mp.szfs / (2 * RHO), mp.szfe / (2 * RHO)
    fi_e = _int_bx_entry(us) + math.sin(ue) - math.sin(us)
    fr_e = _int_by_entry(us) + math.cos(us) - math.cos(ue)
    szl = 0.0
    worst = 0.0
    for k in range(7 * spp):          # entry (2 periods), main (3), exit (2)
        assert o.rk4par(szl, h) == 0
        szl += h
        z = (k + 1) * h
        if z <= mp.szfs:
            fi, fr = _int_bx_entry(z / (2 * RHO)), _int_by_entry(z / (2 * RHO))
        elif z <= mp.szfe:
            u = z / (2 * RHO)
            fi = _int_bx_entry(us) + math.sin(u) - math.sin(us)
            fr = _int_by_entry(us) + math.cos(us) - math.cos(u)
        else:
            u = (z - mp.szfe) / (2 * RHO)
            fi, fr = fi_e + _int_bx_exit(u), fr_e + _int_by_exit(u)
        worst = max(worst, abs(o.parts["px"][0] - gy * n2col * fr), abs(o.parts["py"][0] - gx * n2col * fi))
        if abs(z - mp.szfs) < 0.5 * h:
            # the entry ends on the matched orbit: p_r = -fy n2col cos(u), p_i = fx n2col sin(u)
            assert o.parts["px"][0] == pytest.approx(-gy * n2col * math.cos(us), abs=1e-7)
            assert o.parts["py"][0] == pytest.approx(gx * n2col * math.sin(us), abs=1e-7)
    assert worst < 1e-7, worst
    # the exit leaves the macroparticle with no transverse momentum
    assert abs(o.parts["px"][0]) < 1e-6 and abs(o.parts["py"][0]) < 1e-6
