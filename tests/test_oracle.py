"""Pins of the CPU oracle: the reference's golden vectors, closed forms,
library routines and invariants (no GPU needed)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as orc
from puffin_amd.params import (ModuleParams, PI_REF, UND_ELLIPTICAL, UND_HELICAL,
                               UND_PLANEPOLE, scaling, step_size)
from f1main_case import NSTEPS, initial_particles, load_fixture, module_params


# ---------------------------------------------------------------------------
# Reference-held golden vectors (test/testMPIIntegration.pf:84-127)
# ---------------------------------------------------------------------------

def run_f1main_oracle():
    d, info = load_fixture()
    mp = module_params()
    parts = initial_particles(d["electrons"][:, 6])
    orc.match(mp, 1.0, parts, 0.0, +1)
    o = orc.Oracle(mp, parts, (np.zeros(mp.nz2), np.zeros(mp.nz2)), taper=(1.0, 1.0, 0.0))
    h = step_size(mp.rho, 30)
    szl = 0.0
    for _ in range(NSTEPS):
        assert o.rk4par(szl, h) == 0
        szl = szl + h
    orc.match(mp, o.taper[0], o.parts, szl, -1)
    orc.correct_trans(o.parts)
    return o, d, info


def test_f1main_constants_match_runinfo():
    _, info = load_fixture()
    mp = module_params()
    assert mp.eta == info["eta"]
    assert mp.kappa == info["kappa"]
    assert mp.dz2 == info["sLengthOfElmZ2"]
    assert step_size(mp.rho, 30) == info["sStepSize"]
    # the fixture's step size is only reproduced with pi as a default-real literal
    assert 4 * math.pi * 0.01 / 30 != info["sStepSize"]
    assert NSTEPS * step_size(mp.rho, 30) == pytest.approx(info["zbarTotal"], rel=1e-15)


def test_f1main_golden_particles_and_field():
    o, d, info = run_f1main_oracle()
    e = d["electrons"]
    order = np.argsort(o.parts["z2"])
    for c, k in enumerate(("x", "y", "z2", "px", "py", "gam")):
        got = o.parts[k][order]
        # the reference's own bar is 1e-10 absolute; the restatement is far tighter
        assert np.abs(got - e[:, c]).max() <= 1e-13, k
    assert np.abs(o.ar - d["aperp"][0]).max() <= 1e-13
    assert np.abs(o.ai - d["aperp"][1]).max() <= 1e-13


def test_f1main_golden_power_and_bunching():
    o, d, info = run_f1main_oracle()
    # 1-D 'power' dataset = |A|^2 * transArea (hdf5_puff.f90:207-210)
    pw = o.power() * info["transArea"]
    assert np.abs(pw - d["power"]).max() <= 1e-12 * d["power"].max()
    b = o.bunching(1, 32, 4 * PI_REF * 0.01)
    assert np.abs(b - d["bunching_fundamental"]).max() <= 1e-12 * d["bunching_fundamental"].max()


# ---------------------------------------------------------------------------
# p2 / Gamma (gamma_to_p2.f90)
# ---------------------------------------------------------------------------

def test_p2_closed_form_and_round_trip():
    eta, _, _, _ = scaling(0.01, 1.0, 100.0, 1.0, 1.0)
    # p_perp = 0, Gamma = 1: p2 = (1/sqrt(1 - 1/gamma_r^2) - 1)/eta
    p2 = orc.get_p2(1.0, 0.0, 0.0, eta, 100.0, 1.0)
    assert p2 == pytest.approx((1.0 / math.sqrt(1.0 - 1.0 / 100.0 ** 2) - 1.0) / eta, rel=1e-14)
    rng = np.random.default_rng(1)
    for _ in range(50):
        g = 1.0 + 0.01 * rng.standard_normal()
        pr, pi = rng.standard_normal(2)
        p2 = orc.get_p2(g, pr, pi, eta, 100.0, 1.0)
        assert orc.get_gamma(p2, pr, pi, eta, 100.0, 1.0) == pytest.approx(g, rel=1e-10)
    # resonance: on the matched helical orbit (|p_perp| = 1, Gamma = 1) p2 = 1
    assert orc.get_p2(1.0, 1.0, 0.0, eta, 100.0, 1.0) == pytest.approx(1.0, rel=1e-8)


# ---------------------------------------------------------------------------
# b-fields (bfields.f90) at points where the closed forms are known
# ---------------------------------------------------------------------------

def _mp3d(**kw):
    eta, kappa, fx, fy = scaling(0.01, 1.0, 100.0, kw.pop("ux", 1.0), kw.pop("uy", 1.0),
                                 kw.get("und_type", UND_ELLIPTICAL))
    base = dict(nx=9, ny=9, nz2=16, dx=0.5, dy=0.5, dz2=0.01, rho=0.01, eta=eta, kappa=kappa,
                gamma_r=100.0, aw=1.0, fx=fx, fy=fy)
    base.update(kw)
    return ModuleParams(**base)


def test_bfields_main_section_values():
    mp = _mp3d(ux=0.6, uy=0.8)
    rho = mp.rho
    se = math.sqrt(mp.eta)
    x, y = 0.3, -0.7
    bx, by, bz = orc.bfields(mp, 1, 0.0, x, y)
    assert (bx, by) == (pytest.approx(0.6), pytest.approx(0.0, abs=1e-15))
    assert bz == pytest.approx(se / (2 * rho) * 0.8 * y, rel=1e-14)
    bx, by, bz = orc.bfields(mp, 1, math.pi * rho, x, y)   # szt = pi/2
    assert bx == pytest.approx(0.0, abs=1e-15)
    assert by == pytest.approx(0.8, rel=1e-15)
    assert bz == pytest.approx(-se / (2 * rho) * 0.6 * x, rel=1e-12)


def test_bfields_helical_is_elliptical_unit():
    mh = _mp3d(und_type=UND_HELICAL)
    me = _mp3d(ux=1.0, uy=1.0)
    for place in (0, 1, 2):
        for z in (0.01, 0.05, 0.3):
            assert orc.bfields(mh, place, z, 0.4, 0.2) == orc.bfields(me, place, z, 0.4, 0.2)


def test_bfields_entry_ramp_joins_main_section():
    mp = _mp3d(und_ends=1, ux=0.6, uy=0.8)
    rho = mp.rho
    z = 8 * PI_REF * rho      # sZFS of initUndulator (acc_lattice.f90:882)
    entry = orc.bfields(mp, 0, z, 0.3, -0.2)
    main = orc.bfields(mp, 1, z, 0.3, -0.2)
    assert np.allclose(entry, main, rtol=1e-6, atol=1e-9)
    # below pi*rho the entry x-field is off
    assert orc.bfields(mp, 0, 0.5 * PI_REF * rho, 0.3, -0.2)[0] == 0.0


def test_bfields_planepole_and_focusing():
    mp = _mp3d(und_type=UND_PLANEPOLE)
    se = math.sqrt(mp.eta)
    y = 0.4
    bx, by, bz = orc.bfields(mp, 1, 0.0, 0.1, y)
    assert bx == 0.0 and by == 0.0
    assert bz == pytest.approx(math.sinh(se / (2 * mp.rho) * y), rel=1e-14)
    mf = _mp3d(focusing=1, kbeta_x=0.3, kbeta_y=0.2)
    m0 = _mp3d()
    fx = orc.bfields(mf, 1, 0.02, 0.5, 0.25)
    f0 = orc.bfields(m0, 1, 0.02, 0.5, 0.25)
    sf = math.sqrt(mf.eta)
    assert fx[0] - f0[0] == pytest.approx(sf * 0.2 ** 2 / mf.kappa * 0.25, rel=1e-12)
    assert fx[1] - f0[1] == pytest.approx(-sf * 0.3 ** 2 / mf.kappa * 0.5, rel=1e-12)


# ---------------------------------------------------------------------------
# Interpolation and deposition (system_interpolation.f90)
# ---------------------------------------------------------------------------

def _random_parts(mp, n, rng, margin=1.0):
    hx = (mp.nx - 1) / 2 * mp.dx - margin * mp.dx
    hy = (mp.ny - 1) / 2 * mp.dy - margin * mp.dy
    lz = (mp.nz2 - 2) * mp.dz2
    return dict(x=rng.uniform(-hx, hx, n), y=rng.uniform(-hy, hy, n),
                z2=rng.uniform(0.0, lz, n), px=rng.standard_normal(n),
                py=rng.standard_normal(n), gam=1.0 + 0.01 * rng.standard_normal(n),
                chi=rng.uniform(0.5, 1.5, n))


def test_interp_partition_of_unity_and_linear_reproduction():
    mp = _mp3d(nx=7, ny=5, nz2=12, dx=0.3, dy=0.4, dz2=0.05)
    rng = np.random.default_rng(2)
    X = (np.arange(mp.nx) - (mp.nx - 1) / 2) * mp.dx
    Y = (np.arange(mp.ny) - (mp.ny - 1) / 2) * mp.dy
    Z = np.arange(mp.nz2) * mp.dz2
    zz, yy, xx = np.meshgrid(Z, Y, X, indexing="ij")
    a, b, c, d, e = 0.7, -1.3, 2.1, 0.4, 0.9
    f = a + b * xx + c * yy + d * zz + e * xx * yy * zz    # tri-linear per cell
    parts = _random_parts(mp, 200, rng)
    for i in range(200):
        flag, node, w = orc.interp_one(mp, parts["x"][i], parts["y"][i], parts["z2"][i])
        assert flag == 0
        assert w.sum() == pytest.approx(1.0, abs=1e-14)
        x, y, z = parts["x"][i], parts["y"][i], parts["z2"][i]
        assert (w * f.reshape(-1)[node]).sum() == pytest.approx(a + b * x + c * y + d * z + e * x * y * z,
                                                                abs=1e-12)


def test_deposit_is_adjoint_of_interpolation_and_conserves_charge():
    mp = _mp3d(nx=6, ny=6, nz2=10, dx=0.25, dy=0.25, dz2=0.05)
    rng = np.random.default_rng(3)
    parts = _random_parts(mp, 300, rng)
    ar = rng.standard_normal(mp.nmesh)
    ai = rng.standard_normal(mp.nmesh)
    st, fr, fi, dar, dai = orc.interp_deposit(mp, parts, ar, ai)
    assert st == 0
    p2 = np.array([orc.get_p2(g, r, i, mp.eta, mp.gamma_r, mp.aw)
                   for g, r, i in zip(parts["gam"], parts["px"], parts["py"])])
    dV3 = mp.dx * mp.dy * mp.dz2
    sr = parts["chi"] / dV3 * (1 + mp.eta * p2) * parts["px"] / parts["gam"]
    si = parts["chi"] / dV3 * (1 + mp.eta * p2) * parts["py"] / parts["gam"]
    # <interp(A), s> = <A, deposit(s)>
    assert (fr * sr).sum() == pytest.approx((ar * dar).sum(), rel=1e-12)
    assert (fi * si).sum() == pytest.approx((ai * dai).sum(), rel=1e-12)
    # total current
    assert dar.sum() == pytest.approx(sr.sum(), rel=1e-12)
    assert dai.sum() == pytest.approx(si.sum(), rel=1e-12)


def test_deposit_brute_force_small_mesh():
    mp = _mp3d(nx=4, ny=4, nz2=4, dx=1.0, dy=1.0, dz2=1.0)
    rng = np.random.default_rng(4)
    parts = _random_parts(mp, 5, rng, margin=0.0)
    st, fr, fi, dar, dai = orc.interp_deposit(mp, parts, np.zeros(64), np.zeros(64))
    assert st == 0
    want = np.zeros((4, 4, 4))
    for i in range(5):
        gx = parts["x"][i] + 1.5
        gy = parts["y"][i] + 1.5
        gz = parts["z2"][i]
        p2 = orc.get_p2(parts["gam"][i], parts["px"][i], parts["py"][i], mp.eta, mp.gamma_r, mp.aw)
        s = parts["chi"][i] * (1 + mp.eta * p2) * parts["px"][i] / parts["gam"][i]
        for kz in range(4):
            for ky in range(4):
                for kx in range(4):
                    wx = max(0.0, 1 - abs(gx - kx))
                    wy = max(0.0, 1 - abs(gy - ky))
                    wz = max(0.0, 1 - abs(gz - kz))
                    want[kz, ky, kx] += wx * wy * wz * s
    assert np.allclose(dar.reshape(4, 4, 4), want, rtol=1e-12, atol=1e-14)


def test_interp_bounds_flags():
    mp = _mp3d(nx=5, ny=5, nz2=8, dx=1.0, dy=1.0, dz2=1.0)
    assert orc.interp_one(mp, 2.5, 0.0, 1.0)[0] == 1   # beyond +halfx
    assert orc.interp_one(mp, -2.01, 0.0, 1.0)[0] == 1
    assert orc.interp_one(mp, 0.0, 0.0, 6.99)[0] == 0
    assert orc.interp_one(mp, 0.0, 0.0, 7.0)[0] == 4    # z2node >= NZ2: fatal
    mpp = _mp3d(nx=5, ny=5, nz2=8, dx=1.0, dy=1.0, dz2=1.0, periodic=1)
    f, node, w = orc.interp_one(mpp, 0.0, 0.0, 6.5)     # last periodic cell wraps to plane 0
    assert f == 0 and node[4] // 25 == 0


# ---------------------------------------------------------------------------
# FFT and diffraction (transforms.f90, diffraction.f90, masks.f90)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 11, 12, 33, 43, 64, 128, 129, 257])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    fwd = orc.fft1d(a, -1)
    assert np.allclose(fwd, np.fft.fft(a), rtol=1e-12, atol=1e-12 * max(1, n))
    bwd = orc.fft1d(a, +1)
    assert np.allclose(bwd, np.fft.ifft(a) * n, rtol=1e-12, atol=1e-12 * max(1, n))


def test_fft3d_brute_force_dft():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((8, 5, 5)) + 1j * rng.standard_normal((8, 5, 5))
    want = np.zeros_like(a)
    idx = [np.arange(s) for s in a.shape]
    for k3 in range(8):
        for k2 in range(5):
            for k1 in range(5):
                ph = np.exp(-2j * np.pi * (idx[0][:, None, None] * k3 / 8 + idx[1][None, :, None] * k2 / 5
                                          + idx[2][None, None, :] * k1 / 5))
                want[k3, k2, k1] = (a * ph).sum()
    assert np.allclose(orc.fft3d(a, -1), want, rtol=1e-12, atol=1e-11)


def test_kvalues_match_fftfreq():
    mp = _mp3d(nx=129, ny=33, nz2=128, dx=0.15, dy=0.3, dz2=0.004)
    kx, ky, kz = orc.kvalues(mp)
    assert np.allclose(kx, 2 * np.pi * np.fft.fftfreq(129, d=0.15), rtol=1e-14, atol=1e-14)
    assert np.allclose(ky, 2 * np.pi * np.fft.fftfreq(33, d=0.3), rtol=1e-14, atol=1e-14)
    assert np.allclose(kz, 2 * np.pi * np.fft.fftfreq(128, d=0.004), rtol=1e-14, atol=1e-12)
    mpp = _mp3d(nx=9, ny=9, nz2=129, dx=0.15, dy=0.3, dz2=0.004, periodic=1)
    _, _, kzp = orc.kvalues(mpp)
    assert kzp.size == 128
    assert np.allclose(kzp, 2 * np.pi * np.fft.fftfreq(128, d=0.004), rtol=1e-14, atol=1e-12)


def test_masks_closed_forms():
    mp = _mp3d(nx=41, ny=41, nz2=100, dx=0.2, dy=0.2, dz2=0.01, nbx=16, nby=16, nbz2=37)
    mxy, mz = orc.masks(mp)
    m = mxy.reshape(41, 41)
    # Mask1D2 is 1 on the outermost node, 0 in the interior, cos^2 in between
    assert m[20, 0] == pytest.approx(1.0) and m[20, 40] == pytest.approx(1.0)
    assert m[20, 20] == 0.0
    lb = 15 * 0.2
    assert m[20, 3] == pytest.approx(math.cos(PI_REF * (3 * 0.2) / (2 * lb)) ** 2, rel=1e-14)
    # corner: Mask1D2_x + Maskf4_x * Mask1D2_y with Maskf4 = 1 at its inner edge
    assert m[0, 15] == pytest.approx(math.cos(PI_REF * (15 * 0.2) / (2 * lb)) ** 2 + 1.0, rel=1e-12)
    # z2 mask: zero before the first boundary node, Lf = nx - dx quirk (masks.f90:402)
    fbn = 100 - 37 + 1
    assert np.all(mz[: fbn - 2] == 0.0)
    x0 = (100 - 1 * 0.01) - 36 * 0.01 / 2.0
    assert mz[99] == pytest.approx(math.cos(PI_REF * (99 * 0.01 - x0) / (36 * 0.01)) ** 2, rel=1e-12)
    mpp = _mp3d(nx=41, ny=41, nz2=100, dx=0.2, dy=0.2, dz2=0.01, periodic=1)
    assert np.all(orc.masks(mpp)[1] == 0.0)


def _gauss_case(nx=64, dx=0.25, nz=16, dz=0.05, w0=1.2, q=3, h=0.05):
    mp = ModuleParams(nx=nx, ny=nx, nz2=nz, dx=dx, dy=dx, dz2=dz, rho=0.01, eta=1e-4, kappa=0.5,
                      gamma_r=100.0, aw=1.0, beta=0.0, filter=0, filt_cutoff=0.0, nbz2=3)
    X = (np.arange(nx) - (nx - 1) / 2) * dx
    xx = X[None, :]
    yy = X[:, None]
    kz = 2 * np.pi * q / (nz * dz)
    Z = np.arange(nz) * dz
    a = np.exp(-(xx ** 2 + yy ** 2) / w0 ** 2)[None, :, :] * np.exp(1j * kz * Z)[:, None, None]
    return mp, a, kz, xx, yy, Z


def test_diffraction_gaussian_beam_closed_form():
    """multiplyexp on one kz2 mode is paraxial propagation of a Gaussian:
    A = w0^2/q exp(-r^2/q) with q = w0^2 - 2 i h / kz2 (sign fixes FFTW's
    convention); beta = 0 makes AbsorptionStep the identity."""
    w0, h = 1.2, 0.05
    mp, a, kz, xx, yy, Z = _gauss_case(w0=w0, h=h)
    ar = np.ascontiguousarray(a.real).reshape(-1).copy()
    ai = np.ascontiguousarray(a.imag).reshape(-1).copy()
    o = orc.Oracle(mp, dict(x=np.zeros(1), y=np.zeros(1), z2=np.zeros(1), px=np.zeros(1),
                            py=np.zeros(1), gam=np.ones(1), chi=np.zeros(1)), (ar, ai))
    o.diffraction_step(h)
    got = (o.ar + 1j * o.ai).reshape(a.shape)
    qq = w0 ** 2 - 2j * h / kz
    want = (w0 ** 2 / qq) * np.exp(-(xx ** 2 + yy ** 2) / qq)[None] * np.exp(1j * kz * Z)[:, None, None]
    assert np.abs(got - want).max() < 1e-9
    wrong = (w0 ** 2 / np.conj(qq)) * np.exp(-(xx ** 2 + yy ** 2) / np.conj(qq))[None] * \
        np.exp(1j * kz * Z)[:, None, None]
    assert np.abs(got - wrong).max() > 1e-3


def test_diffraction_conserves_energy_without_absorption_or_filter():
    rng = np.random.default_rng(6)
    mp = ModuleParams(nx=9, ny=7, nz2=17, dx=0.3, dy=0.3, dz2=0.02, rho=0.01, eta=1e-4,
                      kappa=0.5, gamma_r=100.0, aw=1.0, beta=0.0, filter=0, filt_cutoff=0.0,
                      nbx=3, nby=3, nbz2=3)
    ar = rng.standard_normal(mp.nmesh)
    ai = rng.standard_normal(mp.nmesh)
    e0 = (ar ** 2 + ai ** 2).sum()
    o = orc.Oracle(mp, dict(x=np.zeros(1), y=np.zeros(1), z2=np.zeros(1), px=np.zeros(1),
                            py=np.zeros(1), gam=np.ones(1), chi=np.zeros(1)), (ar, ai))
    o.diffraction_step(0.37)
    assert (o.ar ** 2 + o.ai ** 2).sum() == pytest.approx(e0, rel=1e-12)
    # the filter removes the low |kz2| band; with a cutoff energy must drop
    mp2 = ModuleParams(**{**mp.asdict(), "filter": 1, "filt_cutoff": 60.0})
    o2 = orc.Oracle(mp2, o.parts, (ar, ai))
    o2.diffraction_step(0.37)
    assert (o2.ar ** 2 + o2.ai ** 2).sum() < 0.9 * e0


def test_absorption_damps_only_the_boundary():
    mp, a, kz, xx, yy, Z = _gauss_case()
    # periodic mesh: no z2 absorbing layer (diffraction.f90:447); NZ2 = FFT length + 1
    mp = ModuleParams(**{**mp.asdict(), "beta": 1.0, "nbx": 8, "nby": 8, "periodic": 1,
                         "nz2": mp.nz2 + 1})
    a = np.concatenate([a, a[:1]], axis=0)
    ar = np.ascontiguousarray(a.real).reshape(-1).copy()
    ai = np.ascontiguousarray(a.imag).reshape(-1).copy()
    dummy = dict(x=np.zeros(1), y=np.zeros(1), z2=np.zeros(1), px=np.zeros(1), py=np.zeros(1),
                 gam=np.ones(1), chi=np.zeros(1))
    o = orc.Oracle(mp, dummy, (ar, ai))
    o.diffraction_step(0.05)
    mp0 = ModuleParams(**{**mp.asdict(), "beta": 0.0})
    o0 = orc.Oracle(mp0, dummy, (ar, ai))
    o0.diffraction_step(0.05)
    # the Gaussian sits inside the interior: absorption changes it very little
    d = np.abs((o.ar - o0.ar) + 1j * (o.ai - o0.ai))
    assert d.max() < 1e-6


# ---------------------------------------------------------------------------
# RK4 (puffin_mpi_RK4.f90) on the decoupled helical orbit
# ---------------------------------------------------------------------------

def _free_orbit_case(h_div=1, one_d=1):
    """Uncoupled electrons; in 1-D bz = 0 so the orbit is closed-form."""
    eta, kappa, fx, fy = scaling(0.01, 1.0, 100.0, 1.0, 1.0)
    h = step_size(0.01, 30) / h_div
    nt = 1 if one_d else 17
    mp = ModuleParams(nx=nt, ny=nt, nz2=64, dx=1.0 if one_d else 0.05, dy=1.0 if one_d else 0.05,
                      dz2=0.05, rho=0.01, eta=eta, one_d=one_d,
                      kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy, coupling=0,
                      szfs=0.0, szfe=1e9)
    n = 4
    parts = dict(x=np.zeros(n), y=np.zeros(n), z2=np.linspace(0.5, 1.5, n), px=np.zeros(n),
                 py=np.zeros(n), gam=np.ones(n), chi=np.full(n, 1e-3))
    orc.match(mp, 1.0, parts, 0.0, +1)
    return mp, parts, h


def test_rk4_free_orbit_and_fourth_order_convergence():
    errs = []
    for div in (1, 2):
        mp, parts, h = _free_orbit_case(div)
        o = orc.Oracle(mp, parts, (np.zeros(mp.nmesh), np.zeros(mp.nmesh)))
        szl = 0.0
        for _ in range(30 * div):
            assert o.rk4par(szl, h) == 0
            szl += h
        P = o.parts
        # analytic orbit of the uncoupled elliptical undulator with fx = fy = 1:
        # pr = -cos(z/2rho), pi = sin(z/2rho), Gamma constant
        err = max(np.abs(P["px"] + np.cos(szl / (2 * mp.rho))).max(),
                  np.abs(P["py"] - np.sin(szl / (2 * mp.rho))).max())
        assert np.abs(P["gam"] - 1.0).max() == 0.0
        errs.append(err)
    assert errs[0] < 1e-3
    assert 12.0 < errs[0] / errs[1] < 20.0     # RK4: error falls as h^4


def test_rk4_flag_leaves_state_unchanged():
    mp, parts, h = _free_orbit_case(one_d=0)
    parts["x"][1] = 10.0   # outside the transverse mesh
    o = orc.Oracle(mp, parts, (np.zeros(mp.nmesh), np.ones(mp.nmesh)))
    before = {k: v.copy() for k, v in o.parts.items()}
    assert o.rk4par(0.0, h) == 1
    for k in before:
        assert np.array_equal(before[k], o.parts[k])
    assert np.all(o.ai == 1.0)


def test_periodic_wrap():
    mp = _mp3d(nz2=33, dz2=0.1, periodic=1)
    o = orc.Oracle(mp, dict(x=np.zeros(3), y=np.zeros(3), z2=np.array([-0.05, 1.0, 3.25]),
                            px=np.zeros(3), py=np.zeros(3), gam=np.ones(3), chi=np.ones(3)),
                   (np.zeros(mp.nmesh), np.zeros(mp.nmesh)))
    o.wrap_z2()
    assert np.allclose(o.parts["z2"], [3.15, 1.0, 0.05])


# ---------------------------------------------------------------------------
# Slice statistics and beam current (getSliceTwiss / getCurr) against the
# reference's own integrated file (test/expected/f1main_expected_integrated.h5)
# ---------------------------------------------------------------------------

def test_f1main_golden_slice_stats_and_current():
    import os
    from puffin_amd import h5io
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    fi = os.path.join(golden, "f1main_expected_integrated.h5")
    _, _, run = h5io.read_field(os.path.join(golden, "f1main_expected_field.h5"))
    o, d, info = run_f1main_oracle()
    L = 4 * PI_REF * 0.01                          # one slice per resonant wavelength
    st, ss = o.slice_stats(32, L)
    assert st == 0
    st, cur = o.current(33, L)
    assert st == 0
    mine = h5io.slice_datasets(ss, run, cur, L)
    for name, v in mine.items():
        ref = h5io.read_dataset(fi, name)
        # relative 1e-12, or absolute 1e-16 for the x, y moments (|x| ~ 1e-6 in 1-D,
        # themselves matched to 1e-13 absolute by the particle test above)
        assert np.abs(v - ref).max() <= 1e-12 * np.abs(ref).max() + 1e-16, name
    # the reference's constants are single-precision literals: exact ones miss
    exact = cur * run.npk_bar * run.trans_area * 1.60217653e-19 / L * 2.99792458e8 / run.lc
    ref = h5io.read_dataset(fi, "beamCurrent")
    assert np.abs(exact - ref).max() > 1e-9 * ref.max()


def test_slice_stats_closed_forms_and_periodic_wrap():
    """Two macroparticles per slice at known positions: the statistics follow
    by hand; on a periodic mesh slice index nslices wraps to slice 1
    (av_write.f90:507-513) and on a temporal mesh out-of-range particles are
    dropped."""
    mp = _mp3d(periodic=0)
    L = 0.5
    z2 = np.array([0.1, 0.3, 0.6, 0.9, 1.7])         # slices 1, 1, 2, 2, 4 (of 3: dropped)
    p = dict(x=np.array([1.0, 3.0, -1.0, 1.0, 5.0]), y=np.array([2.0, 2.0, 0.0, 4.0, 0.0]), z2=z2,
             px=np.array([0.5, 1.5, 0.0, 0.0, 9.0]), py=np.array([1.0, -1.0, 2.0, 2.0, 0.0]),
             gam=np.array([1.0, 1.2, 0.9, 1.1, 3.0]), chi=np.array([1.0, 1.0, 2.0, 2.0, 1.0]))
    st, s = orc.slice_stats(mp, p, 3, L)
    assert st != 0                                  # the particle past the last slice
    np.testing.assert_allclose(s["sdata"], [2.0, 4.0, 0.0])
    np.testing.assert_allclose(s["aveX"], [2.0, 0.0, 0.0])
    np.testing.assert_allclose(s["sdX"], [1.0, 1.0, 0.0])
    np.testing.assert_allclose(s["aveY"], [2.0, 2.0, 0.0])
    np.testing.assert_allclose(s["sdY"], [0.0, 2.0, 0.0])
    np.testing.assert_allclose(s["avePY"], [0.0, -2.0, 0.0])   # mean of the stored -py
    np.testing.assert_allclose(s["sdpX"], [0.5, 0.0, 0.0])
    np.testing.assert_allclose(s["aveGamma"], [1.1, 1.0, 0.0])
    np.testing.assert_allclose(s["aveDgamma"], [0.1, 0.1, 0.0], rtol=1e-12)
    # slice 1: x = 1, 3 and px = 0.5, 1.5 are perfectly correlated: no emittance
    assert abs(s["eX"][0]) < 1e-7
    mpp = _mp3d(periodic=1)
    st, s = orc.slice_stats(mpp, dict(p, z2=np.array([0.1, 0.3, 0.6, 0.9, 1.4])), 3, L)
    assert st == 0
    np.testing.assert_allclose(s["sdata"], [3.0, 4.0, 0.0])    # slice 3 wrapped onto slice 1
    st, c = orc.current(mp, dict(p, z2=np.array([0.25, 0.5, 0.6, 0.9, 1.0])), 4, L)
    assert st == 0
    # node j at (j - 1) L: 0.25 -> half to nodes 1 and 2; 0.5 -> all to node 2 (ceil);
    # 0.6 -> 0.8 / 0.2 to nodes 2 / 3; 0.9 -> 0.2 / 0.8; 1.0 -> all to node 3
    np.testing.assert_allclose(c, [0.5, 0.5 + 1.0 + 2 * 0.8 + 2 * 0.2, 2 * 0.2 + 2 * 0.8 + 1.0, 0.0], rtol=1e-12)
    assert np.isclose(c.sum(), p["chi"].sum())
