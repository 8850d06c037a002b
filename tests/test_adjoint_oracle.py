"""The oracle's discrete adjoint (orc_rk4par_adjoint, orc_diffraction_adjoint)
pinned without reference data (the reference computes no derivatives):

* diffraction: the dot-product identity <D x, y> = <x, D^H y> to rounding,
  with masks, damping and the spectral filter active, periodic and temporal;
* one RK4 step: central finite differences of orc_rk4par along each state
  component separately (x, y, z2, p_r, p_i, Gamma, Re A, Im A), for every
  undulator type, the entry ramp, focusing, 1-D and periodic meshes;
* the per-stage dJ/dn2col against finite differences of a per-period taper
  table whose slot boundary splits the step's stages;
* a whole module (diffraction half steps, wraps, 48 steps): dJ/d(seed
  amplitude) and dJ/d(table) of the integrated power against central
  differences.
"""
from __future__ import annotations


import numpy as np
import pytest

import oracle as orc
from adjoint_case import OracleModule, small_module
from puffin_amd.params import PI_REF, ModuleParams, scaling, step_size

KEYS = orc.ADJ_KEYS


def _case(periodic=0, und_type=0, focusing=0, und_ends=0, n=600, one_d=False, seed=5):
    rho = 0.01
    eta, kappa, fx, fy = scaling(rho, 1.0, 100.0, 1 / np.sqrt(2), 1 / np.sqrt(2))
    h = step_size(rho, 30)
    nz2 = 33 if periodic else 40
    nx = 1 if one_d else 9
    mp = ModuleParams(nx=nx, ny=nx, nz2=nz2, dx=0.5, dy=0.5, dz2=4 * PI_REF * rho / 16, rho=rho, eta=eta,
                      kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy, periodic=periodic,
                      one_d=int(one_d), und_type=und_type, focusing=focusing, und_ends=und_ends,
                      kbeta_x=0.3, kbeta_y=0.2, kx_und=0.7, ky_und=0.9, nbx=3, nby=3, nbz2=7,
                      szfs=0.0 if not und_ends else 8 * h, szfe=60 * h, filt_cutoff=5.0, beta=1.0)
    rng = np.random.default_rng(seed)
    lz = (nz2 - 1) * mp.dz2
    zero = np.zeros(n)
    p = dict(x=zero.copy() if one_d else rng.uniform(-1.2, 1.2, n),
             y=zero.copy() if one_d else rng.uniform(-1.2, 1.2, n),
             z2=rng.uniform(0.1 * lz, 0.6 * lz, n), px=0.3 * rng.standard_normal(n),
             py=0.3 * rng.standard_normal(n), gam=1 + 0.01 * rng.standard_normal(n),
             chi=np.full(n, 2 * np.pi * lz / n * 50))
    ar = 0.05 * rng.standard_normal(mp.nmesh)
    ai = 0.05 * rng.standard_normal(mp.nmesh)
    if periodic:
        ntr = mp.nx * mp.ny
        ar[(nz2 - 1) * ntr:] = ar[:ntr]
        ai[(nz2 - 1) * ntr:] = ai[:ntr]
    return mp, p, ar, ai, h


CONFIGS = {
    "temporal": {},
    "periodic": dict(periodic=1),
    "one_d": dict(one_d=True),
    "focusing": dict(focusing=1),
    "planepole": dict(und_type=1),
    "curved": dict(und_type=3),
    "und_ends": dict(und_ends=1),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_rk4_step_adjoint_per_component(name):
    kw = CONFIGS[name]
    mp, p, ar, ai, h = _case(**kw)
    rng = np.random.default_rng(9)
    taper = (1.0, 1.0, -0.02)
    sz = 10 * h if kw.get("und_ends") else 3 * h
    lam = {k: rng.standard_normal(p["x"].size) for k in KEYS}
    lar = rng.standard_normal(mp.nmesh)
    lai = rng.standard_normal(mp.nmesh)

    def J(pp, a_r, a_i):
        o = orc.Oracle(mp, pp, (a_r, a_i), taper=taper)
        assert o.rk4par(sz, h) == 0
        return (sum(np.dot(lam[k], o.parts[k]) for k in KEYS) + np.dot(lar, o.ar) + np.dot(lai, o.ai))

    L = {k: lam[k].copy() for k in KEYS}
    LR, LI = lar.copy(), lai.copy()
    st, _ = orc.rk4par_adjoint(mp, taper, p, ar, ai, L, LR, LI, sz, h)
    assert st == 0
    scale = dict(x=1e-6, y=1e-6, z2=1e-7, px=1e-6, py=1e-6, gam=1e-7)
    comps = list(KEYS) + ["ar", "ai"]
    if mp.one_d:
        comps = [c for c in comps if c not in ("x", "y")]
    ntr = mp.nx * mp.ny
    for comp in comps:
        if comp in KEYS:
            d = rng.standard_normal(p["x"].size) * scale[comp]
            pp = dict(p)
            pm = dict(p)
            pp[comp] = p[comp] + d
            pm[comp] = p[comp] - d
            fd = (J(pp, ar, ai) - J(pm, ar, ai)) / 2
            ad = np.dot(L[comp], d)
        else:
            d = 1e-6 * rng.standard_normal(mp.nmesh)
            if mp.periodic:
                d[(mp.nz2 - 1) * ntr:] = 0.0   # the aliased input plane is not read
            if comp == "ar":
                fd = (J(p, ar + d, ai) - J(p, ar - d, ai)) / 2
                ad = np.dot(LR, d)
            else:
                fd = (J(p, ar, ai + d) - J(p, ar, ai - d)) / 2
                ad = np.dot(LI, d)
        assert abs(fd - ad) <= 2e-6 * abs(ad) + 2e-12, (name, comp, fd, ad)


def test_rk4_step_adjoint_n2col_per_stage():
    """dJ/dn2col of the four stages: a table slot boundary between stages 3
    and 4 separates stage 4's derivative from stages 1-3."""
    mp, p, ar, ai, h = _case()
    rng = np.random.default_rng(4)
    sz = 3 * h
    period = (sz + 0.75 * h)          # slot 0: stages 1-3, slot 1: stage 4
    lam = {k: rng.standard_normal(p["x"].size) for k in KEYS}
    lar = rng.standard_normal(mp.nmesh)
    lai = rng.standard_normal(mp.nmesh)
    table = np.array([1.02, 0.97])

    def J(tab):
        orc.set_taper_table(tab, 0.0, period)
        try:
            o = orc.Oracle(mp, p, (ar, ai))
            assert o.rk4par(sz, h) == 0
        finally:
            orc.set_taper_table(None)
        return sum(np.dot(lam[k], o.parts[k]) for k in KEYS) + np.dot(lar, o.ar) + np.dot(lai, o.ai)

    orc.set_taper_table(table, 0.0, period)
    try:
        assert [orc.taper_slot(z) for z in (sz, sz + h / 2, sz + h / 2, sz + h)] == [0, 0, 0, 1]
        L = {k: lam[k].copy() for k in KEYS}
        st, dn = orc.rk4par_adjoint(mp, (1.0, 1.0, 0.0), p, ar, ai, L, lar.copy(), lai.copy(), sz, h)
    finally:
        orc.set_taper_table(None)
    assert st == 0
    eps = 1e-6
    for k, expect in ((0, dn[0] + dn[1] + dn[2]), (1, dn[3])):
        tp = table.copy()
        tm = table.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (J(tp) - J(tm)) / (2 * eps)
        assert abs(fd - expect) <= 1e-6 * abs(expect), (k, fd, expect)


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("filt", [0, 1])
def test_diffraction_adjoint_dot_product(periodic, filt):
    mp, p, _, _, h = _case(periodic=periodic)
    mp.filter = filt
    rng = np.random.default_rng(3 + periodic + 2 * filt)
    x = rng.standard_normal(mp.nmesh) + 1j * rng.standard_normal(mp.nmesh)
    y = rng.standard_normal(mp.nmesh) + 1j * rng.standard_normal(mp.nmesh)
    mxy, mz = orc.masks(mp)
    assert mxy.max() > 0.1                       # the absorbing boundary takes part
    assert periodic or mz.max() > 0.1
    o = orc.Oracle(mp, p, (x.real.copy(), x.imag.copy()))
    o.diffraction_step(7 * h)
    lr, li = y.real.copy(), y.imag.copy()
    orc.diffraction_adjoint(mp, lr, li, 7 * h)
    lhs = np.sum(o.ar * y.real + o.ai * y.imag)
    rhs = np.sum(x.real * lr + x.imag * li)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    if periodic:
        ntr = mp.nx * mp.ny
        assert not lr[(mp.nz2 - 1) * ntr:].any() and not li[(mp.nz2 - 1) * ntr:].any()


@pytest.mark.parametrize("periodic", [1, 0])
def test_module_gradient_against_finite_differences(periodic):
    """J = integrated power after 4 periods (48 steps, diffraction half steps,
    wraps); central differences with h = 1e-6.  The seed amplitude enters
    smoothly; J's dependence on the table carries the tri-linear kinks of
    many particle crossings, so its differences scatter at ~1e-4 of the
    largest component."""
    c = small_module(periodic=periodic)
    om = OracleModule(c)
    J, gt, gs, _ = om.gradient(c["table"], c["seed"])
    eps = 1e-6
    fd = (om.J(c["table"], c["seed"] * (1 + eps)) - om.J(c["table"], c["seed"] * (1 - eps))) / (2 * eps * c["seed"])
    assert abs(fd - gs) <= 1e-6 * abs(gs), (fd, gs)
    for k in range(gt.size):
        tp = c["table"].copy()
        tm = c["table"].copy()
        tp[k] += eps
        tm[k] -= eps
        fdk = (om.J(tp, c["seed"]) - om.J(tm, c["seed"])) / (2 * eps)
        assert abs(fdk - gt[k]) <= 1e-3 * np.abs(gt).max(), (k, fdk, gt[k])


def test_adjoint_catches_planted_mistake():
    """A cotangent with a flipped field sign fails the dot-product check:
    the FD comparison above would see a sign or index slip."""
    mp, p, ar, ai, h = _case(periodic=1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(mp.nmesh)
    y = rng.standard_normal(mp.nmesh)
    o = orc.Oracle(mp, p, (x.copy(), np.zeros(mp.nmesh)))
    o.diffraction_step(7 * h)
    lr, li = y.copy(), np.zeros(mp.nmesh)
    orc.diffraction_adjoint(mp, lr, li, -7 * h)     # planted: wrong direction
    lhs = np.sum(o.ar * y)
    rhs = np.sum(x * lr)
    assert abs(lhs - rhs) > 1e-6 * abs(lhs)
