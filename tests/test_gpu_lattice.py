"""Lattice elements on the device (through the C ABI) against the oracle:
each element alone at FP64 (1e-13 relative to the largest value) and in the
mixed form (x, y, px, py held in FP32: 1e-6), and a short lattice run
through ``run_lattice`` against the oracle's element chain (1e-10)."""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import pytest

import oracle as orc
from lattice_case import LATTICE, LATTICE_ENDS, oracle_lattice
from test_gpu_parity import KEYS, _small_3d, relerr
from puffin_amd import gpu
from puffin_amd.lattice import parse_lattice, run_lattice
from puffin_amd.params import FP64, MIXED, PI_REF, UND_CURVED, UND_ELLIPTICAL, UND_HELICAL, UND_PLANEPOLE

pytestmark = pytest.mark.gpu

TOL = {FP64: 1e-13, MIXED: 1e-6}


def _case(periodic=0, und_type=UND_ELLIPTICAL, n=3000):
    mp, p, (ar, ai), h = _small_3d(periodic, n=n, matched=False)
    rng = np.random.default_rng(11)
    p["px"] = 0.3 * rng.standard_normal(n)
    p["py"] = 0.3 * rng.standard_normal(n)
    if und_type == UND_CURVED:
        k = math.sqrt(mp.eta / (8.0 * mp.rho ** 2))
        mp = dataclasses.replace(mp, kx_und=k, ky_und=k)
    return mp, p, (ar, ai)


def _element(name, g, mp, p):
    """Apply one element to the device state and to ``p`` on the oracle."""
    if name == "drift":
        g.drift(0.37)
        orc.drift(mp, p, 0.37)
    elif name == "chicane":
        lenz2 = mp.dz2 * (mp.nz2 - 1)
        g.chicane(2.5e-3, 0.3 * lenz2)
        orc.chicane(mp, p, 2.5e-3, 0.3 * lenz2)
    elif name == "quad":
        g.quad(2.7, -3.1)
        orc.quad(mp, p, 2.7, -3.1)
    elif name == "modulation":
        g.modulation(7.0, 2e-3)
        orc.modulation(p, 7.0, 2e-3)
    elif name == "correct_trans":
        g.correct_trans()
        orc.correct_trans(p)
    else:
        raise ValueError(name)


@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("name,periodic", [("drift", 0), ("chicane", 0), ("chicane", 1), ("quad", 0),
                                           ("modulation", 1), ("correct_trans", 0)])
def test_element_matches_oracle(name, periodic, prec):
    print(f"[gpu] lattice element {name} periodic={periodic} prec={prec}", flush=True)
    mp, p, _ = _case(periodic)
    with gpu.Integrator(mp, p["x"].size, prec) as g:
        g.upload_particles(p)
        ref = {k: v.copy() for k, v in g.download_particles().items()}   # the stored precision
        ref["chi"] = p["chi"].copy()
        _element(name, g, mp, ref)
        out = g.download_particles()
    for k in KEYS:
        assert relerr(out[k], ref[k]) < TOL[prec], k


@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("und_type", [UND_ELLIPTICAL, UND_PLANEPOLE, UND_CURVED, UND_HELICAL])
def test_match_in_and_out_match_oracle(und_type, prec):
    print(f"[gpu] lattice match und_type={und_type} prec={prec}", flush=True)
    mp, p, _ = _case(0, und_type)
    fx, fy = {UND_ELLIPTICAL: (mp.fx, mp.fy), UND_PLANEPOLE: (0.0, 1.0), UND_CURVED: (0.0, 1.0),
              UND_HELICAL: (1.0, 1.0)}[und_type]
    with gpu.Integrator(mp, p["x"].size, prec) as g:
        g.upload_particles(p)
        g.set_undulator(und_type, fx, fy, 0.0, 0.0, 0.0, 1.0)
        mu = g.mp
        ref = {k: v.copy() for k, v in g.download_particles().items()}
        g.match(0.0, 1.1, +1)
        orc.match(mu, 1.1, ref, 0.0, +1)
        out_in = g.download_particles()
        for k in KEYS:
            assert relerr(out_in[k], ref[k]) < TOL[prec], ("in", k)
        g.match(0.4321, 0.93, -1)
        orc.match(mu, 0.93, ref, 0.4321, -1)
        out = g.download_particles()
    for k in KEYS:
        assert relerr(out[k], ref[k]) < TOL[prec], ("out", k)


def test_element_arguments_are_checked():
    mp, p, _ = _case(0, n=256)
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        with pytest.raises(gpu.PuffinError):
            g.quad(0.0, 1.0)
        with pytest.raises(gpu.PuffinError):
            g.set_undulator(7, 1.0, 1.0)
        with pytest.raises(gpu.PuffinError):
            g.set_undulator(UND_ELLIPTICAL, 0.0, 0.0)
        out = g.download_particles()
    for k in KEYS:
        assert np.array_equal(out[k], p[k]), k


@pytest.mark.parametrize("periodic", [0, 1])
def test_lattice_run_matches_oracle(periodic):
    print(f"[gpu] lattice run periodic={periodic}", flush=True)
    mp, p, (ar, ai), h = _small_3d(periodic, n=2000, matched=False, nz2=None if periodic else 160)
    if not periodic:
        p["z2"] = np.random.default_rng(9).uniform(0.05, 0.35, p["z2"].size)
    els = parse_lattice(LATTICE, mp.rho)
    diff_step = 4.0 * PI_REF * mp.rho
    op, oar, oai, osz, ocs = oracle_lattice(mp, p, (ar, ai), els, diff_step=diff_step)
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        res = run_lattice(g, els, rho=mp.rho, diff_step=diff_step, redist_every=20)
        assert res.status == 0 and res.elements_done == len(els)
        out = g.download_particles()
        gr, gi = g.download_field()
    assert res.sz == osz and res.csteps == ocs
    for k in KEYS:
        assert relerr(out[k], op[k]) < 1e-10, k
    assert relerr(gr, oar) < 1e-10 and relerr(gi, oai) < 1e-10


def test_lattice_run_with_undulator_ends_matches_oracle():
    print("[gpu] lattice run with undulator ends", flush=True)
    mp, p, (ar, ai), h = _small_3d(1, n=2000, matched=False, und_ends=1)
    els = parse_lattice(LATTICE_ENDS, mp.rho)
    diff_step = 4.0 * PI_REF * mp.rho
    op, oar, oai, osz, ocs = oracle_lattice(mp, p, (ar, ai), els, diff_step=diff_step, und_ends=True)
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        res = run_lattice(g, els, rho=mp.rho, diff_step=diff_step, redist_every=20)
        assert res.status == 0
        out = g.download_particles()
        gr, gi = g.download_field()
    for k in KEYS:
        assert relerr(out[k], op[k]) < 1e-10, k
    assert relerr(gr, oar) < 1e-10 and relerr(gi, oai) < 1e-10


def test_one_d_drift_moves_z2_only_and_quad_is_inert():
    """On a 1-D mesh (the reference's own f1main module) driftSection moves z2
    only and Quad leaves the momenta alone (acc_lattice.f90:501, 544)."""
    print("[gpu] lattice 1-D drift and quad", flush=True)
    from f1main_case import initial_particles, load_fixture, module_params
    d, _ = load_fixture()
    mp = module_params()
    p = initial_particles(d["electrons"][:, 6])
    rng = np.random.default_rng(3)
    p["px"] = 0.2 * rng.standard_normal(p["x"].size)
    p["py"] = 0.2 * rng.standard_normal(p["x"].size)
    p["gam"] = 1.0 + 1e-3 * rng.standard_normal(p["x"].size)
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.drift(0.05)
        g.quad(2.0, -2.0)
        out = g.download_particles()
    ref = {k: v.copy() for k, v in p.items()}
    orc.drift(mp, ref, 0.05)
    for k in ("x", "y", "px", "py", "gam"):
        assert np.array_equal(out[k], p[k]), k
    assert relerr(out["z2"], ref["z2"]) < 1e-13
    assert np.abs(out["z2"] - p["z2"]).min() > 0.0
