"""The deterministic deposit (puff_set_deterministic, FP64 form): the same
inputs give the same state bit for bit, whatever the order in which atomics
land and whatever the macroparticles' slot order (re-sorted every step, or
never), and the state still matches the oracle at the 1e-10 parity bar."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as orc
from puffin_amd import gpu
from puffin_amd.configs import case_c0, gaussian_seed, match_in
from puffin_amd.params import FP64, MIXED, PI_REF, ModuleParams, scaling, step_size
from puffin_amd.undulator import und_section

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z2", "px", "py", "gam")


def _dense_case(periodic, n=60000):
    """Many macroparticles per cell (LDS-tile deposits) plus a sparse halo
    (workgroups whose node box overflows the tile deposit with global
    atomics)."""
    rho = 0.01
    eta, kappa, fx, fy = scaling(rho, 1.0, 100.0, 1 / math.sqrt(2), 1 / math.sqrt(2))
    h = step_size(rho, 30)
    mp = ModuleParams(nx=17, ny=17, nz2=65 if periodic else 64, dx=0.4, dy=0.4, dz2=4 * PI_REF * rho / 16,
                      rho=rho, eta=eta, kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy,
                      periodic=periodic, szfs=0.0, szfe=60 * h, nbx=4, nby=4, nbz2=9, filt_cutoff=10.0)
    rng = np.random.default_rng(17 + periodic)
    lz = (mp.nz2 - 1) * mp.dz2
    core = n * 3 // 4
    x = np.concatenate([0.3 * rng.standard_normal(core), rng.uniform(-2.8, 2.8, n - core)])
    y = np.concatenate([0.3 * rng.standard_normal(core), rng.uniform(-2.8, 2.8, n - core)])
    p = dict(x=x, y=y, z2=rng.uniform(0.05, 0.5 * lz, n), px=np.zeros(n), py=np.zeros(n),
             gam=1.0 + 1e-3 * rng.standard_normal(n), chi=np.full(n, 2 * math.pi * lz / n))
    match_in(mp, 1.0, p)
    ar, ai = gaussian_seed(mp, 1e-2, 1e-2, w0=1.5)
    return mp, p, ar, ai, h


def _run(mp, p, ar, ai, h, sort_every, det=True, steps=24):
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.set_deterministic(det)
        g.set_sort_interval(sort_every)
        g.upload_particles(p)
        g.upload_field(ar, ai)
        r = und_section(g, steps, h, diffraction=True, steps4diff=12, periodic=bool(mp.periodic))
        assert r.status == 0
        out = g.download_particles()
        fr, fi = g.download_field()
    return out, fr, fi


@pytest.mark.parametrize("periodic", [0, 1])
def test_deterministic_state_is_bit_identical_across_runs_and_slot_orders(periodic):
    print(f"[gpu] deterministic periodic={periodic}", flush=True)
    mp, p, ar, ai, h = _dense_case(periodic)
    runs = [_run(mp, p, ar, ai, h, sort_every=s) for s in (1, 1, 0, 7)]
    ref = runs[0]
    for out, fr, fi in runs[1:]:
        for k in KEYS:
            assert np.array_equal(out[k], ref[0][k]), k
        assert np.array_equal(fr, ref[1]) and np.array_equal(fi, ref[2])
    # and the deterministic sums are the oracle's to the parity bar
    o = orc.Oracle(mp, p, (ar, ai))
    st, _, _, _ = o.und_section(24, h, True, 12, 60 if periodic else 0)
    assert st == 0
    for k in KEYS:
        e = np.abs(ref[0][k] - o.parts[k]).max() / np.abs(o.parts[k]).max()
        assert e < 1e-10, (k, e)
    scale = np.abs(o.ar + 1j * o.ai).max()
    assert np.abs(ref[1] - o.ar).max() / scale < 1e-10
    assert np.abs(ref[2] - o.ai).max() / scale < 1e-10


def test_deterministic_and_atomic_deposits_agree():
    mp, p, ar, ai, h = _dense_case(0)
    a = _run(mp, p, ar, ai, h, sort_every=30, det=True)
    b = _run(mp, p, ar, ai, h, sort_every=30, det=False)
    for k in KEYS:
        assert np.abs(a[0][k] - b[0][k]).max() <= 1e-12 * np.abs(b[0][k]).max(), k
    scale = np.abs(b[1] + 1j * b[2]).max()
    assert np.abs(a[1] - b[1]).max() <= 1e-12 * scale


def test_c0_whole_module_deterministic_matches_oracle():
    case = case_c0()
    print(f"[gpu] C0 deterministic, {case.n} macroparticles, {case.nsteps} steps", flush=True)
    res = []
    for _ in range(2):
        with gpu.Integrator(case.mp, case.n, FP64) as g:
            g.set_deterministic(True)
            g.upload_particles(case.particles)
            g.upload_field(*case.field)
            r = und_section(g, case.nsteps, case.h, diffraction=True, steps4diff=case.steps4diff,
                            periodic=True)
            assert r.status == 0
            res.append((g.download_particles(), g.download_field()))
    for k in KEYS:
        assert np.array_equal(res[0][0][k], res[1][0][k]), k
    assert np.array_equal(res[0][1][0], res[1][1][0]) and np.array_equal(res[0][1][1], res[1][1][1])
    o = orc.Oracle(case.mp, case.particles, case.field, nthreads=16)
    st, _, _, _ = o.und_section(case.nsteps, case.h, True, case.steps4diff, 60)
    assert st == 0
    for k in KEYS:
        e = np.abs(res[0][0][k] - o.parts[k]).max() / np.abs(o.parts[k]).max()
        print(f"[gpu] C0 deterministic {k} rel err {e:.2e}", flush=True)
        assert e < 1e-10, k


def test_mixed_form_refuses_deterministic_mode():
    mp, p, ar, ai, h = _dense_case(0, n=1000)
    with gpu.Integrator(mp, 1000, MIXED) as g:
        with pytest.raises(gpu.PuffinError):
            g.set_deterministic(True)
