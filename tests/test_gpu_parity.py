"""Parity of the HIP integrator (through the C ABI) with the CPU oracle and
with the reference's own golden vectors.  Tolerances: the reference's own
bar for its 1-D run is 1e-10 absolute (test/testMPIIntegration.pf:102-127);
FP64 parity with the oracle is held at 1e-10 relative to the largest value.
"""
from __future__ import annotations

import functools
import math

import numpy as np
import pytest

import oracle as orc
from f1main_case import NSTEPS, initial_particles, load_fixture, module_params
from puffin_amd import gpu
from puffin_amd.configs import gaussian_seed, match_in
from puffin_amd.params import FP64, MIXED, ModuleParams, scaling, step_size
from puffin_amd.undulator import und_section

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z2", "px", "py", "gam")


@pytest.fixture(scope="module", autouse=True)
def _device():
    n = gpu.device_count()
    assert n >= 1, "no HIP device visible: the gpu tests need an MI355X"
    print(f"[gpu] devices={n}", flush=True)


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


# ---------------------------------------------------------------------------
# the reference's 1-D run (FP64): golden vectors at the reference's own bar
# ---------------------------------------------------------------------------

def test_f1main_golden_vectors_on_gpu():
    d, info = load_fixture()
    mp = module_params()
    parts = initial_particles(d["electrons"][:, 6])
    orc.match(mp, 1.0, parts, 0.0, +1)
    h = step_size(mp.rho, 30)
    with gpu.Integrator(mp, parts["x"].size, FP64) as g:
        g.upload_particles(parts)
        g.upload_field(np.zeros(mp.nmesh), np.zeros(mp.nmesh))
        g.set_taper(1.0, 1.0, 0.0)
        szl = 0.0
        for _ in range(NSTEPS):
            assert g.rk4par(szl, h) == 0
            szl += h
        out = g.download_particles()
        ar, ai = g.download_field()
        n2col, _ = g.get_taper()
    out["chi"] = parts["chi"]
    orc.match(mp, n2col, out, szl, -1)
    orc.correct_trans(out)
    e = d["electrons"]
    order = np.argsort(out["z2"])
    for c, k in enumerate(KEYS):
        assert np.abs(out[k][order] - e[:, c]).max() <= 1e-10, k
    assert np.abs(ar - d["aperp"][0]).max() <= 1e-10
    assert np.abs(ai - d["aperp"][1]).max() <= 1e-10


# ---------------------------------------------------------------------------
# FFT (Transform_MultiD) against numpy
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(9, 7, 8), (33, 33, 128), (129, 5, 16), (5, 6, 257), (11, 4, 43),
                                   (17, 17, 17), (257, 3, 4), (3, 257, 4), (85, 7, 6),
                                   # the line lengths with compile-time plans (x, y and z2);
                                   # 4096-point z2 lines run as 64 x 64 (four-step), in
                                   # x-blocks of 8 (FP64) or 16 (FP32) columns: 17 leaves a
                                   # partial block in both forms
                                   (129, 129, 4), (33, 33, 256), (3, 5, 1024), (2, 3, 4096),
                                   (17, 3, 4096), (257, 257, 4)])
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_transform_matches_numpy(shape, precision):
    nx, ny, nz = shape
    print(f"[gpu] transform shape={shape} precision={precision}", flush=True)
    mp = ModuleParams(nx=nx, ny=ny, nz2=nz, dx=0.1, dy=0.1, dz2=0.01, rho=0.01, eta=1e-4,
                      kappa=0.5, gamma_r=100.0, aw=1.0)
    rng = np.random.default_rng(7)
    a = rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))
    with gpu.Integrator(mp, 1, precision) as g:
        g.upload_field(a.real.reshape(-1).copy(), a.imag.reshape(-1).copy())
        g.transform(-1)
        ar, ai = g.download_field()
        f = (ar + 1j * ai).reshape(nz, ny, nx)
        g.transform(+1)
        br, bi = g.download_field()
        b = (br + 1j * bi).reshape(nz, ny, nx)
    tol = 1e-13 if precision == FP64 else 2e-6
    assert relerr(f, np.fft.fftn(a)) < tol
    assert relerr(b, a * (nx * ny * nz)) < tol


# ---------------------------------------------------------------------------
# DiffractionStep against the oracle
# ---------------------------------------------------------------------------

def _diff_case(periodic: int):
    nz2 = 65 if periodic else 64
    mp = ModuleParams(nx=33, ny=33, nz2=nz2, dx=0.3, dy=0.3, dz2=0.01, rho=0.01, eta=1e-4,
                      kappa=0.5, gamma_r=100.0, aw=1.0, periodic=periodic, nbz2=9,
                      filt_cutoff=15.0, beta=1.0)
    rng = np.random.default_rng(11 + periodic)
    ar = rng.standard_normal(mp.nmesh)
    ai = rng.standard_normal(mp.nmesh)
    if periodic:
        ntr = mp.ntr
        ar[(nz2 - 1) * ntr:] = ar[:ntr]
        ai[(nz2 - 1) * ntr:] = ai[:ntr]
    return mp, ar, ai


@pytest.mark.parametrize("periodic", [0, 1])
def test_diffraction_step_matches_oracle(periodic):
    mp, ar, ai = _diff_case(periodic)
    dummy = dict(x=np.zeros(1), y=np.zeros(1), z2=np.zeros(1), px=np.zeros(1), py=np.zeros(1),
                 gam=np.ones(1), chi=np.zeros(1))
    o = orc.Oracle(mp, dummy, (ar, ai))
    o.diffraction_step(0.0837)
    with gpu.Integrator(mp, 1, FP64) as g:
        g.upload_field(ar, ai)
        g.diffraction_step(0.0837)
        gr, gi = g.download_field()
    assert relerr(gr, o.ar) < 1e-12
    assert relerr(gi, o.ai) < 1e-12


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_diffraction_step_four_step_lines_match_oracle(periodic, precision):
    """4096-point z2 lines (C4's temporal mesh) take the four-step z2 passes:
    the propagator and the damping at q = k1 + 64 k2 between them."""
    print(f"[gpu] diffraction 4096-point z2 lines periodic={periodic} precision={precision}", flush=True)
    nz2 = 4097 if periodic else 4096
    mp = ModuleParams(nx=17, ny=9, nz2=nz2, dx=0.3, dy=0.3, dz2=0.002, rho=0.01, eta=1e-4,
                      kappa=0.5, gamma_r=100.0, aw=1.0, periodic=periodic, nbz2=65,
                      filt_cutoff=15.0, beta=1.0)
    rng = np.random.default_rng(23 + periodic)
    ar = rng.standard_normal(mp.nmesh)
    ai = rng.standard_normal(mp.nmesh)
    if periodic:
        ntr = mp.ntr
        ar[(nz2 - 1) * ntr:] = ar[:ntr]
        ai[(nz2 - 1) * ntr:] = ai[:ntr]
    dummy = dict(x=np.zeros(1), y=np.zeros(1), z2=np.zeros(1), px=np.zeros(1), py=np.zeros(1),
                 gam=np.ones(1), chi=np.zeros(1))
    o = orc.Oracle(mp, dummy, (ar, ai))
    o.diffraction_step(0.0837)
    with gpu.Integrator(mp, 1, precision) as g:
        g.upload_field(ar, ai)
        g.diffraction_step(0.0837)
        gr, gi = g.download_field()
    tol = 1e-12 if precision == FP64 else 2e-5
    assert relerr(gr, o.ar) < tol
    assert relerr(gi, o.ai) < tol


@functools.lru_cache(maxsize=None)
def _prime_case(periodic: int):
    """A 257 x 257 mesh with 8 z2 planes (and the aliased one) and the
    oracle's diffraction step of it."""
    nz2 = 9 if periodic else 8
    mp = ModuleParams(nx=257, ny=257, nz2=nz2, dx=0.1, dy=0.1, dz2=0.01, rho=0.01, eta=1e-4,
                      kappa=0.5, gamma_r=100.0, aw=1.0, periodic=periodic, nbx=40, nby=40, nbz2=3,
                      filt_cutoff=15.0, beta=1.0)
    rng = np.random.default_rng(31 + periodic)
    ar = rng.standard_normal(mp.nmesh)
    ai = rng.standard_normal(mp.nmesh)
    if periodic:
        ntr = mp.ntr
        ar[(nz2 - 1) * ntr:] = ar[:ntr]
        ai[(nz2 - 1) * ntr:] = ai[:ntr]
    dummy = dict(x=np.zeros(1), y=np.zeros(1), z2=np.zeros(1), px=np.zeros(1), py=np.zeros(1),
                 gam=np.ones(1), chi=np.zeros(1))
    o = orc.Oracle(mp, dummy, (ar, ai), nthreads=8)
    o.diffraction_step(0.0837)
    return mp, ar, ai, o.ar, o.ai


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_diffraction_step_prime_lines_match_oracle(periodic, precision):
    """257-point transverse lines (C4's axes) take Rader's algorithm in the x
    and y passes; in the mixed form in place, with the permutations folded
    into the passes' global reads and writes, in the mask split between two
    transforms and in the recombination."""
    print(f"[gpu] diffraction 257 x 257 lines periodic={periodic} precision={precision}", flush=True)
    mp, ar, ai, oar, oai = _prime_case(periodic)
    with gpu.Integrator(mp, 1, precision) as g:
        g.upload_field(ar, ai)
        g.diffraction_step(0.0837)
        gr, gi = g.download_field()
    tol = 1e-12 if precision == FP64 else 2e-5
    assert relerr(gr, oar) < tol
    assert relerr(gi, oai) < tol


# ---------------------------------------------------------------------------
# 3-D RK4 steps against the oracle
# ---------------------------------------------------------------------------

def _small_3d(periodic=0, n=3000, seed_amp=1e-2, und_ends=0, planar=False, nx=17, ny=17, nz2=None,
              matched=True):
    rho = 0.01
    ux, uy = (1.0, 0.0) if planar else (1 / math.sqrt(2), 1 / math.sqrt(2))
    eta, kappa, fx, fy = scaling(rho, 1.0, 100.0, ux, uy)
    h = step_size(rho, 30)
    if nz2 is None:
        nz2 = 65 if periodic else 80
    mp = ModuleParams(nx=nx, ny=ny, nz2=nz2, dx=0.4, dy=0.4, dz2=4 * 3.14159274 * rho / 16,
                      rho=rho, eta=eta, kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy,
                      periodic=periodic, und_ends=und_ends, szfs=0.0 if not und_ends else 8 * 3.14159274 * rho,
                      szfe=60 * h if not und_ends else 60 * h - 8 * 3.14159274 * rho,
                      nbx=4, nby=4, nbz2=9, filt_cutoff=10.0)
    rng = np.random.default_rng(5 + periodic)
    lz = (nz2 - 1) * mp.dz2
    p = dict(x=np.clip(rng.standard_normal(n), -2.5, 2.5), y=np.clip(rng.standard_normal(n), -2.5, 2.5),
             z2=rng.uniform(0.05, 0.6 * lz, n), px=np.zeros(n), py=np.zeros(n),
             gam=1.0 + 1e-3 * rng.standard_normal(n), chi=np.full(n, 2 * math.pi * lz / n))
    if matched:
        match_in(mp, 1.0, p)
    ar, ai = gaussian_seed(mp, seed_amp, seed_amp, w0=1.5)
    return mp, p, (ar, ai), h


@pytest.mark.parametrize("periodic,nz2", [(0, 79), (1, 70)])
def test_rk4_steps_odd_mesh_match_oracle(periodic, nz2):
    """A mesh whose cell counts (13 x 13 x 78 or 69) leave partial sort bricks
    on every axis, with re-sorts between the steps."""
    print(f"[gpu] rk4 odd mesh periodic={periodic} nz2={nz2}", flush=True)
    mp, p, (ar, ai), h = _small_3d(periodic, nx=14, ny=14, nz2=nz2)
    o = orc.Oracle(mp, p, (ar, ai), taper=(1.0, 1.0, -0.01))
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        g.set_taper(1.0, 1.0, -0.01)
        g.set_sort_interval(3)
        szl = 0.0
        for _ in range(12):
            assert o.rk4par(szl, h) == 0
            assert g.rk4par(szl, h) == 0
            szl += h
        out = g.download_particles()
        gr, gi = g.download_field()
    for k in KEYS:
        assert relerr(out[k], o.parts[k]) < 1e-10, k
    assert relerr(gr, o.ar) < 1e-10
    assert relerr(gi, o.ai) < 1e-10


@pytest.mark.parametrize("periodic,und_ends,planar", [(0, 0, False), (1, 0, False), (0, 1, True)])
def test_rk4_steps_match_oracle(periodic, und_ends, planar):
    print(f"[gpu] rk4 periodic={periodic} und_ends={und_ends} planar={planar}", flush=True)
    mp, p, (ar, ai), h = _small_3d(periodic, und_ends=und_ends, planar=planar)
    o = orc.Oracle(mp, p, (ar, ai), taper=(1.0, 1.0, -0.01))
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        g.set_taper(1.0, 1.0, -0.01)
        szl = 0.0
        for _ in range(12):
            assert o.rk4par(szl, h) == 0
            assert g.rk4par(szl, h) == 0
            szl += h
        out = g.download_particles()
        gr, gi = g.download_field()
        assert g.get_taper()[0] == pytest.approx(o.taper[0], rel=1e-14)
    for k in KEYS:
        assert relerr(out[k], o.parts[k]) < 1e-10, k
    assert relerr(gr, o.ar) < 1e-10
    assert relerr(gi, o.ai) < 1e-10


def test_out_of_mesh_flag_keeps_state():
    mp, p, (ar, ai), h = _small_3d(0)
    p["x"][17] = 50.0
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        assert g.rk4par(0.0, h) == 1
        out = g.download_particles()
        gr, gi = g.download_field()
    for k in KEYS:
        assert np.array_equal(out[k], np.asarray(p[k], dtype=np.float64)), k
    assert np.array_equal(gr, ar) and np.array_equal(gi, ai)


def test_temporal_overflow_is_fatal_status():
    mp, p, (ar, ai), h = _small_3d(0)
    p["z2"][3] = (mp.nz2 - 1) * mp.dz2 + 1e-9
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        assert g.rk4par(0.0, h) == 3


# ---------------------------------------------------------------------------
# the UndSection schedule (rk4 + diffraction + wraps) against the oracle's
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("periodic", [0, 1])
def test_und_section_matches_oracle(periodic):
    print(f"[gpu] und_section periodic={periodic}", flush=True)
    mp, p, (ar, ai), h = _small_3d(periodic, n=2000)
    nsteps, s4d = 40, 15
    wm = np.zeros(nsteps, dtype=np.uint8)
    wm[[14, 29, 39]] = 1
    o = orc.Oracle(mp, p, (ar, ai), taper=(1.0, 1.0, 0.0))
    st, opow, cs, szl = o.und_section(nsteps, h, True, s4d, 20, write_mask=wm)
    assert st == 0
    powers = []
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        res = und_section(g, nsteps, h, diffraction=True, steps4diff=s4d, redist_every=20,
                          write=lambda i: powers.append(g.power()), write_every=15,
                          int_write_every=10 ** 6, periodic=bool(periodic))
        assert res.status == 0
        out = g.download_particles()
        gr, gi = g.download_field()
    for k in KEYS:
        assert relerr(out[k], o.parts[k]) < 1e-10, k
    assert relerr(gr, o.ar) < 1e-10 and relerr(gi, o.ai) < 1e-10
    assert len(powers) == 3
    for a, b in zip(powers, opow):
        assert relerr(a, b) < 1e-10


# ---------------------------------------------------------------------------
# diagnostics
# ---------------------------------------------------------------------------

def test_power_and_bunching_match_oracle():
    mp, p, (ar, ai), h = _small_3d(0)
    o = orc.Oracle(mp, p, (ar, ai))
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        pw = g.power()
        b = g.bunching(1, 8, 4 * 3.14159274 * mp.rho)
    assert relerr(pw, o.power()) < 1e-12
    assert relerr(b, o.bunching(1, 8, 4 * 3.14159274 * mp.rho)) < 1e-12


def test_current_outside_the_nodes_is_an_error():
    """getCurr STOPs when a macroparticle lies past the last node; the device
    returns an error instead, and the particles are left as they were."""
    mp, p, _, _ = _small_3d(0, n=2000)
    L = 4 * 3.1415927410125732 * mp.rho
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        with pytest.raises(gpu.PuffinError):
            g.current(2, L)                         # nodes cover [0, L] only
        with pytest.raises(gpu.PuffinError):
            g.slice_stats(0, L)
        out = g.download_particles()
        ok = g.current(int(np.ceil((mp.nz2 - 1) * mp.dz2 / L)) + 1, L)
    for k in KEYS:
        assert np.array_equal(out[k], p[k]), k
    assert abs(ok.sum() - p["chi"].sum()) <= 1e-12 * p["chi"].sum()


def test_tile_fallback_share_is_a_fraction():
    """The share of stage workgroups that deposited with global atomics is
    recorded per synchronised step (and right after a re-sort)."""
    mp, p, (ar, ai), h = _small_3d(0, n=3000)
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        assert g.rk4par(0.0, h) == 0
        assert g.rk4par(h, h) == 0
        last, base = g.tile_fallback()
    assert 0.0 <= last <= 1.0 and 0.0 <= base <= 1.0


@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("periodic", [0, 1])
def test_slice_stats_and_current_match_oracle(periodic, prec):
    """getSliceTwiss's slice statistics and getCurr's binned charge on the
    device against the oracle, after steps that spread p and Gamma."""
    print(f"[gpu] slice stats periodic={periodic} prec={prec}", flush=True)
    mp, p, (ar, ai), h = _small_3d(periodic, n=4000)
    L = 4 * 3.1415927410125732 * mp.rho
    lz = (mp.nz2 - 1) * mp.dz2
    nslices, npts = int(np.ceil(lz / L)), int(np.ceil(lz / L)) + 1
    with gpu.Integrator(mp, p["x"].size, prec) as g:
        g.upload_particles(p)
        g.upload_field(ar, ai)
        szl = 0.0
        for _ in range(6):
            assert g.rk4par(szl, h) == 0
            szl += h
        out = g.download_particles()
        s = g.slice_stats(nslices, L)
        cur = g.current(npts, L)
    # the charge weights as the device holds them (FP32 in the mixed form)
    out["chi"] = p["chi"] if prec == FP64 else p["chi"].astype(np.float32).astype(np.float64)
    st, ref = orc.slice_stats(mp, out, nslices, L)
    assert st == 0
    # Both sides sum the same stored values, in different orders.  Means and
    # charge: 1e-12.  Spreads and emittances are differences of near-equal
    # second moments (Gamma spreads 1e-3 about 1: six digits cancel), so
    # summation order moves them by ~1e-9 relative: 1e-8.
    spread = ("sdX", "sdY", "sdpX", "sdpY", "eX", "eY", "aveDgamma")
    for k, v in ref.items():
        tol = 1e-8 if k in spread else 1e-12
        assert np.abs(s[k] - v).max() <= tol * max(np.abs(v).max(), 1e-30), k
    st, rc = orc.current(mp, out, npts, L)
    assert st == 0
    assert relerr(cur, rc) < 1e-12


# ---------------------------------------------------------------------------
# mixed precision on the small case
# ---------------------------------------------------------------------------

def test_mixed_precision_tracks_fp64():
    mp, p, (ar, ai), h = _small_3d(1, n=4000)
    res = {}
    for prec in (FP64, MIXED):
        with gpu.Integrator(mp, p["x"].size, prec) as g:
            g.upload_particles(p)
            g.upload_field(ar, ai)
            r = und_section(g, 60, h, diffraction=True, steps4diff=30, periodic=True)
            assert r.status == 0
            res[prec] = (g.download_particles(), g.power())
    assert relerr(res[MIXED][1], res[FP64][1]) < 1e-4
    for k in ("z2", "gam"):
        assert relerr(res[MIXED][0][k], res[FP64][0][k]) < 1e-6, k
