"""Synthetic workloads of the integrator (seeded, reproducible).

Each case fixes the module constants (as the reference's setup would leave
them, see params.py), the macroparticles and the initial field.  Random
numbers come from numpy's counter-based Philox generator with seed
20261202 + case index, so a case is a pure function of its name and size.

Initial conditions follow the reference's conventions:
  * p_perp and the transverse offsets at zbar = 0 from matchIn
    (acc_lattice.f90:743-828 with init_conds.f90:29-223);
  * PY holds -p_y (simple_electron_gen.f90:330);
  * z2 quiet start with shot-noise jitter (u - 1/2) dz2_k / sqrt(N_k)
    (noise.f90:84-85), N_k electrons per macroparticle;
  * seed field magx env sin(z2 / 2 rho) in Re, magy env cos(z2 / 2 rho) in Im
    (getSeed, setup_calcs.f90:1438-1580), env = exp(-r^2 / w0^2);
  * charge weight chi = (2 pi sx sy Lz2) / N per macroparticle for a beam
    sampled from a Gaussian transverse and uniform (or Gaussian) z2 profile,
    i.e. peak density 1 in units of the reference density npk.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .params import (PI_REF, ModuleParams, filter_cutoff, gamma_r_for_eta,
                     scaling, step_size, steps_for_diffraction)

SEED_BASE = 20261202


@dataclass
class Case:
    name: str
    mp: ModuleParams
    particles: dict
    field: tuple
    h: float
    nsteps: int
    steps4diff: int
    n2col0: float = 1.0
    undgrad: float = 0.0
    taper_table: np.ndarray | None = None
    taper_period: float = 0.0
    info: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.particles["x"].size)


def match_in(mp: ModuleParams, n2col: float, p: dict, sz: float = 0.0) -> None:
    """matchIn for the elliptical undulator (acc_lattice.f90:743-828)."""
    spx0 = -mp.fy * n2col * math.cos(sz / (2.0 * mp.rho))           # pxOffset
    spy0 = -1.0 * (-mp.fx * n2col * math.sin(sz / (2.0 * mp.rho)))  # -pyOffset
    ncx = mp.aw ** 2
    ncy = 2.0 * mp.aw ** 2 / (mp.fx ** 2 + mp.fy ** 2)
    gj = mp.gamma_r * p["gam"]
    s2 = spx0 ** 2 + spy0 ** 2
    xo = (-mp.fy * 4.0 * mp.kappa * mp.rho ** 2 / math.sqrt(mp.eta) *
          (mp.gamma_r / np.sqrt(gj ** 2 - (1.0 + ncx * s2))) * n2col * math.sin(sz / (2.0 * mp.rho)))
    yo = (mp.fx * 4.0 * mp.kappa * mp.rho ** 2 / math.sqrt(mp.eta) *
          (mp.gamma_r / np.sqrt(gj ** 2 - (1.0 + ncy * s2))) * n2col * math.cos(sz / (2.0 * mp.rho)))
    p["x"] = p["x"] + xo
    p["y"] = p["y"] + yo
    p["px"] = p["px"] + spx0
    p["py"] = p["py"] + spy0


def _truncated_normal(rng, n, sigma, cut):
    out = rng.standard_normal(n) * sigma
    bad = np.abs(out) > cut
    while bad.any():
        out[bad] = rng.standard_normal(int(bad.sum())) * sigma
        bad = np.abs(out) > cut
    return out


def _mesh(nx, nz2, rho, nodes_per_lambda, lx, periodic):
    dz2 = 4.0 * PI_REF * rho / float(nodes_per_lambda - 1)   # calcSamples (setup_calcs.f90:955)
    dx = lx / float(nx - 1)
    return dx, dz2


def _helical_like(name, idx, nx, nz2, n, periods, periodic, seed_amp=0.0, aw_vec=None,
                  nodes_per_lambda=17, lx=12.0, sig_x=1.0, gam_spread=1e-3, rho=0.01,
                  n_per_macro=1.0e4, xcut=4.5):
    """Elliptical (ux = uy = 1/sqrt 2) case, quiet start uniform in z2."""
    rng = np.random.Generator(np.random.Philox(SEED_BASE + idx))
    ux = uy = 1.0 / math.sqrt(2.0)
    aw = 1.0
    gamma_r = gamma_r_for_eta(1.0e-4, aw, ux, uy)
    eta, kappa, fx, fy = scaling(rho, aw, gamma_r, ux, uy)
    dx, dz2 = _mesh(nx, nz2, rho, nodes_per_lambda, lx, periodic)
    h = step_size(rho, 30)
    nsteps = 30 * periods
    lz = dz2 * (nz2 - 1 if periodic else nz2 - 1)
    mp = ModuleParams(nx=nx, ny=nx, nz2=nz2, dx=dx, dy=dx, dz2=dz2, rho=rho, eta=eta, kappa=kappa,
                      gamma_r=gamma_r, aw=aw, fx=fx, fy=fy, periodic=int(periodic),
                      szfs=0.0, szfe=nsteps * h,
                      filt_cutoff=filter_cutoff(rho, nz2, dz2, periodic, 0.3,
                                                per_waves=lz / (4 * PI_REF * rho)))
    # z2 quiet start over the periodic length (or the first 3/4 of a temporal mesh)
    span = lz if periodic else 0.75 * lz
    dzk = span / n
    z2 = (np.arange(n) + 0.5) * dzk
    z2 = z2 + (rng.random(n) - 0.5) * dzk / math.sqrt(n_per_macro)
    p = dict(x=_truncated_normal(rng, n, sig_x, xcut), y=_truncated_normal(rng, n, sig_x, xcut),
             z2=z2, px=np.zeros(n), py=np.zeros(n),
             gam=1.0 + gam_spread * rng.standard_normal(n),
             chi=np.full(n, 2 * math.pi * sig_x * sig_x * span / n))
    n2col0 = 1.0 if aw_vec is None else float(aw_vec[0])
    match_in(mp, n2col0, p)
    ar = np.zeros(mp.nmesh)
    ai = np.zeros(mp.nmesh)
    if seed_amp:
        ar, ai = gaussian_seed(mp, seed_amp / math.sqrt(2.0), seed_amp / math.sqrt(2.0), w0=2.0)
    c = Case(name, mp, p, (ar, ai), h, nsteps, steps_for_diffraction(rho, h), n2col0=n2col0,
             info=dict(gamma_r=gamma_r, sig_x=sig_x, n_per_macro=n_per_macro, seed_amp=seed_amp))
    if aw_vec is not None:
        c.taper_table = np.asarray(aw_vec, dtype=np.float64)
        c.taper_period = 4.0 * PI_REF * rho
    return c


def gaussian_seed(mp: ModuleParams, magx: float, magy: float, w0: float, zc=None, zsig=None):
    """getSeed (setup_calcs.f90:1438-1580) with sigma = w0 / sqrt 2."""
    lx = mp.dx * (mp.nx - 1)
    ly = mp.dy * (mp.ny - 1)
    xs = np.linspace(-lx / 2, lx / 2, mp.nx)
    ys = np.linspace(-ly / 2, ly / 2, mp.ny)
    z2 = np.arange(mp.nz2) * mp.dz2
    sig = w0 / math.sqrt(2.0)
    env_xy = np.exp(-((ys[:, None] / sig) ** 2) / 2.0) * np.exp(-((xs[None, :] / sig) ** 2) / 2.0)
    zenv = np.ones(mp.nz2) if zsig is None else np.exp(-(((z2 - zc) / zsig) ** 2) / 2.0)
    oscx = zenv * np.sin(z2 / (2.0 * mp.rho))
    oscy = zenv * np.cos(z2 / (2.0 * mp.rho))
    ar = (magx * env_xy[None, :, :] * oscx[:, None, None]).reshape(-1)
    ai = (magy * env_xy[None, :, :] * oscy[:, None, None]).reshape(-1)
    return np.ascontiguousarray(ar), np.ascontiguousarray(ai)


def case_c0(n: int = 1 << 18, periods: int = 8) -> Case:
    """Helical-like elliptical FEL, periodic 33 x 33 x (128 + 1), 2^18
    macroparticles, rho = 0.01, eta = 1e-4, 30 steps per period."""
    return _helical_like("c0_helical_periodic_33x33x128", 0, 33, 129, n, periods, True)


def case_c3(n: int = 1 << 20, periods: int = 16, seed: float = 1.0e-3, aw_vec=None) -> Case:
    """Helical-like adjoint case, periodic 33 x 33 x (256 + 1), per-period a_w."""
    rng = np.random.Generator(np.random.Philox(SEED_BASE + 3))
    if aw_vec is None:
        aw_vec = 1.0 + 0.005 * rng.standard_normal(periods)
    return _helical_like("c3_helical_adjoint_33x33x256", 3, 33, 257, n, periods, True,
                         seed_amp=seed, aw_vec=aw_vec)


def case_c1(n: int = 1 << 24, nx: int = 129, nz2: int = 1024, periods: int = 64,
            run_periods: int | None = None, with_field: bool = True) -> Case:
    """Planar (ux = 1, uy = 0) elliptical FEL with natural focusing and a
    linear taper of -1% over 64 periods, rho = 0.005, temporal mesh
    129 x 129 x 1024, Gaussian seed (w0 = 2, amplitude 1e-3), Gaussian
    current sigma = 200 nodes, Gaussian transverse sigma = 1.5 (cut at
    4 sigma), Gamma spread 5e-4.

    Sized so that the 64-period module (1920 steps) stays on the temporal
    mesh, as the reference requires (it STOPs when a macroparticle passes the
    last node, system_interpolation.f90:70-75):
      * 8 nodes per resonant wavelength, so 64 periods of resonant slippage
        are 512 nodes;
      * eta = 1e-6 (gamma_r = 866).  The slippage rate p2 grows with
        aw^2 |p_perp|^2, and the natural-focusing betatron momentum of a
        particle at amplitude a is about sqrt(eta) a / (2 sqrt(2) rho) in
        scaled units.  At eta = 1e-4 a 4-sigma particle slips 5-6 times
        faster than a resonant one; at 1e-6 the fastest slips 64.4 lambda_r
        (515 nodes) over the module;
      * the Gaussian current is cut at +-1.2 sigma (+-240 nodes) and centred
        on node 244, so the beam spans nodes 4..484 and its front ends the
        module at node ~1000 of 1024."""
    rng = np.random.Generator(np.random.Philox(SEED_BASE + 1))
    rho = 0.005
    ux, uy = 1.0, 0.0
    aw = 1.0
    eta_target = 1.0e-6
    gamma_r = gamma_r_for_eta(eta_target, aw, ux, uy)
    eta, kappa, fx, fy = scaling(rho, aw, gamma_r, ux, uy)
    nodes_per_lambda = 8.0
    dz2 = 4.0 * PI_REF * rho / nodes_per_lambda
    lx = 19.2 * (nx - 1) / 128.0
    dx = lx / float(nx - 1)
    h = step_size(rho, 30)
    nsteps = 30 * periods
    undgrad = -0.01 / (periods * 4.0 * PI_REF * rho)
    mp = ModuleParams(nx=nx, ny=nx, nz2=nz2, dx=dx, dy=dx, dz2=dz2, rho=rho, eta=eta, kappa=kappa,
                      gamma_r=gamma_r, aw=aw, fx=fx, fy=fy, periodic=0, szfs=0.0, szfe=nsteps * h,
                      filt_cutoff=filter_cutoff(rho, nz2, dz2, False, 0.3))
    scale = nz2 / 1024.0
    zc = 244 * dz2 * scale
    zs = 200 * dz2 * scale
    zcut = 1.2
    z2 = zc + _truncated_normal(rng, n, zs, zcut * zs)
    sx = 1.5
    xcut = 4.0
    p = dict(x=_truncated_normal(rng, n, sx, xcut * sx), y=_truncated_normal(rng, n, sx, xcut * sx),
             z2=z2, px=np.zeros(n), py=np.zeros(n),
             gam=1.0 + 5.0e-4 * rng.standard_normal(n))
    # Gaussian current: chi = (2 pi sx^2 sqrt(2 pi) sz) / N (peak density 1)
    p["chi"] = np.full(n, 2 * math.pi * sx * sx * math.sqrt(2 * math.pi) * zs / n)
    match_in(mp, 1.0, p)
    ar, ai = gaussian_seed(mp, 0.0, 1.0e-3, w0=2.0) if with_field else (None, None)
    run = nsteps if run_periods is None else 30 * run_periods
    return Case("c1_planar_taper_129x129x1024", mp, p, (ar, ai), h, run, steps_for_diffraction(rho, h),
                n2col0=1.0, undgrad=undgrad,
                info=dict(gamma_r=gamma_r, periods=periods, eta=eta, nodes_per_lambda=nodes_per_lambda,
                          sig_x=sx, x_cut_sigma=xcut, z2_centre=zc, z2_sigma=zs, z2_cut_sigma=zcut,
                          gam_spread=5.0e-4))


def case_c4(n: int = 1 << 27, nx: int = 257, nz2: int = 4096, periods: int = 128,
            run_periods: int | None = None, with_field: bool = True) -> Case:
    """Large planar (ux = 1, uy = 0) FEL for the FP32-mixed form: temporal
    mesh 257 x 257 x 4096, 2^27 macroparticles, Gaussian transverse sigma = 2
    (cut at 4 sigma), flat-top current over 3000 z2 nodes with Gaussian edges
    (sigma 12 nodes, cut at 3 sigma), Gamma spread 1e-3, rho = 0.005, a
    per-period taper table of `periods` values
    n2col_k = 1 - 0.01 (k / (periods - 1))^2, Gaussian seed (w0 = 2,
    amplitude 1e-3).

    Sized so that the 128-period module (3840 steps) stays on the temporal
    mesh: 7 nodes per resonant wavelength and eta = 1e-6 (see case_c1), so
    that the fastest 4-sigma particle slips 131.5 lambda_r = 921 nodes; the
    flat top spans nodes 38..3038 and the front edge ends the module at node
    ~3995 of 4096."""
    rng = np.random.Generator(np.random.Philox(SEED_BASE + 4))
    rho = 0.005
    ux, uy = 1.0, 0.0
    aw = 1.0
    gamma_r = gamma_r_for_eta(1.0e-6, aw, ux, uy)
    eta, kappa, fx, fy = scaling(rho, aw, gamma_r, ux, uy)
    nodes_per_lambda = 7.0
    dz2 = 4.0 * PI_REF * rho / nodes_per_lambda
    dx = 0.1
    h = step_size(rho, 30)
    nsteps = 30 * periods
    mp = ModuleParams(nx=nx, ny=nx, nz2=nz2, dx=dx, dy=dx, dz2=dz2, rho=rho, eta=eta, kappa=kappa,
                      gamma_r=gamma_r, aw=aw, fx=fx, fy=fy, periodic=0, szfs=0.0, szfe=nsteps * h,
                      filt_cutoff=filter_cutoff(rho, nz2, dz2, False, 0.3))
    scale = nz2 / 4096.0
    flat = 3000 * dz2 * scale
    edge = 12 * dz2
    ecut = 3.0
    z0 = 38 * dz2 * scale
    z2 = z0 + flat * rng.random(n) + _truncated_normal(rng, n, edge, ecut * edge)
    sx = 2.0
    xcut = 4.0
    p = dict(x=_truncated_normal(rng, n, sx, xcut * sx), y=_truncated_normal(rng, n, sx, xcut * sx),
             z2=z2, px=np.zeros(n), py=np.zeros(n),
             gam=1.0 + 1.0e-3 * rng.standard_normal(n))
    p["chi"] = np.full(n, 2 * math.pi * sx * sx * flat / n)
    k = np.arange(periods, dtype=np.float64)
    table = 1.0 - 0.01 * (k / max(periods - 1, 1)) ** 2
    match_in(mp, float(table[0]), p)
    ar, ai = gaussian_seed(mp, 0.0, 1.0e-3, w0=2.0) if with_field else (None, None)
    run = nsteps if run_periods is None else 30 * run_periods
    c = Case("c4_planar_large_257x257x4096", mp, p, (ar, ai), h, run, steps_for_diffraction(rho, h),
             n2col0=float(table[0]),
             info=dict(gamma_r=gamma_r, periods=periods, seed_amp=1.0e-3, eta=eta,
                       nodes_per_lambda=nodes_per_lambda, sig_x=sx, x_cut_sigma=xcut,
                       z2_flat=(z0, z0 + flat), z2_edge_sigma=edge, z2_edge_cut_sigma=ecut,
                       gam_spread=1.0e-3))
    c.taper_table = table
    c.taper_period = 4.0 * PI_REF * rho
    return c


CASES = {"c0": case_c0, "c1": case_c1, "c3": case_c3, "c4": case_c4}
