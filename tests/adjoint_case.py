"""Shared pieces of the adjoint tests: a small helical module with a
per-period taper table, and the oracle's module gradient (the same schedule
and objective as puffin_amd.adjoint.ModuleGradient, with the oracle's FP64
steps and their transposes, every step's state kept in host memory)."""
from __future__ import annotations

import math

import numpy as np

import oracle as orc
from puffin_amd.adjoint import integrated_power, module_program, power_weights
from puffin_amd.configs import gaussian_seed, match_in
from puffin_amd.params import PI_REF, ModuleParams, gamma_r_for_eta, scaling

KEYS = orc.ADJ_KEYS


def small_module(periodic: int = 1, n: int = 2000, periods: int = 4, steps_per_period: int = 12,
                 nx: int = 9, nz2: int | None = None, seed_amp: float = 0.1, seed: int = 7,
                 charge: float = 4.0):
    """Helical (ux = uy = 1/sqrt 2) module with a 4-value n2col table."""
    rho = 0.01
    ux = uy = 1.0 / math.sqrt(2.0)
    gamma_r = gamma_r_for_eta(1.0e-4, 1.0, ux, uy)
    eta, kappa, fx, fy = scaling(rho, 1.0, gamma_r, ux, uy)
    lam = 4 * PI_REF * rho
    h = lam / steps_per_period
    dz2 = lam / 8.0
    if nz2 is None:
        nz2 = 33 if periodic else 72
    mp = ModuleParams(nx=nx, ny=nx, nz2=nz2, dx=0.6, dy=0.6, dz2=dz2, rho=rho, eta=eta, kappa=kappa,
                      gamma_r=gamma_r, aw=1.0, fx=fx, fy=fy, periodic=periodic, szfs=0.0,
                      szfe=periods * lam, nbx=3, nby=3, nbz2=7, filt_cutoff=0.5, beta=1.0)
    rng = np.random.default_rng(seed)
    lz = (nz2 - 1) * dz2
    span = lz if periodic else 0.3 * lz
    p = dict(x=np.clip(rng.standard_normal(n), -1.5, 1.5), y=np.clip(rng.standard_normal(n), -1.5, 1.5),
             z2=(np.arange(n) + 0.5) * span / n + (rng.random(n) - 0.5) * span / n * 0.3,
             px=np.zeros(n), py=np.zeros(n), gam=1.0 + 1e-3 * rng.standard_normal(n),
             chi=np.full(n, 2 * math.pi * span / n * charge))
    match_in(mp, 1.0, p)
    shape = gaussian_seed(mp, 1.0, 1.0, w0=1.5)
    table = 1.0 + 0.005 * rng.standard_normal(periods)
    ops = module_program(periods * steps_per_period, h, steps_per_period // 2, diffraction=True,
                         periodic=bool(periodic), redist_every=2 * steps_per_period)
    return dict(mp=mp, parts=p, shape=shape, table=table, seed=seed_amp, ops=ops, h=h,
                period=lam)


class OracleModule:
    """J and its gradient on the CPU oracle for one module."""

    def __init__(self, case, nthreads: int = 4):
        self.c = case
        self.nthreads = nthreads

    def _state(self, table, seed):
        c = self.c
        orc.set_taper_table(table, 0.0, c["period"])
        ar = seed * c["shape"][0]
        ai = seed * c["shape"][1]
        return orc.Oracle(c["mp"], c["parts"], (ar, ai), nthreads=self.nthreads)

    def J(self, table, seed) -> float:
        o = self._state(table, seed)
        try:
            for op in self.c["ops"]:
                self._run(o, op)
        finally:
            orc.set_taper_table(None)
        return integrated_power(self.c["mp"], o.power())

    @staticmethod
    def _run(o, op):
        if op.kind == "rk4":
            st = o.rk4par(op.sz, op.h)
            assert st == 0, st
        elif op.kind == "diff":
            o.diffraction_step(op.h)
        else:
            o.wrap_z2()

    def gradient(self, table, seed):
        c = self.c
        mp = c["mp"]
        o = self._state(table, seed)
        saved = []
        try:
            for op in c["ops"]:
                if op.kind == "rk4":
                    saved.append(({k: o.parts[k].copy() for k in KEYS + ("chi",)}, o.ar.copy(),
                                  o.ai.copy(), o.taper.copy()))
                self._run(o, op)
            J = integrated_power(mp, o.power())
            wz = power_weights(mp)
            ntr = mp.nx * mp.ny
            wx = np.full(mp.nx, mp.dx)
            wx[0] *= 0.5
            wx[-1] *= 0.5
            wy = np.full(mp.ny, mp.dy)
            wy[0] *= 0.5
            wy[-1] *= 0.5
            wxy = (wy[:, None] * wx[None, :]).reshape(-1)
            w3 = (wz[:, None] * wxy[None, :]).reshape(-1)
            lar = 2.0 * w3 * o.ar
            lai = 2.0 * w3 * o.ai
            lam = {k: np.zeros(o.n) for k in KEYS}
            g_table = np.zeros(len(table))
            j = len(saved)
            for op in reversed(c["ops"]):
                if op.kind == "rk4":
                    j -= 1
                    P, ar, ai, taper = saved[j]
                    st, dn = orc.rk4par_adjoint(mp, taper, P, ar, ai, lam, lar, lai, op.sz, op.h,
                                                self.nthreads)
                    assert st == 0, st
                    zs = (op.sz, op.sz + op.h / 2, op.sz + op.h / 2, op.sz + op.h)
                    for s in range(4):
                        g_table[orc.taper_slot(zs[s])] += dn[s]
                elif op.kind == "diff":
                    orc.diffraction_adjoint(mp, lar, lai, op.h, self.nthreads)
        finally:
            orc.set_taper_table(None)
        g_seed = float(np.dot(lar, c["shape"][0]) + np.dot(lai, c["shape"][1]))
        return J, g_table, g_seed, (lam, lar, lai)
