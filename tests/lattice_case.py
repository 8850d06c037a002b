"""A short lattice (two undulator modules with a drift, a quad, a chicane and
an energy modulation between them) and the same element chain run on the
oracle, element by element as puffin_module.f90:58-90 runs it."""
from __future__ import annotations

import dataclasses
import math

import numpy as np

import oracle as orc
from puffin_amd.lattice import Chicane, Drift, Modulation, Quad, Undulator
from puffin_amd.params import PI_REF

LATTICE = """! two modules, the second helical, with the elements between them
UN '' 2 1.0 -0.02 10 0.7071067811865476 0.7071067811865476 0.0 0.0
DR 1.0
QU 3.0 -3.0
CH 0.5 0.5 1.0e-3
MO 4.0 1.0e-4
UN 'helical' 2 1.0 0.0 10 1.0 1.0 0.0 0.0
"""

# with undulator ends the modules must be longer than the two 2-period ends
LATTICE_ENDS = """UN '' 6 1.0 0.0 8 0.7071067811865476 0.7071067811865476 0.0 0.0
DR 0.5
CH 0.0 0.25 0.0
UN '' 6 1.0 0.0 8 0.7071067811865476 0.7071067811865476 0.0 0.0
"""


def oracle_lattice(mp, parts, field, elements, *, diff_step, und_ends=False, redist_every=20):
    """The element chain on the oracle; returns (particles, ar, ai, sz, csteps)."""
    parts = {k: np.array(v, dtype=np.float64) for k, v in parts.items()}
    ar, ai = (np.array(a, dtype=np.float64) for a in field)
    sz, csteps = 0.0, 0
    for el in elements:
        if isinstance(el, Undulator):
            h, ns = el.h, el.nsteps
            if und_ends:
                szfs = 4.0 * PI_REF * mp.rho * 2.0
                szfe = ns * h - 4.0 * PI_REF * mp.rho * 2.0
            else:
                szfs, szfe = 0.0, ns * h
            mu = dataclasses.replace(mp, und_type=el.und_type, fx=el.ux, fy=el.uy, kbeta_x=el.kbeta_x,
                                     kbeta_y=el.kbeta_y, szfs=szfs, szfe=szfe, und_ends=int(und_ends))
            if not und_ends:
                orc.match(mu, el.alpha, parts, 0.0, +1)
            o = orc.Oracle(mu, parts, (ar, ai), taper=(el.alpha, el.alpha, el.taper))
            s4d = int(math.floor(diff_step / h + 0.5))
            st, _, csteps, szl = o.und_section(ns, h, True, s4d, redist_every, csteps=csteps)
            assert st == 0, st
            parts, ar, ai = o.parts, o.ar, o.ai
            sz = sz + szl
            if not und_ends:
                orc.match(mu, float(o.taper[0]), parts, sz, -1)
            orc.correct_trans(parts)
        elif isinstance(el, Quad):
            orc.quad(mp, parts, el.fx, el.fy)
        elif isinstance(el, (Drift, Chicane)):
            if isinstance(el, Drift):
                orc.drift(mp, parts, el.zbar)
            else:
                orc.chicane(mp, parts, el.disp, el.slip)
            if el.zbar > 0.0:
                o = orc.Oracle(mp, parts, (ar, ai))
                o.diffraction_step(el.zbar)
                ar, ai = o.ar, o.ai
            sz += el.zbar
        elif isinstance(el, Modulation):
            orc.modulation(parts, el.wavenum, el.mag)
    return parts, ar, ai, sz, csteps
