"""The reference's element chain over the C ABI.

* ``parse_lattice`` reads a lattice file as readLatt does
  (acc_lattice.f90:243-406): one element per line, keyed by the first two
  letters of its first token (QU, UN, CH, DR, MO); other lines are skipped;
  chicane lengths and slips are scaled by 4 pi rho, drift lengths likewise,
  and an undulator's step size is 4 pi rho / (steps per period).
* ``run_lattice`` walks the elements as the main loop does
  (puffin_module.f90:58-90): undulator modules through ``und_section`` with
  initUndulator's per-module settings (acc_lattice.f90:848-895) and
  UndSection's entry and exit (matchIn without undulator ends, matchOut and
  correctTrans at the exit; undulator.f90:71-86, 337-340); quads, chicanes,
  drifts and energy modulations through the device elements, the field
  diffracting over drifts and chicanes (acc_lattice.f90:447, 512).
"""
from __future__ import annotations

import math
import shlex
from dataclasses import dataclass, field
from typing import Callable, Optional

from .params import PI_REF, UND_CURVED, UND_ELLIPTICAL, UND_HELICAL, UND_PLANEPOLE
from .undulator import SectionResult, und_section

UND_TYPES = {"helical": UND_HELICAL, "planepole": UND_PLANEPOLE, "curved": UND_CURVED}


@dataclass
class Undulator:
    und_type: int
    periods: int
    alpha: float          # a_w / aw0 (mf): the module's n2col0
    taper: float          # d alpha / d zbar (tapers): undgrad
    steps_per_period: int
    ux: float
    uy: float
    kbeta_x: float
    kbeta_y: float
    h: float              # delmz = 4 pi rho / steps_per_period

    @property
    def nsteps(self) -> int:
        return self.periods * self.steps_per_period


@dataclass
class Quad:
    fx: float
    fy: float


@dataclass
class Chicane:
    zbar: float           # 4 pi rho x (periods of the field slipped over)
    slip: float           # 4 pi rho x slippage in resonant wavelengths
    disp: float           # dispersion D


@dataclass
class Drift:
    zbar: float           # 4 pi rho x length in undulator periods


@dataclass
class Modulation:
    wavenum: float
    mag: float


def _float(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def parse_lattice(text: str, rho: float, pi: float = PI_REF) -> list:
    """Elements of a lattice file (readLatt, acc_lattice.f90:243-406).
    ``pi`` is the reference's constant (a default-real literal, widened)."""
    c1 = 2.0 * rho
    elements = []
    for line in text.splitlines():
        head = line.replace(",", " ").split()
        if not head or head[0][:2] not in ("QU", "UN", "CH", "DR", "MO"):
            continue                                  # comments and other lines
        key = head[0][:2]
        toks = shlex.split(line.replace(",", " "), posix=True)   # quoted undulator types
        if key == "QU":
            elements.append(Quad(_float(toks[1]), _float(toks[2])))
        elif key == "UN":
            ztype = toks[1].strip()                       # compared as written
            nw, mf, taper, nperlam = int(toks[2]), _float(toks[3]), _float(toks[4]), int(toks[5])
            ux, uy, kbx, kby = (_float(t) for t in toks[6:10])
            und_type = UND_TYPES.get(ztype, UND_ELLIPTICAL)
            if und_type in (UND_CURVED, UND_PLANEPOLE):   # readLatt's fixed polarisation
                ux, uy = 0.0, 1.0
            elif und_type == UND_HELICAL:
                ux, uy = 1.0, 1.0
            elements.append(Undulator(und_type, nw, mf, taper, nperlam, ux, uy, kbx, kby,
                                      4.0 * pi * rho / float(nperlam)))
        elif key == "CH":
            zbar, slip, disp = _float(toks[1]), _float(toks[2]), _float(toks[3])
            elements.append(Chicane(2.0 * pi * c1 * zbar, 2.0 * pi * c1 * slip, disp))
        elif key == "DR":
            elements.append(Drift(_float(toks[1]) * 4.0 * pi * rho))
        elif key == "MO":
            elements.append(Modulation(_float(toks[1]), _float(toks[2])))
    return elements


@dataclass
class LatticeResult:
    status: int
    sz: float
    csteps: int = 0       # steps over all modules (iCsteps): the redistribution clock
    modules: list = field(default_factory=list)
    elements_done: int = 0


def _nint(x: float) -> int:
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def run_lattice(drv, elements: list, *, rho: float, diffraction: bool = True, diff_step: float,
                und_ends: Optional[bool] = None, periodic: Optional[bool] = None, redist_every: int = 60, sz: float = 0.0,
                write: Optional[Callable[[int, int], None]] = None, write_every: int | None = None,
                pi: float = PI_REF) -> LatticeResult:
    """Run the element chain on the particles and field held by ``drv`` (an
    Integrator).  ``diff_step`` is the zbar length between diffraction steps
    inside undulators (sDiffFrac 4 pi rho); ``write(module, istep)`` is
    called at each write of a module.  ``und_ends`` and ``periodic`` default
    to the integrator's module settings (qUndEnds_G, the field mesh type)."""
    if und_ends is None:
        und_ends = bool(drv.mp.und_ends)
    if periodic is None:
        periodic = bool(drv.mp.periodic)
    res = LatticeResult(status=0, sz=sz)
    for iel, el in enumerate(elements):
        if isinstance(el, Undulator):
            # initUndulator (acc_lattice.f90:848-895)
            h = el.h
            nsteps = el.nsteps
            if und_ends:
                szfs = 4.0 * pi * rho * 2.0
                szfe = nsteps * h - 4.0 * pi * rho * 2.0
            else:
                szfs, szfe = 0.0, nsteps * h
            drv.set_undulator(el.und_type, el.ux, el.uy, el.kbeta_x, el.kbeta_y, szfs, szfe)
            drv.set_taper(el.alpha, el.alpha, el.taper)
            if not und_ends:
                drv.match(0.0, el.alpha, +1)            # matchIn(szl), undulator.f90:86
            steps4diff = _nint(diff_step / h)
            imod = len(res.modules)
            wr = (lambda istep, imod=imod: write(imod, istep)) if write is not None else None
            sec: SectionResult = und_section(drv, nsteps, h, diffraction=diffraction, steps4diff=steps4diff,
                                             redist_every=redist_every, csteps=res.csteps, write=wr,
                                             write_every=write_every, periodic=periodic)
            res.modules.append(sec)
            res.csteps = sec.csteps
            if sec.status != 0:
                res.status = sec.status
                return res
            res.sz = res.sz + sec.sz_local              # sZ = sZ0 + szl (undulator.f90:223)
            if not und_ends:
                n2col, _ = drv.get_taper()
                drv.match(res.sz, n2col, -1)            # matchOut(sZ), undulator.f90:337
            drv.correct_trans()                         # :340
        elif isinstance(el, Quad):
            drv.quad(el.fx, el.fy)
        elif isinstance(el, Chicane):
            drv.chicane(el.disp, el.slip)
            if diffraction and el.zbar > 0.0:
                drv.diffraction_step(el.zbar)
            res.sz += el.zbar
        elif isinstance(el, Drift):
            drv.drift(el.zbar)
            if diffraction:
                drv.diffraction_step(el.zbar)
            res.sz += el.zbar
        elif isinstance(el, Modulation):
            drv.modulation(el.wavenum, el.mag)
        res.elements_done = iel + 1
    return res
