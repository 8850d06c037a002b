"""Module constants of one undulator section, mirroring ``puff_params`` in
include/puffin_gpu.h, plus the scaling relations that produce them.

The scaling follows the reference's setup:
  * calcScaling (setup/setup_calcs.f90:722-803): eta, kappa from rho, aw,
    gamma_r and the undulator type;
  * calcSamples (setup/setup_calcs.f90:833-1094): dz2, step size;
  * passToGlobals (setup/setup_calcs.f90:195-248): filter cutoff, diffraction
    step, boundary widths NBX/NBY/NBZ2;
  * initUndulator (acc_lattice.f90:848-895): sZFS / sZFE of a module.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, asdict

import numpy as np

# puffin_constants.f90:22 declares pi with a default-real literal, so every
# "pi" of the reference outside GetKValues / gaussian is float32(pi) widened.
PI_REF = float(np.float32(math.pi))

UND_ELLIPTICAL, UND_PLANEPOLE, UND_HELICAL, UND_CURVED = 0, 1, 2, 3
FP64, MIXED = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1

STATUS_NAMES = {
    0: "ok",
    1: "macroparticle outside the transverse mesh",
    2: "macroparticle beyond the z2 buffer",
    3: "macroparticle beyond the temporal z2 mesh",
    -1: "invalid argument",
    -2: "HIP error",
    -3: "device allocation failed",
    -4: "invalid state",
}


class PuffParams(ctypes.Structure):
    """ctypes mirror of ``puff_params`` (include/puffin_gpu.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz2", ctypes.c_int32),
        ("periodic", ctypes.c_int32),
        ("one_d", ctypes.c_int32),
        ("und_type", ctypes.c_int32),
        ("und_ends", ctypes.c_int32),
        ("focusing", ctypes.c_int32),
        ("electrons_evolve", ctypes.c_int32),
        ("field_evolve", ctypes.c_int32),
        ("coupling", ctypes.c_int32),
        ("filter", ctypes.c_int32),
        ("nbx", ctypes.c_int32),
        ("nby", ctypes.c_int32),
        ("nbz2", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("dz2", ctypes.c_double),
        ("rho", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("gamma_r", ctypes.c_double),
        ("aw", ctypes.c_double),
        ("fx", ctypes.c_double),
        ("fy", ctypes.c_double),
        ("szfs", ctypes.c_double),
        ("szfe", ctypes.c_double),
        ("kbeta_x", ctypes.c_double),
        ("kbeta_y", ctypes.c_double),
        ("kx_und", ctypes.c_double),
        ("ky_und", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("filt_cutoff", ctypes.c_double),
        ("pi_ref", ctypes.c_double),
    ]


@dataclass
class ModuleParams:
    """Python view of one module's constants (same fields as PuffParams)."""

    nx: int
    ny: int
    nz2: int
    dx: float
    dy: float
    dz2: float
    rho: float
    eta: float
    kappa: float
    gamma_r: float
    aw: float
    fx: float = 1.0
    fy: float = 1.0
    periodic: int = 0
    one_d: int = 0
    und_type: int = UND_ELLIPTICAL
    und_ends: int = 0
    focusing: int = 0
    electrons_evolve: int = 1
    field_evolve: int = 1
    coupling: int = 1
    filter: int = 1
    nbx: int = 16
    nby: int = 16
    nbz2: int = 37
    szfs: float = 0.0
    szfe: float = 1.0e30
    kbeta_x: float = 0.0
    kbeta_y: float = 0.0
    kx_und: float = 0.0
    ky_und: float = 0.0
    beta: float = 1.0
    filt_cutoff: float = 0.0
    pi_ref: float = PI_REF

    def to_ctypes(self) -> PuffParams:
        p = PuffParams()
        for name, _ in PuffParams._fields_:
            if name == "reserved":
                continue
            setattr(p, name, getattr(self, name))
        return p

    @property
    def nzp(self) -> int:
        """FFT length along z2 (transforms.f90:146-150)."""
        return self.nz2 - 1 if self.periodic else self.nz2

    @property
    def ntr(self) -> int:
        return self.nx * self.ny

    @property
    def nmesh(self) -> int:
        return self.nx * self.ny * self.nz2

    def asdict(self) -> dict:
        return asdict(self)


def scaling(rho: float, aw: float, gamma_r: float, ux: float, uy: float,
            und_type: int = UND_ELLIPTICAL):
    """eta, kappa, (fx, fy) from calcScaling (setup_calcs.f90:722-803)."""
    fx, fy = ux, uy
    if und_type in (UND_CURVED, UND_PLANEPOLE):
        aw_rms = aw / math.sqrt(2.0)
        fx, fy = 0.0, 1.0
    elif und_type == UND_HELICAL:
        aw_rms = aw
        fx, fy = 1.0, 1.0
    else:
        aw_rms = aw * math.sqrt(ux ** 2 + uy ** 2) / math.sqrt(2.0)
    betaz = math.sqrt(gamma_r ** 2.0 - 1.0 - aw_rms ** 2.0) / gamma_r
    eta = (1.0 - betaz) / betaz
    kappa = aw / 2.0 / rho / gamma_r
    return eta, kappa, fx, fy


def gamma_r_for_eta(eta: float, aw: float, ux: float, uy: float,
                    und_type: int = UND_ELLIPTICAL) -> float:
    """Invert calcScaling's eta(gamma_r) for a target eta (bisection on the
    closed form; eta falls monotonically with gamma_r)."""
    lo, hi = 2.0, 1.0e7
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        e, _, _, _ = scaling(0.01, aw, mid, ux, uy, und_type)
        if e > eta:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def filter_cutoff(rho: float, nz2: int, dz2: float, periodic: bool,
                  filt_frac: float = 0.3, per_waves: float | None = None) -> float:
    """multiplyexp's kz2 cutoff: 2 pi sFilt / (NZ2 dz2) with
    sFilt = NZ2 dz2 / (4 pi rho) * sFiltFrac (setup_calcs.f90:208-210,
    diffraction.f90:138-139); zero for a periodic mesh of <= 1.1 waves."""
    lenz2 = dz2 * nz2
    lam_r_bar = 4 * PI_REF * rho
    sfilt = lenz2 / lam_r_bar * filt_frac
    cutoff = 2.0 * PI_REF * sfilt / (float(nz2) * dz2)
    if periodic and per_waves is not None and per_waves <= 1.1:
        cutoff = 0.0
    return cutoff


def step_size(rho: float, steps_per_period: int) -> float:
    """sStepSize = 4 pi rho / stepsPerPeriod (setup_calcs.f90:935-939)."""
    return 4.0 * PI_REF * rho / float(steps_per_period)


def steps_for_diffraction(rho: float, h: float, diff_frac: float = 1.0) -> int:
    """iSteps4Diff = nint(diffStep / sStepSize), diffStep = sDiffFrac 4 pi rho
    (setup_calcs.f90:240, undulator.f90:106)."""
    v = diff_frac * 4.0 * PI_REF * rho / h
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))
