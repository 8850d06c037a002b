"""The reference's own 1-D end-to-end case (test/inputs/f1main.in with
beam_file.in and seed_file.in; checked by test/testMPIIntegration.pf:54-127),
rebuilt from its deck:

* scaling: rho = 0.01, aw = 1, gamma_r = 100, ux = uy = 1 (elliptical);
* mesh: 1-D, sFModelLengthZ2 = 4 with 17 nodes per lambda_r gives NZ2 = 511,
  dz2 = 4/510 (setup_calcs.f90:955-1094);
* beam: flat-top of length 2, 20 macroparticles per wave -> 319 cold
  macroparticles at the cell midpoints of [0, 2] (gen_macros.f90:201-205),
  Gamma = 1, p_perp = 0, x = y = 0, then matchIn at zbar = 0
  (acc_lattice.f90:743-828); the charge weights are the fixture's own;
* seed: zero field; 30 steps of 4 pi rho / 30; no diffraction (1-D);
* end of run: matchOut then correctTrans (undulator.f90:347-349), then the
  final dump that the fixture holds (puffin_module.f90:92-97).
"""
from __future__ import annotations

import os

import numpy as np

from puffin_amd.params import ModuleParams, scaling, step_size

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "f1main_expected.npz")
NSTEPS = 30


def load_fixture():
    d = np.load(FIXTURE)
    info = dict(zip([str(k) for k in d["info_keys"]], [float(v) for v in d["info_vals"]]))
    return d, info


def module_params() -> ModuleParams:
    eta, kappa, fx, fy = scaling(0.01, 1.0, 100.0, 1.0, 1.0)
    h = step_size(0.01, 30)
    return ModuleParams(nx=1, ny=1, nz2=511, dx=1.0, dy=1.0, dz2=4.0 / 510, rho=0.01,
                        eta=eta, kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy,
                        one_d=1, szfs=0.0, szfe=NSTEPS * h)


def initial_particles(chi: np.ndarray) -> dict:
    n = 319
    d = 2.0 / n
    z2 = np.array([j * d + d / 2.0 for j in range(n)])
    return dict(x=np.zeros(n), y=np.zeros(n), z2=z2, px=np.zeros(n), py=np.zeros(n),
                gam=np.ones(n), chi=np.array(chi, dtype=np.float64).copy())
