"""Host-side split-step schedule of one undulator module, mirroring the
reference's UndSection (undulator.f90:33-358) over the C ABI primitives:

* wrap z2 when the field layout is recomputed at the start
  (getLocalFieldIndices, undulator.f90:103; para_field.f90:212-215);
* first half diffraction step (undulator.f90:154-164);
* per step: rk4par (undulator.f90:199), zbar += h (:222), and on every
  iSteps4Diff-th or the last step the combined half steps dzdF + dzdS, or
  finish / write / start when a write falls on that step (:229-289);
* writes (qWriteq, io/write_adapter.f90:204-255);
* periodic re-layout every iRedistStp steps (:325-331).

A raised flag ends the module with its status (the reference retries up to
three times after recomputing its window; the whole mesh is the window here,
so a retry cannot succeed).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional


def write_due(istep: int, csteps: int, write_every: int, int_write_every: int, nsteps: int) -> bool:
    """qWriteq without a write-steps file (io/write_adapter.f90:230-239)."""
    return (csteps % int_write_every == 0) or (csteps == nsteps) or (csteps % write_every == 0)


@dataclass
class SectionResult:
    status: int
    steps_done: int
    csteps: int
    sz_local: float
    n_diffractions: int = 0
    diffraction_lengths: list = field(default_factory=list)


def und_section(drv, nsteps: int, h: float, *, diffraction: bool, steps4diff: int,
                redist_every: int = 60, csteps: int = 0, sz_local: float = 0.0,
                write: Optional[Callable[[int], None]] = None,
                write_every: int | None = None, int_write_every: int | None = None,
                periodic: bool = False) -> SectionResult:
    """Run one module.  ``drv`` provides rk4par(sz, h) -> status,
    diffraction_step(h) and wrap_z2().  ``write(istep)`` is called at each
    write, after the field has completed its split step."""
    we = write_every if write_every else 10 ** 12
    iwe = int_write_every if int_write_every else 10 ** 12
    res = SectionResult(status=0, steps_done=0, csteps=csteps, sz_local=sz_local)
    if periodic:
        drv.wrap_z2()
    dzds = 0.0
    if diffraction:
        dzds = float(steps4diff) * h / 2.0
        drv.diffraction_step(dzds)
        res.diffraction_lengths.append(dzds)
    szl = sz_local
    for istep in range(1, nsteps + 1):
        res.csteps += 1
        st = drv.rk4par(szl, h)
        if st != 0:
            res.status = st
            break
        szl = szl + h
        res.steps_done = istep
        wr = write is not None and write_due(istep, res.csteps, we, iwe, nsteps)
        done = False
        if diffraction and ((istep % steps4diff == 0) or (istep == nsteps)):
            dzdf = dzds
            if istep == nsteps:
                dzds = 0.0
            elif istep + steps4diff <= nsteps:
                dzds = float(steps4diff) * h / 2
            else:
                dzds = float(nsteps - istep) * h / 2
            if not wr:
                drv.diffraction_step(dzdf + dzds)
                res.diffraction_lengths.append(dzdf + dzds)
            else:
                drv.diffraction_step(dzdf)
                res.diffraction_lengths.append(dzdf)
                write(istep)
                done = True
                if dzds > 0.0:
                    drv.diffraction_step(dzds)
                    res.diffraction_lengths.append(dzds)
        if wr and not done:
            write(istep)
        if periodic and redist_every > 0 and res.csteps % redist_every == 0:
            drv.wrap_z2()
    res.sz_local = szl
    res.n_diffractions = len(res.diffraction_lengths)
    return res
