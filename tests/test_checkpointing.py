"""Binomial checkpointing of the module gradient (adjoint.BinomialGradient)
on the CPU: the schedule against brute force and the revolve bound, and the
engine's order of operations against a recording driver."""
from __future__ import annotations

from functools import lru_cache
from math import comb

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from puffin_amd.adjoint import BinomialGradient, checkpoint_schedule, module_program
from puffin_amd.params import ModuleParams


@lru_cache(None)
def _brute(n, s):
    if n <= 1:
        return 0
    if s == 0:
        return n * (n - 1) // 2
    return min(m + _brute(n - m, s - 1) + _brute(m, s) for m in range(1, n))


def test_schedule_is_optimal_and_meets_the_revolve_bound():
    F, M = checkpoint_schedule(60, 5)
    for n in range(1, 61):
        for s in range(0, 6):
            assert F[n, s] == _brute(n, s), (n, s)
            if s >= 1 and n >= 2:
                m = M[n, s]
                assert F[n, s] == m + F[n - m, s - 1] + F[m, s]
    # Griewank-Walther: with t the least number of sweeps reversing n steps
    # with s snapshots, at most t n - C(s + t, s + 1) forward steps (revolve
    # counts a final advance that this recursion does not need)
    for s in range(1, 6):
        for n in range(2, 61):
            t = 0
            while comb(s + t, s) < n:
                t += 1
            assert F[n, s] <= t * n - comb(s + t, s + 1), (n, s)


class RecordingDriver:
    """Stands in for Integrator: the state is the count of operations applied
    since the module's start; snapshots map blob addresses to that count."""

    def __init__(self, ops):
        self.ops = ops
        self.mp = ModuleParams(nx=3, ny=3, nz2=4, dx=1.0, dy=1.0, dz2=0.1, rho=0.01, eta=1e-4,
                               kappa=0.5, gamma_r=100.0, aw=1.0)
        self.pos = 0
        self.blobs = {}
        self.log = []
        self.seeded = 0
        self.forward_steps = 0

    def state_bytes(self):
        return 8

    def state_save(self, ptr, hdr=None):
        self.blobs[ptr] = self.pos

    def state_load(self, ptr, hdr=None):
        self.pos = self.blobs[ptr]

    def _expect(self, kind, *args):
        op = self.ops[self.pos]
        assert op.kind == kind, (self.pos, op, kind)
        self.pos += 1

    def rk4par(self, sz, h):
        self._expect("rk4")
        self.forward_steps += 1
        return 0

    def diffraction_step(self, h):
        self._expect("diff")

    def wrap_z2(self):
        self._expect("wrap")

    def power(self):
        assert self.pos == len(self.ops), "J is taken at the module's end"
        return np.ones(self.mp.nz2)

    def adjoint_enable(self):
        pass

    def adjoint_zero(self):
        pass

    def adjoint_seed_power(self, wz, scale):
        self.seeded += 1

    def rk4par_adjoint(self, sz, h):
        op = self.ops[self.pos]
        assert op.kind == "rk4" and (op.sz, op.h) == (sz, h), "adjoint step from its start state"
        self.log.append(("rk4", self.pos))
        return 0, np.zeros(4), np.full(4, -1)

    def diffraction_adjoint(self, h):
        self.log.append(("diff", h))

    def set_deferred(self, on):
        pass

    def deferred_status(self):
        return 0

    def set_taper_table(self, *a):
        pass

    def upload_field_torch(self, ar, ai):
        assert self.pos == 0

    def adjoint_taper_gradient(self, n):
        return np.zeros(n)

    def adjoint_download_field_torch(self):
        return torch.zeros(self.mp.nmesh, dtype=torch.float64), torch.zeros(self.mp.nmesh, dtype=torch.float64)


@pytest.mark.parametrize("nsteps,slots", [(7, 2), (30, 3), (64, 4), (97, 6)])
def test_engine_reverses_every_step_from_its_start_state(nsteps, slots):
    ops = module_program(nsteps, 0.01, 5, diffraction=True, periodic=True, redist_every=10)
    drv = RecordingDriver(ops)
    shape = (np.zeros(drv.mp.nmesh), np.zeros(drv.mp.nmesh))
    eng = BinomialGradient(drv, ops, shape, slots=slots, device="cpu")
    drv.pos = 0
    eng.gradient(np.ones(3), 1.0)
    assert drv.seeded == 1
    # the adjoint steps run from the last step to the first, each preceded by
    # the adjoints of the diffraction steps that followed it
    rk4_pos = [i for i, op in enumerate(ops) if op.kind == "rk4"]
    assert [p for k, p in drv.log if k == "rk4"] == rk4_pos[::-1]
    want = []
    for i in range(len(ops) - 1, -1, -1):
        if ops[i].kind == "diff":
            want.append(("diff", ops[i].h))
        elif ops[i].kind == "rk4":
            want.append(("rk4", i))
    assert drv.log == want
    # forward steps: the schedule's count, plus the run up to the module's end
    # for J and the cotangent seed: the last step once more when its start
    # state sits in a slot, else the whole last slot-less segment again
    F, M = checkpoint_schedule(nsteps, slots - 1)
    n, s = nsteps, slots - 1
    while s > 0 and n > 1:
        n, s = n - int(M[n, s]), s - 1
    assert drv.forward_steps == F[nsteps, slots - 1] + (1 if n == 1 else n)
