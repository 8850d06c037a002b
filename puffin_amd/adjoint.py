"""Reverse-mode gradient of the integrated power of one undulator module with
respect to the per-period undulator strength a_w(z) and the seed amplitude,
behind a ``torch.autograd.Function``.

No reference counterpart: the reference computes no derivatives.  The
forward model is the reference's UndSection schedule (undulator.f90:103-333,
restated in ``undulator.und_section``) with the linear taper replaced by a
per-period table of n2col values (a_w / aw); the output is

    J = sum_iz wz[iz] power(iz),   power as gPowerP (io/av_write.f90:39-202),

i.e. the integral of |A_perp|^2 over the mesh (trapezoid in z2 on a temporal
mesh, one period of planes on a periodic mesh).

The backward pass is the discrete adjoint of the implemented steps
(puff_rk4par_adjoint, puff_diffraction_adjoint) with checkpoint and
recompute: the forward keeps a device snapshot of the state every ``every``
RK4 steps; the backward replays each segment from its snapshot, keeping one
snapshot per step of the segment, and walks the segment's operations in
reverse.  Snapshots are caller-owned device memory (torch uint8 tensors).
When memory allows, the segment's replay keeps each step's stage data
(puff_rk4par_capture), so that the adjoint steps do not recompute forward
stages.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .gpu import STATE_HEADER_BYTES, Integrator
from .params import PI_REF, ModuleParams
from .undulator import und_section

try:
    import torch
except ImportError:   # the gradient engine needs torch for device memory
    torch = None


@dataclass
class Op:
    kind: str          # "rk4", "diff" or "wrap"
    sz: float = 0.0
    h: float = 0.0


class _Recorder:
    """Stand-in driver that records the schedule's calls instead of running them."""

    def __init__(self):
        self.ops: list[Op] = []
        self._sz = 0.0

    def rk4par(self, sz, h):
        self.ops.append(Op("rk4", float(sz), float(h)))
        return 0

    def diffraction_step(self, h):
        self.ops.append(Op("diff", 0.0, float(h)))

    def wrap_z2(self):
        self.ops.append(Op("wrap"))


def module_program(nsteps: int, h: float, steps4diff: int, *, diffraction: bool = True,
                   periodic: bool = False, redist_every: int = 60, sz_local: float = 0.0) -> list[Op]:
    """The operations UndSection performs for one module (no writes)."""
    rec = _Recorder()
    und_section(rec, nsteps, h, diffraction=diffraction, steps4diff=steps4diff,
                redist_every=redist_every, sz_local=sz_local, periodic=periodic)
    return rec.ops


def power_weights(mp: ModuleParams) -> np.ndarray:
    """z2 weights wz of J = sum wz power: trapezoid over the NZ2 nodes of a
    temporal mesh; dz2 per plane over one period of a periodic mesh (its
    aliased last plane weighs 0)."""
    wz = np.full(mp.nz2, mp.dz2)
    if mp.periodic:
        wz[-1] = 0.0
    else:
        wz[0] *= 0.5
        wz[-1] *= 0.5
    return wz


def integrated_power(mp: ModuleParams, power: np.ndarray) -> float:
    return float(np.dot(power_weights(mp), power))


def segments(ops: list[Op], every: int) -> list[tuple[int, int]]:
    """[start, stop) op ranges; a new range starts at every ``every``-th RK4 op."""
    starts = [0]
    nrk = 0
    for i, op in enumerate(ops):
        if op.kind == "rk4":
            if nrk > 0 and nrk % every == 0 and i > 0:
                starts.append(i)
            nrk += 1
    bounds = starts + [len(ops)]
    return [(bounds[k], bounds[k + 1]) for k in range(len(starts)) if bounds[k] < bounds[k + 1]]


class ModuleGradient:
    """Forward run of one module with checkpoints, and its adjoint.

    Two memory plans: when every step's capture fits (``keep_all``, chosen
    automatically against ``keep_budget`` and the free device memory) the
    forward keeps them all and the backward is adjoint steps only; otherwise
    the forward keeps a checkpoint every ``every`` steps and the backward
    replays each segment (with step captures or start-state snapshots)
    before reversing it.

    ``drv`` is an Integrator whose particles hold the module's initial state;
    ``field_shape`` is the seed field of unit amplitude (split Re / Im);
    ``taper_period`` / ``taper_z0`` place the per-period table on zbar."""

    def __init__(self, drv: Integrator, ops: list[Op], field_shape: tuple, *, every: int = 30,
                 taper_period: float | None = None, taper_z0: float = 0.0,
                 capture: bool | None = None, capture_budget: int = 32 << 30,
                 keep_all: bool | None = None, keep_budget: int = 128 << 30):
        if torch is None:
            raise RuntimeError("the gradient engine keeps its snapshots in torch device memory")
        self.drv = drv
        self.mp = drv.mp
        self.ops = ops
        self.every = int(every)
        self.segs = segments(ops, self.every)
        dev = torch.device("cuda", torch.cuda.current_device())
        # the seed field's unit shape stays on the device: each evaluation
        # scales it there and uploads it device to device
        self.shape_r = torch.as_tensor(np.ascontiguousarray(field_shape[0], dtype=np.float64).reshape(-1),
                                       device=dev)
        self.shape_i = torch.as_tensor(np.ascontiguousarray(field_shape[1], dtype=np.float64).reshape(-1),
                                       device=dev)
        self._seed_r = torch.empty_like(self.shape_r)
        self._seed_i = torch.empty_like(self.shape_i)
        self.taper_period = taper_period if taper_period else 4.0 * PI_REF * self.mp.rho
        self.taper_z0 = taper_z0
        self.wz = power_weights(self.mp)
        nb = drv.state_bytes()
        self._blob = lambda: torch.empty(nb, dtype=torch.uint8, device=dev)
        # initial particles (taken from the driver as it is now)
        hdr = lambda: np.zeros(STATE_HEADER_BYTES, dtype=np.uint8)
        self.initial = self._blob()
        self.initial_hdr = hdr()
        drv.state_save(self.initial.data_ptr(), self.initial_hdr)
        nc = drv.capture_bytes()
        nrk4 = sum(1 for op in ops if op.kind == "rk4")
        # Whole-module captures: the forward keeps every step's capture (start
        # state and stage data) and the backward reverses the steps with no
        # replay, when they fit keep_budget and 60 % of the free device memory.
        if keep_all is None:
            free, _ = torch.cuda.mem_get_info(dev)
            keep_all = nrk4 * nc <= min(keep_budget, 0.6 * free)
        self.keep_all = bool(keep_all)
        if self.keep_all:
            self.ckpts, self.ckpt_hdr = [], []
            self.capture = True
            self.steps = [torch.empty(nc, dtype=torch.uint8, device=dev) for _ in range(nrk4)]
            self.step_hdr = [hdr() for _ in range(nrk4)]
        else:
            self.ckpts = [self._blob() for _ in self.segs]
            self.ckpt_hdr = [hdr() for _ in self.segs]
            # per-step data of the segment being reversed: step captures (start
            # state and stage data, no forward stage recomputed by the adjoint
            # step) when they fit the budget, else start-state snapshots
            self.capture = (self.every * nc <= capture_budget) if capture is None else bool(capture)
            if self.capture:
                self.steps = [torch.empty(nc, dtype=torch.uint8, device=dev) for _ in range(self.every)]
            else:
                self.steps = [self._blob() for _ in range(self.every)]
            self.step_hdr = [hdr() for _ in range(self.every)]
        self.table = None
        self.seed = None
        self.J = None
        self.rk4_calls = 0
        self.dev = dev
        self.peak_device_bytes = 0

    def memory_bytes(self) -> int:
        return (1 + len(self.ckpts)) * self.initial.numel() + sum(t.numel() for t in self.steps)

    # ---- forward
    def _set_inputs(self, table, seed: float):
        self.table = np.ascontiguousarray(table, dtype=np.float64)
        self.seed = float(seed)
        self.drv.state_load(self.initial.data_ptr(), self.initial_hdr)
        torch.mul(self.shape_r, self.seed, out=self._seed_r)
        torch.mul(self.shape_i, self.seed, out=self._seed_i)
        self.drv.upload_field_torch(self._seed_r, self._seed_i)
        self.drv.set_taper_table(self.table, self.taper_z0, self.taper_period)

    def _run(self, op: Op):
        if op.kind == "rk4":
            st = self.drv.rk4par(op.sz, op.h)
            self.rk4_calls += 1
            if st != 0:
                raise RuntimeError(f"rk4par status {st} at zbar {op.sz}")
        elif op.kind == "diff":
            self.drv.diffraction_step(op.h)
        else:
            self.drv.wrap_z2()

    def _deferred(self, on: bool):
        """Steps without a host wait each (the status is checked per segment)."""
        self.drv.set_deferred(on)

    def _check_status(self, where: str):
        st = self.drv.deferred_status()
        if st != 0:
            raise RuntimeError(f"{where}: a step raised status {st}")

    def forward(self, table, seed: float) -> float:
        self._set_inputs(table, seed)
        self._deferred(True)
        try:
            if self.keep_all:
                j = 0
                for op in self.ops:
                    if op.kind == "rk4":
                        self.drv.rk4par_capture(op.sz, op.h, self.steps[j].data_ptr(), self.step_hdr[j])
                        self.rk4_calls += 1
                        j += 1
                    else:
                        self._run(op)
            for k, (a, b) in enumerate(self.segs if not self.keep_all else []):
                self.drv.state_save(self.ckpts[k].data_ptr(), self.ckpt_hdr[k])
                for op in self.ops[a:b]:
                    self._run(op)
            self._check_status("forward")
        finally:
            self._deferred(False)
        self.J = integrated_power(self.mp, self.drv.power())
        return self.J

    def evaluate(self, table, seed: float) -> float:
        """J without keeping checkpoints (finite differences)."""
        self._set_inputs(table, seed)
        self._deferred(True)
        try:
            for op in self.ops:
                self._run(op)
            self._check_status("forward")
        finally:
            self._deferred(False)
        return integrated_power(self.mp, self.drv.power())

    # ---- backward
    def backward(self, scale: float = 1.0) -> dict:
        """dJ/d(table) and dJ/d(seed amplitude), times ``scale``.  Must follow
        forward() (the state must be the module's final state)."""
        drv = self.drv
        drv.adjoint_enable()
        drv.adjoint_zero()
        drv.adjoint_seed_power(self.wz, scale)
        # the snapshots, the captures and the adjoint buffers are all live now
        free, total = torch.cuda.mem_get_info(self.dev)
        self.peak_device_bytes = max(self.peak_device_bytes, int(total - free))
        self._deferred(True)
        try:
            if self.keep_all:
                j = len(self.steps)
                for op in reversed(self.ops):
                    if op.kind == "rk4":
                        j -= 1
                        drv.rk4par_adjoint_captured(op.sz, op.h, self.steps[j].data_ptr(), self.step_hdr[j])
                    elif op.kind == "diff":
                        drv.diffraction_adjoint(op.h)
                self._check_status("adjoint")
            for k in range(len(self.segs) - 1 if not self.keep_all else -1, -1, -1):
                a, b = self.segs[k]
                drv.state_load(self.ckpts[k].data_ptr(), self.ckpt_hdr[k])
                slots = {}
                for i in range(a, b):
                    op = self.ops[i]
                    if op.kind == "rk4":
                        j = len(slots)
                        slots[i] = j
                        if self.capture:
                            drv.rk4par_capture(op.sz, op.h, self.steps[j].data_ptr(), self.step_hdr[j])
                            continue
                        drv.state_save(self.steps[j].data_ptr(), self.step_hdr[j])
                    self._run(op)
                for i in range(b - 1, a - 1, -1):
                    op = self.ops[i]
                    if op.kind == "rk4":
                        j = slots[i]
                        if self.capture:
                            drv.rk4par_adjoint_captured(op.sz, op.h, self.steps[j].data_ptr(), self.step_hdr[j])
                        else:
                            drv.state_load(self.steps[j].data_ptr(), self.step_hdr[j])
                            drv.rk4par_adjoint(op.sz, op.h)
                    elif op.kind == "diff":
                        drv.diffraction_adjoint(op.h)
                self._check_status("adjoint replay")
            # per-slot dJ/dn2col accumulated on the device by the adjoint steps
            g_table = drv.adjoint_taper_gradient(self.table.size + 1)[: self.table.size]
        finally:
            self._deferred(False)
        lar, lai = drv.adjoint_download_field_torch()
        g_seed = float(torch.dot(lar, self.shape_r) + torch.dot(lai, self.shape_i))
        return {"table": g_table, "seed": g_seed, "field": (lar, lai)}



def checkpoint_schedule(n: int, s: int):
    """Optimal snapshot placement for reversing n steps with s free snapshot
    slots besides the one holding the start state (the recursion of
    Griewank and Walther's binomial checkpointing, "Algorithm 799: revolve",
    ACM TOMS 26 (2000) 19-45): advance m steps, snapshot, reverse the last
    n - m steps with s - 1 slots, then the first m with s.  Its forward
    steps obey F(k, 0) = k (k - 1) / 2 (replays from the start for each step)
    and F(k, j) = min_m [m + F(k - m, j - 1) + F(m, j)], solved here by
    dynamic programming.  Returns (F, M): F[k, j] forward steps and M[k, j]
    the best m, for k <= n, j <= s."""
    F = np.zeros((n + 1, s + 1), dtype=np.int64)
    M = np.ones((n + 1, s + 1), dtype=np.int64)
    k = np.arange(n + 1, dtype=np.int64)
    F[:, 0] = k * (k - 1) // 2
    for j in range(1, s + 1):
        for kk in range(2, n + 1):
            m = np.arange(1, kk)
            c = m + F[kk - m, j - 1] + F[m, j]
            i = int(np.argmin(c))
            F[kk, j] = c[i]
            M[kk, j] = m[i]
    return F, M


class BinomialGradient:
    """Gradient of one module's integrated power with binomial checkpointing
    over ``slots`` device snapshots, for modules whose every-step captures do
    not fit the device (C4: 3840 steps, 7.6 GB per snapshot, 27 GB per step
    capture).  A unit is one RK4 step and the operations that follow it
    before the next step (diffraction, wrap); the module's leading operations
    (its first half diffraction) come before unit 0.  The reverse sweep of
    Griewank and Walther's recursion restores a snapshot, advances to the
    next one, and reverses the units after it first; each adjoint step
    recomputes its forward stages from the step's start state
    (puff_rk4par_adjoint), so no capture is kept.  The first advance doubles
    as the forward pass: J and the cotangent seed come from the state after
    the last unit, reached once on the way.  The snapshot positions follow
    checkpoint_schedule, which minimises the forward steps."""

    def __init__(self, drv: Integrator, ops: list[Op], field_shape: tuple, *, slots: int,
                 taper_period: float | None = None, taper_z0: float = 0.0, progress=None,
                 progress_every: int = 240, device=None):
        if torch is None:
            raise RuntimeError("the gradient engine keeps its snapshots in torch device memory")
        if slots < 2:
            raise ValueError("binomial checkpointing needs at least two snapshot slots")
        self.drv = drv
        self.mp = drv.mp
        self.taper_period = taper_period if taper_period else 4.0 * PI_REF * self.mp.rho
        self.taper_z0 = taper_z0
        self.wz = power_weights(self.mp)
        first = next(i for i, op in enumerate(ops) if op.kind == "rk4")
        self.lead = ops[:first]
        self.units = []
        for op in ops[first:]:
            if op.kind == "rk4":
                self.units.append([op])
            else:
                self.units[-1].append(op)
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dev = dev
        self.shape_r = torch.as_tensor(np.ascontiguousarray(field_shape[0], dtype=np.float64).reshape(-1),
                                       device=dev)
        self.shape_i = torch.as_tensor(np.ascontiguousarray(field_shape[1], dtype=np.float64).reshape(-1),
                                       device=dev)
        self._seed_r = torch.empty_like(self.shape_r)
        self._seed_i = torch.empty_like(self.shape_i)
        nb = drv.state_bytes()
        self.initial = torch.empty(nb, dtype=torch.uint8, device=dev)
        self.initial_hdr = np.zeros(STATE_HEADER_BYTES, dtype=np.uint8)
        drv.state_save(self.initial.data_ptr(), self.initial_hdr)
        self.blobs = [torch.empty(nb, dtype=torch.uint8, device=dev) for _ in range(slots)]
        self.hdrs = [np.zeros(STATE_HEADER_BYTES, dtype=np.uint8) for _ in range(slots)]
        self.progress = progress
        self.progress_every = progress_every
        self.stats = dict(forward_steps=0, adjoint_steps=0, restores=0, snapshots=0)
        self.sched_F, self.sched_M = checkpoint_schedule(len(self.units), slots - 1)
        self.peak_device_bytes = 0

    def memory_bytes(self) -> int:
        return (1 + len(self.blobs)) * self.initial.numel()

    # ---- primitives
    def _run(self, op: Op):
        if op.kind == "rk4":
            self.drv.rk4par(op.sz, op.h)
            self.stats["forward_steps"] += 1
        elif op.kind == "diff":
            self.drv.diffraction_step(op.h)
        else:
            self.drv.wrap_z2()

    def _advance(self, lo: int, m: int):
        for j in range(lo, lo + m):
            for op in self.units[j]:
                self._run(op)
            self._tick()

    def _save(self, slot: int):
        self.drv.state_save(self.blobs[slot].data_ptr(), self.hdrs[slot])
        self.stats["snapshots"] += 1

    def _load(self, slot: int):
        self.drv.state_load(self.blobs[slot].data_ptr(), self.hdrs[slot])
        self.stats["restores"] += 1

    def _tick(self):
        done = self.stats["forward_steps"] + self.stats["adjoint_steps"]
        if self.progress and done % self.progress_every == 0:
            st = self.drv.deferred_status()
            if st != 0:
                raise RuntimeError(f"a step raised status {st}")
            self.progress(dict(self.stats))

    def _note_memory(self):
        if self.dev.type == "cuda":
            free, total = torch.cuda.mem_get_info(self.dev)
            self.peak_device_bytes = max(self.peak_device_bytes, int(total - free))

    def _seed(self):
        """The state is at the start of the last unit: run it, take J and the
        cotangent of the final state (every buffer of the gradient is
        allocated from here on: the device memory in use is read here)."""
        for op in self.units[-1]:
            self._run(op)
        st = self.drv.deferred_status()
        if st != 0:
            raise RuntimeError(f"forward: a step raised status {st}")
        self.J = integrated_power(self.mp, self.drv.power())
        self.drv.adjoint_enable()
        self.drv.adjoint_zero()
        self.drv.adjoint_seed_power(self.wz, self.scale)
        self._note_memory()

    def _adjoint_unit(self, j: int):
        unit = self.units[j]
        for op in reversed(unit[1:]):
            if op.kind == "diff":
                self.drv.diffraction_adjoint(op.h)
        self.drv.rk4par_adjoint(unit[0].sz, unit[0].h)
        self.stats["adjoint_steps"] += 1
        self._tick()

    def _rev(self, lo: int, n: int, base: int, free: list):
        """Reverse units lo .. lo + n - 1; the state at unit lo is held in slot
        `base`; `free` lists the unused slots."""
        if n == 1:
            self._load(base)
            if lo == len(self.units) - 1:
                self._seed()
                self._load(base)
            self._adjoint_unit(lo)
            return
        if not free:
            # no slot left: replay from the base for each unit (quadratic)
            for j in range(n - 1, -1, -1):
                self._load(base)
                self._advance(lo, j)
                if lo + j == len(self.units) - 1:
                    self._seed()
                    self._load(base)
                    self._advance(lo, j)
                self._adjoint_unit(lo + j)
            return
        m = int(self.sched_M[n, len(free)])
        self._load(base)
        self._advance(lo, m)
        slot = free.pop()
        self._save(slot)
        self._rev(lo + m, n - m, slot, free)
        free.append(slot)
        self._rev(lo, m, base, free)

    # ---- the gradient
    def gradient(self, table, seed: float, scale: float = 1.0) -> dict:
        """J and dJ/d(table), dJ/d(seed amplitude) (times ``scale``)."""
        self.scale = scale
        self.table = np.ascontiguousarray(table, dtype=np.float64)
        drv = self.drv
        drv.state_load(self.initial.data_ptr(), self.initial_hdr)
        torch.mul(self.shape_r, float(seed), out=self._seed_r)
        torch.mul(self.shape_i, float(seed), out=self._seed_i)
        drv.upload_field_torch(self._seed_r, self._seed_i)
        drv.set_taper_table(self.table, self.taper_z0, self.taper_period)
        drv.set_deferred(True)
        try:
            for op in self.lead:
                self._run(op)
            self._save(0)
            self._rev(0, len(self.units), 0, list(range(1, len(self.blobs))))
            for op in reversed(self.lead):
                if op.kind == "diff":
                    drv.diffraction_adjoint(op.h)
            st = drv.deferred_status()
            if st != 0:
                raise RuntimeError(f"adjoint: a step raised status {st}")
            g_table = drv.adjoint_taper_gradient(self.table.size + 1)[: self.table.size]
        finally:
            drv.set_deferred(False)
        lar, lai = drv.adjoint_download_field_torch()
        self._note_memory()
        g_seed = float(torch.dot(lar, self.shape_r) + torch.dot(lai, self.shape_i))
        return {"J": self.J, "table": g_table, "seed": g_seed, "stats": dict(self.stats)}

if torch is not None:

    class IntegratedPower(torch.autograd.Function):
        """J(table, seed) of one module; gradients by the discrete adjoint.

        Inputs: ``table`` (float64 tensor of per-period n2col values, on the
        CPU or the device), ``seed`` (0-d float64 tensor, seed amplitude) and
        the ModuleGradient ``engine``.  Output: 0-d float64 tensor J on
        ``seed``'s device; the gradients come back on the inputs' devices."""

        @staticmethod
        def forward(ctx, table, seed, engine):
            J = engine.forward(table.detach().cpu().numpy(), float(seed))
            ctx.engine = engine
            ctx.devices = (table.device, seed.device)
            return torch.tensor(J, dtype=torch.float64, device=seed.device)

        @staticmethod
        def backward(ctx, gJ):
            g = ctx.engine.backward(float(gJ))
            dt, ds = ctx.devices
            return (torch.from_numpy(g["table"]).to(dt),
                    torch.tensor(g["seed"], dtype=torch.float64, device=ds), None)
