"""Throughput of the integrator on one MI355X: macroparticle-steps/s and the
HBM roofline fraction of the dominant kernel, FP64 and FP32-mixed, on the
planar tapered workload (129 x 129 x 1024 temporal mesh, 2^24
macroparticles).  A step is one RK4 step (rk4par) plus its share of the
diffraction step that runs every iSteps4Diff = 30 steps.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

WORKLOADS = {
    "c1": "planar tapered FEL, temporal mesh 129x129x1024, 2^24 macroparticles",
    "c0": "elliptical (helical-like) FEL, periodic mesh 33x33x128, 2^18 macroparticles",
    "c3": "elliptical FEL with per-period a_w, periodic mesh 33x33x256, 2^20 macroparticles",
}


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


# Algorithmic bytes: per macroparticle per RK4 stage (stage s reads y0, k_{s-1}
# and chi and writes k_s; stage 4 also reads k1, k2 and writes y) and per
# active node per stage; the field combination and the diffraction per node.
def stage_bytes(precision: int):
    if precision == 0:   # FP64: six-vector 48 B, chi 8 B, complex 16 B
        part = [48 + 8 + 48, 48 + 48 + 8 + 48, 48 + 48 + 8 + 48, 4 * 48 + 8 + 48]
        node = [2 * 16, 3 * 16, 3 * 16, 3 * 16]
        comb, diff = 6 * 16, 14 * 16
    else:                # mixed: state 32 B (z2, Gamma FP64), derivative k 24 B (all FP32),
        # chi 4 B, complex 8 B
        part = [32 + 4 + 24, 32 + 24 + 4 + 24, 32 + 24 + 4 + 24, 32 + 3 * 24 + 4 + 32]
        node = [2 * 8, 3 * 8, 3 * 8, 3 * 8]
        comb, diff = 6 * 8, 14 * 8
    return part, node, comb, diff


def run_gpu(case, precision: int, steps: int, warmup: int):
    """Warm up on the module's schedule (diffraction every iSteps4Diff steps),
    then time `steps` RK4 steps and, separately, E = ceil(steps / iSteps4Diff)
    diffraction events, each segment bracketed by device synchronisation.
    A step is one RK4 step plus 1 / iSteps4Diff of a diffraction event:
    ms_per_step = rk4_ms + event_ms / iSteps4Diff, whatever `steps` is."""
    from puffin_amd import gpu
    mp = case.mp
    name = "fp64" if precision == 0 else "mixed"
    log(f"{name}: creating context n={case.n} mesh={mp.nx}x{mp.ny}x{mp.nz2}")
    g = gpu.Integrator(mp, case.n, precision)
    try:
        g.upload_particles(case.particles)
        g.upload_field(*case.field)
        g.set_taper(case.n2col0, case.n2col0, case.undgrad)
        h, s4d = case.h, case.steps4diff
        szl = 0.0
        for istep in range(1, warmup + 1):
            st = g.rk4par(szl, h)
            if st != 0:
                raise RuntimeError(f"rk4par status {st} at warm-up step {istep}")
            szl += h
            if istep % s4d == 0:
                g.diffraction_step(s4d * h)
        g.diffraction_step(s4d * h)
        g.synchronize()
        g.profile(True)
        g.profile_read(reset=True)
        events = max(1, -(-steps // s4d))
        log(f"{name}: timing {steps} RK4 steps and {events} diffraction events")
        t0 = time.perf_counter()
        for istep in range(steps):
            st = g.rk4par(szl, h)
            if st != 0:
                raise RuntimeError(f"rk4par status {st} at timed step {istep}")
            szl += h
        g.synchronize()
        t1 = time.perf_counter()
        for _ in range(events):
            g.diffraction_step(s4d * h)
        g.synchronize()
        t2 = time.perf_counter()
        prof = g.profile_read(reset=True)
        g.profile(False)
        box = g.active_box()
        power = g.power()
        rk4_ms = (t1 - t0) * 1e3 / steps
        event_ms = (t2 - t1) * 1e3 / events
        ms = rk4_ms + event_ms / s4d
        log(f"{name}: {ms:.3f} ms/step ({rk4_ms:.3f} RK4 + {event_ms:.3f} / {s4d} diffraction)")
        return dict(ms_per_step=ms, rk4_ms=rk4_ms, event_ms=event_ms, events=events,
                    timed_s=t2 - t0, prof=prof, box=box, power=power)
    finally:
        g.close()


def roofline(case, precision: int, res: dict, steps: int):
    part, node, comb, diff = stage_bytes(precision)
    b = res["box"]
    n_act = (b[1] - b[0] + 2) * (b[3] - b[2] + 2) * (b[5] - b[4] + 2) if b[1] >= 0 else 0
    n_p = case.n
    kernels = {}
    for s in range(4):
        ms, cnt = res["prof"][f"stage{s + 1}"]
        if cnt:
            kernels[f"stage{s + 1}"] = dict(avg_ms=ms / cnt, bytes=n_p * part[s] + n_act * node[s])
    ms, cnt = res["prof"]["combine"]
    if cnt:
        kernels["combine"] = dict(avg_ms=ms / cnt, bytes=n_act * comb)
    ms, cnt = res["prof"]["diffraction"]
    nmesh_fft = case.mp.nx * case.mp.ny * case.mp.nzp
    if cnt:
        kernels["diffraction"] = dict(avg_ms=ms / cnt, bytes=nmesh_fft * diff)
    for k in kernels.values():
        k["GBps"] = k["bytes"] / (k["avg_ms"] * 1e-3) / 1e9
        k["frac"] = k["GBps"] / HBM_PEAK_GBS
    total_ms = {k: res["prof"][k][0] for k in res["prof"]}
    dom = max((k for k in kernels if k.startswith("stage")), key=lambda k: total_ms[k])
    step_bytes = (n_p * sum(part) + n_act * (sum(node) + comb) + nmesh_fft * diff / case.steps4diff)
    return kernels, dom, n_act, step_bytes


def measured_traffic(precision: int, kernel: str):
    """HBM bytes per dispatch of a stage kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/hbm_traffic_c1_*.md), scaled by
    the copy-kernel calibration recorded there; None when not measured."""
    name = "fp64" if precision == 0 else "mixed"
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_c1_{name}.md")
    if not os.path.exists(path) or not kernel.startswith("stage"):
        return None
    typ = "double" if precision == 0 else "float"
    want = f"`k_stage<{typ}, {kernel[-1]}>`"
    sf = sw = None
    row = None
    for line in open(path):
        if line.startswith("SCALE"):
            parts = dict(p.split("=") for p in line.split()[1:])
            sf, sw = float(parts["fetch"]), float(parts["write"])
        if line.startswith("| " + want):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            row = (float(cols[2]), float(cols[3]))
    if row is None or sf is None:
        return None
    return (row[0] * sf + row[1] * sw) * 1024.0


def integrated_power(power: np.ndarray, dz2: float) -> float:
    return float(np.trapezoid(power, dx=dz2)) if hasattr(np, "trapezoid") else float(np.trapz(power, dx=dz2))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, nsample: int, steps: int):
    """The CPU oracle (FP64 C restatement, OpenMP, atomic deposit as the
    reference's !$OMP ATOMIC) on a sample of the same workload: `steps` RK4
    steps of a sample of the beam on the full mesh, and one diffraction event
    of the full mesh, amortised over iSteps4Diff steps as on the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(case.n, size=min(nsample, case.n), replace=False))
    parts = {k: np.ascontiguousarray(v[idx]) for k, v in case.particles.items()}
    # charge of the sample scaled so the current density matches the full beam
    parts["chi"] = parts["chi"] * (case.n / idx.size)
    orc.set_deposit_atomic(True)
    o = orc.Oracle(case.mp, parts, case.field, taper=(case.n2col0, case.n2col0, case.undgrad),
                   nthreads=threads)
    log(f"cpu baseline: {idx.size} macroparticles, {threads} threads, {cpu_model()}")
    szl = 0.0
    st = o.rk4par(szl, case.h)   # warm-up
    szl += case.h
    t0 = time.perf_counter()
    for _ in range(steps):
        st = o.rk4par(szl, case.h)
        szl += case.h
    t1 = time.perf_counter()
    o.diffraction_step(case.steps4diff * case.h)
    t2 = time.perf_counter()
    orc.set_deposit_atomic(False)
    assert st == 0, st
    step_s = (t1 - t0) / steps + (t2 - t1) / case.steps4diff
    val = idx.size / step_s
    log(f"cpu baseline: {val:.4g} macroparticle-steps/s ({(t1 - t0) / steps:.3f} s per RK4 step, "
        f"{t2 - t1:.2f} s per diffraction event)")
    return dict(value=val, unit="macroparticle-steps/s", cores=threads, kind="port",
                cpu=cpu_model(), rk4_s_per_step=(t1 - t0) / steps, diffraction_s_per_event=t2 - t1,
                sample=(f"{idx.size} of the {case.n} macroparticles on the full "
                        f"{case.mp.nx}x{case.mp.ny}x{case.mp.nz2} mesh: {steps} RK4 steps plus one "
                        f"diffraction event of the full mesh amortised over {case.steps4diff} steps; "
                        f"FP64 C oracle with OpenMP atomics"))


def accuracy_run(case, precision: int, steps: int, progress_every: int | None = None):
    """The module's schedule (UndSection: first half diffraction step, rk4par,
    merged diffraction every iSteps4Diff steps, last half step) for `steps`
    steps; returns the final power(z2), slice bunching |b| and the on-axis
    spectrum.  With progress_every, a write every that many steps logs the
    step (UndSection then splits the diffraction step there, as it does
    around its writes, in both forms alike)."""
    from puffin_amd import gpu
    from puffin_amd.undulator import und_section
    mp = case.mp
    g = gpu.Integrator(mp, case.n, precision)
    try:
        g.upload_particles(case.particles)
        g.upload_field(*case.field)
        g.set_taper(case.n2col0, case.n2col0, case.undgrad)
        if case.taper_table is not None:
            g.set_taper_table(case.taper_table, 0.0, case.taper_period)
        t0 = time.perf_counter()
        wr = None
        if progress_every:
            def wr(istep):
                log(f"accuracy ({'fp64' if precision == 0 else 'mixed'}): step {istep}, "
                    f"{time.perf_counter() - t0:.1f} s")
        res = und_section(g, steps, case.h, diffraction=True, steps4diff=case.steps4diff,
                          periodic=bool(mp.periodic), write=wr, write_every=progress_every)
        g.synchronize()
        wall = time.perf_counter() - t0
        if res.status != 0:
            raise RuntimeError(f"rk4par status {res.status} after {res.steps_done} steps")
        power = g.power()
        lam = 4.0 * math.pi * mp.rho
        nsl = int(math.ceil((mp.nz2 - 1) * mp.dz2 / lam))
        bunch = g.bunching(1, nsl, lam)
        ar, ai = g.download_field()
    finally:
        g.close()
    ntr = mp.nx * mp.ny
    c = (mp.ny // 2) * mp.nx + mp.nx // 2
    axis = ar[c::ntr][: mp.nzp] + 1j * ai[c::ntr][: mp.nzp]
    spec = np.abs(np.fft.fft(axis)) ** 2
    spec = spec[: mp.nzp // 2 + 1]                  # one-sided (plotSpecInt.py:79-100)
    return dict(power=power, bunching=bunch, spectrum=spec, wall_s=wall,
                integrated_power=integrated_power(power, mp.dz2), steps=res.steps_done,
                diffractions=res.n_diffractions)


def precision_report(case, steps: int, progress_every: int | None = None):
    """Error of the FP32-mixed form against FP64 per output after `steps`
    steps of the module's schedule."""
    log(f"accuracy: {steps} steps FP64 and mixed")
    r64 = accuracy_run(case, 0, steps, progress_every)
    r32 = accuracy_run(case, 1, steps, progress_every)

    def err(k):
        a, b = r32[k], r64[k]
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    p0 = integrated_power(np.asarray(case_power0(case)), case.mp.dz2)
    out = dict(steps=steps, periods=steps * case.h / (4.0 * math.pi * case.mp.rho),
               diffraction_events=r64["diffractions"],
               integrated_power_fp64=r64["integrated_power"],
               integrated_power_rel_err=abs(r32["integrated_power"] - r64["integrated_power"]) /
               abs(r64["integrated_power"]),
               power_z_max_rel_err=err("power"), bunching_max_rel_err=err("bunching"),
               spectrum_max_rel_err=err("spectrum"),
               power_gain=r64["integrated_power"] / p0 if p0 > 0 else None,
               wall_s=dict(fp64=r64["wall_s"], mixed=r32["wall_s"]))
    log(f"accuracy: integrated power rel err {out['integrated_power_rel_err']:.3g}, "
        f"power(z) {out['power_z_max_rel_err']:.3g}, |b| {out['bunching_max_rel_err']:.3g}, "
        f"spectrum {out['spectrum_max_rel_err']:.3g}")
    return out


def case_power0(case):
    """power(z2) of the initial field (transverse trapezoid of |A|^2)."""
    mp = case.mp
    ar = case.field[0].reshape(mp.nz2, mp.ny, mp.nx)
    ai = case.field[1].reshape(mp.nz2, mp.ny, mp.nx)
    f = ar * ar + ai * ai
    wx = np.full(mp.nx, mp.dx)
    wx[[0, -1]] *= 0.5
    wy = np.full(mp.ny, mp.dy)
    wy[[0, -1]] *= 0.5
    return np.einsum("zyx,y,x->z", f, wy, wx)


def adjoint_bench(steps_check: bool = True):
    """d(integrated power)/d(per-period a_w, seed amplitude) on the adjoint
    workload (helical, periodic 33 x 33 x 256, 2^20 macroparticles, 16
    periods), under two memory plans: every step's capture kept in the
    forward (no replay), and checkpoints every 30 steps with each segment
    replayed (with step captures) before it is reversed.  Forward and
    gradient wall time, the device memory in use while the snapshots and
    adjoint buffers are live, and central differences (h = 1e-6) of two
    components."""
    import torch
    from puffin_amd import gpu
    from puffin_amd.adjoint import ModuleGradient, module_program
    from puffin_amd.configs import case_c3
    c = case_c3()
    mp = c.mp
    ops = module_program(c.nsteps, c.h, c.steps4diff, diffraction=True, periodic=True, redist_every=60)
    drv = gpu.Integrator(mp, c.n, 0)
    try:
        drv.upload_particles(c.particles)
        drv.upload_field(np.zeros(mp.nmesh), np.zeros(mp.nmesh))
        shape = (c.field[0] / c.info["seed_amp"], c.field[1] / c.info["seed_amp"])
        table, seed = c.taper_table, c.info["seed_amp"]
        plans = {}
        grads = {}
        for name, keep_all in (("captures", None), ("checkpoints_every_30", False)):
            # each engine snapshots the driver's particles as the module's start
            drv.upload_particles(c.particles)
            eng = ModuleGradient(drv, ops, shape, every=30, taper_period=c.taper_period, keep_all=keep_all)
            eng.evaluate(table, seed)             # warm-up
            drv.synchronize()
            t0 = time.perf_counter()
            J = eng.evaluate(table, seed)
            t1 = time.perf_counter()
            eng.forward(table, seed)
            t2 = time.perf_counter()
            g = eng.backward()
            t3 = time.perf_counter()
            plans[name] = dict(
                memory_plan=("every step captured in the forward, no replay" if eng.keep_all
                             else "checkpoints every 30 steps, each segment replayed with step captures"
                             if eng.capture else "checkpoints every 30 steps, segments replayed"),
                J=J, forward_s=t1 - t0, forward_checkpointed_s=t2 - t1, backward_s=t3 - t2,
                gradient_cost_vs_forward=(t3 - t1) / (t1 - t0),
                snapshot_bytes=eng.memory_bytes(), peak_device_bytes=eng.peak_device_bytes,
                grad_seed=g["seed"])
            grads[name] = g
            log(f"adjoint ({name}): forward {t1 - t0:.3f} s, gradient {t3 - t1:.3f} s "
                f"({plans[name]['gradient_cost_vs_forward']:.2f}x), "
                f"{eng.peak_device_bytes / 2**30:.1f} GiB in use")
            if name == "captures" and steps_check:
                eps = 1e-6
                fd_s = ((eng.evaluate(table, seed * (1 + eps)) - eng.evaluate(table, seed * (1 - eps)))
                        / (2 * eps * seed))
                k = int(np.argmax(np.abs(g["table"])))
                tp, tm = table.copy(), table.copy()
                tp[k] += eps
                tm[k] -= eps
                fd_t = (eng.evaluate(tp, seed) - eng.evaluate(tm, seed)) / (2 * eps)
                plans[name]["fd_check"] = dict(
                    h=eps, seed_rel_err=abs(fd_s - g["seed"]) / abs(g["seed"]), table_index=k,
                    table_rel_err=abs(fd_t - g["table"][k]) / np.abs(g["table"]).max())
            del eng, g
            torch.cuda.empty_cache()
        a, b = grads["captures"], grads["checkpoints_every_30"]
        out = dict(workload=WORKLOADS["c3"] + ", 16 periods of 30 steps", plans=plans,
                   plans_table_rel_diff=float(np.abs(a["table"] - b["table"]).max()
                                              / np.abs(a["table"]).max()),
                   plans_seed_rel_diff=abs(a["seed"] - b["seed"]) / abs(a["seed"]))
        # the headline keys are the captures plan's
        for k in ("J", "forward_s", "forward_checkpointed_s", "backward_s", "gradient_cost_vs_forward",
                  "memory_plan", "snapshot_bytes", "peak_device_bytes", "grad_seed"):
            out[k] = plans["captures"][k]
        if "fd_check" in plans["captures"]:
            out["fd_check"] = plans["captures"]["fd_check"]
        out["grad_table"] = [float(v) for v in a["table"]]
        return out
    finally:
        drv.close()


def c1_gradient(precision: int = 0, fd_rel: float = 1e-4):
    """d(integrated power)/d(per-period taper, seed amplitude) of the whole
    C1 module (129 x 129 x 1024, 2^24 macroparticles, 1920 steps on
    UndSection's schedule), the workload the brief quotes its <= 4x gradient
    cost on.  The linear taper of -1% over 64 periods enters as a 64-value
    per-period table (the ramp at period centres, as n2col per period).
    Checkpoints every 30 steps; each segment is replayed with step captures
    before it is reversed.  Times a plain forward, the gradient, and central
    differences of the seed amplitude and of one table entry; the device
    memory is read while every checkpoint, capture and adjoint buffer is
    live."""
    import torch
    from puffin_amd import gpu
    from puffin_amd.adjoint import ModuleGradient, module_program
    from puffin_amd.configs import case_c1
    from puffin_amd.params import PI_REF
    c = case_c1()
    mp = c.mp
    seed = 1.0e-3                      # the Gaussian seed's amplitude (Im part)
    shape = (c.field[0] / seed, c.field[1] / seed)
    period = 4.0 * PI_REF * mp.rho
    periods = c.info["periods"]
    table = c.n2col0 + c.undgrad * period * (np.arange(periods) + 0.5)
    ops = module_program(c.nsteps, c.h, c.steps4diff, diffraction=True, periodic=False)
    name = "fp64" if precision == 0 else "mixed"
    out = dict(workload=WORKLOADS["c1"] + f", whole 64-period module (1920 steps), {name}",
               taper="-1% over 64 periods as a 64-value per-period table (ramp at period centres)")
    drv = gpu.Integrator(mp, c.n, precision)
    try:
        drv.upload_particles(c.particles)
        drv.upload_field(np.zeros(mp.nmesh), np.zeros(mp.nmesh))
        nc = drv.capture_bytes()
        eng = ModuleGradient(drv, ops, shape, every=30, taper_period=period, capture=True,
                             capture_budget=30 * nc, keep_all=False)
        eng.evaluate(table, seed)           # warm-up
        drv.synchronize()
        t0 = time.perf_counter()
        J = eng.evaluate(table, seed)
        t1 = time.perf_counter()
        eng.forward(table, seed)
        t2 = time.perf_counter()
        g = eng.backward()
        t3 = time.perf_counter()
        out.update(J=J, forward_s=t1 - t0, forward_checkpointed_s=t2 - t1, backward_s=t3 - t2,
                   gradient_cost_vs_forward=(t3 - t1) / (t1 - t0),
                   memory_plan="checkpoints every 30 steps, each segment replayed with step captures",
                   checkpoint_bytes=eng.memory_bytes(), capture_bytes_per_step=nc,
                   peak_device_bytes=eng.peak_device_bytes, grad_seed=g["seed"],
                   grad_table_first=float(g["table"][0]), grad_table_last=float(g["table"][-1]))
        log(f"c1 gradient ({name}): forward {t1 - t0:.2f} s, gradient {t3 - t1:.2f} s "
            f"({out['gradient_cost_vs_forward']:.2f}x), {eng.peak_device_bytes / 2**30:.1f} GiB in use")
        fd_s = ((eng.evaluate(table, seed * (1 + fd_rel)) - eng.evaluate(table, seed * (1 - fd_rel)))
                / (2 * fd_rel * seed))
        k = int(np.argmax(np.abs(g["table"])))
        hk = 1e-6
        tp, tm = table.copy(), table.copy()
        tp[k] += hk
        tm[k] -= hk
        fd_t = (eng.evaluate(tp, seed) - eng.evaluate(tm, seed)) / (2 * hk)
        out["fd_check"] = dict(seed_rel_step=fd_rel, fd_seed=fd_s,
                               seed_rel_err=abs(fd_s - g["seed"]) / abs(g["seed"]),
                               table_index=k, table_h=hk, fd_table=fd_t,
                               table_rel_err=abs(fd_t - g["table"][k]) / np.abs(g["table"]).max())
        log(f"c1 gradient ({name}): dJ/dseed {g['seed']:.6e} (fd {fd_s:.6e}), "
            f"dJ/dtable[{k}] {g['table'][k]:.6e} (fd {fd_t:.6e})")
        del eng, g
        torch.cuda.empty_cache()
    finally:
        drv.close()
    return out


def device_memory_used() -> int:
    import torch
    free, total = torch.cuda.mem_get_info()
    return int(total - free)


def diffraction_bench(events: int = 5):
    """Device time of one diffraction event (DiffractionStep: propagator and
    absorption, nine line passes) on the C1 mesh (FP64 and FP32-mixed) and
    the C4 mesh (FP32-mixed), against its algorithmic bytes (14 complex per
    node), after 4 x `events` warm-up events."""
    from puffin_amd import gpu
    from puffin_amd.configs import case_c1, case_c4
    out = {}
    for name, make, precs in (("c1", case_c1, (0, 1)), ("c4", case_c4, (1,))):
        c = make(n=1024)
        mp = c.mp
        for prec in precs:
            g = gpu.Integrator(mp, 1024, prec)
            try:
                g.upload_field(*c.field)
                # warm-up events first: five events timed right after the
                # context's creation took 1.96 ms each on C1 FP64, against
                # 1.75 ms per event inside the step loop
                for _ in range(4 * events):
                    g.diffraction_step(c.steps4diff * c.h)
                g.synchronize()
                g.profile(True)
                g.profile_read(reset=True)
                for _ in range(events):
                    g.diffraction_step(c.steps4diff * c.h)
                ms, cnt = g.profile_read(reset=True)["diffraction"]
            finally:
                g.close()
            nbytes = mp.nx * mp.ny * mp.nzp * 14 * (16 if prec == 0 else 8)
            gbps = nbytes / (ms / cnt * 1e-3) / 1e9
            key = f"{name}_{'fp64' if prec == 0 else 'mixed'}"
            out[key] = dict(mesh=[mp.nx, mp.ny, mp.nz2], ms=ms / cnt, bytes=nbytes, GBps=gbps,
                            frac=gbps / HBM_PEAK_GBS)
            log(f"diffraction {key}: {ms / cnt:.2f} ms ({gbps:.0f} GB/s)")
    return out


def _kinds(prof: dict) -> dict:
    return {k: dict(ms=round(v[0], 3), calls=v[1]) for k, v in prof.items() if v[1]}


def c4_bench(steps: int = 30, adjoint_steps: int = 30, every: int = 4, captures: bool = True):
    """The large FP32-mixed workload (257 x 257 x 4096 temporal mesh, 2^27
    macroparticles, per-period taper table): throughput over `steps` steps
    (one diffraction), device time per kernel kind over as many more, device
    memory, and the gradient of a module of `adjoint_steps` steps (snapshots
    every `every` steps; with step captures, about 27 GB each, the adjoint
    steps replay no forward stage) with its time and memory."""
    from puffin_amd import gpu
    from puffin_amd.adjoint import ModuleGradient, module_program
    from puffin_amd.configs import case_c4
    t = time.perf_counter()
    c = case_c4()
    mp = c.mp
    log(f"c4: workload generated in {time.perf_counter() - t:.1f} s")
    base = device_memory_used()
    drv = gpu.Integrator(mp, c.n, 1)
    out = dict(workload="planar FEL, temporal mesh 257x257x4096, 2^27 macroparticles, FP32-mixed",
               macroparticles=c.n, mesh=[mp.nx, mp.ny, mp.nz2], steps_per_diffraction=c.steps4diff)

    def run(szl, k):
        for istep in range(1, k + 1):
            st = drv.rk4par(szl, c.h)
            if st != 0:
                raise RuntimeError(f"c4 rk4par status {st}")
            szl += c.h
            if istep % c.steps4diff == 0:
                drv.diffraction_step(c.steps4diff * c.h)
        return szl

    try:
        drv.upload_particles(c.particles)
        drv.upload_field(*c.field)
        drv.set_taper_table(c.taper_table, 0.0, c.taper_period)
        out["context_bytes"] = device_memory_used() - base
        szl = run(0.0, 3)
        drv.synchronize()
        t0 = time.perf_counter()
        szl = run(szl, steps)
        drv.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        out.update(ms_per_step=ms, value=c.n / (ms * 1e-3), unit="macroparticle-steps/s", steps=steps)
        drv.profile(True)
        drv.profile_read(reset=True)
        szl = run(szl, steps)
        out["device_ms_by_kind"] = _kinds(drv.profile_read(reset=True))
        drv.profile(False)
        log(f"c4: {ms:.2f} ms/step, context {out['context_bytes'] / 2**30:.1f} GiB")
        # gradient of a module from the initial beam
        drv.upload_particles(c.particles)
        ops = module_program(adjoint_steps, c.h, c.steps4diff, diffraction=True, periodic=False)
        shape = (c.field[0] / c.info["seed_amp"], c.field[1] / c.info["seed_amp"])
        eng = ModuleGradient(drv, ops, shape, every=every, taper_period=c.taper_period,
                             capture=captures)
        table, seed = c.taper_table, c.info["seed_amp"]
        eng.evaluate(table, seed)   # first use of the per-module inputs
        drv.synchronize()
        t0 = time.perf_counter()
        eng.evaluate(table, seed)
        t1 = time.perf_counter()
        J = eng.forward(table, seed)
        t2 = time.perf_counter()
        drv.profile(True)
        drv.profile_read(reset=True)
        g = eng.backward()
        t3 = time.perf_counter()
        prof = _kinds(drv.profile_read(reset=True))
        drv.profile(False)
        out["adjoint"] = dict(steps=adjoint_steps, checkpoint_every=every, J=J,
                              forward_s=t1 - t0, forward_checkpointed_s=t2 - t1, backward_s=t3 - t2,
                              gradient_cost_vs_forward=(t3 - t1) / (t1 - t0),
                              snapshot_bytes=eng.memory_bytes(), captures=eng.capture,
                              total_device_bytes=device_memory_used() - base,
                              backward_device_ms_by_kind=prof,
                              grad_seed=g["seed"], grad_table_first=float(g["table"][0]))
        log(f"c4 adjoint: forward {t1 - t0:.2f} s, checkpointed {t2 - t1:.2f} s, backward {t3 - t2:.2f} s, "
            f"device {out['adjoint']['total_device_bytes'] / 2**30:.1f} GiB")
        del eng
    finally:
        drv.close()
    return out


def c4_full_module(precision: int = 1):
    """The whole 128-period module of the large workload (3840 steps on
    UndSection's schedule, 257 x 257 x 4096 mesh, 2^27 macroparticles):
    status, integrated power every 16 periods and the final outputs."""
    from puffin_amd import gpu
    from puffin_amd.configs import case_c4
    from puffin_amd.undulator import und_section
    c = case_c4()
    mp = c.mp
    out = dict(workload="planar FEL, temporal mesh 257x257x4096, 2^27 macroparticles, whole module",
               steps=c.nsteps, precision="fp64" if precision == 0 else "mixed", powers=[])
    g = gpu.Integrator(mp, c.n, precision)
    try:
        g.upload_particles(c.particles)
        g.upload_field(*c.field)
        g.set_taper_table(c.taper_table, 0.0, c.taper_period)
        t0 = time.perf_counter()

        def write(istep):
            p = integrated_power(g.power(), mp.dz2)
            out["powers"].append([istep, p])
            log(f"c4 module: step {istep}, integrated power {p:.6e}, "
                f"{time.perf_counter() - t0:.1f} s")

        res = und_section(g, c.nsteps, c.h, diffraction=True, steps4diff=c.steps4diff,
                          write=write, write_every=480, int_write_every=480)
        g.synchronize()
        wall = time.perf_counter() - t0
        out.update(status=res.status, steps_done=res.steps_done, wall_s=wall,
                   ms_per_step=wall * 1e3 / max(res.steps_done, 1),
                   diffraction_events=res.n_diffractions)
        if res.status == 0:
            q = g.download_particles()
            out["max_z2_node"] = float(q["z2"].max() / mp.dz2)
            out["integrated_power"] = integrated_power(g.power(), mp.dz2)
    finally:
        g.close()
    log(f"c4 module: status {out['status']}, {out['steps_done']} steps in {out['wall_s']:.1f} s")
    return out


def c4_full_gradient(fd_rel: float = 1e-3, reserve_gb: float = 48.0):
    """d(integrated power)/d(per-period taper table, seed amplitude) of the
    whole 128-period module of the large workload (3840 steps, 2^27
    macroparticles, 257 x 257 x 4096, FP32-mixed) with binomial
    checkpointing over as many device snapshots as fit beside the context and
    the adjoint buffers.  Times a plain forward, the gradient, and two more
    forwards for a central difference on the seed amplitude; the device
    memory in use is read while every snapshot and adjoint buffer is live."""
    import torch
    from puffin_amd import gpu
    from puffin_amd.adjoint import BinomialGradient, checkpoint_schedule, integrated_power as ipow, module_program
    from puffin_amd.configs import case_c4
    from puffin_amd.undulator import und_section
    c = case_c4()
    mp = c.mp
    seed = c.info["seed_amp"]
    shape = (c.field[0] / seed, c.field[1] / seed)
    ops = module_program(c.nsteps, c.h, c.steps4diff, diffraction=True, periodic=False)
    drv = gpu.Integrator(mp, c.n, 1)
    out = dict(workload="planar FEL, temporal mesh 257x257x4096, 2^27 macroparticles, FP32-mixed, "
                        "whole 128-period module (3840 steps)")
    try:
        drv.upload_particles(c.particles)
        drv.upload_field(np.zeros(mp.nmesh), np.zeros(mp.nmesh))

        def forward(s):
            drv.upload_particles(c.particles)
            drv.upload_field(c.field[0] * (s / seed), c.field[1] * (s / seed))
            drv.set_taper_table(c.taper_table, 0.0, c.taper_period)
            drv.synchronize()
            t0 = time.perf_counter()
            res = und_section(drv, c.nsteps, c.h, diffraction=True, steps4diff=c.steps4diff)
            drv.synchronize()
            if res.status != 0:
                raise RuntimeError(f"c4 forward status {res.status}")
            return ipow(mp, drv.power()), time.perf_counter() - t0

        J_fwd, t_fwd = forward(seed)
        log(f"c4 gradient: plain forward {t_fwd:.1f} s, J = {J_fwd:.6e}")
        drv.upload_particles(c.particles)
        blob = drv.state_bytes()
        free, total = torch.cuda.mem_get_info()
        shape_bytes = 4 * mp.nmesh * 8   # unit seed shape and its scaled copy, FP64 device tensors
        slots = int((free - shape_bytes - reserve_gb * 1e9) // blob) - 1
        slots = max(2, min(slots, 40))
        F, _ = checkpoint_schedule(3840, slots - 1)
        log(f"c4 gradient: {slots} snapshot slots of {blob / 1e9:.2f} GB, schedule {F[3840, slots - 1]} "
            f"forward steps for 3840")
        t0 = time.perf_counter()

        def progress(st):
            log(f"c4 gradient: {st['forward_steps']} forward, {st['adjoint_steps']} adjoint steps, "
                f"{time.perf_counter() - t0:.0f} s")

        eng = BinomialGradient(drv, ops, shape, slots=slots, taper_period=c.taper_period, progress=progress,
                               progress_every=480)
        t0 = time.perf_counter()
        g = eng.gradient(c.taper_table, seed)
        t_grad = time.perf_counter() - t0
        log(f"c4 gradient: {t_grad:.1f} s ({t_grad / t_fwd:.2f}x the forward), peak device "
            f"{eng.peak_device_bytes / 2**30:.1f} GiB")
        out.update(forward_s=t_fwd, J_forward=J_fwd, J_gradient_run=g["J"], gradient_s=t_grad,
                   gradient_cost_vs_forward=t_grad / t_fwd, snapshot_slots=slots, snapshot_bytes=blob,
                   schedule_forward_steps=int(F[3840, slots - 1]), stats=g["stats"],
                   peak_device_bytes=eng.peak_device_bytes, device_total_bytes=int(total),
                   grad_seed=g["seed"], grad_table_norm=float(np.linalg.norm(g["table"])),
                   grad_table_first=float(g["table"][0]), grad_table_last=float(g["table"][-1]))
        del eng
        torch.cuda.empty_cache()
        Jp, _ = forward(seed * (1 + fd_rel))
        Jm, _ = forward(seed * (1 - fd_rel))
        fd = (Jp - Jm) / (2 * fd_rel * seed)
        out["fd_check"] = dict(rel_step=fd_rel, fd_seed=fd, seed_rel_err=abs(fd - g["seed"]) / abs(g["seed"]))
        log(f"c4 gradient: dJ/dseed {g['seed']:.6e}, central difference {fd:.6e} "
            f"(rel err {out['fd_check']['seed_rel_err']:.2e})")
    finally:
        drv.close()
    return out


def copy_rate(mib: int = 1024, reps: int = 10) -> float:
    """Achievable streaming rate (TB/s): a device-to-device copy of mib MiB,
    read plus write bytes over the time, timed with events on torch's
    stream (where the copy runs)."""
    import torch
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2.0 * n * 4 / (ms * 1e-3) / 1e12


def calibrate_copy(mib: int):
    """A plain device-to-device copy of known size (reads and writes mib MiB
    each, 5 times), used to scale the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB 5 times")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=1 << 24)
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both")
    ap.add_argument("--cpu-sample", type=int, default=1 << 21)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--case", choices=("c1", "c0", "c3"), default="c1",
                    help="workload (the JSON line is quoted on c1)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--accuracy-periods", type=int, default=-1,
                    help="undulator periods of the FP64-vs-mixed accuracy run "
                         "(-1: the whole module, 0: skip)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient measurement")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the gradient measurement (its JSON object on stdout)")
    ap.add_argument("--no-c4", action="store_true", help="skip the large FP32-mixed workload")
    ap.add_argument("--diffraction-only", action="store_true",
                    help="only the diffraction event on the C1 and C4 meshes (its JSON object on stdout)")
    ap.add_argument("--c4-only", action="store_true",
                    help="only the large FP32-mixed workload (its JSON object on stdout)")
    ap.add_argument("--c4-module", action="store_true",
                    help="only the whole 3840-step module of the large workload (JSON on stdout)")
    ap.add_argument("--c4-accuracy", action="store_true",
                    help="only the FP32-mixed against FP64 error of the whole 3840-step large module (JSON on stdout)")
    ap.add_argument("--c1-gradient", action="store_true",
                    help="only the gradient of the whole C1 module (FP64; JSON on stdout)")
    ap.add_argument("--c4-gradient", action="store_true",
                    help="only the gradient of the whole large module (binomial checkpointing; JSON on stdout)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_bench()), flush=True)
        return
    if args.c4_only:
        print(json.dumps(c4_bench()), flush=True)
        return
    if args.c4_module:
        print(json.dumps(c4_full_module()), flush=True)
        return
    if args.c4_gradient:
        print(json.dumps(c4_full_gradient()), flush=True)
        return
    if args.c1_gradient:
        print(json.dumps(c1_gradient()), flush=True)
        return
    if args.c4_accuracy:
        from puffin_amd.configs import case_c4
        c4 = case_c4()
        out = precision_report(c4, c4.nsteps, progress_every=480)
        out["workload"] = "planar FEL, temporal mesh 257x257x4096, 2^27 macroparticles, whole 128-period module"
        print(json.dumps(out), flush=True)
        return
    if args.diffraction_only:
        print(json.dumps(diffraction_bench()), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path "
              "(the z2 slab decomposition across GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    from puffin_amd.configs import CASES
    t = time.perf_counter()
    case = CASES[args.case](n=args.particles) if args.case == "c1" else CASES[args.case]()
    log(f"workload generated in {time.perf_counter() - t:.1f} s")

    results = {}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        if args.precision not in ("both", name):
            continue
        results[name] = run_gpu(case, prec, args.steps, args.warmup)

    def summary(name, prec):
        r = results[name]
        kernels, dom, n_act, sbytes = roofline(case, prec, r, args.steps)
        k = kernels[dom]
        traffic = (measured_traffic(prec, dom) if case.n == 1 << 24 and args.case == "c1" else None)
        return dict(
            value=case.n / (r["ms_per_step"] * 1e-3),
            ms_per_step=r["ms_per_step"],
            ms_per_step_rk4_only=r["rk4_ms"],
            diffraction_event_ms=r["event_ms"],
            roofline=dict(bound="hbm", kernel=dom, achieved=k["GBps"], peak=HBM_PEAK_GBS, unit="GB/s",
                          frac=k["frac"], traffic=traffic,
                          traffic_source=(f"profiles/hbm_traffic_c1_{name}.md (rocprofv3 --pmc "
                                          "FETCH_SIZE and WRITE_SIZE passes of this workload, "
                                          "scaled by a copy-kernel calibration), not measured in "
                                          "this run") if traffic is not None else None),
            step_roofline=dict(bytes_per_step=sbytes,
                               achieved=sbytes / (r["ms_per_step"] * 1e-3) / 1e9,
                               frac=sbytes / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS),
            kernels={kk: dict(avg_ms=v["avg_ms"], GBps=v["GBps"], frac=v["frac"])
                     for kk, v in kernels.items()},
            active_nodes=n_act,
            integrated_power=integrated_power(r["power"], case.mp.dz2),
        )

    prim_name = "fp64" if "fp64" in results else "mixed"
    prim = summary(prim_name, 0 if prim_name == "fp64" else 1)
    line = dict(
        metric="macroparticle-steps/s",
        value=prim["value"],
        unit="macroparticle-steps/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=prim["ms_per_step"],
        ms_per_step_rk4_only=prim["ms_per_step_rk4_only"],
        diffraction_event_ms=prim["diffraction_event_ms"],
        timing=("ms_per_step = (wall time of `steps` RK4 steps) / steps + (wall time of "
                "ceil(steps / steps_per_diffraction) diffraction events) / events / "
                "steps_per_diffraction; each segment synchronised on the device"),
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f64" if prim_name == "fp64" else "f32-mixed",
        data="synthetic (seeded Philox beam and Gaussian seed field)",
        config=dict(workload=WORKLOADS.get(args.case, case.name) if args.case != "c1" or case.n == 1 << 24
                    else f"{WORKLOADS['c1']} (reduced to {case.n} macroparticles)",
                    mesh=[case.mp.nx, case.mp.ny, case.mp.nz2], macroparticles=case.n,
                    steps_per_diffraction=case.steps4diff, rho=case.mp.rho,
                    precision=prim_name),
        roofline=prim["roofline"],
        step_roofline=prim["step_roofline"],
        kernels=prim["kernels"],
    )
    cr = copy_rate()
    line["copy_rate"] = dict(TBps=cr, what="device copy of 1 GiB, (read + write) bytes / time")
    line["roofline"]["frac_of_copy_rate"] = line["roofline"]["achieved"] / (cr * 1e3)
    if "mixed" in results and prim_name == "fp64":
        mx = summary("mixed", 1)
        perr = abs(mx["integrated_power"] - prim["integrated_power"]) / abs(prim["integrated_power"])
        line["mixed"] = dict(value=mx["value"], ms_per_step=mx["ms_per_step"],
                             ms_per_step_rk4_only=mx["ms_per_step_rk4_only"],
                             diffraction_event_ms=mx["diffraction_event_ms"],
                             speedup_vs_fp64=mx["value"] / prim["value"], roofline=mx["roofline"],
                             step_roofline=mx["step_roofline"], kernels=mx["kernels"],
                             integrated_power_rel_err=perr,
                             power_err_after_steps=args.warmup + args.steps)
    if "mixed" in line and args.case == "c1" and args.accuracy_periods != 0:
        acc_steps = case.nsteps if args.accuracy_periods < 0 else 30 * args.accuracy_periods
        line["mixed"]["accuracy"] = precision_report(case, acc_steps)
    if not args.no_adjoint and args.case == "c1":
        line["adjoint"] = adjoint_bench()
        try:
            line["adjoint_c1"] = c1_gradient(0)
        except Exception as e:   # reported in the line, the C1 measurement stands
            log(f"c1 gradient failed: {e!r}")
            line["adjoint_c1"] = dict(error=repr(e))
    if not args.no_c4 and args.case == "c1":
        try:
            line["c4"] = c4_bench()
        except Exception as e:   # reported in the line, the C1 measurement stands
            log(f"c4 failed: {e!r}")
            line["c4"] = dict(error=repr(e))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(case, args.cpu_sample, args.cpu_steps)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
