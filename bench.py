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


# Algorithmic bytes: per macroparticle per RK4 stage and per
# active node per stage; the field combination and the diffraction per node.
def stage_bytes(precision: int):
    if precision == 0:   # FP64: six-vector 48 B, chi 8 B, complex 16 B
        part = [48 + 8 + 48, 48 + 48 + 8 + 48, 48 + 48 + 8 + 48, 4 * 48 + 8 + 48]
        node = [2 * 16, 3 * 16, 3 * 16, 3 * 16]
        comb, diff = 6 * 16, 14 * 16
    else:                # mixed: six-vector 32 B (z2, Gamma FP64), chi 4 B, complex 8 B
        part = [32 + 4 + 32, 32 + 32 + 4 + 32, 32 + 32 + 4 + 32, 4 * 32 + 4 + 32]
        node = [2 * 8, 3 * 8, 3 * 8, 3 * 8]
        comb, diff = 6 * 8, 14 * 8
    return part, node, comb, diff


def run_gpu(case, precision: int, steps: int, warmup: int):
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
        istep = 0

        def step():
            nonlocal szl, istep
            st = g.rk4par(szl, h)
            if st != 0:
                raise RuntimeError(f"rk4par status {st} at step {istep}")
            szl += h
            istep += 1
            if istep % s4d == 0:
                g.diffraction_step(s4d * h)

        for _ in range(warmup):
            step()
        g.synchronize()
        g.profile(True)
        g.profile_read(reset=True)
        log(f"{name}: timing {steps} steps")
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        g.synchronize()
        t1 = time.perf_counter()
        prof = g.profile_read(reset=True)
        g.profile(False)
        box = g.active_box()
        power = g.power()
        ms = (t1 - t0) * 1e3 / steps
        log(f"{name}: {ms:.3f} ms/step")
        return dict(ms_per_step=ms, prof=prof, box=box, power=power)
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


def cpu_baseline(case, nsample: int, steps: int):
    """The CPU oracle (FP64 C restatement, OpenMP, atomic deposit as the
    reference's !$OMP ATOMIC) on a sample of the same workload."""
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
    log(f"cpu baseline: {idx.size} macroparticles, {threads} threads")
    szl = 0.0
    st = o.rk4par(szl, case.h)   # warm-up
    szl += case.h
    t0 = time.perf_counter()
    for _ in range(steps):
        st = o.rk4par(szl, case.h)
        szl += case.h
    t1 = time.perf_counter()
    orc.set_deposit_atomic(False)
    assert st == 0, st
    val = idx.size * steps / (t1 - t0)
    log(f"cpu baseline: {val:.4g} macroparticle-steps/s")
    return dict(value=val, unit="macroparticle-steps/s", cores=threads, kind="port",
                sample=(f"{idx.size} of the {case.n} macroparticles on the full "
                        f"{case.mp.nx}x{case.mp.ny}x{case.mp.nz2} mesh, {steps} RK4 steps "
                        f"(no diffraction), FP64 C oracle with OpenMP atomics"))


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
    ap.add_argument("--cpu-sample", type=int, default=1 << 24)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--case", choices=("c1", "c0", "c3"), default="c1",
                    help="workload (the JSON line is quoted on c1)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(z2 slab decomposition is described in DESIGN.md, not built)", file=sys.stderr)
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
        return dict(
            value=case.n / (r["ms_per_step"] * 1e-3),
            ms_per_step=r["ms_per_step"],
            roofline=dict(bound="hbm", kernel=dom, achieved=k["GBps"], peak=HBM_PEAK_GBS, unit="GB/s",
                          frac=k["frac"], traffic=measured_traffic(prec, dom) if case.n == 1 << 24 and
                          args.case == "c1" else None),
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
    if "mixed" in results and prim_name == "fp64":
        mx = summary("mixed", 1)
        perr = abs(mx["integrated_power"] - prim["integrated_power"]) / abs(prim["integrated_power"])
        line["mixed"] = dict(value=mx["value"], ms_per_step=mx["ms_per_step"],
                             speedup_vs_fp64=mx["value"] / prim["value"], roofline=mx["roofline"],
                             step_roofline=mx["step_roofline"], kernels=mx["kernels"],
                             integrated_power_rel_err=perr,
                             power_err_after_steps=args.warmup + args.steps)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(case, args.cpu_sample, args.cpu_steps)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
