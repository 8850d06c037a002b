"""The HIP adjoint (puff_rk4par_adjoint, puff_diffraction_adjoint, state
snapshots, ModuleGradient, IntegratedPower) against the oracle's adjoint and
against central finite differences of the HIP forward."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from adjoint_case import OracleModule, small_module
from puffin_amd import FP64, MIXED
from test_adjoint_oracle import CONFIGS, _case

KEYS = orc.ADJ_KEYS
PKEYS = ("x", "y", "z2", "px", "py", "gam")

pytestmark = pytest.mark.gpu


def _gpu():
    from puffin_amd import gpu
    if gpu.device_count() < 1:
        pytest.skip("no HIP device")
    return gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(CONFIGS))
@pytest.mark.parametrize("sort_first", [False, True])
def test_rk4_step_adjoint_matches_oracle(name, sort_first):
    print(f"[gpu] rk4 adjoint {name} sort_first={sort_first}", flush=True)
    gpu = _gpu()
    kw = CONFIGS[name]
    mp, p, ar, ai, h = _case(**kw)
    rng = np.random.default_rng(11)
    taper = (1.0, 1.0, -0.02)
    sz = 10 * h if kw.get("und_ends") else 3 * h
    lam = {k: rng.standard_normal(p["x"].size) for k in KEYS}
    lar = rng.standard_normal(mp.nmesh)
    lai = rng.standard_normal(mp.nmesh)
    L = {k: lam[k].copy() for k in KEYS}
    LR, LI = lar.copy(), lai.copy()
    st, dn = orc.rk4par_adjoint(mp, taper, p, ar, ai, L, LR, LI, sz, h)
    assert st == 0
    with gpu.Integrator(mp, p["x"].size, FP64) as drv:
        drv.upload_particles(p)
        drv.upload_field(ar, ai)
        drv.set_taper(*taper)
        drv.adjoint_enable()
        drv.adjoint_upload({k: lam[k] for k in KEYS}, lar, lai)
        if sort_first:
            drv.sort_now()          # the cotangent follows the new slot order
        st, gdn, slot = drv.rk4par_adjoint(sz, h)
        assert st == 0
        assert list(slot) == [-1] * 4
        g, gr, gi = drv.adjoint_download()
        # the state is unchanged by the adjoint step
        q = drv.download_particles()
        for k in PKEYS:
            assert np.array_equal(q[k], p[k]), k
    for k in KEYS:
        if mp.one_d and k in ("x", "y"):
            continue
        assert _rel(g[k], L[k]) <= 1e-9, (k, _rel(g[k], L[k]))
    assert _rel(gr, LR) <= 1e-9
    assert _rel(gi, LI) <= 1e-9
    assert np.abs(gdn - dn).max() <= 1e-9 * np.abs(dn).max()


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("filt", [0, 1])
def test_diffraction_adjoint_matches_oracle(periodic, filt):
    print(f"[gpu] diffraction adjoint periodic={periodic} filter={filt}", flush=True)
    gpu = _gpu()
    mp, p, _, _, h = _case(periodic=periodic)
    mp.filter = filt
    rng = np.random.default_rng(5)
    y = rng.standard_normal(mp.nmesh) + 1j * rng.standard_normal(mp.nmesh)
    lr, li = y.real.copy(), y.imag.copy()
    orc.diffraction_adjoint(mp, lr, li, 7 * h)
    with gpu.Integrator(mp, p["x"].size, FP64) as drv:
        drv.upload_particles(p)
        drv.adjoint_enable()
        drv.adjoint_upload(None, y.real.copy(), y.imag.copy())
        drv.diffraction_adjoint(7 * h)
        _, gr, gi = drv.adjoint_download()
    assert _rel(gr, lr) <= 1e-12
    assert _rel(gi, li) <= 1e-12


def test_diffraction_adjoint_four_step_lines_matches_oracle():
    """The adjoint event on 4096-point z2 lines (four-step z2 passes)."""
    print("[gpu] diffraction adjoint, 4096-point z2 lines", flush=True)
    gpu = _gpu()
    from puffin_amd.params import ModuleParams
    mp = ModuleParams(nx=9, ny=17, nz2=4096, dx=0.3, dy=0.3, dz2=0.002, rho=0.01, eta=1e-4,
                      kappa=0.5, gamma_r=100.0, aw=1.0, nbz2=65, filt_cutoff=15.0, beta=1.0)
    rng = np.random.default_rng(9)
    y = rng.standard_normal(mp.nmesh) + 1j * rng.standard_normal(mp.nmesh)
    lr, li = y.real.copy(), y.imag.copy()
    orc.diffraction_adjoint(mp, lr, li, 0.0837)
    p = dict(x=np.zeros(1), y=np.zeros(1), z2=np.full(1, 0.1), px=np.zeros(1), py=np.zeros(1),
             gam=np.ones(1), chi=np.zeros(1))
    with gpu.Integrator(mp, 1, FP64) as drv:
        drv.upload_particles(p)
        drv.adjoint_enable()
        drv.adjoint_upload(None, y.real.copy(), y.imag.copy())
        drv.diffraction_adjoint(0.0837)
        _, gr, gi = drv.adjoint_download()
    assert _rel(gr, lr) <= 1e-12
    assert _rel(gi, li) <= 1e-12


def test_state_snapshot_round_trip():
    print("[gpu] state snapshot round trip", flush=True)
    gpu = _gpu()
    import torch
    mp, p, ar, ai, h = _case(periodic=1)
    with gpu.Integrator(mp, p["x"].size, FP64) as drv:
        drv.upload_particles(p)
        drv.upload_field(ar, ai)
        drv.set_sort_interval(2)
        assert drv.rk4par(0.0, h) == 0
        blob = torch.empty(drv.state_bytes(), dtype=torch.uint8, device="cuda")
        drv.state_save(blob.data_ptr())
        ref = drv.download_particles()
        fr, fi = drv.download_field()
        for s in range(5):                 # sorts in between
            assert drv.rk4par((1 + s) * h, h) == 0
        drv.state_load(blob.data_ptr())
        q = drv.download_particles()
        gr, gi = drv.download_field()
        for k in PKEYS:
            assert np.array_equal(q[k], ref[k]), k
        assert np.array_equal(gr, fr) and np.array_equal(gi, fi)
        # replaying the same steps from the snapshot repeats them
        assert drv.rk4par(h, h) == 0
        q1 = drv.download_particles()
        drv.state_load(blob.data_ptr())
        assert drv.rk4par(h, h) == 0
        q2 = drv.download_particles()
        for k in PKEYS:
            assert np.abs(q1[k] - q2[k]).max() <= 1e-12 * np.abs(q1[k]).max() + 1e-300, k


def _engine(gpu, c, precision, every, capture=None, keep_all=False):
    from puffin_amd.adjoint import ModuleGradient
    mp = c["mp"]
    drv = gpu.Integrator(mp, c["parts"]["x"].size, precision)
    drv.upload_particles(c["parts"])
    drv.upload_field(np.zeros(mp.nmesh), np.zeros(mp.nmesh))
    eng = ModuleGradient(drv, c["ops"], c["shape"], every=every, taper_period=c["period"],
                         capture=capture, keep_all=keep_all)
    assert eng.keep_all == bool(keep_all)
    return drv, eng


@pytest.mark.parametrize("capture,keep_all", [(False, False), (True, False), (None, True)])
@pytest.mark.parametrize("periodic", [1, 0])
def test_module_gradient_matches_oracle(periodic, capture, keep_all):
    """Checkpoints every 10 steps (segments end between diffractions and
    wraps); the default sort interval (30 steps) sorts inside segments; the
    adjoint steps replay their forward stages or use the step captures; or
    the forward keeps every step's capture and nothing is replayed (the sorts
    then fall between captured steps across the whole module)."""
    print(f"[gpu] module gradient periodic={periodic} capture={capture} keep_all={keep_all}", flush=True)
    gpu = _gpu()
    c = small_module(periodic=periodic)
    J0, gt0, gs0, _ = OracleModule(c).gradient(c["table"], c["seed"])
    drv, eng = _engine(gpu, c, FP64, every=10, capture=capture, keep_all=keep_all)
    try:
        J = eng.forward(c["table"], c["seed"])
        g = eng.backward()
    finally:
        drv.close()
    assert abs(J - J0) <= 1e-10 * abs(J0)
    assert np.abs(g["table"] - gt0).max() <= 1e-7 * np.abs(gt0).max()
    assert abs(g["seed"] - gs0) <= 1e-7 * abs(gs0)


def test_module_gradient_against_gpu_finite_differences():
    print("[gpu] module gradient vs finite differences", flush=True)
    gpu = _gpu()
    c = small_module(periodic=1, n=4000)
    drv, eng = _engine(gpu, c, FP64, every=12)
    try:
        eng.forward(c["table"], c["seed"])
        g = eng.backward()
        eps = 1e-6
        s = c["seed"]
        fd = (eng.evaluate(c["table"], s * (1 + eps)) - eng.evaluate(c["table"], s * (1 - eps))) / (2 * eps * s)
        assert abs(fd - g["seed"]) <= 1e-6 * abs(g["seed"]), (fd, g["seed"])
        for k in range(c["table"].size):
            tp = c["table"].copy()
            tm = c["table"].copy()
            tp[k] += eps
            tm[k] -= eps
            fdk = (eng.evaluate(tp, s) - eng.evaluate(tm, s)) / (2 * eps)
            assert abs(fdk - g["table"][k]) <= 1e-3 * np.abs(g["table"]).max(), (k, fdk, g["table"][k])
    finally:
        drv.close()


@pytest.mark.parametrize("keep_all", [False, True])
def test_mixed_gradient_tracks_fp64(keep_all):
    print(f"[gpu] mixed-precision gradient keep_all={keep_all}", flush=True)
    gpu = _gpu()
    c = small_module(periodic=1, n=4000)
    out = {}
    for prec in (FP64, MIXED):
        drv, eng = _engine(gpu, c, prec, every=12, capture=(prec == MIXED) if not keep_all else None,
                           keep_all=keep_all)
        try:
            J = eng.forward(c["table"], c["seed"])
            out[prec] = (J, eng.backward())
        finally:
            drv.close()
    (J64, g64), (J32, g32) = out[FP64], out[MIXED]
    assert abs(J32 - J64) <= 1e-4 * abs(J64)
    assert np.abs(g32["table"] - g64["table"]).max() <= 1e-3 * np.abs(g64["table"]).max()
    assert abs(g32["seed"] - g64["seed"]) <= 1e-3 * abs(g64["seed"])


def test_integrated_power_autograd_function():
    print("[gpu] IntegratedPower autograd", flush=True)
    gpu = _gpu()
    import torch
    from puffin_amd.adjoint import IntegratedPower
    c = small_module(periodic=1)
    drv, eng = _engine(gpu, c, FP64, every=12)
    try:
        table = torch.tensor(c["table"], dtype=torch.float64, requires_grad=True)
        seed = torch.tensor(c["seed"], dtype=torch.float64, requires_grad=True)
        J = IntegratedPower.apply(table, seed, eng)
        (3.0 * J).backward()
        ref = eng.forward(c["table"], c["seed"])
        g = eng.backward(3.0)
    finally:
        drv.close()
    assert abs(float(J.detach()) - ref) <= 1e-12 * abs(ref)
    assert np.allclose(table.grad.numpy(), g["table"], rtol=1e-9, atol=0)
    assert abs(float(seed.grad) - g["seed"]) <= 1e-9 * abs(g["seed"])


@pytest.mark.parametrize("slots", [3, 6])
@pytest.mark.parametrize("periodic", [1, 0])
def test_binomial_checkpointing_gradient_matches_oracle(periodic, slots):
    """BinomialGradient (few snapshot slots, adjoint steps that recompute their
    forward stages) against the oracle's module gradient."""
    print(f"[gpu] binomial gradient periodic={periodic} slots={slots}", flush=True)
    gpu = _gpu()
    from puffin_amd.adjoint import BinomialGradient
    c = small_module(periodic=periodic)
    J0, gt0, gs0, _ = OracleModule(c).gradient(c["table"], c["seed"])
    mp = c["mp"]
    drv = gpu.Integrator(mp, c["parts"]["x"].size, FP64)
    try:
        drv.upload_particles(c["parts"])
        drv.upload_field(np.zeros(mp.nmesh), np.zeros(mp.nmesh))
        eng = BinomialGradient(drv, c["ops"], c["shape"], slots=slots, taper_period=c["period"])
        g = eng.gradient(c["table"], c["seed"])
    finally:
        drv.close()
    assert abs(g["J"] - J0) <= 1e-10 * abs(J0)
    assert np.abs(g["table"] - gt0).max() <= 1e-7 * np.abs(gt0).max()
    assert abs(g["seed"] - gs0) <= 1e-7 * abs(gs0)
    assert g["stats"]["adjoint_steps"] == sum(1 for op in c["ops"] if op.kind == "rk4")
