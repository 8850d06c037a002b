"""The PyTorch-ROCm side of the C ABI: particles and field handed over as
torch device tensors (PUFF_MEM_DEVICE), the context running on torch's own
stream (puff_set_stream), results downloaded into torch tensors.  With the
deterministic deposit (FP64) that is bit for bit what the host-pointer path
gives; the mixed form's atomic FP32 flushes land in any order, so its two
runs agree to FP32 rounding."""
from __future__ import annotations

import math

import numpy as np
import pytest

from puffin_amd import gpu
from puffin_amd.configs import gaussian_seed, match_in
from puffin_amd.params import FP64, MIXED, PI_REF, ModuleParams, scaling, step_size

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z2", "px", "py", "gam")


def _case(periodic=0, n=5000):
    rho = 0.01
    eta, kappa, fx, fy = scaling(rho, 1.0, 100.0, 1 / math.sqrt(2), 1 / math.sqrt(2))
    h = step_size(rho, 30)
    mp = ModuleParams(nx=17, ny=17, nz2=81 if periodic else 80, dx=0.4, dy=0.4, dz2=4 * PI_REF * rho / 16,
                      rho=rho, eta=eta, kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy,
                      periodic=periodic, szfs=0.0, szfe=60 * h, nbx=4, nby=4, nbz2=9, filt_cutoff=10.0)
    rng = np.random.default_rng(5 + periodic)
    lz = (mp.nz2 - 1) * mp.dz2
    p = dict(x=np.clip(rng.standard_normal(n), -2.5, 2.5), y=np.clip(rng.standard_normal(n), -2.5, 2.5),
             z2=rng.uniform(0.05, 0.6 * lz, n), px=np.zeros(n), py=np.zeros(n),
             gam=1.0 + 1e-3 * rng.standard_normal(n), chi=np.full(n, 2 * math.pi * lz / n))
    match_in(mp, 1.0, p)
    ar, ai = gaussian_seed(mp, 1e-2, 1e-2, w0=1.5)
    return mp, p, ar, ai, h


def _run(g, h, steps=12):
    szl = 0.0
    for k in range(steps):
        assert g.rk4par(szl, h) == 0
        szl += h
        if (k + 1) % 6 == 0:
            g.diffraction_step(6 * h)


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("periodic", [0, 1])
def test_torch_tensors_on_torch_stream_match_host_path(precision, periodic):
    import torch
    print(f"[gpu] torch boundary precision={precision} periodic={periodic}", flush=True)
    mp, p, ar, ai, h = _case(periodic)
    det = precision == FP64
    with gpu.Integrator(mp, p["x"].size, precision) as g:
        if det:
            g.set_deterministic(True)
        g.upload_particles(p)
        g.upload_field(ar, ai)
        _run(g, h)
        ref = g.download_particles()
        rr, ri = g.download_field()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        tp = {k: torch.from_numpy(p[k].copy()).to("cuda") for k in KEYS + ("chi",)}
        tar = torch.from_numpy(ar.copy()).to("cuda")
        tai = torch.from_numpy(ai.copy()).to("cuda")
        with gpu.Integrator(mp, p["x"].size, precision) as g:
            if det:
                g.set_deterministic(True)
            g.use_torch_stream()
            g.upload_particles_torch(tp)
            g.upload_field_torch(tar, tai)
            _run(g, h)
            out = g.download_particles_torch()
            dr, di = g.download_field_torch()
            # the downloads are complete on return: torch work on them needs no wait
            total = float(sum(out[k].sum() for k in KEYS))
    assert math.isfinite(total)
    for k in KEYS:
        assert isinstance(out[k], torch.Tensor) and out[k].is_cuda
        got = out[k].cpu().numpy()
        if det:
            assert np.array_equal(got, ref[k]), k
        else:
            assert np.abs(got - ref[k]).max() <= 1e-5 * max(np.abs(ref[k]).max(), 1.0), k
    if det:
        assert np.array_equal(dr.cpu().numpy(), rr) and np.array_equal(di.cpu().numpy(), ri)
    else:
        scale = max(np.abs(rr).max(), np.abs(ri).max())
        assert np.abs(dr.cpu().numpy() - rr).max() <= 1e-5 * scale
        assert np.abs(di.cpu().numpy() - ri).max() <= 1e-5 * scale


def test_torch_upload_rejects_wrong_tensors():
    import torch
    mp, p, ar, ai, h = _case()
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        with pytest.raises(ValueError):
            g.upload_field_torch(torch.zeros(mp.nmesh, dtype=torch.float32, device="cuda"),
                                 torch.zeros(mp.nmesh, dtype=torch.float32, device="cuda"))
        with pytest.raises(ValueError):
            g.upload_field_torch(torch.zeros(mp.nmesh, dtype=torch.float64),
                                 torch.zeros(mp.nmesh, dtype=torch.float64))
