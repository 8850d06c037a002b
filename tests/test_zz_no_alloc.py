"""puff_slice_stats_noalloc (device arrays in, device array out, static
scratch) against puff_slice_stats (the context path) on the same
macroparticles, through the C ABI: two calls in a row on one stream, the
second with fewer macroparticles and slices, so the second must clear what
the first left in the shared scratch."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from puffin_amd import gpu
from puffin_amd.params import FP64, PI_REF, ModuleParams

pytestmark = pytest.mark.gpu

SPREAD = ("sdX", "sdY", "sdpX", "sdpY", "eX", "eY", "aveDgamma")


def _beam(n, nz2, dz2, seed):
    rng = np.random.default_rng(seed)
    lz = (nz2 - 1) * dz2
    return dict(x=rng.normal(0.0, 1.0, n), y=rng.normal(0.0, 1.0, n), z2=rng.uniform(0.0, lz, n),
                px=rng.normal(0.0, 0.5, n), py=rng.normal(0.3, 0.5, n),
                gam=1.0 + 1e-3 * rng.standard_normal(n), chi=rng.uniform(0.5, 1.5, n) * 1e-3)


def _noalloc(mp, p, nslices, slice_len, stream):
    import torch
    dev = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to("cuda") for k in
           ("x", "y", "z2", "px", "py", "gam", "chi")}
    out = torch.full((13 * nslices,), float("nan"), dtype=torch.float64, device="cuda")
    params = mp.to_ctypes()
    n = p["x"].size
    print(f"[gpu] slice_stats_noalloc n={n} nslices={nslices}", flush=True)
    st = gpu.lib().puff_slice_stats_noalloc(
        ctypes.byref(params), n, *[ctypes.c_void_p(dev[k].data_ptr()) for k in
                                   ("x", "y", "z2", "px", "py", "gam", "chi")],
        nslices, slice_len, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
    assert st == 0, gpu.lib().puff_last_error()
    stream.synchronize()
    o = out.cpu().numpy().reshape(13, nslices)
    return {k: o[i].copy() for i, k in enumerate(gpu.Integrator.SLICE_KEYS)}


def _context(mp, p, nslices, slice_len):
    with gpu.Integrator(mp, p["x"].size, FP64) as g:
        g.upload_particles(p)
        return g.slice_stats(nslices, slice_len)


@pytest.mark.parametrize("periodic", [0, 1])
def test_noalloc_slice_stats_matches_context_path(periodic):
    import torch
    assert gpu.device_count() >= 1
    rho = 0.01
    dz2 = 4 * PI_REF * rho / 16
    L = 4 * PI_REF * rho
    stream = torch.cuda.current_stream()
    cases = []
    for n, nz2 in ((300_000, 3201), (7_000, 401)):
        mp = ModuleParams(nx=33, ny=33, nz2=nz2, dx=0.4, dy=0.4, dz2=dz2, rho=rho, eta=1e-4,
                          kappa=0.5, gamma_r=100.0, aw=1.0, periodic=periodic)
        nslices = int(math.ceil((nz2 - 1) * dz2 / L))
        cases.append((mp, _beam(n, nz2, dz2, seed=n + periodic), nslices))
    # two calls in a row, the second smaller, then the context path
    got = [_noalloc(mp, p, nslices, L, stream) for mp, p, nslices in cases]
    ref = [_context(mp, p, nslices, L) for mp, p, nslices in cases]
    for (mp, p, nslices), g, r in zip(cases, got, ref):
        assert g["sdata"].sum() > 0
        for k, v in r.items():
            assert np.isfinite(g[k]).all(), k
            tol = 1e-8 if k in SPREAD else 1e-12
            assert np.abs(g[k] - v).max() <= tol * max(np.abs(v).max(), 1e-30), (p["x"].size, k)


def test_noalloc_refuses_oversized_request():
    import torch
    mp = ModuleParams(nx=3, ny=3, nz2=9, dx=1.0, dy=1.0, dz2=0.1, rho=0.01, eta=1e-4, kappa=0.5,
                      gamma_r=100.0, aw=1.0)
    out = torch.zeros(13, dtype=torch.float64, device="cuda")
    params = mp.to_ctypes()
    lim = gpu.lib().puff_slice_stats_noalloc
    null = ctypes.c_void_p(0)
    st = lim(ctypes.byref(params), 0, null, null, null, null, null, null, null,
             (1 << 12) + 1, 0.1, ctypes.c_void_p(out.data_ptr()), null)
    assert st == -1
    assert b"nslices" in gpu.lib().puff_last_error()
