"""The Fortran host side (puffin_amd/fortran): drop-in rk4par / diffractIM
through iso_c_binding, driven by a Fortran restatement of UndSection's
schedule, against the CPU oracle."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from puffin_amd.configs import gaussian_seed, match_in
from puffin_amd.params import ModuleParams, PI_REF, scaling, step_size

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "puffin_amd", "fortran", "und_section")
KEYS = ("x", "y", "z2", "px", "py", "gam")


def _case(periodic=1, n=2000, matched=True):
    rho = 0.01
    eta, kappa, fx, fy = scaling(rho, 1.0, 100.0, 1 / np.sqrt(2), 1 / np.sqrt(2))
    h = step_size(rho, 30)
    nz2 = 65 if periodic else 80
    mp = ModuleParams(nx=17, ny=17, nz2=nz2, dx=0.4, dy=0.4, dz2=4 * PI_REF * rho / 16, rho=rho,
                      eta=eta, kappa=kappa, gamma_r=100.0, aw=1.0, fx=fx, fy=fy, periodic=periodic,
                      szfs=0.0, szfe=60 * h, nbx=4, nby=4, nbz2=9, filt_cutoff=10.0)
    rng = np.random.default_rng(21)
    lz = (nz2 - 1) * mp.dz2
    p = dict(x=np.clip(rng.standard_normal(n), -2.5, 2.5), y=np.clip(rng.standard_normal(n), -2.5, 2.5),
             z2=rng.uniform(0.05, 0.6 * lz, n), px=np.zeros(n), py=np.zeros(n),
             gam=1.0 + 1e-3 * rng.standard_normal(n), chi=np.full(n, 2 * np.pi * lz / n))
    if matched:
        match_in(mp, 1.0, p)
    ar, ai = gaussian_seed(mp, 1e-2, 1e-2, w0=1.5)
    return mp, p, (ar, ai), h


def write_input(path, mp, p, field, h, nsteps, s4d, redist, write_every, taper=(1.0, 1.0, 0.0),
                elements=None):
    prm = mp.to_ctypes()
    with open(path, "wb") as f:
        f.write(ctypes.string_at(ctypes.addressof(prm), ctypes.sizeof(prm)))
        f.write(np.array([p["x"].size, nsteps, s4d, redist, 1, mp.periodic, write_every,
                          int(elements is not None)], dtype=np.int64).tobytes())
        f.write(np.array([h, taper[0], taper[1], taper[2]], dtype=np.float64).tobytes())
        f.write(np.array(elements if elements is not None else [0.0] * 7, dtype=np.float64).tobytes())
        for k in KEYS + ("chi",):
            f.write(np.ascontiguousarray(p[k], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(field[0], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(field[1], dtype=np.float64).tobytes())


def read_output(path, n, mp):
    raw = open(path, "rb").read()
    nw = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    off = 8
    arr = np.frombuffer(raw[off:], dtype=np.float64)
    out = {}
    for i, k in enumerate(KEYS):
        out[k] = arr[i * n:(i + 1) * n].copy()
    o = 6 * n
    ar = arr[o:o + mp.nmesh].copy()
    ai = arr[o + mp.nmesh:o + 2 * mp.nmesh].copy()
    o += 2 * mp.nmesh
    powers = arr[o:o + nw * mp.nz2].reshape(nw, mp.nz2).copy()
    status = int(np.frombuffer(raw[-8:], dtype=np.int64)[0])
    return out, (ar, ai), powers, status


def test_fortran_driver_built():
    assert os.path.exists(EXE), "build the Fortran driver (make -C puffin_amd)"


@pytest.mark.gpu
@pytest.mark.parametrize("periodic", [0, 1])
def test_fortran_und_section_matches_oracle(tmp_path, periodic):
    print(f"[gpu] fortran und_section periodic={periodic}", flush=True)
    mp, p, field, h = _case(periodic)
    nsteps, s4d, redist, we = 40, 15, 20, 15
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, mp, p, field, h, nsteps, s4d, redist, we)
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out, (ar, ai), powers, status = read_output(fout, p["x"].size, mp)
    assert status == 0
    o = orc.Oracle(mp, p, field)
    wm = np.zeros(nsteps, dtype=np.uint8)
    wm[[14, 29, 39]] = 1
    st, opow, _, _ = o.und_section(nsteps, h, True, s4d, redist, write_mask=wm)
    assert st == 0
    for k in KEYS:
        assert np.abs(out[k] - o.parts[k]).max() <= 1e-10 * np.abs(o.parts[k]).max(), k
    scale = np.abs(o.ar + 1j * o.ai).max()
    assert np.abs(ar - o.ar).max() <= 1e-10 * scale
    assert np.abs(ai - o.ai).max() <= 1e-10 * scale
    assert powers.shape == opow.shape
    assert np.abs(powers - opow).max() <= 1e-10 * np.abs(opow).max()


@pytest.mark.gpu
def test_fortran_lattice_elements_match_oracle(tmp_path):
    """matchIn before the arrays reach the device, the module, then matchOut,
    correctTrans, drift (with its diffraction), quad, chicane and energy
    modulation through the Fortran drop-ins."""
    print("[gpu] fortran lattice elements", flush=True)
    mp, p, field, h = _case(1, matched=False)
    nsteps, s4d, redist = 30, 15, 20
    lenz2 = mp.dz2 * (mp.nz2 - 1)
    el = [0.2, 2.7, -3.1, 2.5e-3, 0.3 * lenz2, 7.0, 2e-3]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, mp, p, field, h, nsteps, s4d, redist, 0, elements=el)
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out, (ar, ai), _, status = read_output(fout, p["x"].size, mp)
    assert status == 0
    raw = np.frombuffer(open(fout, "rb").read(), dtype=np.float64)
    nsl = int(np.ceil(mp.dz2 * (mp.nz2 - 1) / (4 * PI_REF * mp.rho)))
    tail = raw[-1 - (6 * nsl + nsl + 1):-1]          # before the status word
    assert int(raw[-2 - (6 * nsl + nsl + 1)].view(np.int64)) == nsl
    f_sd, f_ag, f_adg, f_apx, f_sdpx, f_b1 = tail[:6 * nsl].reshape(6, nsl)
    f_cur = tail[6 * nsl:]
    q = {k: v.copy() for k, v in p.items()}
    orc.match(mp, 1.0, q, 0.0, +1)
    o = orc.Oracle(mp, q, field)
    st, _, _, szl = o.und_section(nsteps, h, True, s4d, redist)
    assert st == 0
    orc.match(mp, float(o.taper[0]), o.parts, szl, -1)
    orc.correct_trans(o.parts)
    orc.drift(mp, o.parts, el[0])
    o.diffraction_step(el[0])
    orc.quad(mp, o.parts, el[1], el[2])
    orc.chicane(mp, o.parts, el[3], el[4])
    orc.modulation(o.parts, el[5], el[6])
    for k in KEYS:
        assert np.abs(out[k] - o.parts[k]).max() <= 1e-10 * np.abs(o.parts[k]).max(), k
    scale = np.abs(o.ar + 1j * o.ai).max()
    assert np.abs(ar - o.ar).max() <= 1e-10 * scale
    assert np.abs(ai - o.ai).max() <= 1e-10 * scale
    # getSliceTwiss and getCurr drop-ins on the final state
    L = 4 * PI_REF * mp.rho
    st, ss = orc.slice_stats(mp, o.parts, nsl, L)
    assert st == 0
    for got, k in ((f_sd, "sdata"), (f_ag, "aveGamma"), (f_apx, "avePX")):
        assert np.abs(got - ss[k]).max() <= 1e-10 * np.abs(ss[k]).max(), k
    for got, k in ((f_adg, "aveDgamma"), (f_sdpx, "sdpX")):
        assert np.abs(got - ss[k]).max() <= 1e-7 * np.abs(ss[k]).max(), k
    assert np.abs(f_b1 - o.bunching(1, nsl, L)).max() <= 1e-10 * np.abs(f_b1).max()
    st, cur = orc.current(mp, o.parts, nsl + 1, L)
    assert st == 0
    q_e, c = float(np.float32(1.60217653e-19)), float(np.float32(2.99792458e8))
    np.testing.assert_allclose(f_cur, cur * q_e / L * c, rtol=1e-10)
