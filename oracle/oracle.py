"""ctypes front end of the CPU FP64 oracle (oracle/puffin_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product (puffin_amd/).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from puffin_amd.params import ModuleParams

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        build()
        l = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        l.orc_get_p2.restype = ctypes.c_double
        l.orc_get_p2.argtypes = [ctypes.c_double] * 6
        l.orc_get_gamma.restype = ctypes.c_double
        l.orc_get_gamma.argtypes = [ctypes.c_double] * 6
        l.orc_bfields.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _dp, _dp, _dp]
        l.orc_interp_one.restype = ctypes.c_int
        l.orc_interp_one.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i64p, _dp]
        l.orc_rk4par.restype = ctypes.c_int
        l.orc_rk4par.argtypes = [P, _dp, ctypes.c_int64] + [_dp] * 9 + [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        l.orc_interp_deposit.restype = ctypes.c_int
        l.orc_interp_deposit.argtypes = [P, ctypes.c_int64] + [_dp] * 13
        l.orc_wrap_z2.argtypes = [P, ctypes.c_int64, _dp]
        l.orc_fft1d.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        l.orc_fft3d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        l.orc_kvalues.argtypes = [P, _dp, _dp, _dp]
        l.orc_mask_xy.argtypes = [P, _dp]
        l.orc_mask_z2.argtypes = [P, _dp]
        l.orc_diffraction_step.restype = ctypes.c_int
        l.orc_diffraction_step.argtypes = [P, _dp, _dp, ctypes.c_double, ctypes.c_int]
        l.orc_power.argtypes = [P, _dp, _dp, ctypes.c_int, _dp]
        l.orc_bunching.restype = ctypes.c_int
        l.orc_bunching.argtypes = [P, ctypes.c_int64, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, _dp]
        l.orc_slice_stats.restype = ctypes.c_int
        l.orc_slice_stats.argtypes = [P, ctypes.c_int64] + [_dp] * 7 + [ctypes.c_int, ctypes.c_double, _dp]
        l.orc_current.restype = ctypes.c_int
        l.orc_current.argtypes = [P, ctypes.c_int64, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp]
        l.orc_match.argtypes = [P, ctypes.c_double, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp,
                                ctypes.c_double, ctypes.c_int]
        l.orc_set_deposit_atomic.argtypes = [ctypes.c_int]
        l.orc_set_taper_table.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        l.orc_taper_slot.restype = ctypes.c_int
        l.orc_taper_slot.argtypes = [ctypes.c_double]
        l.orc_bfields_grad.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _dp]
        l.orc_rk4par_adjoint.restype = ctypes.c_int
        l.orc_rk4par_adjoint.argtypes = ([P, _dp, ctypes.c_int64] + [_dp] * 9 + [_dp] * 8 +
                                         [ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int])
        l.orc_diffraction_adjoint.restype = ctypes.c_int
        l.orc_diffraction_adjoint.argtypes = [P, _dp, _dp, ctypes.c_double, ctypes.c_int]
        l.orc_correct_trans.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, _dp]
        l.orc_drift.argtypes = [P, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double]
        l.orc_chicane.argtypes = [P, ctypes.c_int64, _dp, _dp, ctypes.c_double, ctypes.c_double]
        l.orc_quad.argtypes = [P, ctypes.c_int64, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double]
        l.orc_modulation.argtypes = [ctypes.c_int64, _dp, _dp, ctypes.c_double, ctypes.c_double]
        l.orc_und_section.restype = ctypes.c_int
        l.orc_und_section.argtypes = ([P, _dp, ctypes.c_int64] + [_dp] * 9 +
                                      [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, _i64p, ctypes.POINTER(ctypes.c_uint8), _dp, _dp,
                                       ctypes.c_int])
        _lib = l
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int64:
        return a.ctypes.data_as(_i64p)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    raise TypeError(a.dtype)


class Oracle:
    """State holder mirroring the reference's module globals for one module."""

    def __init__(self, mp: ModuleParams, particles: dict, field: tuple,
                 taper=(1.0, 1.0, 0.0), nthreads: int = 1):
        self.mp = mp
        self._p = mp.to_ctypes()
        self.pp = ctypes.byref(self._p)
        self.parts = {k: np.ascontiguousarray(particles[k], dtype=np.float64).copy()
                      for k in ("x", "y", "z2", "px", "py", "gam", "chi")}
        self.ar = np.ascontiguousarray(field[0], dtype=np.float64).copy()
        self.ai = np.ascontiguousarray(field[1], dtype=np.float64).copy()
        assert self.ar.size == mp.nmesh and self.ai.size == mp.nmesh
        self.taper = np.array(taper, dtype=np.float64)
        self.nthreads = nthreads
        self.n = self.parts["x"].size

    def _pargs(self):
        P = self.parts
        return [_ptr(P[k]) for k in ("x", "y", "z2", "px", "py", "gam", "chi")]

    def rk4par(self, sz: float, h: float) -> int:
        return lib().orc_rk4par(self.pp, _ptr(self.taper), self.n, *self._pargs(),
                                _ptr(self.ar), _ptr(self.ai), sz, h, self.nthreads)

    def diffraction_step(self, h: float) -> int:
        return lib().orc_diffraction_step(self.pp, _ptr(self.ar), _ptr(self.ai), h, self.nthreads)

    def wrap_z2(self):
        lib().orc_wrap_z2(self.pp, self.n, _ptr(self.parts["z2"]))

    def power(self, f32_accumulate: bool = False) -> np.ndarray:
        out = np.zeros(self.mp.nz2)
        lib().orc_power(self.pp, _ptr(self.ar), _ptr(self.ai), int(f32_accumulate), _ptr(out))
        return out

    def bunching(self, harmonic: int, nslices: int, slice_len: float) -> np.ndarray:
        out = np.zeros(nslices)
        lib().orc_bunching(self.pp, self.n, _ptr(self.parts["z2"]), _ptr(self.parts["chi"]),
                           harmonic, nslices, slice_len, _ptr(out))
        return out

    def slice_stats(self, nslices: int, slice_len: float):
        """getSliceTwiss's per-slice outputs: (status, dict of arrays)."""
        return slice_stats(self.mp, self.parts, nslices, slice_len)

    def current(self, npts: int, sam_len: float):
        """getCurr's binned charge (before unit scalings): (status, array)."""
        return current(self.mp, self.parts, npts, sam_len)

    def und_section(self, nsteps: int, h: float, diffraction: bool, steps4diff: int,
                    redist_every: int, csteps: int = 0, write_mask=None, sz_local: float = 0.0):
        wm = None
        nw = 0
        if write_mask is not None:
            wm = np.ascontiguousarray(write_mask, dtype=np.uint8)
            nw = int(wm.sum())
        powers = np.zeros(max(nw, 1) * self.mp.nz2)
        cs = np.array([csteps], dtype=np.int64)
        szl = np.array([sz_local], dtype=np.float64)
        st = lib().orc_und_section(self.pp, _ptr(self.taper), self.n, *self._pargs(),
                                   _ptr(self.ar), _ptr(self.ai), nsteps, h, int(diffraction),
                                   steps4diff, redist_every, _ptr(cs),
                                   _ptr(wm) if wm is not None else None, _ptr(powers),
                                   _ptr(szl), self.nthreads)
        return st, powers[: nw * self.mp.nz2].reshape(nw, self.mp.nz2), int(cs[0]), float(szl[0])


# ---- free functions used by the tests ------------------------------------

def get_p2(gam, pr, pi, eta, gamma_r, aw):
    return lib().orc_get_p2(gam, pr, pi, eta, gamma_r, aw)


def get_gamma(p2, pr, pi, eta, gamma_r, aw):
    return lib().orc_get_gamma(p2, pr, pi, eta, gamma_r, aw)


def bfields(mp: ModuleParams, place: int, sz: float, x: float, y: float):
    p = mp.to_ctypes()
    bx, by, bz = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_bfields(ctypes.byref(p), place, sz, x, y, ctypes.byref(bx), ctypes.byref(by),
                      ctypes.byref(bz))
    return bx.value, by.value, bz.value


def interp_one(mp: ModuleParams, x, y, z2):
    p = mp.to_ctypes()
    node = np.zeros(8, dtype=np.int64)
    w = np.zeros(8)
    f = lib().orc_interp_one(ctypes.byref(p), x, y, z2, _ptr(node), _ptr(w))
    return f, node, w


def interp_deposit(mp: ModuleParams, parts: dict, ar, ai):
    p = mp.to_ctypes()
    n = parts["x"].size
    fr = np.zeros(n)
    fi = np.zeros(n)
    dar = np.zeros(mp.nmesh)
    dai = np.zeros(mp.nmesh)
    args = [np.ascontiguousarray(parts[k], dtype=np.float64) for k in ("x", "y", "z2", "px", "py", "gam", "chi")]
    st = lib().orc_interp_deposit(ctypes.byref(p), n, *[_ptr(a) for a in args],
                                  _ptr(np.ascontiguousarray(ar, dtype=np.float64)),
                                  _ptr(np.ascontiguousarray(ai, dtype=np.float64)),
                                  _ptr(fr), _ptr(fi), _ptr(dar), _ptr(dai))
    return st, fr, fi, dar, dai


def fft1d(a: np.ndarray, sign: int) -> np.ndarray:
    buf = np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64).copy()
    lib().orc_fft1d(_ptr(buf), a.size, sign)
    return buf.view(np.complex128)


def fft3d(a: np.ndarray, sign: int, nthreads: int = 1) -> np.ndarray:
    """a has shape (n3, n2, n1) (z2, y, x)."""
    n3, n2, n1 = a.shape
    buf = np.ascontiguousarray(a, dtype=np.complex128).view(np.float64).reshape(-1).copy()
    lib().orc_fft3d(_ptr(buf), n1, n2, n3, sign, nthreads)
    return buf.view(np.complex128).reshape(n3, n2, n1)


def kvalues(mp: ModuleParams):
    p = mp.to_ctypes()
    kx = np.zeros(mp.nx)
    ky = np.zeros(mp.ny)
    kz = np.zeros(mp.nzp)
    lib().orc_kvalues(ctypes.byref(p), _ptr(kx), _ptr(ky), _ptr(kz))
    return kx, ky, kz


def masks(mp: ModuleParams):
    p = mp.to_ctypes()
    mxy = np.zeros(mp.ntr)
    mz = np.zeros(mp.nzp)
    lib().orc_mask_xy(ctypes.byref(p), _ptr(mxy))
    lib().orc_mask_z2(ctypes.byref(p), _ptr(mz))
    return mxy, mz


def match(mp: ModuleParams, n2col: float, parts: dict, sz: float, sign: int):
    p = mp.to_ctypes()
    n = parts["x"].size
    lib().orc_match(ctypes.byref(p), n2col, n, _ptr(parts["x"]), _ptr(parts["y"]),
                    _ptr(parts["px"]), _ptr(parts["py"]), _ptr(parts["gam"]), sz, sign)


SLICE_KEYS = ("sdata", "aveX", "aveY", "avePX", "avePY", "sdX", "sdY", "sdpX", "sdpY", "eX", "eY",
              "aveGamma", "aveDgamma")


def slice_stats(mp: ModuleParams, parts: dict, nslices: int, slice_len: float):
    p = mp.to_ctypes()
    out = np.zeros(13 * nslices)
    P = {k: np.ascontiguousarray(parts[k], dtype=np.float64) for k in ("x", "y", "z2", "px", "py", "gam", "chi")}
    st = lib().orc_slice_stats(ctypes.byref(p), P["x"].size, *[_ptr(P[k]) for k in ("x", "y", "z2", "px", "py",
                                                                                       "gam", "chi")],
                               nslices, slice_len, _ptr(out))
    o = out.reshape(13, nslices)
    return st, {k: o[i].copy() for i, k in enumerate(SLICE_KEYS)}


def current(mp: ModuleParams, parts: dict, npts: int, sam_len: float):
    p = mp.to_ctypes()
    out = np.zeros(npts)
    z2 = np.ascontiguousarray(parts["z2"], dtype=np.float64)
    chi = np.ascontiguousarray(parts["chi"], dtype=np.float64)
    st = lib().orc_current(ctypes.byref(p), z2.size, _ptr(z2), _ptr(chi), npts, sam_len, _ptr(out))
    return st, out


def set_deposit_atomic(on: bool):
    """Shared-grid atomic deposit (timing runs) instead of private grids."""
    lib().orc_set_deposit_atomic(int(on))


def drift(mp: ModuleParams, parts: dict, dz: float):
    """driftSection's particle part (acc_lattice.f90:478-518), in place."""
    p = mp.to_ctypes()
    lib().orc_drift(ctypes.byref(p), parts["x"].size, _ptr(parts["x"]), _ptr(parts["y"]), _ptr(parts["z2"]),
                    _ptr(parts["px"]), _ptr(parts["py"]), _ptr(parts["gam"]), dz)


def chicane(mp: ModuleParams, parts: dict, disp: float, slip: float):
    """disperse's particle part (acc_lattice.f90:420-463), in place."""
    p = mp.to_ctypes()
    lib().orc_chicane(ctypes.byref(p), parts["z2"].size, _ptr(parts["z2"]), _ptr(parts["gam"]), disp, slip)


def quad(mp: ModuleParams, parts: dict, qfx: float, qfy: float):
    """Quad (acc_lattice.f90:531-561), in place."""
    p = mp.to_ctypes()
    lib().orc_quad(ctypes.byref(p), parts["x"].size, _ptr(parts["x"]), _ptr(parts["y"]), _ptr(parts["px"]),
                   _ptr(parts["py"]), qfx, qfy)


def modulation(parts: dict, wavenum: float, mag: float):
    """bModulation (acc_lattice.f90:575-585), in place."""
    lib().orc_modulation(parts["z2"].size, _ptr(parts["z2"]), _ptr(parts["gam"]), wavenum, mag)


def correct_trans(parts: dict):
    n = parts["x"].size
    lib().orc_correct_trans(n, _ptr(parts["x"]), _ptr(parts["y"]), _ptr(parts["px"]),
                            _ptr(parts["py"]), _ptr(parts["chi"]))


# ---- taper table and the discrete adjoint ------------------------------------

def set_taper_table(values, z0: float = 0.0, period: float = 1.0):
    """Per-period n2col table used in place of the linear taper (None clears)."""
    if values is None:
        lib().orc_set_taper_table(None, 0, 0.0, 1.0)
        return
    v = np.ascontiguousarray(values, dtype=np.float64)
    lib().orc_set_taper_table(_ptr(v), v.size, z0, period)


def taper_slot(sz: float) -> int:
    return lib().orc_taper_slot(sz)


def bfields_grad(mp: ModuleParams, place: int, sz: float, x: float, y: float):
    p = mp.to_ctypes()
    g = np.zeros(6)
    lib().orc_bfields_grad(ctypes.byref(p), place, sz, x, y, _ptr(g))
    return g


ADJ_KEYS = ("x", "y", "z2", "px", "py", "gam")


def rk4par_adjoint(mp: ModuleParams, taper, parts: dict, ar, ai, lam: dict, lar, lai,
                   sz: float, h: float, nthreads: int = 1):
    """Cotangent of one rk4par step pulled back to the start of the step.

    parts / ar / ai: the state at the start of the step (unchanged).
    lam / lar / lai: cotangent of the end state, replaced in place by the
    cotangent of the start state.  Returns (status, dJ/dn2col per stage)."""
    p = mp.to_ctypes()
    t = np.ascontiguousarray(taper, dtype=np.float64)
    n = parts["x"].size
    P = [np.ascontiguousarray(parts[k], dtype=np.float64) for k in ADJ_KEYS + ("chi",)]
    A = [np.ascontiguousarray(ar, dtype=np.float64), np.ascontiguousarray(ai, dtype=np.float64)]
    for k in ADJ_KEYS:
        assert lam[k].dtype == np.float64 and lam[k].flags["C_CONTIGUOUS"]
    dn = np.zeros(4)
    st = lib().orc_rk4par_adjoint(ctypes.byref(p), _ptr(t), n, *[_ptr(a) for a in P],
                                  *[_ptr(a) for a in A], *[_ptr(lam[k]) for k in ADJ_KEYS],
                                  _ptr(lar), _ptr(lai), sz, h, _ptr(dn), nthreads)
    return st, dn


def diffraction_adjoint(mp: ModuleParams, lr: np.ndarray, li: np.ndarray, h: float,
                        nthreads: int = 1) -> int:
    """Apply the adjoint of one diffraction step to the cotangent, in place."""
    p = mp.to_ctypes()
    return lib().orc_diffraction_adjoint(ctypes.byref(p), _ptr(lr), _ptr(li), h, nthreads)
