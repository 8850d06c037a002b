"""Puffin-format HDF5 files through libpuffin_h5.so (include/puffin_h5.h).

Particle files hold the (N, 7) ``electrons`` table (x, y, z2, px, py, Gamma,
chi-bar; readH5Beamfile, io/h5_in.f90:416-720), field files the
component-major ``aperp`` dataset (readH5FieldfileSingleDump,
io/h5_in.f90:726-1143), integrated files the power and bunching datasets
(wr_h5, io/hdf5_puff.f90:173-427); all carry ``runInfo`` and the position
attributes.  The library is host code over the HDF5 C library; a missing
library raises at first use.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import ModuleParams, PI_REF

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)

KEYS = ("x", "y", "z2", "px", "py", "gam", "chi")


class H5Info(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("zbar_total", ctypes.c_double), ("zbar_local", ctypes.c_double),
        ("zbar_inter", ctypes.c_double), ("z_total", ctypes.c_double), ("z_local", ctypes.c_double),
        ("z_inter", ctypes.c_double),
        ("istep", ctypes.c_int32), ("icsteps", ctypes.c_int32), ("iund_cr", ctypes.c_int32),
        ("ichic_cr", ctypes.c_int32), ("idrift_cr", ctypes.c_int32), ("iquad_cr", ctypes.c_int32),
        ("imodulation_cr", ctypes.c_int32), ("il", ctypes.c_int32), ("iwrite_cr", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz2", ctypes.c_int32),
        ("field_mesh", ctypes.c_int32), ("n_steps", ctypes.c_int32), ("num_spatial_dims", ctypes.c_int32),
        ("iscale", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz2", ctypes.c_double),
        ("rho", ctypes.c_double), ("eta", ctypes.c_double), ("kappa", ctypes.c_double),
        ("gamma_r", ctypes.c_double), ("aw", ctypes.c_double), ("lg", ctypes.c_double), ("lc", ctypes.c_double),
        ("lambda_w", ctypes.c_double), ("lambda_r", ctypes.c_double), ("step_size", ctypes.c_double),
        ("npk_bar", ctypes.c_double), ("trans_area", ctypes.c_double), ("trans_area_si", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class H5Error(RuntimeError):
    pass


def lib_path() -> str:
    return os.path.join(_HERE, "libpuffin_h5.so")


def lib():
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            raise H5Error(f"{p} is not built (make -C puffin_amd libpuffin_h5.so)")
        L = ctypes.CDLL(p)
        ip = ctypes.POINTER(H5Info)
        c = ctypes.c_char_p
        L.puff_h5_last_error.restype = ctypes.c_char_p
        L.puff_h5_last_error.argtypes = []
        L.puff_h5_write_electrons.argtypes = [c, ip, ctypes.c_int64] + [_dp] * 7
        L.puff_h5_count_electrons.argtypes = [c, ctypes.POINTER(ctypes.c_int64)]
        L.puff_h5_read_electrons.argtypes = [c, ctypes.c_int64] + [_dp] * 7 + [ip]
        L.puff_h5_write_field.argtypes = [c, ip, _dp, _dp]
        L.puff_h5_read_field_info.argtypes = [c, ip]
        L.puff_h5_read_field.argtypes = [c, ctypes.c_int64, _dp, _dp]
        L.puff_h5_write_integrated.argtypes = [c, ip, ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_int32, _dp]
        L.puff_h5_read_dataset_1d.argtypes = [c, c, ctypes.c_int64, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.puff_h5_append_1d.argtypes = [c, ip, c, ctypes.c_int64, _dp]
        for name in ("puff_h5_write_electrons", "puff_h5_count_electrons", "puff_h5_read_electrons",
                     "puff_h5_write_field", "puff_h5_read_field_info", "puff_h5_read_field",
                     "puff_h5_write_integrated", "puff_h5_read_dataset_1d", "puff_h5_append_1d"):
            getattr(L, name).restype = ctypes.c_int
        _LIB = L
    return _LIB


def _check(st: int, where: str):
    if st != 0:
        raise H5Error(f"{where}: {lib().puff_h5_last_error().decode()} (status {st})")


def _col(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def info_from_module(mp: ModuleParams, *, time: float = 0.0, zbar_local: float = 0.0, istep: int = 0,
                     icsteps: int = 0, n_steps: int = 0, step_size: float = 0.0, lambda_w: float = 0.0,
                     npk_bar: float = 0.0, trans_area: float = 0.0, counters: dict | None = None) -> H5Info:
    """A record for ``mp``'s mesh and scaling at machine position ``time``
    (zbar).  With the undulator period ``lambda_w`` (m) given, the lengths in
    metres follow the reference's scalings: lambda_r = eta lambda_w,
    Lg = lambda_w / (4 pi rho), Lc = lambda_r / (4 pi rho), z = zbar Lg.
    ``trans_area`` is the 1-D transverse area (ata_G) that scales the 1-D
    power; ``counters`` sets element counters by field name."""
    f = H5Info()
    f.time = f.zbar_total = time
    f.zbar_local = zbar_local
    f.istep = istep
    f.icsteps = icsteps
    f.nx, f.ny, f.nz2 = (1, 1, mp.nz2) if mp.one_d else (mp.nx, mp.ny, mp.nz2)
    f.field_mesh = mp.periodic
    f.n_steps = n_steps
    f.num_spatial_dims = 1 if mp.one_d else 3
    f.iscale = 1
    f.dx, f.dy, f.dz2 = (1.0, 1.0, mp.dz2) if mp.one_d else (mp.dx, mp.dy, mp.dz2)
    f.rho, f.eta, f.kappa, f.gamma_r, f.aw = mp.rho, mp.eta, mp.kappa, mp.gamma_r, mp.aw
    f.step_size = step_size
    f.npk_bar = npk_bar
    f.trans_area = trans_area
    if lambda_w > 0.0:
        f.lambda_w = lambda_w
        f.lambda_r = mp.eta * lambda_w
        f.lg = lambda_w / (4.0 * PI_REF * mp.rho)
        f.lc = f.lambda_r / (4.0 * PI_REF * mp.rho)
        f.z_total = time * f.lg
        f.z_local = zbar_local * f.lg
        f.trans_area_si = trans_area * f.lg * f.lc
    for k, v in (counters or {}).items():
        setattr(f, k, v)
    return f


def write_electrons(path: str, parts: dict, info: H5Info):
    cols = [_col(parts[k]) for k in KEYS]
    n = cols[0].size
    assert all(c.size == n for c in cols)
    _check(lib().puff_h5_write_electrons(os.fsencode(path), ctypes.byref(info), n, *[_p(c) for c in cols]),
           "write_electrons")


def read_electrons(path: str):
    """(particles dict with x, y, z2, px, py, gam, chi; H5Info)."""
    n = ctypes.c_int64()
    _check(lib().puff_h5_count_electrons(os.fsencode(path), ctypes.byref(n)), "read_electrons")
    cols = {k: np.empty(n.value) for k in KEYS}
    info = H5Info()
    _check(lib().puff_h5_read_electrons(os.fsencode(path), n.value, *[_p(cols[k]) for k in KEYS],
                                        ctypes.byref(info)), "read_electrons")
    return cols, info


def write_field(path: str, ar: np.ndarray, ai: np.ndarray, info: H5Info):
    ar, ai = _col(ar), _col(ai)
    nmesh = info.nz2 * (1 if info.num_spatial_dims == 1 else info.nx * info.ny)
    if ar.size != nmesh or ai.size != nmesh:
        raise H5Error(f"write_field: {ar.size} nodes for a mesh of {nmesh}")
    _check(lib().puff_h5_write_field(os.fsencode(path), ctypes.byref(info), _p(ar), _p(ai)), "write_field")


def read_field(path: str):
    """(ar, ai, H5Info), each part flat with x fastest."""
    info = H5Info()
    _check(lib().puff_h5_read_field_info(os.fsencode(path), ctypes.byref(info)), "read_field")
    nmesh = info.nz2 * (1 if info.num_spatial_dims == 1 else info.nx * info.ny)
    ar, ai = np.empty(nmesh), np.empty(nmesh)
    _check(lib().puff_h5_read_field(os.fsencode(path), nmesh, _p(ar), _p(ai)), "read_field")
    return ar, ai, info


def write_integrated(path: str, info: H5Info, power: np.ndarray, bunching: np.ndarray | None = None):
    """``power`` per z2 node as the device computes it; ``bunching`` is
    (nharm, nslices), harmonics 1..nharm."""
    p = _col(power)
    b = np.zeros((0, 1)) if bunching is None else np.ascontiguousarray(np.atleast_2d(bunching), dtype=np.float64)
    _check(lib().puff_h5_write_integrated(os.fsencode(path), ctypes.byref(info), p.size, _p(p), b.shape[0],
                                          b.shape[1], _p(b)), "write_integrated")


def append_dataset(path: str, info: H5Info, name: str, values):
    v = _col(values)
    _check(lib().puff_h5_append_1d(os.fsencode(path), ctypes.byref(info), name.encode(), v.size, _p(v)),
           "append_dataset")


# The reference's physical constants are default-real literals
# (common/puffin_constants.f90:23-26): single precision, widened.
Q_E = float(np.float32(1.60217653e-19))
C_LIGHT = float(np.float32(2.99792458e8))


def slice_datasets(stats: dict, info: H5Info, current=None, sam_len: float = 0.0) -> dict:
    """The integrated file's slice datasets (hdf5_puff.f90:229-416, scaled
    forms) from a slice_stats() dict and, when given, the binned charge of
    current() with its unit factors (av_write.f90:365-368):
    npk_bar (x transArea in 1-D) x q_e / sam_len x c / Lc."""
    g = np.where(stats["aveGamma"] == 0.0, 1.0, stats["aveGamma"])
    emf = 2.0 * info.kappa / np.sqrt(info.eta) / g
    out = {"Slice Charge": stats["sdata"], "meanXbar": stats["aveX"], "meanYbar": stats["aveY"],
           "meanPXbar": stats["avePX"], "meanPYbar": stats["avePY"], "meanGamma": stats["aveGamma"],
           "meanSigmaGamma": stats["aveDgamma"], "sigmaXbar": stats["sdX"], "sigmaYbar": stats["sdY"],
           "sigmaPxbar": stats["sdpX"], "sigmaPybar": stats["sdpY"],
           "emittanceXbar": emf * stats["eX"], "emittanceYbar": emf * stats["eY"]}
    if current is not None:
        cur = np.asarray(current, dtype=np.float64) * info.npk_bar
        if info.num_spatial_dims == 1:
            cur = cur * info.trans_area
        out["beamCurrent"] = cur * Q_E / sam_len * C_LIGHT / info.lc
    return out


def read_dataset(path: str, name: str) -> np.ndarray:
    n = ctypes.c_int64()
    _check(lib().puff_h5_read_dataset_1d(os.fsencode(path), name.encode(), 0, None, ctypes.byref(n)),
           "read_dataset")
    out = np.empty(n.value)
    _check(lib().puff_h5_read_dataset_1d(os.fsencode(path), name.encode(), out.size, _p(out), ctypes.byref(n)),
           "read_dataset")
    return out
