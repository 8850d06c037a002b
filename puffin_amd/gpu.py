"""Python front end of the C ABI (include/puffin_gpu.h).

``Integrator`` holds one device context: the macroparticles and the A_perp
mesh of one undulator module, exactly the module globals the reference's
rk4par / DiffractionStep read (deriv_globals.f90:94,128-133;
puffin_mpi_RK4.f90:24).  Every call goes to libpuffin_gpu.so; there is no CPU
fallback: a missing library or device raises.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from .params import FP64, MEM_DEVICE, MEM_HOST, STATUS_NAMES, ModuleParams, PuffParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpuffin_gpu.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# Exported symbols and their signatures (include/puffin_gpu.h).
SIGNATURES = {
    "puff_last_error": (ctypes.c_char_p, []),
    "puff_device_count": (ctypes.c_int, []),
    "puff_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(PuffParams), ctypes.c_int64, ctypes.c_int32]),
    "puff_destroy": (ctypes.c_int, [_vp]),
    "puff_set_stream": (ctypes.c_int, [_vp, _vp]),
    "puff_synchronize": (ctypes.c_int, [_vp]),
    "puff_upload_particles": (ctypes.c_int, [_vp, ctypes.c_int32] + [_vp] * 7),
    "puff_download_particles": (ctypes.c_int, [_vp, ctypes.c_int32] + [_vp] * 6),
    "puff_upload_field": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _vp]),
    "puff_download_field": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _vp]),
    "puff_set_taper": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "puff_get_taper": (ctypes.c_int, [_vp, _dp, _dp]),
    "puff_set_taper_table": (ctypes.c_int, [_vp, _dp, ctypes.c_int32, ctypes.c_double, ctypes.c_double]),
    "puff_set_sort_interval": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "puff_set_deterministic": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "puff_rk4par": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "puff_diffraction_step": (ctypes.c_int, [_vp, ctypes.c_double]),
    "puff_transform": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "puff_wrap_z2": (ctypes.c_int, [_vp]),
    "puff_drift": (ctypes.c_int, [_vp, ctypes.c_double]),
    "puff_chicane": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "puff_quad": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "puff_modulation": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "puff_correct_trans": (ctypes.c_int, [_vp]),
    "puff_match": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32]),
    "puff_set_undulator": (ctypes.c_int, [_vp, ctypes.c_int32] + [ctypes.c_double] * 6),
    "puff_power": (ctypes.c_int, [_vp, _dp]),
    "puff_bunching": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _dp]),
    "puff_tile_fallback": (ctypes.c_int, [_vp, _dp, _dp]),
    "puff_slice_stats": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double, _dp]),
    "puff_current": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double, _dp]),
    "puff_slice_stats_noalloc": (ctypes.c_int, [ctypes.POINTER(PuffParams), ctypes.c_int64] + [_vp] * 7 +
                                 [ctypes.c_int32, ctypes.c_double, _vp, _vp]),
    "puff_particle_count": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "puff_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "puff_profile_read": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32,
                                         ctypes.c_int32]),
    "puff_active_box": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "puff_state_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "puff_state_save": (ctypes.c_int, [_vp, _vp]),
    "puff_state_load": (ctypes.c_int, [_vp, _vp]),
    "puff_state_save_ex": (ctypes.c_int, [_vp, _vp, _vp]),
    "puff_state_load_ex": (ctypes.c_int, [_vp, _vp, _vp]),
    "puff_sort_now": (ctypes.c_int, [_vp]),
    "puff_adjoint_enable": (ctypes.c_int, [_vp]),
    "puff_adjoint_zero": (ctypes.c_int, [_vp]),
    "puff_adjoint_seed_power": (ctypes.c_int, [_vp, _dp, ctypes.c_double]),
    "puff_rk4par_adjoint": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _dp,
                                           ctypes.POINTER(ctypes.c_int32)]),
    "puff_diffraction_adjoint": (ctypes.c_int, [_vp, ctypes.c_double]),
    "puff_capture_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "puff_rk4par_capture": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "puff_rk4par_adjoint_captured": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp, _dp,
                                                    ctypes.POINTER(ctypes.c_int32)]),
    "puff_set_deferred": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "puff_deferred_status": (ctypes.c_int, [_vp]),
    "puff_adjoint_taper_gradient": (ctypes.c_int, [_vp, _dp, ctypes.c_int32]),
    "puff_adjoint_upload": (ctypes.c_int, [_vp, ctypes.c_int32] + [_vp] * 8),
    "puff_adjoint_download": (ctypes.c_int, [_vp, ctypes.c_int32] + [_vp] * 8),
}

PROFILE_KINDS = ("stage1", "stage2", "stage3", "stage4", "combine", "diffraction", "step_begin",
                 "adj_stages", "adj_combine", "adj_diffraction", "adj_replay", "sort")


class PuffinError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = f"{where}: {STATUS_NAMES.get(status, 'error')} (status {status})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


def lib() -> ctypes.CDLL:
    """Load libpuffin_gpu.so (built by `make -C puffin_amd`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C puffin_amd`")
        # torch bundles its own HIP runtime under the same soname
        # (libamdhip64.so.7); whichever loads first serves the whole process.
        # Load torch's first when torch is present, so that the library and
        # torch's device memory and streams share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def last_error() -> str:
    return lib().puff_last_error().decode(errors="replace")


def device_count() -> int:
    return int(lib().puff_device_count())


def _check(status: int, where: str) -> int:
    if status < 0:
        raise PuffinError(status, where, last_error())
    return status


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data


PARTICLE_KEYS = ("x", "y", "z2", "px", "py", "gam")
STATE_HEADER_BYTES = 256   # PUFF_STATE_HEADER_BYTES


class Integrator:
    """One module's particles and field on the current HIP device."""

    def __init__(self, mp: ModuleParams, n: int, precision: int = FP64):
        self.mp = mp
        self.n = int(n)
        self.precision = precision
        self._params = mp.to_ctypes()
        h = _vp()
        _check(lib().puff_create(ctypes.byref(h), ctypes.byref(self._params), self.n, precision),
               "puff_create")
        self._h = h

    # ---- lifetime
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().puff_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # ---- data exchange (host numpy arrays, reference layout)
    def upload_particles(self, p: dict):
        arrs = [np.ascontiguousarray(p[k], dtype=np.float64) for k in PARTICLE_KEYS + ("chi",)]
        for a in arrs:
            assert a.size == self.n
        _check(lib().puff_upload_particles(self._h, MEM_HOST, *[_ptr(a) for a in arrs]),
               "puff_upload_particles")

    def download_particles(self) -> dict:
        out = {k: np.empty(self.n) for k in PARTICLE_KEYS}
        _check(lib().puff_download_particles(self._h, MEM_HOST, *[_ptr(out[k]) for k in PARTICLE_KEYS]),
               "puff_download_particles")
        return out

    def upload_field(self, ar: np.ndarray, ai: np.ndarray):
        ar = np.ascontiguousarray(ar, dtype=np.float64).reshape(-1)
        ai = np.ascontiguousarray(ai, dtype=np.float64).reshape(-1)
        assert ar.size == self.mp.nmesh and ai.size == self.mp.nmesh
        _check(lib().puff_upload_field(self._h, MEM_HOST, _ptr(ar), _ptr(ai)), "puff_upload_field")

    def download_field(self):
        ar = np.empty(self.mp.nmesh)
        ai = np.empty(self.mp.nmesh)
        _check(lib().puff_download_field(self._h, MEM_HOST, _ptr(ar), _ptr(ai)), "puff_download_field")
        return ar, ai

    # device-pointer variants (torch tensors, FP64, reference layout)
    def upload_particles_device(self, ptrs: list[int]):
        _check(lib().puff_upload_particles(self._h, MEM_DEVICE, *ptrs), "puff_upload_particles")

    def download_particles_device(self, ptrs: list[int]):
        _check(lib().puff_download_particles(self._h, MEM_DEVICE, *ptrs), "puff_download_particles")

    def upload_field_device(self, ar_ptr: int, ai_ptr: int):
        _check(lib().puff_upload_field(self._h, MEM_DEVICE, ar_ptr, ai_ptr), "puff_upload_field")

    def download_field_device(self, ar_ptr: int, ai_ptr: int):
        _check(lib().puff_download_field(self._h, MEM_DEVICE, ar_ptr, ai_ptr), "puff_download_field")

    def set_stream(self, stream_ptr: int | None):
        _check(lib().puff_set_stream(self._h, stream_ptr), "puff_set_stream")

    # torch tensors on the context's device (FP64, reference layout).  The
    # uploads first wait for torch's current stream, so that a tensor torch
    # has just produced is complete before the context's stream reads it;
    # downloads return once the copies are done.
    @staticmethod
    def _torch_ready(tensors):
        import torch
        for t in tensors:
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError("expected contiguous float64 device tensors")
        torch.cuda.current_stream().synchronize()
        return [t.data_ptr() for t in tensors]

    def use_torch_stream(self):
        """Run the context's work on torch's current stream."""
        import torch
        self.set_stream(torch.cuda.current_stream().cuda_stream)

    def upload_particles_torch(self, p: dict):
        ts = [p[k] for k in PARTICLE_KEYS + ("chi",)]
        for t in ts:
            assert t.numel() == self.n
        self.upload_particles_device(self._torch_ready(ts))

    def download_particles_torch(self) -> dict:
        import torch
        out = {k: torch.empty(self.n, dtype=torch.float64, device="cuda") for k in PARTICLE_KEYS}
        self.download_particles_device(self._torch_ready([out[k] for k in PARTICLE_KEYS]))
        return out

    def upload_field_torch(self, ar, ai):
        assert ar.numel() == self.mp.nmesh and ai.numel() == self.mp.nmesh
        self.upload_field_device(*self._torch_ready([ar, ai]))

    def download_field_torch(self):
        import torch
        ar = torch.empty(self.mp.nmesh, dtype=torch.float64, device="cuda")
        ai = torch.empty_like(ar)
        self.download_field_device(*self._torch_ready([ar, ai]))
        return ar, ai

    def synchronize(self):
        _check(lib().puff_synchronize(self._h), "puff_synchronize")

    # ---- taper
    def set_taper(self, n2col: float, n2col0: float, undgrad: float):
        _check(lib().puff_set_taper(self._h, n2col, n2col0, undgrad), "puff_set_taper")

    def get_taper(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().puff_get_taper(self._h, ctypes.byref(a), ctypes.byref(b)), "puff_get_taper")
        return a.value, b.value

    def set_taper_table(self, values, z0: float, period: float):
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().puff_set_taper_table(self._h, v.ctypes.data_as(_dp), v.size, z0, period),
               "puff_set_taper_table")

    def set_deterministic(self, on: bool = True):
        """Fixed-point deposit: results independent of the atomics' order (FP64)."""
        _check(lib().puff_set_deterministic(self._h, int(on)), "puff_set_deterministic")

    def set_sort_interval(self, steps: int):
        _check(lib().puff_set_sort_interval(self._h, steps), "puff_set_sort_interval")

    # ---- hot path
    def rk4par(self, sz: float, h: float) -> int:
        """One RK4 step; returns the status (0 ok, 1..3 flag raised, state kept)."""
        return _check(lib().puff_rk4par(self._h, sz, h), "puff_rk4par")

    def diffraction_step(self, h: float):
        _check(lib().puff_diffraction_step(self._h, h), "puff_diffraction_step")

    def transform(self, sign: int):
        _check(lib().puff_transform(self._h, sign), "puff_transform")

    def wrap_z2(self):
        _check(lib().puff_wrap_z2(self._h), "puff_wrap_z2")

    # ---- lattice elements (acc_lattice.f90)
    def drift(self, dz: float):
        """driftSection's particle part; the diffraction over dz is a separate call."""
        _check(lib().puff_drift(self._h, dz), "puff_drift")

    def chicane(self, disp: float, slip: float):
        """disperse's particle part (slip in zbar units, as after readLatt)."""
        _check(lib().puff_chicane(self._h, disp, slip), "puff_chicane")

    def quad(self, fx: float, fy: float):
        _check(lib().puff_quad(self._h, fx, fy), "puff_quad")

    def modulation(self, wavenum: float, mag: float):
        _check(lib().puff_modulation(self._h, wavenum, mag), "puff_modulation")

    def correct_trans(self):
        _check(lib().puff_correct_trans(self._h), "puff_correct_trans")

    def match(self, sz: float, n2col: float, sign: int):
        """matchIn (sign +1) or matchOut (sign -1) at zbar sz."""
        _check(lib().puff_match(self._h, sz, n2col, int(sign)), "puff_match")

    def set_undulator(self, und_type: int, fx: float, fy: float, kbeta_x: float = 0.0, kbeta_y: float = 0.0,
                      szfs: float = 0.0, szfe: float = 0.0):
        _check(lib().puff_set_undulator(self._h, int(und_type), fx, fy, kbeta_x, kbeta_y, szfs, szfe),
               "puff_set_undulator")
        self.mp = dataclasses.replace(self.mp, und_type=int(und_type), fx=fx, fy=fy, kbeta_x=kbeta_x,
                                      kbeta_y=kbeta_y, szfs=szfs, szfe=szfe)

    # ---- instrumentation
    def profile(self, on: bool = True):
        _check(lib().puff_profile_enable(self._h, int(on)), "puff_profile_enable")

    def profile_read(self, reset: bool = False) -> dict:
        k = len(PROFILE_KINDS)
        ms = np.zeros(k)
        cnt = np.zeros(k, dtype=np.int64)
        _check(lib().puff_profile_read(self._h, ms.ctypes.data_as(_dp),
                                       cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), k, int(reset)),
               "puff_profile_read")
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(PROFILE_KINDS)}

    def active_box(self):
        b = (ctypes.c_int32 * 6)()
        _check(lib().puff_active_box(self._h, b), "puff_active_box")
        return tuple(b)

    # ---- diagnostics
    def power(self) -> np.ndarray:
        out = np.empty(self.mp.nz2)
        _check(lib().puff_power(self._h, out.ctypes.data_as(_dp)), "puff_power")
        return out

    def bunching(self, harmonic: int, nslices: int, slice_len: float) -> np.ndarray:
        out = np.empty(nslices)
        _check(lib().puff_bunching(self._h, harmonic, nslices, slice_len, out.ctypes.data_as(_dp)),
               "puff_bunching")
        return out

    def tile_fallback(self):
        """(last step, first step after the last re-sort): share of stage
        workgroups that deposited with global atomics."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().puff_tile_fallback(self._h, ctypes.byref(a), ctypes.byref(b)), "puff_tile_fallback")
        return a.value, b.value

    SLICE_KEYS = ("sdata", "aveX", "aveY", "avePX", "avePY", "sdX", "sdY", "sdpX", "sdpY", "eX", "eY",
                  "aveGamma", "aveDgamma")

    def slice_stats(self, nslices: int, slice_len: float) -> dict:
        """getSliceTwiss's per-slice outputs (charge, means, r.m.s., emittances,
        mean and r.m.s. Gamma), keyed by the reference's variable names."""
        out = np.empty(13 * nslices)
        _check(lib().puff_slice_stats(self._h, nslices, slice_len, out.ctypes.data_as(_dp)), "puff_slice_stats")
        o = out.reshape(13, nslices)
        return {k: o[i].copy() for i, k in enumerate(self.SLICE_KEYS)}

    def current(self, npts: int, sam_len: float) -> np.ndarray:
        """getCurr's binned charge per node, before its unit factors."""
        out = np.empty(npts)
        _check(lib().puff_current(self._h, npts, sam_len, out.ctypes.data_as(_dp)), "puff_current")
        return out

    # ---- state snapshots (device blobs in the slot order)
    def state_bytes(self) -> int:
        b = ctypes.c_int64()
        _check(lib().puff_state_bytes(self._h, ctypes.byref(b)), "puff_state_bytes")
        return int(b.value)

    def state_save(self, dev_ptr: int, header: np.ndarray | None = None):
        """Snapshot into device memory at dev_ptr; ``header`` (uint8[256]),
        when given, receives a host copy of the blob's header."""
        hp = None
        if header is not None:
            assert header.dtype == np.uint8 and header.size >= STATE_HEADER_BYTES
            hp = header.ctypes.data
        _check(lib().puff_state_save_ex(self._h, dev_ptr, hp), "puff_state_save")

    def state_load(self, dev_ptr: int, header: np.ndarray | None = None):
        """Restore a snapshot; with the header of its save, without a wait."""
        hp = None
        if header is not None:
            assert header.dtype == np.uint8 and header.size >= STATE_HEADER_BYTES
            hp = header.ctypes.data
        _check(lib().puff_state_load_ex(self._h, dev_ptr, hp), "puff_state_load")

    def sort_now(self):
        _check(lib().puff_sort_now(self._h), "puff_sort_now")

    # ---- discrete adjoint
    def adjoint_enable(self):
        _check(lib().puff_adjoint_enable(self._h), "puff_adjoint_enable")

    def adjoint_zero(self):
        _check(lib().puff_adjoint_zero(self._h), "puff_adjoint_zero")

    def adjoint_seed_power(self, wz: np.ndarray, scale: float = 1.0):
        w = np.ascontiguousarray(wz, dtype=np.float64)
        assert w.size == self.mp.nz2
        _check(lib().puff_adjoint_seed_power(self._h, w.ctypes.data_as(_dp), scale),
               "puff_adjoint_seed_power")

    def rk4par_adjoint(self, sz: float, h: float):
        """Pull the cotangent back through one step from the current state.
        Returns (status, dJ/dn2col per stage, taper-table slot per stage)."""
        dn = np.zeros(4)
        slot = (ctypes.c_int32 * 4)()
        st = _check(lib().puff_rk4par_adjoint(self._h, sz, h, dn.ctypes.data_as(_dp), slot),
                    "puff_rk4par_adjoint")
        return st, dn, np.array(list(slot), dtype=np.int64)

    def diffraction_adjoint(self, h: float):
        _check(lib().puff_diffraction_adjoint(self._h, h), "puff_diffraction_adjoint")

    def adjoint_upload(self, lam: dict | None, lar=None, lai=None):
        arrs = [np.ascontiguousarray(lam[k], dtype=np.float64) if lam is not None else None
                for k in PARTICLE_KEYS]
        fr = None if lar is None else np.ascontiguousarray(lar, dtype=np.float64).reshape(-1)
        fi = None if lai is None else np.ascontiguousarray(lai, dtype=np.float64).reshape(-1)
        _check(lib().puff_adjoint_upload(self._h, MEM_HOST, *[_ptr(a) for a in arrs], _ptr(fr), _ptr(fi)),
               "puff_adjoint_upload")

    def adjoint_download_field_torch(self):
        """The field cotangent into torch device tensors (split Re / Im)."""
        import torch
        lar = torch.empty(self.mp.nmesh, dtype=torch.float64, device="cuda")
        lai = torch.empty_like(lar)
        ptrs = self._torch_ready([lar, lai])
        _check(lib().puff_adjoint_download(self._h, MEM_DEVICE, None, None, None, None, None, None,
                                           ptrs[0], ptrs[1]), "puff_adjoint_download")
        return lar, lai

    def adjoint_download(self):
        out = {k: np.empty(self.n) for k in PARTICLE_KEYS}
        lar = np.empty(self.mp.nmesh)
        lai = np.empty(self.mp.nmesh)
        _check(lib().puff_adjoint_download(self._h, MEM_HOST, *[_ptr(out[k]) for k in PARTICLE_KEYS],
                                           _ptr(lar), _ptr(lai)), "puff_adjoint_download")
        return out, lar, lai

    # ---- deferred status
    def set_deferred(self, on: bool):
        _check(lib().puff_set_deferred(self._h, int(on)), "puff_set_deferred")

    def deferred_status(self) -> int:
        return _check(lib().puff_deferred_status(self._h), "puff_deferred_status")

    def adjoint_taper_gradient(self, n: int) -> np.ndarray:
        out = np.zeros(n)
        _check(lib().puff_adjoint_taper_gradient(self._h, out.ctypes.data_as(_dp), n),
               "puff_adjoint_taper_gradient")
        return out

    # ---- step capture
    def capture_bytes(self) -> int:
        b = ctypes.c_int64()
        _check(lib().puff_capture_bytes(self._h, ctypes.byref(b)), "puff_capture_bytes")
        return int(b.value)

    def rk4par_capture(self, sz: float, h: float, dev_ptr: int, header: np.ndarray) -> int:
        """rk4par that keeps the step's start state and stage data at dev_ptr
        (capture_bytes() bytes) and its header in ``header`` (uint8[256])."""
        assert header.dtype == np.uint8 and header.size >= STATE_HEADER_BYTES
        return _check(lib().puff_rk4par_capture(self._h, sz, h, dev_ptr, header.ctypes.data),
                      "puff_rk4par_capture")

    def rk4par_adjoint_captured(self, sz: float, h: float, dev_ptr: int, header: np.ndarray):
        dn = np.zeros(4)
        slot = (ctypes.c_int32 * 4)()
        st = _check(lib().puff_rk4par_adjoint_captured(self._h, sz, h, dev_ptr, header.ctypes.data,
                                                       dn.ctypes.data_as(_dp), slot),
                    "puff_rk4par_adjoint_captured")
        return st, dn, np.array(list(slot), dtype=np.int64)
