"""seisacomod2d_amd -- MI355X (gfx950) forward modelling for SeisAcoMod2D inputs.

Python front end of the C ABI in include/sam2d.h (libsam2d.so, built in-tree by
``python -m seisacomod2d_amd.build``).  It mirrors the reference's boundary,
``modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp)`` (src/src_cpu/modelling.h:106):
a grid/time description (mod_t + wave_t + job_t fields), the model, and per-shot
source/receiver coordinates in metres (geo2d_t); out come the pressure
seismograms, trace-major, as the reference writes them.

There is no CPU fallback: every computation runs in the HIP library, and
importing the package without a built library raises.
"""
import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

__all__ = ["Grid", "Shot", "Modeller", "SAM2DError", "lib", "library_path", "time_axis"]

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def library_path():
    return os.path.join(_PKG, "libsam2d.so")


class SAM2DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sam2d error {code}: {msg}")
        self.code = code


class _Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_float), ("dz", ctypes.c_float),
                ("rec_len", ctypes.c_float), ("fd_dt", ctypes.c_float),
                ("real_dt", ctypes.c_float), ("dom_freq", ctypes.c_float),
                ("npml", ctypes.c_int), ("fdop", ctypes.c_int)]


class _Shot(ctypes.Structure):
    _fields_ = [("sx", ctypes.c_float), ("sz", ctypes.c_float),
                ("nrec", ctypes.c_int), ("rec_xz", ctypes.POINTER(ctypes.c_float))]


class _Dims(ctypes.Structure):
    _fields_ = [("fd_nt", ctypes.c_int), ("real_nt", ctypes.c_int), ("dt_factor", ctypes.c_int),
                ("nts", ctypes.c_int), ("nnx", ctypes.c_int), ("nnz", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("cell_updates", ctypes.c_double), ("launches", ctypes.c_longlong),
                ("used_graph", ctypes.c_int), ("variant", ctypes.c_int),
                ("steps_per_launch", ctypes.c_int)]


def lib():
    """Load libsam2d.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m seisacomod2d_amd.build`")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        f32p, i32p = P(ctypes.c_float), P(ctypes.c_int)
        vp = ctypes.c_void_p
        L.sam2d_create.argtypes = [P(_Grid), f32p, f32p, P(vp)]
        L.sam2d_check_stability.argtypes = [P(_Grid), ctypes.c_float, ctypes.c_float]
        L.sam2d_get_dims.argtypes = [vp, P(_Dims)]
        L.sam2d_forward.argtypes = [vp, P(_Shot), ctypes.c_int, f32p]
        L.sam2d_forward_dev.argtypes = [vp, P(_Shot), ctypes.c_int, vp]
        L.sam2d_forward_noalloc.argtypes = [vp, P(_Shot), vp]
        L.sam2d_forward_batched.argtypes = [vp, P(_Shot), ctypes.c_int, i32p, f32p]
        L.sam2d_forward_windows.argtypes = [vp, P(_Shot), ctypes.c_int, i32p, f32p]
        L.sam2d_gradient.argtypes = [vp, P(_Shot), ctypes.c_int, i32p, f32p, ctypes.c_int,
                                     f32p, f32p, f32p]
        L.sam2d_forward_checkpointed.argtypes = [vp, P(_Shot), ctypes.c_int, i32p, ctypes.c_int,
                                                 f32p]
        L.sam2d_gradient_checkpointed.argtypes = [vp, f32p, f32p, f32p]
        L.sam2d_set_option.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.sam2d_get_option.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_int)]
        L.sam2d_autotune.argtypes = [vp, i32p, ctypes.c_int, ctypes.c_int, i32p]
        L.sam2d_num_variants.restype = ctypes.c_int
        L.sam2d_variant_name.argtypes = [ctypes.c_int]
        L.sam2d_variant_name.restype = ctypes.c_char_p
        L.sam2d_get_stats.argtypes = [vp, P(_Stats)]
        L.sam2d_destroy.argtypes = [vp]
        L.sam2d_destroy.restype = None
        L.sam2d_last_error.restype = ctypes.c_char_p
        L.sam2d_version.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise SAM2DError(rc, lib().sam2d_last_error().decode())


@dataclass
class Grid:
    """Job-card parameters of one modelling job (json_parser.cpp:141-265) plus the
    sponge width the reference fixes at compile time (cpumodelling.cpp:15)."""
    nx: int
    nz: int
    dx: float
    dz: float
    rec_len: float
    fd_dt: float
    real_dt: float
    dom_freq: float
    npml: int = 32
    fdop: int = 4

    def _c(self):
        return _Grid(self.nx, self.nz, self.dx, self.dz, self.rec_len, self.fd_dt,
                     self.real_dt, self.dom_freq, self.npml, self.fdop)


@dataclass
class Shot:
    """One shot gather: source (x, z) and receivers (nrec, 2) in metres."""
    sx: float
    sz: float
    rec_xz: np.ndarray

    def __post_init__(self):
        self.rec_xz = np.ascontiguousarray(np.asarray(self.rec_xz, np.float32).reshape(-1, 2))

    @property
    def nrec(self):
        return self.rec_xz.shape[0]


def _shots_c(shots: Sequence[Shot]):
    arr = (_Shot * max(1, len(shots)))()
    for i, s in enumerate(shots):
        arr[i].sx, arr[i].sz, arr[i].nrec = s.sx, s.sz, s.nrec
        arr[i].rec_xz = s.rec_xz.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    return arr


def _windows_c(windows, nshots):
    if windows is None:
        return None, None
    w = np.ascontiguousarray(np.asarray(windows, np.int32).reshape(nshots, 2))
    return w, w.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def time_axis(grid: Grid, vp=None):
    """(fd_nt, real_nt, dt_factor, nts, nnx, nnz) without creating device state."""
    g = grid
    # Evaluated by the library on a 1x1 context would need a device; mirror the
    # float arithmetic of cpumodelling.cpp:172-188 here instead.
    f = np.float32
    fd_nt = int(np.ceil(f(g.rec_len) / f(g.fd_dt)))
    real_nt = int(np.ceil(f(g.rec_len) / f(g.real_dt)))
    dtf = int(f(g.real_dt) / f(g.fd_dt))
    td = float(np.sqrt(f(6.0))) / (3.141592654 * float(f(g.dom_freq)))
    nts = min(int(5.0 * td / float(f(g.fd_dt))) + 1, fd_nt)
    return fd_nt, real_nt, dtf, nts, g.nx + 2 * g.npml, g.nz + 2 * g.npml


class Modeller:
    """A modelling context: the model resident on the GPU, coefficient fields
    built once; forward / batched / gradient calls run shots against it."""

    def __init__(self, grid: Grid, vp: np.ndarray, rho: np.ndarray, check_stability=False):
        L = lib()
        self.grid = grid
        vp = np.ascontiguousarray(vp, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        if vp.shape != (grid.nx, grid.nz) or rho.shape != (grid.nx, grid.nz):
            raise ValueError(f"model must be (nx, nz) = ({grid.nx}, {grid.nz}) x-major")
        self._g = grid._c()
        if check_stability:
            _check(L.sam2d_check_stability(ctypes.byref(self._g), float(vp.min()), float(vp.max())))
        h = ctypes.c_void_p()
        _check(L.sam2d_create(ctypes.byref(self._g), vp.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                              rho.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(h)))
        self._h = h
        d = _Dims()
        _check(L.sam2d_get_dims(self._h, ctypes.byref(d)))
        self.fd_nt, self.real_nt, self.dt_factor = d.fd_nt, d.real_nt, d.dt_factor
        self.nts, self.nnx, self.nnz = d.nts, d.nnx, d.nnz

    def close(self):
        if getattr(self, "_h", None):
            lib().sam2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_option(self, key: str, value: int):
        _check(lib().sam2d_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key: str) -> int:
        v = ctypes.c_int(0)
        _check(lib().sam2d_get_option(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def stats(self):
        s = _Stats()
        _check(lib().sam2d_get_stats(self._h, ctypes.byref(s)))
        return dict(kernel_ms=s.kernel_ms, wall_ms=s.wall_ms, cell_updates=s.cell_updates,
                    launches=s.launches, used_graph=bool(s.used_graph), variant=s.variant,
                    steps_per_launch=s.steps_per_launch)

    def _out(self, shots):
        n = sum(s.nrec for s in shots) * self.real_nt
        return np.zeros(max(n, 1), np.float32)

    def _split(self, flat, shots):
        out, off = [], 0
        for s in shots:
            k = s.nrec * self.real_nt
            out.append(flat[off:off + k].reshape(s.nrec, self.real_nt))
            off += k
        return out

    def forward(self, shots: List[Shot]) -> List[np.ndarray]:
        """Per-shot loop over the full model; one (nrec, real_nt) gather per shot."""
        seis = self._out(shots)
        sc = _shots_c(shots)
        _check(lib().sam2d_forward(self._h, sc, len(shots),
                                   seis.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return self._split(seis, shots)

    def forward_dev(self, shots: List[Shot], seis_dev_ptr: int):
        """Per-shot loop with the seismograms written to a device buffer."""
        sc = _shots_c(shots)
        _check(lib().sam2d_forward_dev(self._h, sc, len(shots), ctypes.c_void_p(seis_dev_ptr)))

    def forward_noalloc(self, shot: Shot, seis_dev_ptr: int):
        """One shot on the full model with no device allocation (working memory
        in the library's static scratch array), seismogram to a device buffer."""
        _check(lib().sam2d_forward_noalloc(self._h, _shots_c([shot]), ctypes.c_void_p(seis_dev_ptr)))

    def forward_windows(self, shots, windows) -> List[np.ndarray]:
        """Per-shot loop, each shot on its aperture window (first column, width)."""
        seis = self._out(shots)
        w, wp = _windows_c(windows, len(shots))
        _check(lib().sam2d_forward_windows(self._h, _shots_c(shots), len(shots), wp,
                                           seis.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return self._split(seis, shots)

    def forward_batched(self, shots, windows=None) -> List[np.ndarray]:
        """All shots in one fused launch per step (time loop as a hipGraph)."""
        seis = self._out(shots)
        w, wp = _windows_c(windows, len(shots))
        _check(lib().sam2d_forward_batched(self._h, _shots_c(shots), len(shots), wp,
                                           seis.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return self._split(seis, shots)

    def gradient(self, shots, adj: List[np.ndarray], ckpt_k=100, windows=None, want_seis=False):
        """dL/dvp, dL/drho (nx, nz) for adjoint sources adj = dL/dseis per shot."""
        a = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float32).ravel() for x in adj]))
        gvp = np.zeros((self.grid.nx, self.grid.nz), np.float32)
        grho = np.zeros_like(gvp)
        seis = self._out(shots) if want_seis else None
        w, wp = _windows_c(windows, len(shots))
        fp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if x is not None else None
        # the library drops the checkpoints of forward_checkpointed
        self.checkpoint_id = getattr(self, "checkpoint_id", 0) + 1
        _check(lib().sam2d_gradient(self._h, _shots_c(shots), len(shots), wp, fp(a), int(ckpt_k),
                                    fp(seis), fp(gvp), fp(grho)))
        if want_seis:
            return gvp, grho, self._split(seis, shots)
        return gvp, grho

    def forward_checkpointed(self, shots, ckpt_k=100, windows=None) -> List[np.ndarray]:
        """Seismograms of the shots (the bits of forward_windows) from the
        gradient's forward sweep; its checkpoints stay on the device for
        gradient_checkpointed."""
        seis = self._out(shots)
        w, wp = _windows_c(windows, len(shots))
        self.checkpoint_id = getattr(self, "checkpoint_id", 0) + 1
        _check(lib().sam2d_forward_checkpointed(self._h, _shots_c(shots), len(shots), wp,
                                                int(ckpt_k),
                                                seis.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return self._split(seis, shots)

    def gradient_checkpointed(self, adj: List[np.ndarray]):
        """dL/dvp, dL/drho from the checkpoints of the last forward_checkpointed
        call, for adjoint sources adj of its shots (the bits of gradient).
        checkpoint_id names the call whose checkpoints are held."""
        a = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float32).ravel() for x in adj]))
        gvp = np.zeros((self.grid.nx, self.grid.nz), np.float32)
        grho = np.zeros_like(gvp)
        fp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        _check(lib().sam2d_gradient_checkpointed(self._h, fp(a), fp(gvp), fp(grho)))
        return gvp, grho

    def tuning(self):
        """ms per step of every variant from the last autotune (0 = not timed)."""
        n = lib().sam2d_num_variants()
        out = (ctypes.c_double * n)()
        lib().sam2d_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        _check(lib().sam2d_get_tuning(self._h, out, n))
        return list(out)

    def autotune(self, windows=None, nshots=1, steps=8):
        best = ctypes.c_int(-1)
        w, wp = _windows_c(windows, nshots) if windows is not None else (None, None)
        _check(lib().sam2d_autotune(self._h, wp, nshots, steps, ctypes.byref(best)))
        return best.value


def variants():
    L = lib()
    return [L.sam2d_variant_name(i).decode() for i in range(L.sam2d_num_variants())]


def coefficients(grid: Grid, vp, rho, wx0=0, w=None):
    """Host-side FP32 fields the kernels read (no device needed): dict of
    KX, KZ, BX, BZ as (nnx, nnz) and prof_x (4, nnx), prof_z (4, nnz)."""
    L = lib()
    w = grid.nx - wx0 if w is None else w
    nnx, nnz = w + 2 * grid.npml, grid.nz + 2 * grid.npml
    vp = np.ascontiguousarray(vp, np.float32)
    rho = np.ascontiguousarray(rho, np.float32)
    out = {k: np.zeros((nnx, nnz), np.float32) for k in ("KX", "KZ", "BX", "BZ")}
    out["prof_x"] = np.zeros((4, nnx), np.float32)
    out["prof_z"] = np.zeros((4, nnz), np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    g = grid._c()
    L.sam2d_coefficients.argtypes = [ctypes.POINTER(_Grid)] + [ctypes.POINTER(ctypes.c_float)] * 2 + [
        ctypes.c_int, ctypes.c_int] + [ctypes.POINTER(ctypes.c_float)] * 6
    _check(L.sam2d_coefficients(ctypes.byref(g), fp(vp), fp(rho), wx0, w, fp(out["KX"]), fp(out["KZ"]),
                                fp(out["BX"]), fp(out["BZ"]), fp(out["prof_x"]), fp(out["prof_z"])))
    return out


def source_increments(grid: Grid, vp, rho, sx, sz):
    """Per-step source increments fd_dt*w[it]*kappa_s/(dx*dz) as float32 (host only)."""
    L = lib()
    nts = time_axis(grid)[3]
    amp = np.zeros(max(nts, 1), np.float32)
    vp = np.ascontiguousarray(vp, np.float32)
    rho = np.ascontiguousarray(rho, np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    g = grid._c()
    L.sam2d_source_increments.argtypes = [ctypes.POINTER(_Grid)] + [ctypes.POINTER(ctypes.c_float)] * 2 + [
        ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
    _check(L.sam2d_source_increments(ctypes.byref(g), fp(vp), fp(rho), sx, sz, fp(amp)))
    return amp[:nts]
