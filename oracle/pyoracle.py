"""ctypes front end of the CPU checkers in oracle/liboracle.so: the FP64
oracle of the reference (sam2d_oracle.c) and the FP32 model of the product's
kernels (sam2d_fp32.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker.  The product package
never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float),
        ("rec_len", ctypes.c_float), ("fd_dt", ctypes.c_float),
        ("real_dt", ctypes.c_float), ("fm", ctypes.c_float),
        ("npml", ctypes.c_int), ("reset_quirk", ctypes.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _LIB = ctypes.CDLL(path)
    return _LIB


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t=ctypes.c_float):
    return a.ctypes.data_as(ctypes.POINTER(t))


def params(nx, nz, dx, dz, rec_len, fd_dt, real_dt, fm, npml=32, reset_quirk=0):
    return OrcParams(nx, nz, dx, dz, rec_len, fd_dt, real_dt, fm, npml, reset_quirk)


def shape(P):
    """(fd_nt, real_nt, dt_factor, nts, nnx, nnz)"""
    out = np.zeros(6, np.int32)
    lib().orc_shape(ctypes.byref(P), _ptr(out, ctypes.c_int))
    return tuple(int(v) for v in out)


def time_params(rec_len, fd_dt, real_dt, fm, dx, dz):
    i = [ctypes.c_int() for _ in range(4)]
    d = [ctypes.c_double() for _ in range(2)]
    L = lib()
    L.orc_time_params.argtypes = [ctypes.c_float] * 6 + [ctypes.POINTER(ctypes.c_int)] * 4 + [
        ctypes.POINTER(ctypes.c_double)] * 2
    L.orc_time_params(rec_len, fd_dt, real_dt, fm, dx, dz, *[ctypes.byref(v) for v in i],
                      *[ctypes.byref(v) for v in d])
    return dict(fd_nt=i[0].value, real_nt=i[1].value, dt_factor=i[2].value, nts=i[3].value,
                dtdx=d[0].value, dtdz=d[1].value)


def expand(a, npml):
    a = _f32(a)
    nx1, nz1 = a.shape
    b = np.zeros((nx1 + 2 * npml, nz1 + 2 * npml), np.float32)
    lib().orc_expand(_ptr(a), ctypes.c_int(nx1), ctypes.c_int(nz1), ctypes.c_int(npml), _ptr(b))
    return b


def stagger(vpe, rhoe):
    vpe, rhoe = _f32(vpe), _f32(rhoe)
    nnx, nnz = vpe.shape
    kappa = np.zeros((nnx, nnz), np.float64)
    bue = np.zeros((nnx, nnz), np.float32)
    bu1 = np.zeros_like(bue)
    bu2 = np.zeros_like(bue)
    lib().orc_stagger(_ptr(vpe), _ptr(rhoe), ctypes.c_int(nnz), ctypes.c_int(nnx),
                      _ptr(kappa, ctypes.c_double), _ptr(bue), _ptr(bu1), _ptr(bu2))
    return kappa, bue, bu1, bu2


def ricker(fm, dt, nts):
    w = np.zeros(nts, np.float64)
    L = lib()
    L.orc_ricker.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.orc_ricker(fm, dt, nts, _ptr(w, ctypes.c_double))
    return w


def pml(kind, fac, npml, n):
    damp = np.zeros(n + 2 * npml, np.float32)
    damp1 = np.zeros_like(damp)
    f = lib().orc_pml_a if kind == "a" else lib().orc_pml_b
    f.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                  ctypes.POINTER(ctypes.c_float)]
    f(fac, npml, n, _ptr(damp), _ptr(damp1))
    return damp, damp1


def dminus(f):
    f = np.ascontiguousarray(f, np.float64)
    out = np.zeros_like(f)
    lib().orc_dminus(_ptr(f, ctypes.c_double), ctypes.c_int(f.size), _ptr(out, ctypes.c_double))
    return out


def dplus(f):
    f = np.ascontiguousarray(f, np.float64)
    out = np.zeros_like(f)
    lib().orc_dplus(_ptr(f, ctypes.c_double), ctypes.c_int(f.size), _ptr(out, ctypes.c_double))
    return out


def setup(P, vp, rho):
    """Coefficient fields and profiles exactly as the reference builds them."""
    fd_nt, real_nt, dtf, nts, nnx, nnz = shape(P)
    vp, rho = _f32(vp), _f32(rho)
    N = nnx * nnz
    fields = [np.zeros((nnx, nnz), np.float64) for _ in range(5)]
    px8 = np.zeros((4, nnx), np.float32)
    pz8 = np.zeros((4, nnz), np.float32)
    wlt = np.zeros(max(nts, 1), np.float64)
    lib().orc_setup(ctypes.byref(P), _ptr(vp), _ptr(rho),
                    *[_ptr(f, ctypes.c_double) for f in fields], _ptr(px8), _ptr(pz8),
                    _ptr(wlt, ctypes.c_double))
    names = ["kappa", "kappaw1", "kappaw2", "buw1", "buw2"]
    out = dict(zip(names, fields))
    out.update(damp2a=px8[0], damp2a1=px8[1], damp2b=px8[2], damp2b1=px8[3],
               damp1a=pz8[0], damp1a1=pz8[1], damp1b=pz8[2], damp1b1=pz8[3], wlt=wlt[:nts])
    return out


def forward(P, vp, rho, src_xz, rec_xz_list, src_amp=1.0):
    """FP64 forward.  src_xz: (ns, 2) metres; rec_xz_list: list of (ng, 2).
    Returns a list of (ng, real_nt) float32 gathers."""
    fd_nt, real_nt, dtf, nts, nnx, nnz = shape(P)
    vp, rho = _f32(vp), _f32(rho)
    assert vp.shape == (P.nx, P.nz) and rho.shape == (P.nx, P.nz)
    src = _f32(src_xz).reshape(-1, 2)
    ns = src.shape[0]
    nrec = np.array([len(r) for r in rec_xz_list], np.int32)
    rec = _f32(np.concatenate([np.asarray(r, np.float32).reshape(-1, 2) for r in rec_xz_list]))
    seis = np.zeros(int(nrec.sum()) * real_nt, np.float32)
    L = lib()
    L.orc_forward.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(ctypes.c_float),
                              ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_float), ctypes.c_double,
                              ctypes.POINTER(ctypes.c_float)]
    rc = L.orc_forward(ctypes.byref(P), _ptr(vp), _ptr(rho), ns, _ptr(src), _ptr(nrec, ctypes.c_int),
                       _ptr(rec), float(src_amp), _ptr(seis))
    if rc != 0:
        raise RuntimeError(f"orc_forward failed: {rc}")
    out, off = [], 0
    for ng in nrec:
        out.append(seis[off:off + ng * real_nt].reshape(ng, real_nt))
        off += ng * real_nt
    return out


def time_steps(P, vp, rho, nsteps):
    """(seconds, threads) of nsteps FP64 reference-structured steps (setup excluded)."""
    vp, rho = _f32(vp), _f32(rho)
    L = lib()
    L.orc_time_steps.restype = ctypes.c_double
    L.orc_time_steps.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(ctypes.c_float),
                                 ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int)]
    th = ctypes.c_int()
    t = L.orc_time_steps(ctypes.byref(P), _ptr(vp), _ptr(rho), int(nsteps), ctypes.byref(th))
    return t, th.value


def dminus_T(g):
    g = np.ascontiguousarray(g, np.float64)
    out = np.zeros_like(g)
    lib().orc_dminus_T(_ptr(g, ctypes.c_double), ctypes.c_int(g.size), _ptr(out, ctypes.c_double))
    return out


def dplus_T(g):
    g = np.ascontiguousarray(g, np.float64)
    out = np.zeros_like(g)
    lib().orc_dplus_T(_ptr(g, ctypes.c_double), ctypes.c_int(g.size), _ptr(out, ctypes.c_double))
    return out


def _geom(src_xz, rec_xz_list, adj_list=None):
    src = _f32(src_xz).reshape(-1, 2)
    nrec = np.array([len(r) for r in rec_xz_list], np.int32)
    rec = _f32(np.concatenate([np.asarray(r, np.float32).reshape(-1, 2) for r in rec_xz_list]))
    adj = None
    if adj_list is not None:
        adj = _f32(np.concatenate([np.asarray(a, np.float32).ravel() for a in adj_list]))
    return src, nrec, rec, adj


def functional(P, vp, rho, src_xz, rec_xz_list, adj_list):
    """J = sum adj * p at the recorded samples (FP64 pressure, double sum)."""
    vp, rho = _f32(vp), _f32(rho)
    src, nrec, rec, adj = _geom(src_xz, rec_xz_list, adj_list)
    L = lib()
    L.orc_functional.restype = ctypes.c_double
    L.orc_functional.argtypes = [ctypes.POINTER(OrcParams)] + [ctypes.POINTER(ctypes.c_float)] * 2 + [
        ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    return L.orc_functional(ctypes.byref(P), _ptr(vp), _ptr(rho), src.shape[0], _ptr(src),
                            _ptr(nrec, ctypes.c_int), _ptr(rec), _ptr(adj))


def gradient(P, vp, rho, src_xz, rec_xz_list, adj_list, padded=False, ckpt_k=0):
    """FP64 dJ/dvp, dJ/drho (nx, nz) of J = sum adj * p; with padded=True also
    the padded-grid dJ/dkappa, dJ/dbu1, dJ/dbu2.  ckpt_k > 0 checkpoints the
    forward state every ckpt_k steps and recomputes the segments (memory
    about (4 fd_nt / ckpt_k + 3 ckpt_k) N doubles instead of 3 N fd_nt); the
    result is the same bits."""
    fd_nt, real_nt, dtf, nts, nnx, nnz = shape(P)
    vp, rho = _f32(vp), _f32(rho)
    src, nrec, rec, adj = _geom(src_xz, rec_xz_list, adj_list)
    gvp = np.zeros((P.nx, P.nz), np.float64)
    grho = np.zeros_like(gvp)
    gk, gb1, gb2 = (np.zeros((nnx, nnz), np.float64) for _ in range(3))
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_gradient_k.argtypes = [ctypes.POINTER(OrcParams)] + [ctypes.POINTER(ctypes.c_float)] * 2 + [
        ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int] + [dp] * 5
    rc = L.orc_gradient_k(ctypes.byref(P), _ptr(vp), _ptr(rho), src.shape[0], _ptr(src),
                          _ptr(nrec, ctypes.c_int), _ptr(rec), _ptr(adj), int(ckpt_k),
                          _ptr(gvp, ctypes.c_double), _ptr(grho, ctypes.c_double),
                          _ptr(gk, ctypes.c_double), _ptr(gb1, ctypes.c_double),
                          _ptr(gb2, ctypes.c_double))
    if rc != 0:
        raise RuntimeError(f"orc_gradient failed: {rc}")
    if padded:
        return gvp, grho, dict(kappa=gk, bu1=gb1, bu2=gb2)
    return gvp, grho


def set_f64_buoyancy(on):
    """Test-only: evaluate 1/rho and the staggered buoyancy averages in FP64."""
    lib().orc_set_f64_buoyancy(ctypes.c_int(1 if on else 0))


def set_reset_plant(cells=-1, fields=31):
    """Test-only: what the reset between shots clears in reset_quirk mode
    (cells < 0: the reference's first nnz cells; fields: mask of p, px, pz,
    vx, vz), to plant a slip that the reference's two-shot output rejects."""
    L = lib()
    L.orc_set_reset_plant.argtypes = [ctypes.c_long, ctypes.c_int]
    L.orc_set_reset_plant(int(cells), int(fields))


def fp32_forward(fields, fd_nt, nts, amp, src_cell, gcells, dt_factor, real_nt, state=None):
    """The product kernels' FP32 arithmetic on the CPU (sam2d_fp32.c).
    fields: KX, KZ, BX, BZ (nnx, nnz) and prof_x (4, nnx), prof_z (4, nnz) as
    the library builds them; amp: source increments; src_cell, gcells: padded
    cell indices ix*nnz + iz.  state: (px, pz, vx, vz) to start from (updated
    in place), default zero, in the kernels' representation (outside the
    sponge frame px holds p and pz is 0).  Returns (seis (ng, real_nt), state)."""
    KX = _f32(fields["KX"])
    nnx, nnz = KX.shape
    f = {k: _f32(fields[k]) for k in ("KZ", "BX", "BZ", "prof_x", "prof_z")}
    if state is None:
        state = tuple(np.zeros((nnx, nnz), np.float32) for _ in range(4))
    g = np.ascontiguousarray(gcells, np.int64)
    seis = np.zeros((len(g), real_nt), np.float32)
    amp = _f32(np.asarray(amp).ravel() if len(amp) else np.zeros(1))
    L = lib()
    fp = ctypes.POINTER(ctypes.c_float)
    L.f32_forward.argtypes = [ctypes.c_int, ctypes.c_int] + [fp] * 6 + [ctypes.c_int, ctypes.c_int, fp,
                              ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_long)] + [fp] * 5
    rc = L.f32_forward(nnx, nnz, _ptr(KX), _ptr(f["KZ"]), _ptr(f["BX"]), _ptr(f["BZ"]),
                       _ptr(f["prof_x"]), _ptr(f["prof_z"]), int(fd_nt), int(nts), _ptr(amp),
                       int(src_cell), int(dt_factor), int(real_nt), len(g), _ptr(g, ctypes.c_long),
                       _ptr(seis), *[_ptr(a) for a in state])
    if rc != 0:
        raise RuntimeError(f"f32_forward failed: {rc}")
    return seis, state
