"""CPU tests of the product library's host side: the C ABI exports, and the
FP32 coefficient fields / source increments it builds, against the FP64
oracle's setup (which is pinned to the reference in test_oracle.py)."""
import ctypes
import os
import re

import numpy as np
import pytest

import seisacomod2d_amd as sam
from conftest import ROOT, load_golden


def declared_symbols(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = txt.split("#ifdef __cplusplus\n}")[0]          # the C part
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(\w+)\s*\(", txt, flags=re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    # include/sam2d.h is libsam2d.so's ABI, include/sam2d_modelling.h
    # libsam2d_host.so's; each library exports every function its header
    # declares (loading them needs no GPU).
    L = sam.lib()
    syms = declared_symbols("sam2d.h")
    assert "sam2d_forward" in syms and "sam2d_gradient" in syms and "sam2d_create" in syms
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    H = ctypes.CDLL(os.path.join(ROOT, "seisacomod2d_amd", "libsam2d_host.so"))
    hsyms = declared_symbols("sam2d_modelling.h")
    assert "sam2d_modelling_module" in hsyms and "write_shot_output" in hsyms
    missing = [s for s in hsyms if not hasattr(H, s)]
    assert not missing, missing
    assert hasattr(H, "_Z16modelling_moduleiP5Geo2d")      # the C++ seam


def test_version_and_variants():
    L = sam.lib()
    assert b"gfx950" in L.sam2d_version()
    v = sam.variants()
    assert len(v) == L.sam2d_num_variants() >= 4
    prefixes = ("twostep", "threestep", "fourstep", "fivestep")
    assert all(name.startswith(("waves",) + tuple(p + "_waves" for p in prefixes)) for name in v)
    for p in prefixes:
        assert any(name.startswith(p) for name in v), p
    assert len(set(v)) == len(v)


def _grid_from(g, npml=32):
    return sam.Grid(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                    g["fm"], npml)


@pytest.mark.parametrize("case,npml", [("hetero_explicit", 32), ("rect_dtfactor", 32),
                                       ("salt_small", 40), ("tiny_twolayer", 7)])
def test_coefficients_match_oracle_setup(oracle, case, npml):
    g = load_golden(case)
    grid = _grid_from(g, npml)
    c = sam.coefficients(grid, g["vp"], g["rho"])
    P = oracle.params(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                      g["fm"], npml)
    o = oracle.setup(P, g["vp"], g["rho"])
    # Profiles are FP32 in the reference: identical.
    np.testing.assert_array_equal(c["prof_x"][0], o["damp2a"])
    np.testing.assert_array_equal(c["prof_x"][1], o["damp2a1"])
    np.testing.assert_array_equal(c["prof_x"][2], o["damp2b"])
    np.testing.assert_array_equal(c["prof_x"][3], o["damp2b1"])
    np.testing.assert_array_equal(c["prof_z"][0], o["damp1a"])
    np.testing.assert_array_equal(c["prof_z"][1], o["damp1a1"])
    np.testing.assert_array_equal(c["prof_z"][2], o["damp1b"])
    np.testing.assert_array_equal(c["prof_z"][3], o["damp1b1"])
    # The kernels form kappaw2 = KX * damp2a1[ix] etc.  Where the profile is 1
    # (interior) KX is the reference's FP64 coefficient rounded to one of the
    # two floats around it (unbiased rounding, test below); in the sponge the
    # product is within 2 float ulps.
    px1, pz1 = c["prof_x"][1][:, None], c["prof_z"][1][None, :]
    bx1, bz1 = c["prof_x"][3][:, None], c["prof_z"][3][None, :]
    for ours, ref in [(c["KX"] * px1, o["kappaw2"]), (c["KZ"] * pz1, o["kappaw1"]),
                      (c["BX"] * bx1, o["buw2"]), (c["BZ"] * bz1, o["buw1"])]:
        inner = (px1 == 1) & (pz1 == 1) & (bx1 == 1) & (bz1 == 1)
        inner = np.broadcast_to(inner, ref.shape)
        lo, hi = _bracket(ref[inner])
        v = ours[inner]
        assert np.all((v == lo) | (v == hi))
        np.testing.assert_allclose(ours, ref, rtol=2.5e-7, atol=0)


def _bracket(t):
    """The two floats around each FP64 value (equal where t is a float)."""
    t = np.asarray(t, np.float64)
    n = t.astype(np.float32)
    lo = np.where(n.astype(np.float64) > t, np.nextafter(n, np.float32(-np.inf)), n)
    hi = np.where(lo.astype(np.float64) == t, lo, np.nextafter(lo, np.float32(np.inf)))
    return lo, hi


@pytest.mark.parametrize("vp,rho", [(1500.0, 310.0 * 1500.0 ** 0.25), (2750.5, 1333.3)])
def test_coefficient_rounding_is_unbiased(oracle, vp, rho):
    # A homogeneous model: every interior cell has the same FP64 coefficient
    # t.  Round-to-nearest would store the same float everywhere (a
    # systematic error of up to 2^-24 relative); the fields hold the upper
    # neighbour of t in a fraction (t - lo)/(hi - lo) of the cells, so the
    # mean stored value is t.  Binomial bounds at 5 sigma.
    nx, nz = 300, 200
    grid = sam.Grid(nx, nz, 7.62, 7.62, 0.1, 0.0005, 0.0005, 15.0, 32)
    vpa = np.full((nx, nz), vp, np.float32)
    rhoa = np.full((nx, nz), rho, np.float32)
    c = sam.coefficients(grid, vpa, rhoa)
    P = oracle.params(nx, nz, 7.62, 7.62, 0.1, 0.0005, 0.0005, 15.0, 32)
    o = oracle.setup(P, vpa, rhoa)
    sl = (slice(40, -40), slice(40, -40))           # interior, away from the edge rows
    for ours, ref in [(c["KX"], o["kappaw2"]), (c["BX"], o["buw2"]), (c["BZ"], o["buw1"])]:
        t = ref[sl]
        assert np.all(t == t.flat[0])
        lo, hi = _bracket(t.flat[0])
        assert lo != hi
        frac = (t.flat[0] - float(lo)) / (float(hi) - float(lo))
        v = ours[sl]
        assert np.all((v == lo) | (v == hi))
        n = v.size
        up = np.count_nonzero(v == hi)
        assert abs(up - frac * n) <= 5 * np.sqrt(n * frac * (1 - frac)) + 1, (up, frac * n)
        mean_err = (v.astype(np.float64).mean() - t.flat[0]) / (float(hi) - float(lo))
        assert abs(mean_err) <= 5 * 0.5 / np.sqrt(n)
        # neighbouring cells are not all rounded alike (no stripes along z or x)
        assert 0 < np.count_nonzero(v[:, 0] == hi) < v.shape[0]
        assert 0 < np.count_nonzero(v[0, :] == hi) < v.shape[1]


def test_source_increments_match_reference_formula(oracle):
    g = load_golden("hetero_explicit")
    grid = _grid_from(g)
    sx, sz = g["src_xz"][0]
    amp = sam.source_increments(grid, g["vp"], g["rho"], sx, sz)
    P = oracle.params(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                      g["fm"], 32)
    o = oracle.setup(P, g["vp"], g["rho"])
    f = np.float32
    ix, iz = int(f(sx) * (f(1) / f(grid.dx))), int(f(sz) * (f(1) / f(grid.dz)))
    k = o["kappa"][ix + 32, iz + 32]
    expect = (np.float64(f(grid.fd_dt)) * o["wlt"] * k / np.float64(f(grid.dx) * f(grid.dz))).astype(np.float32)
    np.testing.assert_array_equal(amp, expect)


def test_windowed_coefficients_are_the_cropped_model(oracle):
    g = load_golden("salt_small")
    grid = _grid_from(g)
    wx0, w = 37, 90
    c = sam.coefficients(grid, g["vp"], g["rho"], wx0, w)
    grid_w = sam.Grid(w, g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"], g["fm"], 32)
    cw = sam.coefficients(grid_w, g["vp"][wx0:wx0 + w], g["rho"][wx0:wx0 + w])
    for k in c:
        np.testing.assert_array_equal(c[k], cw[k])


def test_bad_arguments_return_status_not_abort():
    g = sam.Grid(10, 10, 10, 10, 0.1, 0.001, 0.001, 10)
    vp = np.full((10, 10), 1500, np.float32)
    with pytest.raises(sam.SAM2DError) as e:
        sam.coefficients(g, vp, vp, 5, 10)
    assert e.value.code == -6
    L = sam.lib()
    gc = g._c()
    gc.dx = 10.0
    # CFL: dt > 0.606 h / vmax -> -5 (check_stability.cpp:30-43)
    gc.fd_dt = 0.0041
    assert L.sam2d_check_stability(ctypes.byref(gc), 1500.0, 1500.0) == -5
    gc.fd_dt = 0.004
    assert L.sam2d_check_stability(ctypes.byref(gc), 1500.0, 1500.0) == 0
    # dispersion: 2 f > vmin / (5 hmax) -> -5; exactly at the limit passes
    gc.dom_freq = 15.0
    assert L.sam2d_check_stability(ctypes.byref(gc), 1500.0, 1500.0) == 0
    gc.dom_freq = 15.01
    assert L.sam2d_check_stability(ctypes.byref(gc), 1500.0, 1500.0) == -5
