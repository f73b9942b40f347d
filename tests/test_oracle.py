"""Pins the FP64 CPU oracle (oracle/sam2d_oracle.c).

Each function of the oracle is checked by something other than a restatement
of its own formula: the compiled reference's own output (golden seismograms),
closed forms, exactness of the difference operators on polynomials, and
invariants of the discrete scheme (linearity, exact mirror symmetry).
"""
import math

import numpy as np
import pytest

from conftest import golden_cases, load_golden


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference_bitwise(oracle, case):
    g = load_golden(case)
    P = oracle.params(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                      g["fm"], 32)
    seis = oracle.forward(P, g["vp"], g["rho"], g["src_xz"], [g["rec_xz"]])[0]
    seis = seis[g["trace_idx"]]
    assert seis.shape == g["seis"].shape
    # Same FP64 operation order as the compiled reference: identical bits.
    assert np.array_equal(seis.view(np.int32), g["seis"].view(np.int32))


def test_time_params_against_reference_log_values(oracle):
    # Values printed by the reference executable ("fd_nt", "real_nt",
    # "DT factor", "nts") for the cards in tests/golden/make_goldens.py.
    assert oracle.shape(oracle.params(100, 80, 10, 10, 0.3, 0.001, 0.001, 15)) == (
        300, 300, 1, 260, 164, 144)
    # Float quotients: 0.45 / 0.0008f = 562.5 -> 563 steps; ceil(0.45/0.0016) = 282.
    tp = oracle.time_params(0.45, 0.0008, 0.0016, 12, 10, 10)
    assert (tp["fd_nt"], tp["real_nt"], tp["dt_factor"]) == (563, 282, 2)
    # nts = (int)(5 td / dt) + 1 with td = sqrt(6)/(pi f): 15 Hz, 1 ms -> 260.
    td = math.sqrt(6.0) / (math.pi * 15.0)
    assert oracle.time_params(2.0, 0.001, 0.001, 15, 10, 10)["nts"] == int(5 * td / 0.001) + 1
    # dtdx is the float product fd_dt * (1/dx) widened to double.
    tp = oracle.time_params(1.0, 0.001, 0.001, 15, 7.62, 10.0)
    f = np.float32
    assert tp["dtdx"] == float(f(0.001) * (f(1.0) / f(7.62)))
    assert tp["dtdz"] == float(f(0.001) * (f(1.0) / f(10.0)))


def test_ricker_closed_form(oracle):
    fm, dt, nts = 15.0, 0.001, 260
    w = oracle.ricker(fm, dt, nts)
    t0 = 1.5 * math.sqrt(6.0) / (math.pi * fm)
    t = np.arange(nts, dtype=np.float32) * np.float32(dt)
    a = math.pi * fm * (t.astype(np.float64) - t0)
    expect = (1 - 2 * a * a) * np.exp(-a * a)
    np.testing.assert_allclose(w, expect, rtol=1e-6, atol=1e-12)
    # Peak of the Ricker at t0 = 1.5 sqrt(6)/(pi f) with value 1.
    i0 = int(round(t0 / dt))
    assert abs(w[i0] - 1.0) < 1e-3 and np.argmax(w) == i0


@pytest.mark.parametrize("npml,n", [(32, 50), (40, 7), (5, 3)])
def test_sponge_profiles_closed_form(oracle, npml, n):
    fac = 0.05
    da, da1 = oracle.pml("a", fac, npml, n)
    db, db1 = oracle.pml("b", fac, npml, n)
    L = n + 2 * npml
    i = np.arange(L)
    # pressure nodes: distance into the sponge in cells
    xa = np.where(i < npml, npml - 1 - i, np.where(i >= npml + n, i - npml - n, 0)).astype(float)
    # velocity nodes (half-integer positions)
    xb = np.where(i < npml - 1, npml - 1.5 - i, np.where(i >= npml + n, i - npml - n + 0.5, 0.0))
    ina = (i >= npml) & (i < npml + n)
    inb = (i >= npml - 1) & (i < npml + n)
    np.testing.assert_allclose(da, np.where(ina, 1.0, np.exp(-(fac * xa) ** 2)), rtol=6e-7)
    np.testing.assert_allclose(da1, np.where(ina, 1.0, np.exp(-0.5 * (fac * xa) ** 2)), rtol=6e-7)
    np.testing.assert_allclose(db, np.where(inb, 1.0, np.exp(-(fac * xb) ** 2)), rtol=6e-7)
    np.testing.assert_allclose(db1, np.where(inb, 1.0, np.exp(-0.5 * (fac * xb) ** 2)), rtol=6e-7)
    # scale = sqrt(decay)
    np.testing.assert_allclose(da1.astype(float) ** 2, da, rtol=1e-6)


def test_expand_replicates_nearest_interior_cell(oracle):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 4)).astype(np.float32)
    P = 3
    b = oracle.expand(a, P)
    assert b.shape == (11, 10)
    for ix in range(11):
        for iz in range(10):
            mx = min(max(ix - P, 0), 4)
            mz = min(max(iz - P, 0), 3)
            assert b[ix, iz] == a[mx, mz]


def test_stagger_rules(oracle):
    rng = np.random.default_rng(1)
    vpe = rng.uniform(1500, 4500, (7, 6)).astype(np.float32)
    rhoe = rng.uniform(1000, 2500, (7, 6)).astype(np.float32)
    rhoe[2, 3] = 0.0  # non-positive density -> buoyancy 1e10
    kappa, bue, bu1, bu2 = oracle.stagger(vpe, rhoe)
    np.testing.assert_array_equal(kappa, vpe.astype(float) ** 2 * rhoe.astype(float))
    b = np.where(rhoe > 0, np.float32(1) / np.where(rhoe > 0, rhoe, 1), np.float32(1e10)).astype(np.float32)
    np.testing.assert_array_equal(bue, b)
    # bu1: average along z; last z-row 0 (copied from bu2 before it exists)
    np.testing.assert_array_equal(bu1[:, :-1], np.float32(0.5) * (b[:, 1:] + b[:, :-1]))
    assert np.all(bu1[:, -1] == 0)
    # bu2: average along x; last x-column = bu1's column nnx-2
    np.testing.assert_array_equal(bu2[:-1, :], np.float32(0.5) * (b[1:, :] + b[:-1, :]))
    np.testing.assert_array_equal(bu2[-1, :], bu1[-2, :])


def test_difference_operators_exact_on_cubics(oracle):
    n = 40
    x = np.arange(n, dtype=np.float64)
    a0, a1 = 1.125, float(np.float32(-1.0) / np.float32(24.0))
    for coeffs in [(1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1), (3, -2, 0.5, 0.25)]:
        f = sum(c * x ** k for k, c in enumerate(coeffs))
        df = lambda y: sum(k * c * y ** (k - 1) for k, c in enumerate(coeffs) if k > 0)
        dm = oracle.dminus(f)
        dp = oracle.dplus(f)
        # 4th order interior: derivative at the half point, exact for cubics up
        # to the float rounding of a1 = -1/24 (relative 2e-9).
        im = np.arange(2, n - 1)
        ip = np.arange(1, n - 2)
        np.testing.assert_allclose(dm[im], df(im - 0.5) + (a1 + 1 / 24) * 0, rtol=1e-7, atol=1e-6 * abs(f).max())
        np.testing.assert_allclose(dp[ip], df(ip + 0.5), rtol=1e-7, atol=1e-6 * abs(f).max())
        # edges: 2nd order differences and zeros
        assert dm[0] == 0 and dm[1] == f[1] - f[0] and dm[n - 1] == f[n - 1] - f[n - 2]
        assert dp[n - 1] == 0 and dp[0] == f[1] - f[0] and dp[n - 2] == f[n - 1] - f[n - 2]
    # consistency: a0 + 3 a1 = 1 up to the float rounding of a1
    assert abs(a0 + 3 * a1 - 1.0) < 1e-8


def test_operators_are_negative_transposes_in_the_interior(oracle):
    # D- and D+ on a line: <D- f, g> = -<f, D+ g> for f, g vanishing near the
    # ends (summation by parts), which fixes the stencil signs and offsets.
    n = 30
    rng = np.random.default_rng(2)
    f = np.zeros(n); g = np.zeros(n)
    f[5:25] = rng.standard_normal(20); g[5:25] = rng.standard_normal(20)
    lhs = np.dot(oracle.dminus(f), g)
    rhs = -np.dot(f, oracle.dplus(g))
    assert abs(lhs - rhs) < 1e-12 * (1 + abs(lhs))


def _homog(oracle, nx=81, nz=61, T=0.03):
    vp = np.full((nx, nz), 2000, np.float32)
    rho = np.full((nx, nz), 1800, np.float32)
    P = oracle.params(nx, nz, 10, 10, T, 0.001, 0.001, 20, 32)
    return P, vp, rho


def test_forward_linear_in_source_amplitude(oracle):
    P, vp, rho = _homog(oracle, T=0.1)
    rec = np.stack([np.arange(0, 810, 10.0), np.full(81, 100.0)], 1)
    s1 = oracle.forward(P, vp, rho, [[400, 300]], [rec], src_amp=1.0)[0]
    s2 = oracle.forward(P, vp, rho, [[400, 300]], [rec], src_amp=2.0)[0]
    # A power-of-two scale commutes exactly with every FP64 rounding step; the
    # final cast to float32 is exact too, except for subnormal samples.
    normal = np.abs(s1) >= np.finfo(np.float32).tiny
    np.testing.assert_array_equal(s2[normal], 2 * s1[normal])
    assert np.abs(s2 - 2 * s1).max() <= 2 * np.finfo(np.float32).smallest_subnormal


def test_forward_mirror_symmetry(oracle):
    # Homogeneous medium, source on the centre column: the pressure is
    # mirror-symmetric in x (the operators' differences negate exactly under
    # reflection) while the wavefield has not reached the asymmetric last
    # velocity column; after that, only sponge-damped terms differ.
    nx = 81
    P, vp, rho = _homog(oracle, nx=nx, T=0.02)
    xs = 40 * 10.0
    rec = np.stack([np.arange(nx) * 10.0, np.full(nx, 200.0)], 1)
    s = oracle.forward(P, vp, rho, [[xs, 300]], [rec])[0]
    np.testing.assert_array_equal(s, s[::-1])
    P2, vp, rho = _homog(oracle, nx=nx, T=0.3)
    s = oracle.forward(P2, vp, rho, [[xs, 300]], [rec])[0].astype(np.float64)
    assert np.abs(s - s[::-1]).max() <= 1e-6 * np.abs(s).max()


def _two_shot(oracle, g):
    P = oracle.params(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                      g["fm"], 32, reset_quirk=1)
    return oracle.forward(P, g["vp"], g["rho"], g["src_xz"], [g["rec_xz"], g["rec_xz"]])


def test_two_shot_carryover_matches_reference_bitwise(oracle):
    # The compiled reference ran two shots on one card; the output file it
    # left holds shot 2, which started from shot 1's final state with only the
    # first nnz cells (model column 0 of the sponge) cleared
    # (cpumodelling.cpp:343-348, 367-374).  The oracle's reset_quirk mode gives
    # the same bits; the fresh-state mode does not.
    g = load_golden("two_shot_carryover")
    assert np.array_equal(_two_shot(oracle, g)[1].view(np.int32), g["seis"].view(np.int32))
    P = oracle.params(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                      g["fm"], 32)
    fresh = oracle.forward(P, g["vp"], g["rho"], g["src_xz"], [g["rec_xz"], g["rec_xz"]])[1]
    assert np.abs(fresh - g["seis"]).max() > 0.05 * np.abs(g["seis"]).max()


@pytest.mark.parametrize("cells,fields,rejected", [
    (None, 0b00110, True),      # clears px, pz everywhere, keeps vx, vz
    (None, 0b11000, True),      # clears vx, vz everywhere, keeps px, pz
    ("N", 31, True),            # fresh state
    ("half", 31, True),         # half the grid
    ("nnx", 31, False),         # nnx cells: still inside the damped sponge
    (0, 31, False),             # nothing cleared
])
def test_two_shot_golden_rejects_planted_resets(oracle, cells, fields, rejected):
    # Slips planted in the oracle's reset.  The ones that touch the interior
    # change shot 2 and the reference's file rejects them.  Which sponge cells
    # of column 0 are cleared cannot be seen at all: a state there is damped
    # by the sponge (factor exp(-(0.05*31)^2) ~ 0.09 per step at column 0,
    # more than 150 orders of magnitude before it reaches column 32) so the
    # output is the same bits either way, and the reference's nnz-cell rule is
    # pinned exactly as far as it is observable.
    g = load_golden("two_shot_carryover")
    nnz, nnx = g["nz"] + 64, g["nx"] + 64
    n = {None: -1, "N": nnz * nnx, "half": nnz * nnx // 2, "nnx": nnx}.get(cells, cells)
    if fields != 31:
        n = nnz * nnx
    oracle.set_reset_plant(n, fields)
    try:
        s2 = _two_shot(oracle, g)[1]
    finally:
        oracle.set_reset_plant()
    same = np.array_equal(s2.view(np.int32), g["seis"].view(np.int32))
    assert same != rejected


def test_reset_quirk_only_matters_after_first_shot(oracle):
    P, vp, rho = _homog(oracle, T=0.05)
    rec = np.stack([np.arange(0, 810, 10.0), np.full(81, 100.0)], 1)
    shots = [[300, 300], [500, 300]]
    fresh = oracle.forward(P, vp, rho, shots, [rec, rec])
    P.reset_quirk = 1
    quirk = oracle.forward(P, vp, rho, shots, [rec, rec])
    np.testing.assert_array_equal(fresh[0], quirk[0])
    assert not np.array_equal(fresh[1], quirk[1])
