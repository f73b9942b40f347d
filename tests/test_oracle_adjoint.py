"""Pins the FP64 oracle adjoint (orc_gradient): transposition identities of
the difference operators and central finite differences of the functional
J = sum adj * p in random model directions (10 seeded directions)."""
import numpy as np
import pytest


def test_operator_transposes_exact(oracle):
    rng = np.random.default_rng(3)
    for n in (4, 5, 9, 33):
        f, g = rng.standard_normal(n), rng.standard_normal(n)
        # <D f, g> = <f, D^T g> for the full operators, edges included
        assert abs(np.dot(oracle.dminus(f), g) - np.dot(f, oracle.dminus_T(g))) < 1e-12
        assert abs(np.dot(oracle.dplus(f), g) - np.dot(f, oracle.dplus_T(g))) < 1e-12
        # matrix form, column by column
        I = np.eye(n)
        Dm = np.stack([oracle.dminus(I[k]) for k in range(n)], 1)
        DmT = np.stack([oracle.dminus_T(I[k]) for k in range(n)], 1)
        np.testing.assert_array_equal(DmT, Dm.T)


def _small_case(seed=4, nx=36, nz=28, T=0.16):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, nx)[:, None]; z = np.linspace(0, 1, nz)[None, :]
    vp = (1500 + 300 * z + 60 * np.sin(6 * x) + 15 * rng.standard_normal((nx, nz))).astype(np.float32)
    rho = (1600 + 200 * z + 40 * np.cos(5 * x) + 10 * rng.standard_normal((nx, nz))).astype(np.float32)
    return vp, rho


def _geometry(oracle, P, nx, dx=10.0):
    rec = np.stack([np.arange(0, nx) * dx, np.full(nx, 30.0)], 1)
    rec2 = np.stack([np.full(8, 30.0), np.arange(8) * 30.0 + 5], 1)   # a vertical array
    src = [[150.0, 50.0], [280.0, 120.0]]
    return src, [rec, rec2]


@pytest.mark.parametrize("npml", [6, 12])
def test_gradient_matches_central_differences(oracle, npml):
    nx, nz = 36, 28
    vp, rho = _small_case(nx=nx, nz=nz)
    P = oracle.params(nx, nz, 10.0, 10.0, 0.16, 0.001, 0.002, 25.0, npml)
    src, recs = _geometry(oracle, P, nx)
    fd_nt, real_nt = oracle.shape(P)[:2]
    rng = np.random.default_rng(11)
    adj = [rng.standard_normal((len(r), real_nt)).astype(np.float32) for r in recs]
    # Model values stay inside one binade, [1024, 2048), where float32
    # spacing is 2^-13: perturbations on a 2^-8 grid make vp +- d and
    # rho +- d exact, so the applied step is exactly d.  Steps of ~0.03 keep
    # the O(d^2) term of the central difference below 1e-6 relative.
    assert 1100 <= min(vp.min(), rho.min()) and max(vp.max(), rho.max()) < 2000
    # The reference forms b = 1/rho and its staggered averages in float; a
    # 0.03 step in rho moves b by only ~150 float ulps, so the functional is
    # differenced with that chain in FP64 (the adjoint formulas are the same).
    oracle.set_f64_buoyancy(True)
    try:
        gvp, grho = oracle.gradient(P, vp, rho, src, recs, adj)
        assert np.abs(gvp).max() > 0 and np.abs(grho).max() > 0
        zero = np.zeros_like(vp)
        for k in range(10):
            dv = (np.round(rng.standard_normal(vp.shape) * 8) / 256).astype(np.float32)
            dr = (np.round(rng.standard_normal(rho.shape) * 8) / 256).astype(np.float32)
            for m, d in ((vp, dv), (rho, dr)):
                assert np.array_equal((m + d) - m, d) and np.array_equal(m - (m - d), d)
            # vp alone, rho alone (each relative 1e-6), then both together,
            # measured against the size of the two parts (they can cancel).
            parts = []
            for a, b in ((dv, zero), (zero, dr), (dv, dr)):
                Jp = oracle.functional(P, vp + a, rho + b, src, recs, adj)
                Jm = oracle.functional(P, vp - a, rho - b, src, recs, adj)
                fd = (Jp - Jm) / 2.0
                an = np.sum(gvp * a) + np.sum(grho * b)
                scale = abs(an) if len(parts) < 2 else sum(abs(p) for p in parts)
                assert abs(fd - an) <= 1e-6 * scale, (k, len(parts), fd, an)
                parts.append(an)
    finally:
        oracle.set_f64_buoyancy(False)


def test_gradient_float_buoyancy_agrees_loosely(oracle):
    # With the reference's FP32 buoyancy the same directions agree to the
    # float-rounding noise of b (~1e-3 relative at these steps).
    nx, nz = 36, 28
    vp, rho = _small_case(nx=nx, nz=nz)
    P = oracle.params(nx, nz, 10.0, 10.0, 0.16, 0.001, 0.002, 25.0, 8)
    src, recs = _geometry(oracle, P, nx)
    real_nt = oracle.shape(P)[1]
    rng = np.random.default_rng(12)
    adj = [rng.standard_normal((len(r), real_nt)).astype(np.float32) for r in recs]
    gvp, grho = oracle.gradient(P, vp, rho, src, recs, adj)
    dr = (np.round(rng.standard_normal(rho.shape) * 64) / 256).astype(np.float32)
    fd = (oracle.functional(P, vp, rho + dr, src, recs, adj)
          - oracle.functional(P, vp, rho - dr, src, recs, adj)) / 2
    an = np.sum(grho * dr)
    assert abs(fd - an) <= 1e-2 * abs(an)


@pytest.mark.parametrize("k", [1, 7, 32, 10000])
def test_checkpointed_gradient_equals_full_storage(oracle, k):
    # Segments recomputed from checkpoints every k steps repeat the first
    # sweep's steps exactly: the gradient is the same bits for every k
    # (k = 1: every state a checkpoint; k >= fd_nt: one segment).
    nx, nz = 36, 28
    vp, rho = _small_case(nx=nx, nz=nz)
    P = oracle.params(nx, nz, 10.0, 10.0, 0.16, 0.001, 0.002, 25.0, 8)
    src, recs = _geometry(oracle, P, nx)
    real_nt = oracle.shape(P)[1]
    rng = np.random.default_rng(13)
    adj = [rng.standard_normal((len(r), real_nt)).astype(np.float32) for r in recs]
    full = oracle.gradient(P, vp, rho, src, recs, adj, padded=True)
    ck = oracle.gradient(P, vp, rho, src, recs, adj, padded=True, ckpt_k=k)
    for a, b in zip(full[:2], ck[:2]):
        assert np.array_equal(a, b)
    for f in ("kappa", "bu1", "bu2"):
        assert np.array_equal(full[2][f], ck[2][f])
