"""Synthetic earth models and acquisition layouts of the benchmark workloads.

All models are float32, x-major (nx, nz) -- value (ix, iz) at [ix, iz], the
layout of the reference's model files (src/src_cpu/model.cpp:37-40).  Seeds
are fixed, so every call returns the same arrays (numpy's PCG64 Generator).
Coordinates are metres, (x, z) pairs as crd2d_t (src/src_cpu/modelling.h:35-38).
"""
import numpy as np


def two_layer(nx, nz, iz_interface=None, v=(1500.0, 3000.0), r=(1000.0, 2000.0)):
    """vp = v[0] and rho = r[0] above row iz_interface (default nz // 2),
    v[1] and r[1] from it down."""
    iz = nz // 2 if iz_interface is None else iz_interface
    vp = np.full((nx, nz), v[0], np.float32)
    vp[:, iz:] = v[1]
    rho = np.full((nx, nz), r[0], np.float32)
    rho[:, iz:] = r[1]
    return vp, rho


def surface_spread_shots(nx, h, nshots, zs=20.0):
    """nshots sources evenly spaced along x at depth zs, sx_k = h*floor((k+1/2)*nx/nshots),
    each with nx receivers at every cell of the surface row z = zs."""
    rec = np.stack([np.arange(nx) * h, np.full(nx, zs)], 1).astype(np.float32)
    return [(h * int((k + 0.5) * nx / nshots), zs, rec) for k in range(nshots)]


def salt_model(nx=3201, nz=1201, h=7.62, seed=2026):
    """Water to z = 1500 m, then 1800 + 0.5 (z - 1500) m/s, a seeded random
    polygon salt body at 4500 m/s; Gardner density rho = 310 vp^0.25."""
    rng = np.random.default_rng(seed)
    z = (np.arange(nz) * h)[None, :]
    x = (np.arange(nx) * h)[:, None]
    vp = np.where(z < 1500.0, 1500.0, 1800.0 + 0.5 * (z - 1500.0)) + 0.0 * x
    cx, cz = nx * h * rng.uniform(0.35, 0.65), nz * h * rng.uniform(0.55, 0.7)
    R = 0.18 * nx * h
    ang = np.arctan2(z - cz, x - cx)
    rad = R * np.ones_like(ang)
    for k in range(1, 6):
        rad += R * rng.uniform(-0.12, 0.12) * np.cos(k * ang + rng.uniform(0, 2 * np.pi))
    rad *= np.where(z < cz, 1.0, 0.6)            # flatter base
    vp = np.where(np.hypot(x - cx, z - cz) < rad, 4500.0, vp)
    vp = vp.astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    return vp, rho


def salt_shots(n, nx=3201, h=7.62, seed=7, zr=15.24):
    """n shots with seeded aperture windows 800..2400 cells wide, centred on
    evenly spaced sources and clamped to the model; receivers at every cell
    of the window (at the cell centres, x = (ix + 1/2) h, so the float
    truncation of x/h lands on ix).  Returns [(sx, sz, rec_xz)] and the
    windows [(first column, width)]."""
    rng = np.random.default_rng(seed)
    out, wins = [], []
    for k in range(n):
        w = int(rng.integers(800, 2401))
        cx = int((k + 0.5) * nx / n)
        wx0 = min(max(cx - w // 2, 0), nx - w)
        rec = np.stack([(wx0 + np.arange(w)) * h + 0.5 * h, np.full(w, zr)], 1).astype(np.float32)
        out.append(((cx + 0.5) * h, zr, rec))
        wins.append((wx0, w))
    return out, wins


def anomaly_model(nx=1201, nz=601, h=10.0, seed=3):
    """Background vp = 2000 + 0.6 z plus six seeded Gaussian anomalies
    (+-300 m/s, sigma 150 m); rho = 2000.  Returns (true vp, background vp, rho)."""
    rng = np.random.default_rng(seed)
    z = (np.arange(nz) * h)[None, :]
    x = (np.arange(nx) * h)[:, None]
    bg = (2000.0 + 0.6 * z + 0.0 * x)
    true = bg.copy()
    for k in range(6):
        cx, cz = rng.uniform(0.15, 0.85) * nx * h, rng.uniform(0.25, 0.75) * nz * h
        true += rng.choice([-300.0, 300.0]) * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * 150.0 ** 2))
    rho = np.full((nx, nz), 2000.0, np.float32)
    return true.astype(np.float32), bg.astype(np.float32), rho
