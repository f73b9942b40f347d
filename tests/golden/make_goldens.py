"""Write golden seismograms with the compiled reference executable.

    make -C oracle ref && python tests/golden/make_goldens.py

Runs oracle/_ref/SeisAcoMod2D_MULTI_CORE_CPU_V1.0.exe (src/src_cpu of the
reference, built from its sources by oracle/Makefile.ref) on small job cards
in a scratch directory, then stores each case's inputs and the reference's
output as tests/golden/<case>.npz:

  params   nx, nz, dx, dz, rec_len, fd_dt, real_dt, fm (float32 as parsed)
  vp, rho  models (nx, nz) float32, x-major (omitted for analytic models)
  src_xz   (1, 2) source in metres
  rec_xz   (ng, 2) receivers in metres, as the reference generated them
           (its test_geometry.txt dump of shot 0, create_geometry.cpp:446-459)
  seis     (ng, real_nt) float32 pressure seismogram (the reference's
           <job>_seismogram.bin), possibly decimated in traces: trace_idx
  geometry_text  the shot-file line or explicit-geometry lines of the card

The reference fixes npml = 32 (cpumodelling.cpp:15), so every golden is at
npml = 32.  Runs are single-shot: multi-shot runs of the reference overwrite
one output file and carry state between shots (cpumodelling.cpp:343-374).
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

sys.path.insert(0, ROOT)
from oracle import refrun  # noqa: E402


def run_reference(nx, nz, dx, dz, T, fd_dt, real_dt, fm, vp, rho, shot_line=None, geom_lines=None,
                  threads=8):
    with tempfile.TemporaryDirectory() as d:
        flag = refrun.write_case(d, nx, nz, dx, dz, vp, rho, shot_line=shot_line,
                                 geom_lines=geom_lines)
        _, r = refrun.run_card(d, nx, nz, dx, dz, T, fd_dt, real_dt, fm, flag, threads)
        seis = np.fromfile(os.path.join(d, "out", "card_seismogram.bin"), np.float32)
        geo = np.loadtxt(os.path.join(d, "test_geometry.txt"), dtype=np.float64, ndmin=2)
        log = r.stdout
    src = geo[0, ::-1].astype(np.float32)[None, :]
    rec = geo[1:, ::-1].astype(np.float32)
    ng = rec.shape[0]
    run_reference.geometry_text = shot_line if shot_line else "\n".join(geom_lines)
    return src, rec, seis.reshape(ng, -1), log


def params_arr(nx, nz, dx, dz, T, fd_dt, real_dt, fm):
    return np.array([nx, nz, dx, dz, T, fd_dt, real_dt, fm], np.float64)


def two_layer(nx, nz, iz_if):
    vp = np.full((nx, nz), 1500, np.float32); vp[:, iz_if:] = 3000
    rho = np.full((nx, nz), 1000, np.float32); rho[:, iz_if:] = 2000
    return vp, rho


def smooth_random(nx, nz, seed):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, nx)[:, None]; z = np.linspace(0, 1, nz)[None, :]
    vp = 1800 + 900 * z + 150 * np.sin(7 * x + 3 * z) + 60 * rng.standard_normal((nx, nz))
    rho = 1700 + 500 * z + 80 * np.cos(5 * x - 4 * z) + 30 * rng.standard_normal((nx, nz))
    return vp.astype(np.float32), rho.astype(np.float32)


def salt(nx, nz, h, seed):
    """Water layer, gradient sediments, a seeded polygon body at 4500 m/s,
    Gardner density rho = 310 vp^0.25."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[None, :] * h
    x = np.arange(nx)[:, None] * h
    zw = 0.25 * nz * h
    vp = np.where(z < zw, 1500.0, 1800.0 + 0.5 * (z - zw)) + 0 * x
    cx, cz, R = 0.5 * nx * h, 0.6 * nz * h, 0.22 * min(nx, nz) * h
    ang = np.arctan2(z - cz, x - cx)
    k = np.arange(1, 4)
    a = rng.uniform(-0.25, 0.25, 3); ph = rng.uniform(0, 2 * np.pi, 3)
    rad = R * (1 + sum(a[i] * np.cos(k[i] * ang + ph[i]) for i in range(3)))
    vp = np.where(np.hypot(x - cx, z - cz) < rad, 4500.0, vp)
    rho = 310.0 * vp ** 0.25
    return vp.astype(np.float32), rho.astype(np.float32)


def save(name, **kw):
    kw["geometry_text"] = np.array(run_reference.geometry_text)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **kw)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


def main(with_c0=True):
    if not refrun.available():
        sys.exit(f"{refrun.EXE} missing: make -C oracle ref")

    # 1. two-layer, split spread from a shot file, dt_factor 1
    nx, nz, h = 100, 80, 10.0
    vp, rho = two_layer(nx, nz, 40)
    src, rec, seis, _ = run_reference(nx, nz, h, h, 0.3, 0.001, 0.001, 15, vp, rho,
                                      shot_line="500.000000,20.000000,100,0,10.000000,1")
    save("tiny_twolayer", params=params_arr(nx, nz, h, h, 0.3, 0.001, 0.001, 15), vp=vp, rho=rho,
         src_xz=src, rec_xz=rec, seis=seis, trace_idx=np.arange(seis.shape[0]))

    # 2. dx != dz, DATA DT = 4 x FD DT, end-on-right spread near the left edge
    nx, nz = 120, 90
    vp, rho = two_layer(nx, nz, 50)
    src, rec, seis, _ = run_reference(nx, nz, 15.0, 10.0, 0.4, 0.001, 0.004, 10, vp, rho,
                                      shot_line="150.000000,30.000000,80,0,15.000000,3")
    save("rect_dtfactor", params=params_arr(nx, nz, 15.0, 10.0, 0.4, 0.001, 0.004, 10), vp=vp,
         rho=rho, src_xz=src, rec_xz=rec, seis=seis, trace_idx=np.arange(seis.shape[0]))

    # 3. heterogeneous vp and rho, explicit geometry with receivers at several
    #    depths, at the model edges and on the source cell
    nx, nz, h = 150, 100, 10.0
    vp, rho = smooth_random(nx, nz, 7)
    sx, sz = 740.0, 300.0
    gx = list(range(0, nx * 10, 30)) + [1490.0, 0.0, 740.0, 740.0]
    gz = [50.0] * (len(gx) - 4) + [990.0, 990.0, 300.0, 0.0]
    lines = [f"{sx:f},{sz:f},{x:f},{z:f}" for x, z in zip(gx, gz)]
    src, rec, seis, _ = run_reference(nx, nz, h, h, 0.45, 0.0008, 0.0016, 12, vp, rho,
                                      geom_lines=lines)
    save("hetero_explicit", params=params_arr(nx, nz, h, h, 0.45, 0.0008, 0.0016, 12), vp=vp,
         rho=rho, src_xz=src, rec_xz=rec, seis=seis, trace_idx=np.arange(seis.shape[0]))

    # 4. salt-type body at 7.62 m, dt 0.5 ms (shot-file geometry, float spacing)
    nx, nz, h = 200, 120, 7.62
    vp, rho = salt(nx, nz, h, 11)
    src, rec, seis, _ = run_reference(nx, nz, h, h, 0.35, 0.0005, 0.0005, 15, vp, rho,
                                      shot_line="761.999939,7.620000,150,0,7.620000,1")
    save("salt_small", params=params_arr(nx, nz, h, h, 0.35, 0.0005, 0.0005, 15), vp=vp, rho=rho,
         src_xz=src, rec_xz=rec, seis=seis, trace_idx=np.arange(seis.shape[0]))

    if with_c0:
        # 5. config 0 at the reference's npml = 32: 1000 x 1000, 2 s, 1000
        #    receivers every cell at z = 20 m.  Stored decimated (every 20th
        #    trace) with the SHA-256 of the full output.
        nx = nz = 1000
        vp, rho = two_layer(nx, nz, 500)
        src, rec, seis, _ = run_reference(nx, nz, 10.0, 10.0, 2.0, 0.001, 0.001, 15, vp, rho,
                                          shot_line="5000.000000,20.000000,1000,0,10.000000,1")
        idx = np.arange(0, seis.shape[0], 20)
        save("c0_twolayer_npml32", params=params_arr(nx, nz, 10.0, 10.0, 2.0, 0.001, 0.001, 15),
             src_xz=src, rec_xz=rec, seis=seis[idx], trace_idx=idx,
             sha256=np.array(hashlib.sha256(seis.tobytes()).hexdigest()))


if __name__ == "__main__":
    main(with_c0="--no-c0" not in sys.argv)
