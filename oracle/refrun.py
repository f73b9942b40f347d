"""Run the compiled reference (oracle/_ref/SeisAcoMod2D_MULTI_CORE_CPU_V1.0.exe,
built from src/src_cpu by oracle/Makefile.ref) on a job card in a scratch
directory.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_goldens.py to write the
golden seismograms and by bench.py's cpu_baseline leg to time the reference's
CPU path.  The product library never calls it.

The card uses the reference's keys (json_parser.cpp:141-265); models are raw
float32, x-major (model.cpp:37-40); the output is
<Output Directory>/<job>_seismogram.bin (cpumodelling.cpp:367-374).
"""
import os
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_ref", "SeisAcoMod2D_MULTI_CORE_CPU_V1.0.exe")

CARD = """"Model" : "comment"
    "Velocity" : "vp.bin"
    "Density" : "rho.bin"
    "NX" : "{nx}"
    "NZ" : "{nz}"
    "DX" : "{dx}"
    "DZ" : "{dz}"
    "Time" : "{T}"
    "FD DT" : "{fd_dt}"
    "DATA DT" : "{real_dt}"
    "Frequency" : "{fm}"
    "Read source receiver flag" : "{flag}"
    "if choice=0, Geometry file" : "geom.txt"
    "if choice=1, Shot file" : "shots.txt"
    "Output Directory" : "out"
    "Free surface" : "0"
    "Job type New or Restart" : "New"
    "Order of FD Operator" : "4"
"""


def available():
    return os.path.exists(EXE)


def write_case(d, nx, nz, dx, dz, vp, rho, shot_line=None, geom_lines=None):
    """Model files, shot file (flag 1) or explicit geometry (flag 0) in d;
    returns the geometry flag."""
    np.asarray(vp, np.float32).tofile(os.path.join(d, "vp.bin"))
    np.asarray(rho, np.float32).tofile(os.path.join(d, "rho.bin"))
    if shot_line:
        open(os.path.join(d, "shots.txt"), "w").write(shot_line + "\n")
        return 1
    open(os.path.join(d, "geom.txt"), "w").write("\n".join(geom_lines) + "\n")
    return 0


def run_card(d, nx, nz, dx, dz, T, fd_dt, real_dt, fm, flag, threads, timeout=3600):
    """Write card.json for these parameters into d and run the reference there;
    returns (wall seconds, CompletedProcess)."""
    open(os.path.join(d, "card.json"), "w").write(CARD.format(
        nx=nx, nz=nz, dx=dx, dz=dz, T=T, fd_dt=fd_dt, real_dt=real_dt, fm=fm, flag=flag))
    os.makedirs(os.path.join(d, "out"), exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.perf_counter()
    r = subprocess.run([EXE, "card.json"], cwd=d, env=env, capture_output=True, text=True,
                       timeout=timeout)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
    return wall, r


def time_per_step(nx, nz, dx, vp, rho, shot_line, t_short, t_long, fd_dt, fm, threads):
    """Seconds per time step of the reference on this model, from the difference
    of two runs with record lengths t_short and t_long (removes setup and I/O)."""
    with tempfile.TemporaryDirectory() as d:
        flag = write_case(d, nx, nz, dx, dx, vp, rho, shot_line=shot_line)
        w1, _ = run_card(d, nx, nz, dx, dx, t_short, fd_dt, fd_dt, fm, flag, threads)
        w2, _ = run_card(d, nx, nz, dx, dx, t_long, fd_dt, fd_dt, fm, flag, threads)
    steps = round((t_long - t_short) / fd_dt)
    return (w2 - w1) / steps, steps, (w1, w2)
