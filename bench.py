"""Benchmark of the forward-modelling hot path on one MI355X.

Workload (the headline configuration): plain FP32 forward on a two-layer model
at 5000 x 5000 cells, dx = dz = 10 m (interface at z = 2500 m; vp 1500/3000 m/s,
rho 1000/2000 kg/m^3), dt = 1 ms, 4000 steps, sponge 50 cells (padded grid
5100 x 5100 = 26.01 M cells), 15 Hz Ricker, 64 shots evenly spaced along x at
z = 20 m, 5000 surface receivers per shot, shots run one after another.

One bench "step" is one shot: the full 4000-step time loop over the padded grid
plus its 5000-trace gather.  The metric is padded cell-updates per second.
Inputs are resident in HBM before the timed region (the context uploads the
model and builds the coefficient fields once); seismograms are written to a
device buffer.

    python bench.py [--steps K] [--warmup W] [--shots-per-job 64] [--no-cpu]

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX = NZ = 5000
H = 10.0
NPML = 50
DT = 0.001
T_REC = 4.0
FM = 15.0
NSHOTS = 64
BYTES_PER_CELL = 44          # px, pz, vx, vz read + written (32) + KX, BX, BZ read (12)
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E peak, GB/s


def model():
    vp = np.full((NX, NZ), 1500, np.float32); vp[:, NZ // 2:] = 3000
    rho = np.full((NX, NZ), 1000, np.float32); rho[:, NZ // 2:] = 2000
    return vp, rho


def shots(sam):
    rec = np.stack([np.arange(NX) * H, np.full(NX, 20.0)], 1).astype(np.float32)
    out = []
    for k in range(NSHOTS):
        sx = 10.0 * int((k + 0.5) * NX / NSHOTS)
        out.append(sam.Shot(sx, 20.0, rec))
    return out


class _Hip:
    """Minimal HIP runtime access for a device output buffer."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        rc = self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
        return p.value

    def free(self, p):
        self.L.hipFree(ctypes.c_void_p(p))

    def sync(self):
        self.L.hipDeviceSynchronize()


def cpu_baseline(vp, rho, seconds_target=15.0):
    """FP64 reference-structured CPU restatement (the oracle) on the same grid,
    a bounded number of time steps; cells/s of the step loop alone."""
    from oracle import pyoracle
    threads = int(os.environ.get("BENCH_CPU_THREADS", min(16, os.cpu_count() or 1)))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    P = pyoracle.params(NX, NZ, H, H, T_REC, DT, DT, FM, NPML)
    cells = (NX + 2 * NPML) * (NZ + 2 * NPML)
    t1, th = pyoracle.time_steps(P, vp, rho, 4)
    per = t1 / 4
    n = max(4, min(400, int(seconds_target / max(per, 1e-6))))
    t, th = pyoracle.time_steps(P, vp, rho, n)
    return {"value": round(cells * n / t / 1e9, 4), "unit": "Gcell-updates/s", "cores": th,
            "kind": "port",
            "sample": f"FP64 oracle, reference loop structure (9 sweeps/step), {n} steps of the "
                      f"5100x5100 padded grid, {th} OpenMP threads, setup excluded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="profiling runs only: shorten each shot to this many time steps")
    ap.add_argument("--copy-probe", action="store_true",
                    help="profiling calibration: one 1 GiB device-to-device copy, then exit")
    ap.add_argument("--tuning", action="store_true", help="print per-variant autotune times")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on a single GPU; --gpus must be 1 "
              "(shots are independent, multi-GPU scaling is a shot queue per device).",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    import seisacomod2d_amd as sam

    vp, rho = model()
    t_rec = T_REC if args.profile_steps <= 0 else args.profile_steps * DT
    grid = sam.Grid(NX, NZ, H, H, t_rec, DT, DT, FM, npml=NPML)
    all_shots = shots(sam)
    hip = _Hip()
    with sam.Modeller(grid, vp, rho) as m:
        if args.no_graph:
            m.set_option("graph", 0)
        if args.variant >= 0:
            m.set_option("variant", args.variant)
        else:
            best = m.autotune(steps=6)
            print(f"# autotune picked {sam.variants()[best]}", file=sys.stderr, flush=True)
            if args.tuning:
                for name, ms in zip(sam.variants(), m.tuning()):
                    print(f"#   {name:24s} {ms * 1e3:8.1f} us/step", file=sys.stderr)
        ng = all_shots[0].nrec
        per_shot = ng * m.real_nt
        seis = hip.malloc(per_shot * 4)
        cells = m.nnx * m.nnz
        for i in range(args.warmup):
            m.forward_dev([all_shots[i % NSHOTS]], seis)
        hip.sync()
        k_ms = 0.0
        launches = 0
        t0 = time.perf_counter()
        for i in range(args.steps):
            m.forward_dev([all_shots[(args.warmup + i) % NSHOTS]], seis)
            st = m.stats()
            k_ms += st["kernel_ms"]
            launches += st["launches"]
            print(f"# shot {i + 1}/{args.steps} {st['wall_ms']:.1f} ms", file=sys.stderr, flush=True)
        hip.sync()
        wall = time.perf_counter() - t0
        variant = sam.variants()[m.stats()["variant"]]
        used_graph = m.stats()["used_graph"]
        hip.free(seis)

    cell_updates = float(cells) * m.fd_nt * args.steps
    gcells = cell_updates / wall / 1e9
    launch_ms = k_ms / max(launches, 1)
    achieved = BYTES_PER_CELL * cells / (launch_ms * 1e-3) / 1e9   # GB/s per launch
    out = {
        "metric": "Gcell-updates/s (padded cells x time steps), plain FP32 forward",
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic two-layer model (analytic), 64 shot positions cycled",
        "config": {"workload": "config1_plain_forward_5000x5000_nt4000_npml50",
                   "grid": [NX, NZ], "padded": [m.nnx, m.nnz], "nt": m.fd_nt, "receivers": ng,
                   "shots_per_job": NSHOTS, "kernel_variant": variant, "hipgraph": used_graph},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_step", "bytes_per_cell_update": BYTES_PER_CELL,
                     "avg_launch_us": round(launch_ms * 1e3, 2)},
        "job_seconds_64_shots": round(wall / args.steps * NSHOTS, 2),
    }
    traffic = _measured_traffic()
    if traffic:
        out["roofline"]["traffic"] = traffic
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(vp, rho)
    print(json.dumps(out), flush=True)


def copy_probe(nbytes=1 << 30):
    """A plain device-to-device copy of known size (a blit kernel), so that
    rocprofv3's FETCH_SIZE / WRITE_SIZE can be scaled to bytes."""
    hip = _Hip()
    a, b = hip.malloc(nbytes), hip.malloc(nbytes)
    for _ in range(3):
        rc = hip.L.hipMemcpy(ctypes.c_void_p(b), ctypes.c_void_p(a), ctypes.c_size_t(nbytes), 3)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
    hip.sync()
    hip.free(a)
    hip.free(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "repeats": 3}))


def _measured_traffic():
    """HBM bytes per k_step launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), as recorded in
    profiles/k_step_config1_pmc.txt ("traffic_bytes_per_launch: <n>")."""
    path = os.path.join(ROOT, "profiles", "k_step_config1_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


if __name__ == "__main__":
    main()
