"""Benchmark of the forward-modelling hot path on one MI355X.

Default run, one JSON line:
  * headline (value): plain FP32 forward on a two-layer model at 5000 x 5000
    cells, dx = dz = 10 m (interface at z = 2500 m; vp 1500/3000 m/s, rho
    1000/2000 kg/m^3), dt = 1 ms, 4000 steps, sponge 50 cells (padded grid
    5100 x 5100 = 26.01 M cells), 15 Hz Ricker, 64 shots evenly spaced along x
    at z = 20 m, 5000 surface receivers per shot, shots run one after another.
    One bench "step" is one shot: the full 4000-step time loop over the padded
    grid plus its 5000-trace gather.  The metric is padded cell-updates per
    second.  Inputs are resident in HBM before the timed region (the context
    uploads the model and builds the coefficient fields once); seismograms are
    written to a device buffer.
  * "batched": the salt-type 3201 x 1201 model (7.62 m, 0.5 ms, 8000 steps,
    sponge 60), 64 shots with aperture windows 800-2400 cells: shots/s of the
    batched launch against the per-shot loop on the same windows.
  * "gradient": checkpointed adjoint gradient on the salt model, 32 windowed
    shots, 8000 steps, checkpoints every 100 steps: seconds per shot and the
    ratio to the batched forward and to the per-shot forward of the same shots;
    also the same gradient in the two halves an FWI iteration runs (modelling
    that keeps the checkpoints, then the backward pass from them).
  * "cpu_baseline": the compiled reference (src_cpu) on this host.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-extra]

Separate workloads (each prints its own JSON line):
    --workload batched    the batched leg alone (--quick: 16 shots, 1 s)
    --workload gradient   gradients of the 1201 x 601 anomaly model (16 shots,
                          3000 steps, k = 100) and of the salt model (32
                          windowed shots, 8000 steps, k = 50 / 100 / 200)

--gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from seisacomod2d_amd import synthetic  # noqa: E402

NX = NZ = 5000
H = 10.0
NPML = 50
DT = 0.001
T_REC = 4.0
FM = 15.0
NSHOTS = 64
# Algorithmic bytes per cell of one pass over the grid: px, pz, vx, vz read and
# written (32) and the three per-cell coefficients kappa, bx, bz read (12).
# A launch advances several steps with the states between them in registers,
# so these bytes count once per launch.  The multi-step kernels move fewer:
# outside the sponge frame a cell's state is (p, vx, vz), 12 bytes read and
# written (the split px, pz only matters in the frame, which adds pz: 8
# bytes), and they read the buoyancy 1/rho (4) in place of bx and bz, forming
# both staggered coefficients in registers, plus KX (4; KZ is KX when
# dx == dz): 32 B per cell, 40 in the frame.
BYTES_PER_CELL = 44
KERNEL_LEAST_BYTES_PER_CELL = 32
KERNEL_LEAST_FRAME_EXTRA = 8
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E peak, GB/s
METRIC = "Gcell-updates/s + %HBM peak; shots/s batched vs plain; gradient s/shot"
CPU_SHARE = 16               # host threads one GPU's job may use on the GPU hosts


def model():
    return synthetic.two_layer(NX, NZ)


def shots(sam):
    return [sam.Shot(sx, sz, rec) for sx, sz, rec in synthetic.surface_spread_shots(NX, H, NSHOTS)]


class _Hip:
    """Minimal HIP runtime access for a device output buffer."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.L = ctypes.CDLL("libamdhip64.so.7")     # the runtime libsam2d.so links

    def malloc(self, nbytes):
        p = self.ct.c_void_p()
        rc = self.L.hipMalloc(self.ct.byref(p), self.ct.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
        return p.value

    def free(self, p):
        self.L.hipFree(self.ct.c_void_p(p))

    def sync(self):
        self.L.hipDeviceSynchronize()


def host_cpus():
    """What this host offers: logical CPUs, the process's affinity set and the
    physical cores among them (distinct (package, core) pairs)."""
    logical = os.cpu_count() or 1
    aff = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(logical))
    phys = set()
    for c in aff:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            phys.add((open(base + "physical_package_id").read().strip(),
                      open(base + "core_id").read().strip()))
        except OSError:
            phys.add(("?", str(c)))
    return {"logical_cpus": logical, "affinity_cpus": len(aff), "physical_cores_in_affinity": len(phys)}


def cpu_baseline(vp, rho, seconds_target=15.0):
    """The reference's CPU path timed on this host: the compiled src_cpu build
    (oracle/_ref, "reference") when present, else the FP64 oracle in the
    reference's loop structure ("port").  Cells/s of the time loop alone, at
    min(CPU share, affinity) OpenMP threads."""
    cpus = host_cpus()
    threads = int(os.environ.get("BENCH_CPU_THREADS", min(CPU_SHARE, cpus["affinity_cpus"])))
    from oracle import refrun
    if refrun.available():
        try:
            # The reference fixes its sponge at 32 cells (cpumodelling.cpp:15):
            # its padded grid is 5064 x 5064.  Two runs, 20 and 320 steps,
            # timed by difference (setup and file I/O cancel).
            sx = NX * H / 2
            per, steps, walls = refrun.time_per_step(
                NX, NZ, H, vp, rho, f"{sx:.6f},20.000000,{NX},0,{H:.6f},1", 0.02, 0.32, DT, FM, threads)
            cells = (NX + 64) * (NZ + 64)
            return dict({"value": round(cells / per / 1e9, 4), "unit": "Gcell-updates/s",
                         "cores": threads, "kind": "reference",
                         "sample": f"compiled reference src_cpu (FP64, MPI singleton, {threads} OpenMP "
                                   f"threads), 5000x5000 two-layer model at its fixed sponge of 32 "
                                   f"(5064x5064 padded), 1 shot, 5000 receivers; per-step time from "
                                   f"two runs of 20 and 320 steps ({walls[0]:.1f} s, {walls[1]:.1f} s)"},
                        **cpus)
        except Exception as e:                      # fall back to the port
            print(f"# reference CPU run failed ({e}); timing the FP64 port", file=sys.stderr)
    from oracle import pyoracle
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    P = pyoracle.params(NX, NZ, H, H, T_REC, DT, DT, FM, NPML)
    cells = (NX + 2 * NPML) * (NZ + 2 * NPML)
    t1, th = pyoracle.time_steps(P, vp, rho, 4)
    per = t1 / 4
    n = max(4, min(400, int(seconds_target / max(per, 1e-6))))
    t, th = pyoracle.time_steps(P, vp, rho, n)
    return dict({"value": round(cells * n / t / 1e9, 4), "unit": "Gcell-updates/s", "cores": th,
                 "kind": "port",
                 "sample": f"FP64 oracle, reference loop structure (9 sweeps/step), {n} steps of the "
                           f"5100x5100 padded grid, {th} OpenMP threads, setup excluded"}, **cpus)


def measured_traffic(kernel, config):
    """HBM bytes per launch of the dominant kernel from rocprofv3 --pmc, as
    recorded in profiles/<kernel>_config1_pmc.txt, one line per profiled
    configuration: "profiled: <variant> chunk <c> order <o>
    traffic_bytes_per_launch <n>".  Returns (bytes, the configuration they
    were measured at): the run's own configuration when it was profiled, else
    the first profiled one."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_config1_pmc.txt")
    if not os.path.exists(path):
        return None, None
    found = []
    for line in open(path):
        if line.startswith("profiled:") and "traffic_bytes_per_launch" in line:
            cfg, n = line.split(":", 1)[1].split("traffic_bytes_per_launch")
            try:
                found.append((cfg.strip(), float(n)))
            except ValueError:
                pass
    # A variant's name carries its nominal march length (_chunkNN), which the
    # run's own "chunk" overrides: variants that differ only in it are the
    # same kernel.
    norm = lambda c: re.sub(r"_chunk\d+", "", c)
    for cfg, n in found:
        if norm(cfg) == norm(config):
            return n, cfg
    return (found[0][1], found[0][0]) if found else (None, None)


# --------------------------------------------------------------- config 1
def bench_plain(args, sam):
    if not str(args.variant).lstrip("-").isdigit():
        args.variant = sam.variants().index(args.variant)
    args.variant = int(args.variant)
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
            m.set_option("tile_order", max(args.tile_order, 0))
            m.set_option("chunk", max(args.chunk, 0))
        else:
            best = m.autotune(steps=6)
            if args.tile_order >= 0:
                m.set_option("tile_order", args.tile_order)
            if args.chunk >= 0:
                m.set_option("chunk", args.chunk)
            print(f"# autotune picked {sam.variants()[best]}, tile order "
                  f"{m.get_option('tile_order')}, chunk {m.get_option('chunk')}",
                  file=sys.stderr, flush=True)
            if args.tuning:
                for name, ms in zip(sam.variants(), m.tuning()):
                    print(f"#   {name:28s} {ms * 1e3:8.1f} us/step", file=sys.stderr)
        ng = all_shots[0].nrec
        seis = hip.malloc(ng * m.real_nt * 4)
        cells = m.nnx * m.nnz
        for i in range(args.warmup):
            m.forward_dev([all_shots[i % NSHOTS]], seis)
        hip.sync()
        k_ms, launches = 0.0, 0
        t0 = time.perf_counter()
        for i in range(args.steps):
            m.forward_dev([all_shots[(args.warmup + i) % NSHOTS]], seis)
            st = m.stats()
            k_ms += st["kernel_ms"]
            launches += st["launches"]
            print(f"# shot {i + 1}/{args.steps} {st['wall_ms']:.1f} ms", file=sys.stderr, flush=True)
        hip.sync()
        wall = time.perf_counter() - t0
        variant = sam.variants()[st["variant"]]
        tile_order = m.get_option("tile_order")
        chunk = m.get_option("chunk")
        spl = st["steps_per_launch"]
        used_graph = st["used_graph"]
        hip.free(seis)

    cell_updates = float(cells) * m.fd_nt * args.steps
    gcells = cell_updates / wall / 1e9
    # kernel_ms is HIP events around the time loop on the library's stream;
    # one launch advances spl time steps
    launch_ms = k_ms / max(launches, 1)
    # Roofline of the dominant kernel.  The algorithmic bytes of a launch are
    # one pass over the grid, 44 B per cell, whatever the number of steps the
    # launch advances (the steps between live in registers); frac is that
    # over the launch time, against the HBM peak.  kernel_least_bytes is what
    # the multi-step kernel itself must move (32 B per cell, 40 in the sponge
    # frame), and its rate.
    # effective_gbs counts 44 B per cell-update, the traffic a
    # one-step-per-pass kernel would need for the same work.
    hbm_bytes = BYTES_PER_CELL * cells
    frame_cells = cells - NX * NZ            # the sponge frame (profiles not all 1)
    least = (KERNEL_LEAST_BYTES_PER_CELL * cells + KERNEL_LEAST_FRAME_EXTRA * frame_cells
             if spl > 1 else BYTES_PER_CELL * cells)
    achieved = hbm_bytes / (launch_ms * 1e-3) / 1e9
    kname = "k_step" if spl == 1 else "k_step%d" % spl
    this_cfg = f"{variant} chunk {chunk} order {tile_order}"
    traffic, traffic_cfg = measured_traffic(kname, this_cfg)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": kname, "steps_per_launch": spl,
            "algorithmic_bytes_per_launch": hbm_bytes,
            "bytes_per_cell": BYTES_PER_CELL,
            "kernel_least_bytes_per_launch": least,
            "kernel_least_gbs": round(least / (launch_ms * 1e-3) / 1e9, 1),
            "avg_launch_us": round(launch_ms * 1e3, 2),
            "effective_gbs": round(hbm_bytes * spl / (launch_ms * 1e-3) / 1e9, 1),
            "traffic_profiled_config": traffic_cfg,
            "traffic_config_matches_run": traffic_cfg is not None and
                re.sub(r"_chunk\d+", "", traffic_cfg) == re.sub(r"_chunk\d+", "", this_cfg)}
    if traffic:
        dram = traffic / (launch_ms * 1e-3) / 1e9
        roof["dram_gbs"] = round(dram, 1)
        roof["dram_frac"] = round(dram / HBM_PEAK_GBS, 4)
        roof["traffic_over_algorithmic"] = round(traffic / hbm_bytes, 4)
        roof["traffic_over_kernel_least"] = round(traffic / least, 4)
    return {
        "metric": METRIC,
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
                   "shots_per_job": NSHOTS, "kernel_variant": variant, "tile_order": tile_order,
                   "chunk": chunk, "hipgraph": used_graph, "profile_run": args.profile_steps > 0},
        "roofline": roof,
        "job_seconds_64_shots": round(wall / args.steps * NSHOTS, 2),
    }


# --------------------------------------------------------------- config 2
def bench_c1_batched(args, sam, nb=8):
    """The C1 job's shots through the batched forward, nb shots per launch:
    the same kernel over nb times the tiles, so the tuner can take long
    marches (less x-halo) and each launch drains once for nb shots.  Reported
    beside the per-shot loop, which stays the headline value."""
    vp, rho = model()
    t_rec = T_REC if args.profile_steps <= 0 else args.profile_steps * DT
    grid = sam.Grid(NX, NZ, H, H, t_rec, DT, DT, FM, npml=NPML)
    sh = shots(sam)[:nb]
    variant = str(args.variant)
    variant = int(variant) if variant.lstrip("-").isdigit() else sam.variants().index(variant)
    with sam.Modeller(grid, vp, rho) as m:
        if variant >= 0:                       # a fixed configuration (profiling runs)
            best = variant
            m.set_option("variant", variant)
            m.set_option("tile_order", max(args.tile_order, 0))
            m.set_option("chunk", max(args.chunk, 0))
        else:
            best = m.autotune(nshots=nb, steps=6)
        m.forward_batched(sh)                  # warm: the time loop's graph captured
        t0 = time.perf_counter()
        m.forward_batched(sh)
        wall = time.perf_counter() - t0
        st = m.stats()
        cells = m.nnx * m.nnz
        chunk, order = m.get_option("chunk"), m.get_option("tile_order")
        fd_nt = m.fd_nt
    us = st["kernel_ms"] * 1e3 / st["launches"]
    shot_launch_us = us / nb
    frac = BYTES_PER_CELL * cells / (shot_launch_us * 1e-6) / 1e9 / HBM_PEAK_GBS
    return {"workload": f"config1_5000x5000_nt{fd_nt}_npml50_batched_{nb}_shots",
            "shots": nb, "gcells_wall": round(cells * fd_nt * nb / wall / 1e9, 2),
            "gcells_kernel": round(cells * fd_nt * nb / (st["kernel_ms"] * 1e-3) / 1e9, 2),
            "launch_us": round(us, 2), "us_per_shot_launch": round(shot_launch_us, 2),
            "frac_44B": round(frac, 4), "variant": sam.variants()[best], "chunk": chunk,
            "tile_order": order}


def salt_shots(sam, n, seed=7):
    sh, wins = synthetic.salt_shots(n, seed=seed)
    return [sam.Shot(sx, sz, rec) for sx, sz, rec in sh], wins


def bench_batched(args, sam, nshots=64, t_rec=4.0, batch_sizes=(0, 8, 16)):
    vp, rho = synthetic.salt_model()
    grid = sam.Grid(3201, 1201, 7.62, 7.62, t_rec, 0.0005, 0.0005, 15.0, npml=60)
    sh, wins = salt_shots(sam, nshots)
    cells = sum((w + 120) * (1201 + 120) for _, w in wins)
    with sam.Modeller(grid, vp, rho) as m:
        # per-shot loop on the windows, variant tuned on one window; one
        # untimed batched pass sets up every window's coefficients first, as
        # the timed calls find them
        m.forward_batched(sh, wins)
        m.autotune(windows=wins[:1], nshots=1, steps=6)
        m.forward_windows(sh[:1], wins[:1])                     # warm
        t0 = time.perf_counter()
        m.forward_windows(sh, wins)
        t_plain = time.perf_counter() - t0
        vplain = sam.variants()[m.stats()["variant"]]
        print(f"# per-shot loop: {t_plain:.2f} s ({vplain})", file=sys.stderr, flush=True)
        # batched: variant tuned on the resident batch, for a few batch sizes
        t_batch, best = None, None
        for mb in batch_sizes:
            m.set_option("max_batch", mb)
            m.autotune(windows=wins, nshots=nshots, steps=6)
            m.forward_batched(sh, wins)            # warm: graphs captured and cached
            t0 = time.perf_counter()
            m.forward_batched(sh, wins)
            t = time.perf_counter() - t0
            v = sam.variants()[m.stats()["variant"]]
            print(f"# batched max_batch={mb}: {t:.2f} s ({v})", file=sys.stderr, flush=True)
            if t_batch is None or t < t_batch:
                t_batch, best = t, (mb, v)
        nt = m.fd_nt
    return {
        "workload": f"config2_batched_3201x1201_nt{nt}_{nshots}shots_windows800-2400",
        "shots": nshots, "batched_shots_per_s": round(nshots / t_batch, 3),
        "plain_shots_per_s": round(nshots / t_plain, 3),
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "batched_gcells": round(cells * nt / t_batch / 1e9, 2),
        "plain_gcells": round(cells * nt / t_plain / 1e9, 2),
        "batched_max_batch": best[0], "batched_variant": best[1], "plain_variant": vplain,
    }


# --------------------------------------------------------------- configs 3, 4
def _grad_case(sam, grid, vtrue, v0, rho, sh, wins, ks, label, no_graph=False, adj_steps=2,
               fixed=None, budget_gb=0):
    """Gradient time against the forward modelling of the same shots: the
    batched forward (autotuned, the fastest forward path) and the per-shot
    loop.  fixed: (variant, chunk, tile order) to use instead of the tuner
    (profiling runs)."""
    def tune(m, **kw):
        if fixed is None:
            m.autotune(**kw)
        else:
            m.set_option("variant", fixed[0])
            m.set_option("chunk", fixed[1])
            m.set_option("tile_order", fixed[2])
    res = []
    with sam.Modeller(grid, vtrue, rho) as mt:
        if no_graph:
            mt.set_option("graph", 0)
        obs = mt.forward_batched(sh, wins)
    with sam.Modeller(grid, v0, rho) as m:
        if no_graph:
            m.set_option("graph", 0)
        m.set_option("adjoint_steps", adj_steps)
        if budget_gb > 0:
            m.set_option("hbm_budget_gb", budget_gb)
        # per-shot loop with a variant tuned on one shot; one untimed batched
        # pass sets up every window's coefficients first, as the timed
        # batched and gradient calls find them
        tune(m, windows=wins[:1] if wins else None, nshots=1, steps=6)
        m.forward_batched(sh, wins)
        m.forward_windows(sh[:1], wins[:1]) if wins else m.forward(sh[:1])   # warm
        t0 = time.perf_counter()
        m.forward_windows(sh, wins) if wins else m.forward(sh)
        t_plain = time.perf_counter() - t0
        # batched forward and gradient with a variant tuned on the batch
        tune(m, windows=wins, nshots=len(sh), steps=6)
        d = m.forward_batched(sh, wins)                          # warm
        t0 = time.perf_counter()
        d = m.forward_batched(sh, wins)
        t_fwd = time.perf_counter() - t0
        adj = [a - b for a, b in zip(d, obs)]
        for k in ks:
            t0 = time.perf_counter()
            gvp, grho = m.gradient(sh, adj, ckpt_k=k, windows=wins)
            t_g = time.perf_counter() - t0
            # The same gradient in its two halves, as an FWI iteration runs it:
            # modelling that keeps the checkpoints, then the backward pass.
            fwi = {}
            try:
                t0 = time.perf_counter()
                m.forward_checkpointed(sh, ckpt_k=k, windows=wins)
                t_fc = time.perf_counter() - t0
                t0 = time.perf_counter()
                m.gradient_checkpointed(adj)
                t_b = time.perf_counter() - t0
                fwi = {"forward_checkpointed_s_per_shot": round(t_fc / len(sh), 4),
                       "backward_s_per_shot": round(t_b / len(sh), 4),
                       "fwi_iteration_s_per_shot": round((t_fc + t_b) / len(sh), 4),
                       "backward_ratio_to_batched_forward": round(t_b / t_fwd, 3),
                       "backward_ratio_to_per_shot_forward": round(t_b / t_plain, 3)}
            except sam.SAM2DError as e:            # the shots exceed one resident batch
                fwi = {"forward_checkpointed": str(e)}
            res.append(dict({"workload": label, "ckpt_k": k, "shots": len(sh),
                        "adjoint_steps": adj_steps,
                        "forward_batched_s_per_shot": round(t_fwd / len(sh), 4),
                        "forward_per_shot_loop_s_per_shot": round(t_plain / len(sh), 4),
                        "gradient_s_per_shot": round(t_g / len(sh), 4),
                        "gradient_s_per_batch": round(t_g, 3),
                        "ratio_to_batched_forward": round(t_g / t_fwd, 3),
                        "ratio_to_per_shot_forward": round(t_g / t_plain, 3),
                        "misfit": float(sum(0.5 * np.sum(a.astype(np.float64) ** 2) for a in adj)),
                        "grad_norm_vp": float(np.linalg.norm(gvp))}, **fwi))
            print(f"# {label} k={k}: gradient {t_g:.2f} s, batched forward {t_fwd:.2f} s, "
                  f"per-shot forward {t_plain:.2f} s, backward from checkpoints "
                  f"{fwi.get('backward_s_per_shot', float('nan')) * len(sh):.2f} s",
                  file=sys.stderr, flush=True)
    return res


def _fixed(args, sam):
    if str(args.variant) in ("-1", ""):
        return None
    v = args.variant if str(args.variant).lstrip("-").isdigit() else sam.variants().index(args.variant)
    return (int(v), max(args.chunk, 0), max(args.tile_order, 0))


def gradient_c4(args, sam, nshots=32, ks=(100,)):
    from scipy.ndimage import gaussian_filter
    vs, rs = synthetic.salt_model()
    v4 = gaussian_filter(vs.astype(np.float64), 20.0).astype(np.float32)
    t_rec = 4.0 if args.profile_steps <= 0 else args.profile_steps * 0.0005
    grid4 = sam.Grid(3201, 1201, 7.62, 7.62, t_rec, 0.0005, 0.0005, 15.0, npml=60)
    sh4, w4 = salt_shots(sam, nshots, seed=11)
    return _grad_case(sam, grid4, vs, v4, rs, sh4, w4, list(ks),
                      "config4_gradient_3201x1201_nt8000_windows", args.no_graph, args.adjoint_steps,
                      _fixed(args, sam), args.hbm_budget_gb)


def bench_gradient(args, sam):
    out = []
    vtrue, v0, rho = synthetic.anomaly_model()
    n3 = 4 if args.quick else 16
    t3 = 3.0 if args.profile_steps <= 0 else args.profile_steps * 0.001
    grid = sam.Grid(1201, 601, 10.0, 10.0, t3, 0.001, 0.001, 10.0, npml=32)
    rec = np.stack([np.arange(1201) * 10.0, np.full(1201, 10.0)], 1).astype(np.float32)
    sh = [sam.Shot(10.0 * int((k + 0.5) * 1201 / n3), 10.0, rec) for k in range(n3)]
    if not args.c4_only:
        out += _grad_case(sam, grid, vtrue, v0, rho, sh, None, [100],
                          "config3_gradient_1201x601_nt3000", args.no_graph, args.adjoint_steps,
                          _fixed(args, sam), args.hbm_budget_gb)
    ks = [int(k) for k in args.ckpt_k.split(",")] if args.ckpt_k else (
        [200] if args.quick else [50, 100, 200])
    out += gradient_c4(args, sam, args.shots or (4 if args.quick else 32), ks)
    return {"metric": METRIC, "unit": "s/shot", "higher_is_better": False, "dtype": "f32",
            "results": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched and gradient legs of the default run")
    ap.add_argument("--variant", default="-1",
                    help="kernel variant: index or name (-1: autotuned)")
    ap.add_argument("--chunk", type=int, default=-1,
                    help="march length of the multi-step kernels (-1: autotuned, or the "
                         "variant's with --variant; 0 = the variant's)")
    ap.add_argument("--tile-order", type=int, default=-1,
                    help="chunk-column stride of the multi-step kernels (-1: autotuned, or 0 "
                         "with --variant)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2, choices=(1, 2),
                    help="gradient workload: backward steps per adjoint pass")
    ap.add_argument("--hbm-budget-gb", type=int, default=0,
                    help="gradient workload: device memory a gradient batch may use "
                         "(0: the library's default)")
    ap.add_argument("--workload", default="plain",
                    choices=["plain", "batched", "gradient", "config1-batched"])
    ap.add_argument("--quick", action="store_true", help="batched/gradient: fewer shots")
    ap.add_argument("--max-batch", type=int, default=0)
    ap.add_argument("--c4-only", action="store_true", help="gradient workload: the salt model only")
    ap.add_argument("--ckpt-k", default="", help="gradient workload: checkpoint intervals, e.g. 100")
    ap.add_argument("--shots", type=int, default=0,
                    help="gradient workload: salt-model shots; config1-batched: shots per launch")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="profiling runs only: shorten each shot (each gradient record) to "
                         "this many time steps")
    ap.add_argument("--tuning", action="store_true", help="print per-variant autotune times")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on a single GPU; --gpus must be 1 "
              "(shots are independent, multi-GPU scaling is a shot queue per device).",
              file=sys.stderr)
        sys.exit(2)
    import seisacomod2d_amd as sam
    if args.workload == "batched":
        mbs = (args.max_batch,) if args.max_batch else (0, 8, 16)
        res = bench_batched(args, sam, 16 if args.quick else 64, 1.0 if args.quick else 4.0, mbs)
        print(json.dumps(dict({"metric": METRIC, "value": res["batched_shots_per_s"],
                               "unit": "shots/s", "n_gpus": 1, "higher_is_better": True,
                               "dtype": "f32", "data": "synthetic salt-type model (seeded)"},
                              **res)), flush=True)
        return
    if args.workload == "gradient":
        print(json.dumps(bench_gradient(args, sam)), flush=True)
        return
    if args.workload == "config1-batched":
        print(json.dumps(bench_c1_batched(args, sam, args.shots or 8)), flush=True)
        return
    out = bench_plain(args, sam)
    if not args.no_extra and args.profile_steps <= 0:
        out["batched"] = bench_batched(args, sam)
        out["config1_batched"] = bench_c1_batched(args, sam)
        g = gradient_c4(args, sam)[0]
        out["gradient"] = g
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(*model())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
