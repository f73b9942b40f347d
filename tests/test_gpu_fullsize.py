"""Properties at the benchmark configurations' full sizes (GPU).

The FP64 oracle cannot run a full 5000 x 5000 x 4000-step shot or a 3201 x
1201 x 8000-step gradient in test time, so at these sizes the tests check
what holds independently of size: the tuned multi-step path gives the same
bits as the single-step kernel (which the oracle pins at smaller sizes and on
config 1's full grid over its first steps), every trace is finite and carries
the direct wave, and the gradient is the same bits for any checkpoint
interval and for the one- and two-step adjoints; the whole batched workload
equals its per-shot loop, and the whole gradient workload is the same bits in
one batch, in budget-split batches and as its two halves.
"""
import numpy as np
import pytest

import seisacomod2d_amd as sam

pytestmark = pytest.mark.gpu


def test_config1_full_record_tuned_equals_single_step():
    nx = nz = 5000
    vp = np.full((nx, nz), 1500, np.float32); vp[:, 250:] = 3000
    rho = np.full((nx, nz), 1000, np.float32); rho[:, 250:] = 2000
    rec = np.stack([np.arange(0, nx, 7) * 10.0, np.full(len(range(0, nx, 7)), 20.0)], 1)
    grid = sam.Grid(nx, nz, 10.0, 10.0, 4.0, 0.001, 0.001, 15.0, npml=50)
    names = sam.variants()
    shot = sam.Shot(10.0 * 2460, 20.0, rec)
    print("[gpu] config1 full record, tuned vs single-step", flush=True)
    with sam.Modeller(grid, vp, rho) as m:
        best = m.autotune(steps=6)
        tuned = m.forward([shot])[0]
        print(f"[gpu] tuned {names[best]} order {m.get_option('tile_order')} "
              f"chunk {m.get_option('chunk')}", flush=True)
        m.set_option("variant", names.index("waves4_chunk32_pf2"))
        single = m.forward([shot])[0]
    assert tuned.shape == (len(rec), 4000)
    assert np.all(np.isfinite(tuned))
    assert np.array_equal(tuned.view(np.int32), single.view(np.int32))
    # the direct wave (1500 m/s) reaches every receiver within 5 km of the
    # source inside the 4 s record
    near = np.abs(rec[:, 0] - shot.sx) < 5000.0
    assert np.all(np.abs(tuned[near]).max(axis=1) > 0)


def test_config4_window_full_length_gradient_invariants():
    # One 2400-column window of the config 2/4 shape (3201 x 1201, dx 7.62,
    # sponge 60, 8000 steps), smooth model with a fast inclusion.
    nx, nz, h = 3201, 1201, 7.62
    x = np.arange(nx)[:, None] * h; z = np.arange(nz)[None, :] * h
    vp = (1500 + 0.5 * z + 2500 * (np.hypot(x - 12000, z - 5500) < 1500)).astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    grid = sam.Grid(nx, nz, h, h, 4.0, 0.0005, 0.0005, 15.0, npml=60)
    wx0, w = 400, 2400
    rec = np.stack([(wx0 + np.arange(w)) * h + 0.5 * h, np.full(w, 15.24)], 1).astype(np.float32)
    shot = sam.Shot((wx0 + w // 2 + 0.5) * h, 15.24, rec)
    rng = np.random.default_rng(12)
    print("[gpu] config4 window gradient, k = 200 / 57, one- and two-step adjoint", flush=True)
    with sam.Modeller(grid, vp, rho) as m:
        m.autotune(windows=[(wx0, w)], nshots=1, steps=6)
        d = m.forward_windows([shot], [(wx0, w)])[0]
        adj = [(d * (1 + 0.1 * rng.standard_normal(d.shape))).astype(np.float32)]
        g200 = m.gradient([shot], adj, ckpt_k=200, windows=[(wx0, w)])
        g57 = m.gradient([shot], adj, ckpt_k=57, windows=[(wx0, w)])
        m.set_option("adjoint_steps", 1)
        g1 = m.gradient([shot], adj, ckpt_k=200, windows=[(wx0, w)])
    assert np.all(np.isfinite(g200[0])) and np.all(np.isfinite(g200[1]))
    assert np.abs(g200[0]).max() > 0 and np.abs(g200[1]).max() > 0
    # nothing outside the window is touched
    assert not np.any(g200[0][:wx0]) and not np.any(g200[0][wx0 + w:])
    for g in (g57, g1):
        assert np.array_equal(g[0], g200[0]) and np.array_equal(g[1], g200[1])


def _salt_batch(n, seed):
    from seisacomod2d_amd import synthetic
    vp, rho = synthetic.salt_model()
    sh, wins = synthetic.salt_shots(n, seed=seed)
    return vp, rho, [sam.Shot(sx, sz, r) for sx, sz, r in sh], wins


def test_config2_full_batch_equals_per_shot_loop_bitwise():
    # The whole batched workload: 64 windowed shots of the salt-type model,
    # 8000 steps, all resident at once (about 6.4 GB), against the per-shot
    # loop over the same windows, every trace bit for bit.
    vp, rho, shots, wins = _salt_batch(64, 7)
    grid = sam.Grid(3201, 1201, 7.62, 7.62, 4.0, 0.0005, 0.0005, 15.0, npml=60)
    print("[gpu] config2 full batch, 64 shots x 8000 steps", flush=True)
    with sam.Modeller(grid, vp, rho) as m:
        m.autotune(windows=wins, nshots=len(shots), steps=6)
        batched = m.forward_batched(shots, wins)
        loop = m.forward_windows(shots, wins)
    assert len(batched) == 64
    for s, (a, b) in enumerate(zip(batched, loop)):
        assert a.shape == (shots[s].nrec, 8000)
        assert np.array_equal(a.view(np.int32), b.view(np.int32)), s
        assert np.all(np.isfinite(a)) and np.abs(a).max() > 0


def test_config4_full_gradient_batch_split_by_the_memory_budget_bitwise():
    # The gradient workload at its full size (32 windowed shots, 8000 steps,
    # k = 100, 189 GB when resident at once) gives the same bits when a
    # 64 GiB budget splits it into equal batches, and when its two halves
    # (modelling with checkpoints, backward) run as separate calls.
    vp, rho, shots, wins = _salt_batch(32, 11)
    from scipy.ndimage import gaussian_filter
    v0 = gaussian_filter(vp.astype(np.float64), 20.0).astype(np.float32)
    grid = sam.Grid(3201, 1201, 7.62, 7.62, 4.0, 0.0005, 0.0005, 15.0, npml=60)
    print("[gpu] config4 full gradient, one batch / 64 GiB batches / two halves", flush=True)
    with sam.Modeller(grid, vp, rho) as mt:
        obs = mt.forward_batched(shots, wins)
    with sam.Modeller(grid, v0, rho) as m:
        d = m.forward_batched(shots, wins)
        adj = [a - b for a, b in zip(d, obs)]
        one = m.gradient(shots, adj, ckpt_k=100, windows=wins)
        seis = m.forward_checkpointed(shots, ckpt_k=100, windows=wins)
        halves = m.gradient_checkpointed(adj)
        m.set_option("hbm_budget_gb", 64)
        split = m.gradient(shots, adj, ckpt_k=100, windows=wins)
    for a, b in zip(seis, d):
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert np.all(np.isfinite(one[0])) and np.abs(one[0]).max() > 0
    for g in (halves, split):
        assert np.array_equal(g[0].view(np.int32), one[0].view(np.int32))
        assert np.array_equal(g[1].view(np.int32), one[1].view(np.int32))


def test_config1_batched_shots_equal_single_step_bitwise():
    # The C1 shots through the batched forward as the benchmark runs them
    # (eight per launch, the three-step kernel with 193-column marches, tile
    # order 9), 300 steps, against the single-step kernel one shot at a time,
    # every trace bit for bit.
    from seisacomod2d_amd import synthetic
    vp, rho = synthetic.two_layer(5000, 5000)
    shots = [sam.Shot(sx, sz, rec) for sx, sz, rec in
             synthetic.surface_spread_shots(5000, 10.0, 64)][:8]
    grid = sam.Grid(5000, 5000, 10.0, 10.0, 0.3, 0.001, 0.001, 15.0, npml=50)
    print("[gpu] config1 batched, 8 shots x 300 steps", flush=True)
    with sam.Modeller(grid, vp, rho) as m:
        m.set_option("variant", sam.variants().index("threestep_waves4_chunk256_pf4"))
        m.set_option("chunk", 193)
        m.set_option("tile_order", 9)
        batched = m.forward_batched(shots)
        m.set_option("variant", sam.variants().index("waves4_chunk32_pf2"))
        m.set_option("chunk", 0)
        m.set_option("tile_order", 0)
        loop = [m.forward([s])[0] for s in shots]
    for s, (a, b) in enumerate(zip(batched, loop)):
        assert np.array_equal(a.view(np.int32), b.view(np.int32)), s
        assert np.all(np.isfinite(a)) and np.abs(a).max() > 0
