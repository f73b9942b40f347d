"""GPU parity of the checkpointed adjoint gradient (sam2d_gradient) against
the FP64 oracle adjoint (pinned by finite differences in
test_oracle_adjoint.py) and central differences of the FP64 functional.

Tolerances: the GPU runs the FP32 scheme and accumulates the gradient in
FP32, the oracle is FP64; relative L2 error of dJ/dvp and dJ/drho <= 1e-4,
directional derivatives <= 1e-4 of their magnitude.  Runs that do identical
FP32 arithmetic (any checkpoint interval, any batch split) are compared bit
for bit.
"""
import numpy as np
import pytest

import seisacomod2d_amd as sam

pytestmark = pytest.mark.gpu


def _case(nx=48, nz=36, seed=4):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, nx)[:, None]; z = np.linspace(0, 1, nz)[None, :]
    vp = (1500 + 300 * z + 60 * np.sin(6 * x) + 15 * rng.standard_normal((nx, nz))).astype(np.float32)
    rho = (1600 + 200 * z + 40 * np.cos(5 * x) + 10 * rng.standard_normal((nx, nz))).astype(np.float32)
    return vp, rho


def _setup(npml=8, T=0.2, nx=48, nz=36):
    vp, rho = _case(nx, nz)
    grid = sam.Grid(nx, nz, 10.0, 10.0, T, 0.001, 0.002, 25.0, npml)
    rec = np.stack([np.arange(nx) * 10.0, np.full(nx, 30.0)], 1)
    rec2 = np.stack([np.full(8, 30.0), np.arange(8) * 30.0 + 5], 1)
    shots = [sam.Shot(150.0, 50.0, rec), sam.Shot(380.0, 120.0, rec2)]
    return grid, vp, rho, shots


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle_grad(oracle, grid, vp, rho, shots, adj):
    P = oracle.params(grid.nx, grid.nz, grid.dx, grid.dz, grid.rec_len, grid.fd_dt, grid.real_dt,
                      grid.dom_freq, grid.npml)
    return P, oracle.gradient(P, vp, rho, [[s.sx, s.sz] for s in shots], [s.rec_xz for s in shots], adj)


@pytest.mark.parametrize("npml,T", [(8, 0.2), (2, 0.35)])
def test_gradient_vs_fp64_oracle(oracle, npml, T):
    # npml = 2: the wave reaches the last rows and columns, whose edge rules
    # a thick sponge damps away
    grid, vp, rho, shots = _setup(npml=npml, T=T)
    rng = np.random.default_rng(1)
    with sam.Modeller(grid, vp, rho) as m:
        adj = [rng.standard_normal((s.nrec, m.real_nt)).astype(np.float32) for s in shots]
        gvp, grho = m.gradient(shots, adj, ckpt_k=13)
    P, (ovp, orho) = _oracle_grad(oracle, grid, vp, rho, shots, adj)
    print(f"[gpu] gradient rel L2: vp {_rel(gvp, ovp):.2e} rho {_rel(grho, orho):.2e}", flush=True)
    # about 1e-6 measured; BZ of row nnz-2 halved in the recomputation gave 4e-5
    # with npml = 2
    assert _rel(gvp, ovp) <= 1e-5 and _rel(grho, orho) <= 1e-5


def test_gradient_directional_derivatives_fp64(oracle):
    # Ten seeded directions: <g_gpu, d> against central differences of the
    # FP64 functional J = sum adj * p (exact float steps, FP64 buoyancy chain).
    grid, vp, rho, shots = _setup()
    rng = np.random.default_rng(2)
    with sam.Modeller(grid, vp, rho) as m:
        adj = [rng.standard_normal((s.nrec, m.real_nt)).astype(np.float32) for s in shots]
        gvp, grho = m.gradient(shots, adj, ckpt_k=25)
    P = oracle.params(grid.nx, grid.nz, grid.dx, grid.dz, grid.rec_len, grid.fd_dt, grid.real_dt,
                      grid.dom_freq, grid.npml)
    src = [[s.sx, s.sz] for s in shots]
    recs = [s.rec_xz for s in shots]
    oracle.set_f64_buoyancy(True)
    try:
        for k in range(10):
            dv = (np.round(rng.standard_normal(vp.shape) * 8) / 256).astype(np.float32)
            dr = (np.round(rng.standard_normal(rho.shape) * 8) / 256).astype(np.float32)
            for a, b in ((dv, 0 * dr), (0 * dv, dr)):
                fd = (oracle.functional(P, vp + a, rho + b, src, recs, adj)
                      - oracle.functional(P, vp - a, rho - b, src, recs, adj)) / 2
                an = float(np.sum(gvp.astype(np.float64) * a) + np.sum(grho.astype(np.float64) * b))
                assert abs(fd - an) <= 1e-4 * abs(fd), (k, fd, an)
    finally:
        oracle.set_f64_buoyancy(False)


def test_checkpoint_interval_and_batching_bitwise():
    grid, vp, rho, shots = _setup(T=0.15)
    rng = np.random.default_rng(3)
    with sam.Modeller(grid, vp, rho) as m:
        adj = [rng.standard_normal((s.nrec, m.real_nt)).astype(np.float32) for s in shots]
        ref = m.gradient(shots, adj, ckpt_k=m.fd_nt)        # one segment, no recompute
        for k in (1, 7, 50):
            g = m.gradient(shots, adj, ckpt_k=k)
            assert np.array_equal(g[0], ref[0]) and np.array_equal(g[1], ref[1]), k
        m.set_option("max_batch", 1)
        g = m.gradient(shots, adj, ckpt_k=7)
        assert np.array_equal(g[0], ref[0]) and np.array_equal(g[1], ref[1])


def test_multistep_forward_sweep_bitwise():
    # With a multi-step variant selected and a checkpoint interval that is a
    # multiple of its steps per launch, the forward sweep of the gradient runs
    # several time steps per launch; the checkpoints, the recorded gathers and
    # so the gradient are the same bits.
    grid, vp, rho, shots = _setup(T=0.15)
    rng = np.random.default_rng(6)
    names = sam.variants()
    with sam.Modeller(grid, vp, rho) as m:
        adj = [rng.standard_normal((s.nrec, m.real_nt)).astype(np.float32) for s in shots]
        m.set_option("variant", names.index("waves2_chunk32_pf2"))
        ref = m.gradient(shots, adj, ckpt_k=10, want_seis=True)
        for name in ("twostep_waves2_chunk32", "twostep_waves4_chunk64_pf4",
                     "threestep_waves1_chunk32", "fourstep_waves2_chunk64",
                     "fivestep_waves1_chunk32"):
            m.set_option("variant", names.index(name))
            for k in (12, 10, 7, 3):          # not a multiple: single-step sweep
                g = m.gradient(shots, adj, ckpt_k=k, want_seis=True)
                assert np.array_equal(g[0], ref[0]) and np.array_equal(g[1], ref[1]), (name, k)
                for a, b in zip(g[2], ref[2]):
                    assert np.array_equal(a, b), (name, k)
        # The three-step recomputation stores p = px + pz for its intermediate
        # states; both adjoint kernels read those slots.  The chunk-column
        # order of the forward sweep and the recomputation changes nothing.
        m.set_option("variant", names.index("threestep_waves2_chunk64"))
        for steps, xo, chunk in ((1, 0, 0), (2, 3, 40), (1, 5, 24), (2, 6, 43), (2, 2, 42)):
            m.set_option("adjoint_steps", steps)
            m.set_option("tile_order", xo)
            m.set_option("chunk", chunk)
            for k in (12, 8, 7):
                g = m.gradient(shots, adj, ckpt_k=k)
                assert np.array_equal(g[0], ref[0]) and np.array_equal(g[1], ref[1]), (steps, xo, k)


def _deep_setup(dz=10.0, T=0.16):
    # nz = 100 (116 padded rows): the one-step adjoint works on strips of 58
    # rows, the two-step adjoint on strips of 52.  Receivers sit on padded rows
    # on both sides of every strip boundary (within 6 rows), where a strip
    # needs the residual of a row it does not own.
    nx, nz, npml = 40, 100, 8
    vp, rho = _case(nx, nz)
    grid = sam.Grid(nx, nz, 10.0, dz, T, 0.001, 0.002, 25.0, npml)
    rows = [38, 40, 42, 43, 44, 47, 48, 49, 50, 52, 90, 93, 96, 99]
    rec = np.array([[(2 + 2 * i) * 10.0 + 1.0, iz * dz + 1.0] for i, iz in enumerate(rows)])
    rec2 = np.stack([np.arange(nx) * 10.0 + 1.0, np.full(nx, 49 * dz + 1.0)], 1)
    shots = [sam.Shot(150.0, 30 * dz + 1.0, rec), sam.Shot(260.0, 70 * dz + 1.0, rec2)]
    return grid, vp, rho, shots


@pytest.mark.parametrize("steps", [1, 2])
def test_receiver_rows_at_strip_boundaries_vs_fp64_oracle(oracle, steps):
    grid, vp, rho, shots = _deep_setup()
    rng = np.random.default_rng(8)
    with sam.Modeller(grid, vp, rho) as m:
        m.set_option("adjoint_steps", steps)
        adj = [rng.standard_normal((s.nrec, m.real_nt)).astype(np.float32) for s in shots]
        gvp, grho = m.gradient(shots, adj, ckpt_k=11)
    P, (ovp, orho) = _oracle_grad(oracle, grid, vp, rho, shots, adj)
    print(f"[gpu] steps {steps} rel L2: vp {_rel(gvp, ovp):.2e} rho {_rel(grho, orho):.2e}",
          flush=True)
    assert _rel(gvp, ovp) <= 1e-4 and _rel(grho, orho) <= 1e-4


@pytest.mark.parametrize("dz", [10.0, 8.0])
def test_two_step_adjoint_bitwise_equal_one_step(dz):
    # The two-step adjoint forms every product and sum of two one-step passes
    # in the same order, so the gradient is the same bits for any tile shape
    # and checkpoint interval (odd segments end with a one-step pass).
    grid, vp, rho, shots = _deep_setup(dz)
    rng = np.random.default_rng(9)
    names = sam.variants()
    with sam.Modeller(grid, vp, rho) as m:
        adj = [rng.standard_normal((s.nrec, m.real_nt)).astype(np.float32) for s in shots]
        m.set_option("adjoint_variant", names.index("waves2_chunk32_pf2"))
        m.set_option("adjoint_steps", 1)
        ref = m.gradient(shots, adj, ckpt_k=13)
        m.set_option("adjoint_steps", 2)
        for name in ("waves1_chunk16_pf2", "waves2_chunk32_pf2", "waves4_chunk64_pf4",
                     "waves8_chunk128_pf2_nts", "waves1_chunk256_pf4"):
            m.set_option("adjoint_variant", names.index(name))
            for k in (13, 6, 1):
                g = m.gradient(shots, adj, ckpt_k=k)
                assert np.array_equal(g[0], ref[0]) and np.array_equal(g[1], ref[1]), (name, k)
        # the chunk-column order of the adjoint tiles permutes independent tiles
        for xo in (2, 3, 5):
            m.set_option("adjoint_tile_order", xo)
            for steps in (1, 2):
                m.set_option("adjoint_steps", steps)
                g = m.gradient(shots, adj, ckpt_k=13)
                assert np.array_equal(g[0], ref[0]) and np.array_equal(g[1], ref[1]), (xo, steps)


def test_gradient_forward_seismograms_match_forward():
    grid, vp, rho, shots = _setup(T=0.12)
    rng = np.random.default_rng(4)
    with sam.Modeller(grid, vp, rho) as m:
        fwd = m.forward(shots)
        adj = [rng.standard_normal(f.shape).astype(np.float32) for f in fwd]
        gvp, grho, seis = m.gradient(shots, adj, ckpt_k=9, want_seis=True)
    for a, b in zip(seis, fwd):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("windows", [None, [(5, 40), (30, 45)]])
def test_forward_checkpointed_and_backward_equal_the_gradient_bitwise(windows):
    # sam2d_forward_checkpointed + sam2d_gradient_checkpointed are the two
    # halves of sam2d_gradient: the same seismograms as the forward call and
    # the same gradient bits, the backward repeatable with other sources and
    # unaffected by modelling calls in between; a sam2d_gradient call drops
    # the checkpoints.
    grid, vp, rho, shots = _setup(nx=80, T=0.15)
    if windows is not None:
        shots = []
        for wx0, w in windows:
            rec = np.stack([(wx0 + np.arange(w)) * 10.0 + 5, np.full(w, 30.0)], 1)
            shots.append(sam.Shot((wx0 + w // 2) * 10.0 + 5, 60.0, rec))
    rng = np.random.default_rng(8)
    with sam.Modeller(grid, vp, rho) as m:
        ref_seis = m.forward(shots) if windows is None else m.forward_windows(shots, windows)
        adj1 = [rng.standard_normal(f.shape).astype(np.float32) for f in ref_seis]
        adj2 = [rng.standard_normal(f.shape).astype(np.float32) for f in ref_seis]
        ref1 = m.gradient(shots, adj1, ckpt_k=13, windows=windows)
        ref2 = m.gradient(shots, adj2, ckpt_k=13, windows=windows)
        seis = m.forward_checkpointed(shots, ckpt_k=13, windows=windows)
        for a, b in zip(seis, ref_seis):
            assert np.array_equal(a.view(np.int32), b.view(np.int32))
        g1 = m.gradient_checkpointed(adj1)
        m.forward(shots)                      # other modelling in between
        g2 = m.gradient_checkpointed(adj2)
        for g, ref in ((g1, ref1), (g2, ref2)):
            assert np.array_equal(g[0].view(np.int32), ref[0].view(np.int32))
            assert np.array_equal(g[1].view(np.int32), ref[1].view(np.int32))
        m.gradient(shots, adj1, ckpt_k=13, windows=windows)
        with pytest.raises(sam.SAM2DError):
            m.gradient_checkpointed(adj1)


def test_forward_checkpointed_refuses_shots_beyond_the_budget():
    # 8.1 M cells, 50 checkpoints of 16 B per cell: about 8 GB against 1 GiB
    big = sam.Grid(4000, 2000, 10.0, 10.0, 0.5, 0.001, 0.002, 25.0, 8)
    vpb = np.full((4000, 2000), 2000.0, np.float32)
    with sam.Modeller(big, vpb, vpb) as m:
        m.set_option("hbm_budget_gb", 1)
        rec = np.stack([np.arange(100) * 10.0, np.full(100, 30.0)], 1)
        with pytest.raises(sam.SAM2DError):
            m.forward_checkpointed([sam.Shot(500.0, 50.0, rec)], ckpt_k=10)


def test_windowed_gradient_vs_oracle_on_cropped_models(oracle):
    grid, vp, rho, _ = _setup(nx=80, T=0.15)
    f32 = np.float32
    wins = [(5, 40), (30, 45)]
    shots, oshots = [], []
    for wx0, w in wins:
        rec = np.stack([(wx0 + np.arange(w)) * 10.0 + 5, np.full(w, 30.0)], 1)
        shots.append(sam.Shot((wx0 + w // 2) * 10.0 + 5, 60.0, rec))
    rng = np.random.default_rng(5)
    with sam.Modeller(grid, vp, rho) as m:
        adj = [rng.standard_normal((s.nrec, m.real_nt)).astype(np.float32) for s in shots]
        gvp, grho = m.gradient(shots, adj, ckpt_k=16, windows=wins)
    ovp = np.zeros_like(gvp, dtype=np.float64)
    orho = np.zeros_like(ovp)
    for s, (wx0, w), a in zip(shots, wins, adj):
        P = oracle.params(w, grid.nz, grid.dx, grid.dz, grid.rec_len, grid.fd_dt, grid.real_dt,
                          grid.dom_freq, grid.npml)
        rec = s.rec_xz.copy(); rec[:, 0] -= wx0 * 10.0
        gv, gr = oracle.gradient(P, vp[wx0:wx0 + w], rho[wx0:wx0 + w], [[s.sx - wx0 * 10.0, s.sz]], [rec], [a])
        ovp[wx0:wx0 + w] += gv
        orho[wx0:wx0 + w] += gr
    assert _rel(gvp, ovp) <= 1e-4 and _rel(grho, orho) <= 1e-4
    del f32, oshots


def test_autograd_function_l2_misfit(oracle):
    import torch
    from seisacomod2d_amd.autograd import ModellerCache, Seismograms
    grid, vp, rho, shots = _setup(T=0.15)
    cache = ModellerCache(grid)
    # observed data from a perturbed model
    vp_true = vp.copy(); vp_true[20:28, 10:18] += 40.0
    with sam.Modeller(grid, vp_true, rho) as mt:
        obs = torch.from_numpy(np.concatenate(mt.forward(shots), 0))
    tvp = torch.tensor(vp, requires_grad=True)
    trho = torch.tensor(rho, requires_grad=True)
    seis = Seismograms.apply(tvp, trho, cache, shots, 11, None)
    loss = 0.5 * ((seis - obs) ** 2).sum()
    loss.backward()
    resid = (seis.detach() - obs).numpy()
    adj, off = [], 0
    for s in shots:
        adj.append(resid[off:off + s.nrec]); off += s.nrec
    P, (ovp, orho) = _oracle_grad(oracle, grid, vp, rho, shots, adj)
    assert _rel(tvp.grad.numpy(), ovp) <= 1e-4 and _rel(trho.grad.numpy(), orho) <= 1e-4
    cache.close()
