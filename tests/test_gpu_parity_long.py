"""GPU parity at the configurations' record lengths, against what the compiled
reference wrote and against the checkpointed FP64 oracle adjoint.

  * full records: the 5000 x 5000 two-layer model over 4 s (4000 steps) and a
    3201 x 1201 salt-model window over 4 s at 0.5 ms (8000 steps), both as the
    reference ran them (its 32-cell sponge, one shot); every 100th / 60th
    trace of its output, stored in tests/golden/;
  * the reference's carry-over between shots (cpumodelling.cpp:343-348): its
    two-shot output file (shot 2) against the reset_quirk mode;
  * gradients over whole records: 3000 steps on a reduced 1201 x 601 anomaly
    model and 8000 steps on a reduced salt model, against the FP64 oracle
    adjoint with checkpoints (pinned by finite differences and by the
    full-storage adjoint in test_oracle_adjoint.py).

Tolerances: relative L2 error per trace <= 1e-5 for seismograms (traces
below 1e-3 of the gather's largest trace norm held to 1e-5 of that norm,
conftest.FLOOR; each test prints its worst error and how many traces were
floored); relative L2 error of dJ/dvp and dJ/drho <= 1e-4 for gradients.
"""
import numpy as np
import pytest

import seisacomod2d_amd as sam
from conftest import CARRYOVER_FLOOR, assert_seis_close, floored_traces, golden_cases, load_golden
from seisacomod2d_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _grid(g, npml=32):
    return sam.Grid(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                    g["fm"], npml)


@pytest.mark.parametrize("case", golden_cases("long"))
@pytest.mark.parametrize("tuned", [False, True])
def test_full_record_vs_reference(case, tuned):
    g = load_golden(case)
    print(f"[gpu] {case}: {g['nx']}x{g['nz']} tuned={tuned}", flush=True)
    with sam.Modeller(_grid(g), g["vp"], g["rho"]) as m:
        if tuned:
            m.autotune(steps=3)
        shot = sam.Shot(float(g["src_xz"][0, 0]), float(g["src_xz"][0, 1]), g["rec_xz"])
        seis = m.forward([shot])[0]
        steps = m.fd_nt
    ref = g["seis"]
    e = assert_seis_close(seis[g["trace_idx"]], ref, TOL, case)
    print(f"[gpu] {case}: {steps} steps, {len(ref)} traces, worst trace error {e.max():.2e}, "
          f"median {np.median(e):.2e}, floored traces {floored_traces(ref)}", flush=True)


def test_reset_quirk_vs_reference_two_shot_file():
    # The reference's output file after a two-shot card holds shot 2, which
    # carried shot 1's final wavefield.  Both shots in one call, and one shot
    # per call (how modelling_module runs them), give it.
    g = load_golden("two_shot_carryover")
    shots = [sam.Shot(float(x), float(z), g["rec_xz"]) for x, z in g["src_xz"]]
    with sam.Modeller(_grid(g), g["vp"], g["rho"]) as m:
        m.set_option("reset_quirk", 1)
        both = m.forward(shots)
        m.set_option("reset_quirk", 1)        # restart the chain
        one = [m.forward([s])[0] for s in shots]
        m.set_option("reset_quirk", 0)
        fresh = m.forward(shots)
    e = assert_seis_close(both[1], g["seis"], TOL, "two shots, one call", CARRYOVER_FLOOR)
    assert_seis_close(one[1], g["seis"], TOL, "two shots, two calls", CARRYOVER_FLOOR)
    assert np.array_equal(both[0], fresh[0])        # shot 1 starts fresh either way
    assert np.abs(fresh[1] - g["seis"]).max() > 0.05 * np.abs(g["seis"]).max()
    print(f"[gpu] two-shot carry-over: worst trace error {e.max():.2e}, "
          f"floored {floored_traces(g['seis'], CARRYOVER_FLOOR)}", flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _long_gradient(oracle, grid, vtrue, v0, rho, shots, k):
    with sam.Modeller(grid, vtrue, rho) as mt:
        obs = mt.forward(shots)
    with sam.Modeller(grid, v0, rho) as m:
        d = m.forward(shots)
        adj = [a - b for a, b in zip(d, obs)]
        gvp, grho = m.gradient(shots, adj, ckpt_k=k)
        nt = m.fd_nt
    P = oracle.params(grid.nx, grid.nz, grid.dx, grid.dz, grid.rec_len, grid.fd_dt, grid.real_dt,
                      grid.dom_freq, grid.npml)
    ovp, orho = oracle.gradient(P, v0, rho, [[s.sx, s.sz] for s in shots],
                                [s.rec_xz for s in shots], adj, ckpt_k=k)
    return nt, _rel(gvp, ovp), _rel(grho, orho)


def test_gradient_3000_steps_vs_checkpointed_fp64_oracle(oracle):
    # The anomaly configuration reduced to 241 x 121 cells (10 m), whole 3 s
    # record at 1 ms, 10 Hz, two surface shots, residual of the background
    # model against the model with anomalies.
    vtrue, v0, rho = synthetic.anomaly_model(nx=241, nz=121)
    grid = sam.Grid(241, 121, 10.0, 10.0, 3.0, 0.001, 0.001, 10.0, npml=32)
    rec = np.stack([np.arange(241) * 10.0, np.full(241, 10.0)], 1).astype(np.float32)
    shots = [sam.Shot(600.0, 10.0, rec), sam.Shot(1800.0, 10.0, rec)]
    nt, evp, erho = _long_gradient(oracle, grid, vtrue, v0, rho, shots, 100)
    print(f"[gpu] gradient over {nt} steps: rel L2 vp {evp:.2e} rho {erho:.2e}", flush=True)
    assert nt == 3000 and evp <= 1e-4 and erho <= 1e-4


def test_gradient_8000_steps_vs_checkpointed_fp64_oracle(oracle):
    # The salt-type configuration reduced to 200 x 120 cells (7.62 m), whole
    # 4 s record at 0.5 ms (8000 steps), 15 Hz, the smoothed model as the
    # starting model, checkpoints every 200 steps.
    from scipy.ndimage import gaussian_filter
    vtrue, rho = synthetic.salt_model(nx=200, nz=120, h=7.62, seed=5)
    v0 = gaussian_filter(vtrue.astype(np.float64), 6.0).astype(np.float32)
    grid = sam.Grid(200, 120, 7.62, 7.62, 4.0, 0.0005, 0.0005, 15.0, npml=32)
    rec = np.stack([np.arange(200) * 7.62, np.full(200, 15.24)], 1).astype(np.float32)
    shots = [sam.Shot(99 * 7.62, 15.24, rec)]
    nt, evp, erho = _long_gradient(oracle, grid, vtrue, v0, rho, shots, 200)
    print(f"[gpu] gradient over {nt} steps: rel L2 vp {evp:.2e} rho {erho:.2e}", flush=True)
    assert nt == 8000 and evp <= 1e-4 and erho <= 1e-4
