"""The GPU's seismograms against the FP32 model of its kernels
(oracle/sam2d_fp32.c), bit for bit: the model restates the kernels'
arithmetic on the CPU from the same FP32 fields, so every kernel variant,
single- or multi-step, must give its exact bits.  Together with the model's
own parity against the FP64 oracle (test_oracle_fp32.py) this pins the
kernels' arithmetic cell by cell, not only within a tolerance."""
import numpy as np
import pytest

import seisacomod2d_amd as sam
from conftest import load_golden
from test_oracle_fp32 import fp32_model_forward

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["hetero_explicit", "salt_small", "rect_dtfactor", "tiny_twolayer"])
@pytest.mark.parametrize("variant", [None, "threestep_waves1_chunk64", "fivestep_waves2_chunk32"])
def test_gpu_bits_equal_fp32_model(oracle, case, variant):
    g = load_golden(case)
    print(f"[gpu] fp32 model {case} variant={variant}", flush=True)
    grid = sam.Grid(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                    g["fm"], 32)
    with sam.Modeller(grid, g["vp"], g["rho"]) as m:
        if variant:
            m.set_option("variant", sam.variants().index(variant))
        seis = m.forward([sam.Shot(float(g["src_xz"][0, 0]), float(g["src_xz"][0, 1]), g["rec_xz"])])[0]
    model = fp32_model_forward(oracle, g)
    assert np.array_equal(seis.view(np.int32), model.view(np.int32))


@pytest.mark.parametrize("variant", [None, "twostep_waves1_chunk16", "threestep_waves1_chunk64",
                                     "fivestep_waves2_chunk32"])
def test_gpu_bits_equal_fp32_model_thin_sponge(oracle, variant):
    """A 3-cell sponge on a 20 x 30 grid, receivers at every cell: the wave
    reaches the last rows within the record, so the bottom edge rules
    (BZ of the last two rows, the one-sided stencils) show at the receivers
    bit for bit."""
    nx, nz, dx, dz = 20, 30, 10.0, 8.0
    vp = np.full((nx, nz), 1800, np.float32)
    vp[:, nz // 2:] = 2600
    vp[nx // 3: nx // 2, nz // 4: nz // 2] = 3200
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    rec = np.stack([ix.ravel() * dx, iz.ravel() * dz], 1).astype(np.float32)
    g = dict(nx=nx, nz=nz, dx=dx, dz=dz, rec_len=0.2, fd_dt=0.001, real_dt=0.001, fm=15.0,
             vp=vp, rho=rho, src_xz=np.array([[(nx // 2) * dx, 2 * dz]]), rec_xz=rec)
    print(f"[gpu] fp32 model thin sponge variant={variant}", flush=True)
    grid = sam.Grid(nx, nz, dx, dz, 0.2, 0.001, 0.001, 15.0, 3)
    with sam.Modeller(grid, vp, rho) as m:
        if variant:
            m.set_option("variant", sam.variants().index(variant))
        seis = m.forward([sam.Shot(float(g["src_xz"][0, 0]), float(g["src_xz"][0, 1]), rec)])[0]
    model = fp32_model_forward(oracle, g, npml=3)
    assert np.array_equal(seis.view(np.int32), model.view(np.int32))
