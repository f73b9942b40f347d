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
