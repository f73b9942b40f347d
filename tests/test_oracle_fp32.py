"""The FP32 model of the product's kernels (oracle/sam2d_fp32.c) against the
FP64 oracle and the reference's goldens, and the FP32 error budget.

The model takes the FP32 fields the library builds (sam2d_coefficients) and
runs the kernels' arithmetic on the CPU; test_gpu_fp32_model.py checks that
the GPU gives its bits.  Here it shows what sets the FP32 error of long
records: rounding every coefficient to the nearest float gives all the cells
of a homogeneous region the same relative error, a systematic change of the
wave speed whose phase error grows along the record; the library's unbiased
rounding (sam2d_setup.cpp) removes it.
"""
import numpy as np
import pytest

import seisacomod2d_amd as sam
from conftest import assert_seis_close, golden_cases, load_golden, trace_errors


def _cells(g, nnz, npml, xz):
    f = np.float32
    ix = [int(f(x) * (f(1) / f(g["dx"]))) + npml for x, _ in xz]
    iz = [int(f(z) * (f(1) / f(g["dz"]))) + npml for _, z in xz]
    return [a * nnz + b for a, b in zip(ix, iz)]


def fp32_model_forward(oracle, g, fields=None, npml=32):
    """Seismogram of the golden's shot from the FP32 model of the kernels."""
    grid = sam.Grid(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                    g["fm"], npml)
    fields = fields or sam.coefficients(grid, g["vp"], g["rho"])
    fd_nt, real_nt, dtf, nts, nnx, nnz = sam.time_axis(grid)
    sx, sz = (float(v) for v in g["src_xz"][0])
    amp = sam.source_increments(grid, g["vp"], g["rho"], sx, sz)
    src = _cells(g, nnz, npml, [(sx, sz)])[0]
    seis, _ = oracle.fp32_forward(fields, fd_nt, nts, amp, src, _cells(g, nnz, npml, g["rec_xz"]),
                                  dtf, real_nt)
    return seis


@pytest.mark.parametrize("case", golden_cases())
def test_fp32_model_vs_reference_goldens(oracle, case):
    g = load_golden(case)
    seis = fp32_model_forward(oracle, g)[g["trace_idx"]]
    assert_seis_close(seis, g["seis"], 1e-5, case)


def _water_case(nx=300, nz=100, h=7.62):
    # The salt-type configuration's water layer (vp 1500 m/s, Gardner density)
    # over a whole 4 s record at 0.5 ms: 8000 steps, surface spread.
    vp = np.full((nx, nz), 1500.0, np.float32)
    rho = np.full((nx, nz), np.float32(310.0 * 1500.0 ** 0.25), np.float32)
    rec = np.stack([np.arange(nx) * h, np.full(nx, 2 * h)], 1).astype(np.float32)
    return dict(nx=nx, nz=nz, dx=h, dz=h, rec_len=4.0, fd_dt=0.0005, real_dt=0.0005, fm=15.0,
                vp=vp, rho=rho, src_xz=np.array([[20 * h, 2 * h]], np.float32), rec_xz=rec)


def test_unbiased_coefficient_rounding_removes_the_long_record_drift(oracle):
    g = _water_case()
    P = oracle.params(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                      g["fm"], 32)
    ref = oracle.forward(P, g["vp"], g["rho"], g["src_xz"], [g["rec_xz"]])[0]
    grid = sam.Grid(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                    g["fm"], 32)
    fields = sam.coefficients(grid, g["vp"], g["rho"])
    unbiased = trace_errors(fp32_model_forward(oracle, g, fields), ref)
    # the same fields rounded to nearest from the oracle's FP64 coefficients
    tp = oracle.time_params(g["rec_len"], g["fd_dt"], g["real_dt"], g["fm"], g["dx"], g["dz"])
    kappa, _, bu1, bu2 = oracle.stagger(oracle.expand(g["vp"], 32), oracle.expand(g["rho"], 32))
    near = dict(fields)
    near["KX"] = near["KZ"] = (tp["dtdx"] * kappa).astype(np.float32)
    near["BX"] = (tp["dtdx"] * bu2.astype(np.float64)).astype(np.float32)
    near["BZ"] = (tp["dtdz"] * bu1.astype(np.float64)).astype(np.float32)
    nearest = trace_errors(fp32_model_forward(oracle, g, near), ref)
    print(f"8000 steps, water: worst trace error {unbiased.max():.2e} unbiased, "
          f"{nearest.max():.2e} round-to-nearest")
    assert unbiased.max() < 3e-6
    assert nearest.max() > 3 * unbiased.max()
