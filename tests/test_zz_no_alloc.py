"""GPU test of sam2d_forward_noalloc, the entry point that allocates no
device memory (its working buffers live in the library's static scratch
array), called through the C ABI: its seismograms equal sam2d_forward_dev's
bit for bit, on two calls in a row of which the second has the smaller
input, with a single-step and a multi-step kernel variant; a shot beyond the
array's size is refused with a status."""
import ctypes

import numpy as np
import pytest

import seisacomod2d_amd as sam
from conftest import assert_seis_close, load_golden

pytestmark = pytest.mark.gpu


class _Dev:
    """Caller-owned device buffers (hipMalloc by the test, not by the library),
    through the HIP runtime libsam2d.so is bound to: its soname, once the
    library is loaded (a process that imported torch also holds torch's own
    copy of the runtime under the plain name libamdhip64.so)."""

    def __init__(self):
        sam.lib()
        self.L = ctypes.CDLL("libamdhip64.so.7")

    def alloc(self, n):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(4 * n)) == 0
        return p.value

    def get(self, p, n):
        out = np.empty(n, np.float32)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p),
                                ctypes.c_size_t(4 * n), 2) == 0          # device to host
        return out

    def free(self, p):
        self.L.hipFree(ctypes.c_void_p(p))


def _case(name):
    g = load_golden(name)
    grid = sam.Grid(g["nx"], g["nz"], g["dx"], g["dz"], g["rec_len"], g["fd_dt"], g["real_dt"],
                    g["fm"], 32)
    shot = sam.Shot(float(g["src_xz"][0, 0]), float(g["src_xz"][0, 1]), g["rec_xz"])
    return g, grid, shot


def _both(m, shot, dev, order):
    """(noalloc, forward_dev) seismograms, the entry points called in `order`."""
    n = shot.nrec * m.real_nt
    a, b = dev.alloc(n), dev.alloc(n)
    try:
        for which in order:
            if which == "noalloc":
                m.forward_noalloc(shot, a)
            else:
                m.forward_dev([shot], b)
        return dev.get(a, n).reshape(shot.nrec, -1), dev.get(b, n).reshape(shot.nrec, -1)
    finally:
        dev.free(a)
        dev.free(b)


@pytest.mark.parametrize("variant", ["waves4_chunk32_pf2", "threestep_waves1_chunk64"])
def test_noalloc_equals_forward_dev_on_two_calls(variant):
    dev = _Dev()
    names = sam.variants()
    # first call: the larger grid
    g, grid, shot = _case("hetero_explicit")
    print(f"[gpu] noalloc {variant}: hetero_explicit {grid.nx}x{grid.nz}", flush=True)
    with sam.Modeller(grid, g["vp"], g["rho"]) as m:
        m.set_option("variant", names.index(variant))
        x, y = _both(m, shot, dev, ["noalloc", "dev"])
        assert m.stats()["variant"] == names.index(variant)
    assert np.array_equal(x.view(np.int32), y.view(np.int32))
    assert_seis_close(x[g["trace_idx"]], g["seis"], 1e-5, "noalloc hetero_explicit")
    # second call: a smaller grid and fewer receivers, on the same scratch array
    g, grid, shot = _case("tiny_twolayer")
    shot = sam.Shot(shot.sx, shot.sz, shot.rec_xz[::3])
    print(f"[gpu] noalloc {variant}: tiny_twolayer {grid.nx}x{grid.nz}", flush=True)
    with sam.Modeller(grid, g["vp"], g["rho"]) as m:
        m.set_option("variant", names.index(variant))
        x, y = _both(m, shot, dev, ["dev", "noalloc"])
        x2, _ = _both(m, shot, dev, ["noalloc"])
    assert np.array_equal(x.view(np.int32), y.view(np.int32))
    assert np.array_equal(x2.view(np.int32), y.view(np.int32))
    assert_seis_close(x, g["seis"][::3], 1e-5, "noalloc tiny_twolayer")


def test_noalloc_refuses_a_shot_beyond_the_scratch_array():
    # 2600 x 2600 cells plus the sponge need 32 B per cell, more than the array
    nx = nz = 2600
    grid = sam.Grid(nx, nz, 10.0, 10.0, 0.01, 0.001, 0.001, 15.0, 32)
    vp = np.full((nx, nz), 2000.0, np.float32)
    rec = np.array([[100.0, 20.0]], np.float32)
    dev = _Dev()
    with sam.Modeller(grid, vp, vp) as m:
        out = dev.alloc(m.real_nt)                 # a valid output buffer all the same
        try:
            with pytest.raises(sam.SAM2DError) as e:
                m.forward_noalloc(sam.Shot(500.0, 20.0, rec), out)
        finally:
            dev.free(out)
    assert e.value.code == -6 and "scratch" in str(e.value)
