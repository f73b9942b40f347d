"""CPU tests of the drop-in seam and the trace output.

The reference's backend seam is the C++ function
void modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp) (src/src_cpu/modelling.h:106).
oracle/Makefile.ref links the reference's own front-end objects against
libsam2d_host.so in place of cpumodelling.o and cpu_modelling_kernels.o
(oracle/_ref/SeisAcoMod2D_MI355X_dropin.exe).  Here, without a GPU, the linked
program must run the reference's front end on a card and reach this build's
modelling_module, which reports that no device is present and ends with the
exit status of the reference's Abort(-20) (gpumodelling.cu:133-143).  The GPU
run of the same program is in test_gpu_dropin.py.

SU output: the trace headers written by write_shot_output (SEG-Y trace
identification header, the layout of src/src_cpu/segy.h:27-310) are read back
and checked against the geometry, and the samples against the raw
<job>_seismogram.bin bit for bit.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

DROPIN = os.path.join(ROOT, "oracle", "_ref", "SeisAcoMod2D_MI355X_dropin.exe")
HOST = os.path.join(ROOT, "seisacomod2d_amd", "libsam2d_host.so")


def _nm(path, defined):
    args = ["nm", "-D", "--defined-only", path] if defined else ["nm", "-u", path]
    return subprocess.run(args, capture_output=True, text=True, check=True).stdout


@pytest.fixture(scope="module")
def dropin():
    if not os.path.exists(DROPIN):
        pytest.skip("drop-in not linked (needs the reference sources: make -C oracle ref)")
    return DROPIN


def test_dropin_takes_modelling_module_from_the_host_library(dropin):
    seam = "_Z16modelling_moduleiP5Geo2d"        # modelling_module(int, Geo2d*)
    assert seam in _nm(dropin, defined=False)      # the reference's objects call it
    assert seam in _nm(HOST, defined=True)         # libsam2d_host.so defines it
    ldd = subprocess.run(["ldd", dropin], capture_output=True, text=True, check=True).stdout
    assert "libsam2d_host.so" in ldd and "libsam2d.so" in ldd and "not found" not in ldd


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: see test_gpu_dropin.py")
def test_dropin_runs_reference_front_end_into_the_seam(dropin, tmp_path):
    from golden.make_reference_cards import SMALL, write_inputs
    d = str(tmp_path)
    write_inputs(d, SMALL, "two_layer", 1, "500.000000,20.000000,40,0,10.000000,1", None)
    r = subprocess.run([dropin, "card.json"], cwd=d, capture_output=True, text=True, timeout=120)
    # the reference's front end ran: its own banner and stability report ...
    assert "Shot to be handled for rank 0 is : 1" in r.stdout
    assert "STABILITY CONDITION FULFILLED" in r.stdout
    # ... then this build's modelling_module, which needs a device
    assert "no HIP device available" in r.stderr
    assert r.returncode == (-20) & 0xFF
    assert os.path.exists(os.path.join(d, "test_geometry.txt"))


# --------------------------------------------------------------- SU output
class _Crd(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("z", ctypes.c_float)]


class _Geo(ctypes.Structure):
    _fields_ = [("nsrc", ctypes.c_int), ("nrec", ctypes.POINTER(ctypes.c_int)),
                ("src2d_sp", ctypes.POINTER(_Crd)), ("rec2d_sp", ctypes.POINTER(ctypes.POINTER(_Crd)))]


class _Job(ctypes.Structure):
    _fields_ = [("adj", ctypes.c_int), ("geomflag", ctypes.c_int), ("fdop", ctypes.c_int),
                ("frsf", ctypes.c_int), ("jbtype", ctypes.c_char * 8),
                ("shotfile", ctypes.c_char * 224), ("geomfile", ctypes.c_char * 224),
                ("logfile", ctypes.c_char * 224), ("jobname", ctypes.c_char * 256),
                ("tmppath", ctypes.c_char * 512)]


class _Opts(ctypes.Structure):
    _fields_ = [("npml", ctypes.c_int), ("per_shot_files", ctypes.c_int),
                ("write_restart", ctypes.c_int), ("reset_quirk", ctypes.c_int),
                ("su_output", ctypes.c_int)]


SU_HEADER = np.dtype({
    "names": ["tracl", "tracr", "fldr", "tracf", "trid", "offset", "gelev", "selev", "sdepth",
              "scalel", "scalco", "sx", "gx", "counit", "ns", "dt"],
    "formats": ["<i4", "<i4", "<i4", "<i4", "<i2", "<i4", "<i4", "<i4", "<i4", "<i2", "<i2",
                "<i4", "<i4", "<i2", "<u2", "<u2"],
    "offsets": [0, 4, 8, 12, 28, 36, 40, 44, 48, 68, 70, 72, 80, 88, 114, 116],
    "itemsize": 240})


def read_su(path, ns):
    rec = np.dtype([("h", SU_HEADER), ("s", "<f4", (ns,))])
    return np.fromfile(path, rec)


def _geometry(shots):
    keep = []
    nrec = (ctypes.c_int * len(shots))(*[len(r) for _, r in shots])
    src = (_Crd * len(shots))(*[_Crd(*s) for s, _ in shots])
    rows = (ctypes.POINTER(_Crd) * len(shots))()
    for i, (_, r) in enumerate(shots):
        arr = (_Crd * len(r))(*[_Crd(*p) for p in r])
        keep.append(arr)
        rows[i] = ctypes.cast(arr, ctypes.POINTER(_Crd))
    geo = _Geo(len(shots), ctypes.cast(nrec, ctypes.POINTER(ctypes.c_int)),
               ctypes.cast(src, ctypes.POINTER(_Crd)), rows)
    return geo, (nrec, src, rows, keep)


@pytest.mark.parametrize("per_shot", [0, 1])
def test_su_output_headers_and_samples(tmp_path, per_shot):
    L = ctypes.CDLL(HOST)
    L.write_shot_output.argtypes = [ctypes.c_int, ctypes.POINTER(_Geo), ctypes.POINTER(ctypes.c_float),
                                    ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_longlong)]
    job = ctypes.POINTER(_Job).in_dll(L, "job_sp").contents
    opts = _Opts.in_dll(L, "sam2d_opts")
    job.tmppath = (str(tmp_path) + "/").encode()
    job.jobname = b"card_"
    saved = (opts.su_output, opts.per_shot_files)
    opts.su_output, opts.per_shot_files = 1, per_shot
    rng = np.random.default_rng(3)
    h = 7.62
    shots = [((40 * h, 2 * h), [(i * h, 3 * h) for i in range(10, 15)]),
             ((90 * h, 2 * h), [(i * h, 4 * h + 0.5) for i in range(80, 87)])]
    geo, keep = _geometry(shots)
    ns, dt = 333, 0.0005
    seis = [rng.standard_normal((len(r), ns)).astype(np.float32) for _, r in shots]
    n = ctypes.c_longlong(0)
    try:
        for is_, s in enumerate(seis):
            rc = L.write_shot_output(is_, ctypes.byref(geo), s.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     ns, dt, 1 if is_ == 0 else 0, ctypes.byref(n))
            assert rc == 0
    finally:
        opts.su_output, opts.per_shot_files = saved
    d = str(tmp_path)
    if per_shot:
        files = [os.path.join(d, f"card_shot{k}_seismogram") for k in range(2)]
        traces = [read_su(f + ".su", ns) for f in files]
        bins = [np.fromfile(f + ".bin", np.float32).reshape(-1, ns) for f in files]
    else:
        su = read_su(os.path.join(d, "card_seismogram.su"), ns)
        traces = [su[:5], su[5:]]
        # the reference's single .bin is rewritten by every shot: it holds shot 2
        bins = [None, np.fromfile(os.path.join(d, "card_seismogram.bin"), np.float32).reshape(-1, ns)]
        assert len(su) == 12 and np.array_equal(su["h"]["tracl"], np.arange(1, 13))
    for k, ((sx, sz), rec) in enumerate(shots):
        t = traces[k]
        hd = t["h"]
        rec = np.asarray(rec, np.float32)
        assert np.array_equal(t["s"].view(np.int32), seis[k].view(np.int32))
        if bins[k] is not None:
            assert np.array_equal(t["s"].view(np.int32), bins[k].view(np.int32))
        assert np.all(hd["fldr"] == k + 1) and np.array_equal(hd["tracf"], np.arange(1, len(rec) + 1))
        assert np.all(hd["trid"] == 1) and np.all(hd["ns"] == ns) and np.all(hd["dt"] == 500)
        assert np.all(hd["scalco"] == -100) and np.all(hd["scalel"] == -100) and np.all(hd["counit"] == 1)
        cm = lambda v: np.round(np.asarray(v, np.float64) * 100).astype(np.int64)  # noqa: E731
        assert np.all(hd["sx"] == cm(np.float32(sx))) and np.all(hd["sdepth"] == cm(np.float32(sz)))
        assert np.array_equal(hd["gx"], cm(rec[:, 0])) and np.array_equal(hd["gelev"], -cm(rec[:, 1]))
        assert np.array_equal(hd["offset"], cm(rec[:, 0]) - cm(np.float32(sx)))
        # coordinates recovered to within the 1 cm scalar
        assert np.all(np.abs(hd["gx"] / 100.0 - rec[:, 0]) <= 0.005 + 1e-6)
