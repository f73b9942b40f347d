"""CPU tests of the reference-shaped host layer (libsam2d_host.so,
include/sam2d_modelling.h) and the job-card driver: job card, shot-file and
explicit geometry, model files, stability check -- against what the compiled
reference itself wrote and returned: its test_geometry.txt dumps (golden cases
and tests/golden/reference_cards.json) and its exit statuses."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_cases, load_golden

HOST = os.path.join(ROOT, "seisacomod2d_amd", "libsam2d_host.so")

CARD = """"Model" : "comment"
    "Velocity" : "{vp}"
    "Density" : "{rho}"
    "NX" : "{nx}"
    "NZ" : "{nz}"
"comment" : "NY=1 2D,    NY>1 3D"
    "DX" : "{dx}"
    "DZ" : "{dz}"
    "Time" : "{T}"
    "FD DT" : "{fd_dt}"
    "DATA DT" : "{real_dt}"
    "Frequency" : "{fm}"
    "Read source receiver flag" : "{flag}"
    "if choice=0, Geometry file" : "{geom}"
    "if choice=1, Shot file" : "{shots}"
    "Output Directory" : "{out}"
    "Free surface" : "0"
    "Job type New or Restart" : "{jobtype}"
    "Order of FD Operator" : "4"
"""


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST):
        pytest.fail("libsam2d_host.so is not built")
    L = ctypes.CDLL(HOST)
    for f in ("read_job_card", "read_source2d", "read_recv2d", "write_geometry_dump",
              "check_stability"):
        getattr(L, f).restype = ctypes.c_int
    return L


def _write_card(d, g, flag, text, jobtype="New", extra=""):
    vpf, rhof = os.path.join(d, "vp.bin"), os.path.join(d, "rho.bin")
    g["vp"].astype(np.float32).tofile(vpf)
    g["rho"].astype(np.float32).tofile(rhof)
    shots, geom = os.path.join(d, "shots.txt"), os.path.join(d, "geom.txt")
    open(shots if flag else geom, "w").write(text + "\n")
    card = os.path.join(d, "card.json")
    open(card, "w").write(CARD.format(vp=vpf, rho=rhof, nx=g["nx"], nz=g["nz"], dx=g["dx"], dz=g["dz"],
                                      T=g["rec_len"], fd_dt=g["fd_dt"], real_dt=g["real_dt"],
                                      fm=g["fm"], flag=flag, geom=geom, shots=shots, out=d,
                                      jobtype=jobtype) + extra)
    return card


def test_host_library_exports_the_declared_functions(host):
    txt = open(os.path.join(ROOT, "include", "sam2d_modelling.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    c_part = txt.split("#ifdef __cplusplus\n}")[0]
    names = re.findall(r"^\s*(?:int|void)\s+(\w+)\s*\(", c_part, flags=re.M)
    assert "sam2d_modelling_module" in names and "read_job_card" in names
    for n in names:
        assert hasattr(host, n), n
    for n in ("mod_sp", "wave_sp", "geo2d_sp", "job_sp", "sam2d_opts"):
        ctypes.c_void_p.in_dll(host, n)
    # the reference's seam, C++ linkage: void modelling_module(int, Geo2d *)
    assert "void modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp);" in txt
    assert hasattr(host, "_Z16modelling_moduleiP5Geo2d")


@pytest.mark.parametrize("case", golden_cases())
def test_geometry_matches_reference_dump(host, tmp_path, case):
    g = load_golden(case)
    text = str(np.load(os.path.join(ROOT, "tests", "golden", case + ".npz"))["geometry_text"])
    flag = 0 if text.count("\n") else 1
    card = _write_card(str(tmp_path), g, flag, text)
    assert host.read_job_card(card.encode()) == 0
    rc = host.read_source2d(os.path.join(str(tmp_path), "shots.txt").encode()) if flag else \
        host.read_recv2d(os.path.join(str(tmp_path), "geom.txt").encode())
    assert rc == 0
    dump = os.path.join(str(tmp_path), "geometry.txt")
    assert host.write_geometry_dump(dump.encode()) == 0
    geo = np.loadtxt(dump, ndmin=2)
    np.testing.assert_array_equal(geo[0, ::-1].astype(np.float32), g["src_xz"][0])
    np.testing.assert_array_equal(geo[1:, ::-1].astype(np.float32), g["rec_xz"])


# ---------------------------------------------------------------------------
# Cards the compiled reference was run on (tests/golden/make_reference_cards.py):
# the geometry it generated and the code it aborted with.  The job-card driver
# runs the same flow (--check stops before the modelling, which needs a GPU)
# and must end with the reference's exit status and write the same
# test_geometry.txt.
CARDS = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_cards.json")))["cases"]
EXE = os.path.join(ROOT, "seisacomod2d_amd", "SeisAcoMod2D_MI355X.exe")


@pytest.mark.parametrize("case", CARDS, ids=[c["name"] for c in CARDS])
def test_cli_matches_reference_on_cards(tmp_path, case):
    from golden.make_reference_cards import write_inputs
    d = str(tmp_path)
    write_inputs(d, case["grid"], case["model"], case["flag"], case["text"], case["card_edit"])
    # the job status is checked inside modelling_module, before the GPU is
    # touched; every other abort happens before modelling
    args = [] if case["name"].startswith("abort_restart") else ["--check"]
    r = subprocess.run([EXE, "card.json"] + args, cwd=d, capture_output=True, text=True, timeout=120)
    if case["exit_status"] < 0:
        # The reference crashed (explicit_two_shots_uneven: master_workload
        # copies nrec[0] receivers for every shot, workload.cpp:144); the
        # driver reads every shot's own receivers and runs.
        assert r.returncode == 0, r.stderr[-1000:]
    else:
        assert r.returncode == case["exit_status"], (r.returncode, case, r.stderr[-1000:])
        if case["abort_code"] is not None:
            assert r.returncode == case["abort_code"] & 0xFF
    dump = os.path.join(d, "test_geometry.txt")
    if case["geometry"] is None:
        assert not os.path.exists(dump)
    else:
        assert open(dump).read() == case["geometry"]


def test_reference_cards_cover_every_abort_site():
    # Every abort code the reference's front end can raise on its inputs is
    # among the recorded cards (json_parser.cpp, create_geometry.cpp,
    # read_receiver.cpp, model.cpp, check_stability.cpp, cpumodelling.cpp).
    codes = {c["abort_code"] for c in CARDS if c["abort_code"] is not None}
    assert codes >= {-2, -3, -4, -5, -6, -10, -11, -12, -13, -14, -15, -16, -17, -18, -19, -20,
                     -21, -22, -24, -25, -26}
