"""GPU end-to-end test of the job-card driver SeisAcoMod2D_MI355X.exe: the
reference's card, shot file and model files in, the reference's seismogram
file out (<out>/<job>_seismogram.bin, trace-major float32), compared with the
compiled reference's golden output (relative L2 per trace <= 1e-5)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import CARRYOVER_FLOOR, ROOT, assert_seis_close, load_golden
from test_host_cli import _write_card

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "seisacomod2d_amd", "SeisAcoMod2D_MI355X.exe")


@pytest.mark.parametrize("case", ["tiny_twolayer", "hetero_explicit", "rect_dtfactor"])
def test_cli_reproduces_reference_output_file(tmp_path, case):
    g = load_golden(case)
    text = str(np.load(os.path.join(ROOT, "tests", "golden", case + ".npz"))["geometry_text"])
    flag = 0 if text.count("\n") else 1
    card = _write_card(str(tmp_path), g, flag, text)
    r = subprocess.run([EXE, card], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = np.fromfile(os.path.join(str(tmp_path), "card_seismogram.bin"), np.float32)
    out = out.reshape(g["rec_xz"].shape[0], -1)[g["trace_idx"]]
    assert_seis_close(out, g["seis"], 1e-5, case)
    # the restart file the reference only reads is written: one shot done
    assert open(os.path.join(str(tmp_path), "card_checkpoint_rank0.txt")).read().strip() == "1"


def test_cli_restart_and_abort_codes(tmp_path):
    g = load_golden("tiny_twolayer")
    d = str(tmp_path)
    card = _write_card(d, g, 1, "300.000000,20.000000,50,0,10.000000,1\n600.000000,20.000000,50,0,10.000000,1",
                       jobtype="Restart")
    # Restart without the file -> the reference's Abort(-10), exit status 246
    r = subprocess.run([EXE, card], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == (-10) & 0xFF
    open(os.path.join(d, "card_checkpoint_rank0.txt"), "w").write("1\n")
    r = subprocess.run([EXE, card, "--per-shot-files"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert not os.path.exists(os.path.join(d, "card_shot0_seismogram.bin"))   # skipped by the restart
    assert os.path.exists(os.path.join(d, "card_shot1_seismogram.bin"))
    # CFL violation -> Abort(-5), exit status 251
    txt = open(card).read().replace('"FD DT" : "0.001"', '"FD DT" : "0.004"').replace(
        '"DATA DT" : "0.001"', '"DATA DT" : "0.004"')
    open(card, "w").write(txt)
    r = subprocess.run([EXE, card], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == (-5) & 0xFF


def test_cli_su_output_matches_binary_and_geometry(tmp_path):
    from test_dropin import read_su
    g = load_golden("tiny_twolayer")
    d = str(tmp_path)
    text = "300.000000,20.000000,50,0,10.000000,1\n600.000000,30.000000,50,0,10.000000,1"
    card = _write_card(d, g, 1, text)
    r = subprocess.run([EXE, card, "--su"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    nt = round(g["rec_len"] / g["real_dt"])
    su = read_su(os.path.join(d, "card_seismogram.su"), nt)
    assert len(su) == 100 and np.array_equal(su["h"]["fldr"], np.repeat([1, 2], 50))
    # the .bin holds the last shot, as the reference leaves it
    last = np.fromfile(os.path.join(d, "card_seismogram.bin"), np.float32).reshape(50, nt)
    assert np.array_equal(su["s"][50:].view(np.int32), last.view(np.int32))
    assert np.all(su["h"]["sx"][:50] == 30000) and np.all(su["h"]["sx"][50:] == 60000)
    assert np.all(su["h"]["sdepth"][50:] == 3000) and np.all(su["h"]["ns"] == nt)


def test_cli_reset_quirk_reproduces_reference_two_shot_file(tmp_path):
    # Two shots on one card with --reset-quirk: the output file the driver
    # leaves (shot 2, which carried shot 1's wavefield) against the file the
    # compiled reference left for the same card.
    g = load_golden("two_shot_carryover")
    d = str(tmp_path)
    lines = [f"{sx:f},{sz:f},{x:f},{z:f}" for sx, sz in g["src_xz"] for x, z in g["rec_xz"]]
    card = _write_card(d, g, 0, "\n".join(lines))
    r = subprocess.run([EXE, card, "--reset-quirk"], cwd=d, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = np.fromfile(os.path.join(d, "card_seismogram.bin"), np.float32).reshape(g["seis"].shape)
    assert_seis_close(out, g["seis"], 1e-5, "--reset-quirk", CARRYOVER_FLOOR)
