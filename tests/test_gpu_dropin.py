"""GPU run of the reference's own front end linked against this build's
modelling_module (oracle/_ref/SeisAcoMod2D_MI355X_dropin.exe, see
test_dropin.py): the program the reference's Makefile would produce with
libsam2d_host.so in place of cpumodelling.o and cpu_modelling_kernels.o.
Its output file must match the file the compiled reference wrote for the
same card (relative L2 per trace <= 1e-5)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import CARRYOVER_FLOOR, ROOT, assert_seis_close, load_golden
from test_host_cli import _write_card

pytestmark = pytest.mark.gpu
DROPIN = os.path.join(ROOT, "oracle", "_ref", "SeisAcoMod2D_MI355X_dropin.exe")


def _run(d, card, env=None):
    if not os.path.exists(DROPIN):
        pytest.skip("drop-in not linked (needs the reference sources: make -C oracle ref)")
    print(f"[gpu] drop-in on {card}", flush=True)
    e = dict(os.environ, OMP_NUM_THREADS="4", **(env or {}))
    r = subprocess.run([DROPIN, card], cwd=d, capture_output=True, text=True, timeout=300, env=e)
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
    return r


@pytest.mark.parametrize("case", ["tiny_twolayer", "hetero_explicit", "rect_dtfactor"])
def test_dropin_reproduces_reference_output_file(tmp_path, case):
    g = load_golden(case)
    d = str(tmp_path)
    text = str(np.load(os.path.join(ROOT, "tests", "golden", case + ".npz"))["geometry_text"])
    card = _write_card(d, g, 0 if text.count("\n") else 1, text)
    r = _run(d, card)
    assert "Shot to be handled for rank 0 is : 1" in r.stdout     # the reference's master ran
    out = np.fromfile(os.path.join(d, "card_seismogram.bin"), np.float32)
    out = out.reshape(g["rec_xz"].shape[0], -1)[g["trace_idx"]]
    e = assert_seis_close(out, g["seis"], 1e-5, case)
    print(f"[gpu] drop-in {case}: worst trace error {e.max():.2e}", flush=True)


def test_dropin_two_shot_carryover(tmp_path):
    # SAM2D_RESET_QUIRK=1: the reference's partial reset between shots; the
    # file left behind holds shot 2, as the reference's does.
    g = load_golden("two_shot_carryover")
    d = str(tmp_path)
    lines = [f"{sx:f},{sz:f},{x:f},{z:f}" for sx, sz in g["src_xz"] for x, z in g["rec_xz"]]
    card = _write_card(d, g, 0, "\n".join(lines))
    _run(d, card, {"SAM2D_RESET_QUIRK": "1"})
    out = np.fromfile(os.path.join(d, "card_seismogram.bin"), np.float32).reshape(g["seis"].shape)
    e = assert_seis_close(out, g["seis"], 1e-5, "drop-in two shots", CARRYOVER_FLOOR)
    print(f"[gpu] drop-in two-shot carry-over: worst trace error {e.max():.2e}", flush=True)
