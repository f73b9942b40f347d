"""Record how the compiled reference treats job cards: the receiver geometry it
generates and the abort code it ends with.

    make -C oracle ref && python tests/golden/make_reference_cards.py

Runs oracle/_ref/SeisAcoMod2D_MULTI_CORE_CPU_V1.0.exe (src/src_cpu of the
reference, built from its sources by oracle/Makefile.ref) on each card below
in a scratch directory and writes tests/golden/reference_cards.json:

  name         case name
  grid         nx, nz, dx, dz, T, fd_dt, real_dt, fm of the card
  model        "two_layer", "uniform", or a broken model file ("missing", "short")
  flag         1: shot file, 0: explicit geometry file
  text         the shot-file or geometry-file contents ("" = empty file,
               null = no file at all)
  card_edit    [old, new] text replacement applied to the card, or null
               (["card.json", null]: no job card written)
  exit_status  the process exit status of the reference
  abort_code   the code it passed to MPI::COMM_WORLD.Abort (from its stderr
               line "MPI_Abort(MPI_COMM_WORLD, <code>)"), null when it ran through
  geometry     its test_geometry.txt (create_geometry.cpp:446-459: shot 0 as
               "z x" lines, source first), null when it aborted before writing it

Valid cards use a record of two steps: the reference then models them in a
fraction of a second and exits 0.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import refrun  # noqa: E402

SMALL = dict(nx=100, nz=80, dx=10.0, dz=10.0, T=0.002, fd_dt=0.001, real_dt=0.001, fm=10.0)
SIGSBEE = dict(nx=3201, nz=1201, dx=7.62, dz=7.62, T=0.001, fd_dt=0.0005, real_dt=0.0005, fm=15.0)

# (name, grid, model, flag, text, card_edit)
CASES = [
    # receiver generation from shot files (create_geometry.cpp:265-444)
    ("split_near0_centre", SMALL, "two_layer", 1, "500.000000,20.000000,40,0,10.000000,1", None),
    ("split_near0_left", SMALL, "two_layer", 1, "100.000000,20.000000,40,0,10.000000,1", None),
    ("split_near0_right", SMALL, "two_layer", 1, "950.000000,20.000000,40,0,10.000000,1", None),
    ("split_near10_centre", SMALL, "two_layer", 1, "500.000000,20.000000,40,10.000000,10.000000,1", None),
    ("split_near10_left", SMALL, "two_layer", 1, "100.000000,20.000000,40,10.000000,10.000000,1", None),
    ("split_near10_right", SMALL, "two_layer", 1, "900.000000,20.000000,40,10.000000,10.000000,1", None),
    ("endonleft_near0", SMALL, "two_layer", 1, "800.000000,20.000000,50,0,10.000000,2", None),
    ("endonleft_near20", SMALL, "two_layer", 1, "800.000000,20.000000,50,20.000000,10.000000,2", None),
    ("endonright_near0", SMALL, "two_layer", 1, "100.000000,30.000000,40,0,20.000000,3", None),
    ("endonright_near20", SMALL, "two_layer", 1, "100.000000,20.000000,50,20.000000,10.000000,3", None),
    ("two_shots_split", SMALL, "two_layer", 1,
     "300.000000,20.000000,30,0,10.000000,1\n700.000000,40.000000,30,0,10.000000,1", None),
    ("sigsbee_float_spacing", SIGSBEE, "uniform", 1, "9997.439453,7.620000,401,0,15.240000,1", None),
    ("explicit_two_shots", SMALL, "two_layer", 0,
     "\n".join([f"300,20,{x},30" for x in range(0, 600, 50)] +
               [f"600,40,{x},50" for x in range(100, 940, 70)]), None),
    # Receiver counts that differ between shots: master_workload copies
    # nrec[0] receivers for every shot (workload.cpp:144), so shot 2's 13th
    # receiver is uninitialised memory and the reference crashes while
    # modelling (SIGSEGV); its shot-0 geometry dump is still written.
    ("explicit_two_shots_uneven", SMALL, "two_layer", 0,
     "\n".join([f"300,20,{x},30" for x in range(0, 600, 50)] +
               [f"600,40,{x},50" for x in range(100, 990, 70)]), None),
    # shot-file aborts (create_geometry.cpp:15-260, 326-432)
    ("abort_geometry_type", SMALL, "two_layer", 1, "500.000000,20.000000,100,0,10.000000,4", None),
    ("abort_nrec_zero", SMALL, "two_layer", 1, "500.000000,20.000000,0,0,10.000000,1", None),
    ("abort_src_offgrid_x", SMALL, "two_layer", 1, "505.000000,20.000000,10,0,10.000000,1", None),
    ("abort_src_offgrid_z", SMALL, "two_layer", 1, "500.000000,25.000000,10,0,10.000000,1", None),
    ("abort_spacing_offgrid", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,15.000000,1", None),
    ("abort_src_outside_x", SMALL, "two_layer", 1, "2000.000000,20.000000,10,0,10.000000,1", None),
    ("abort_src_outside_z", SMALL, "two_layer", 1, "500.000000,2000.000000,10,0,10.000000,1", None),
    ("abort_five_fields", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000", None),
    ("abort_seven_fields", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1,7", None),
    ("abort_missing_field", SMALL, "two_layer", 1, "500.000000,,10,0,10.000000,1", None),
    ("abort_mixed_types", SMALL, "two_layer", 1,
     "300.000000,20.000000,10,0,10.000000,1\n500.000000,20.000000,10,0,10.000000,3", None),
    ("abort_split_too_wide", SMALL, "two_layer", 1, "500.000000,20.000000,200,0,10.000000,1", None),
    ("abort_endonleft_outside", SMALL, "two_layer", 1, "10.000000,20.000000,50,0,10.000000,2", None),
    ("abort_endonright_outside", SMALL, "two_layer", 1, "900.000000,20.000000,50,0,10.000000,3", None),
    ("abort_endonright_long", SMALL, "two_layer", 1, "100.000000,30.000000,50,0,20.000000,3", None),
    ("abort_blank_line", SMALL, "two_layer", 1,
     "300.000000,20.000000,10,0,10.000000,1\n\n500.000000,20.000000,10,0,10.000000,1", None),
    ("abort_shot_file_empty", SMALL, "two_layer", 1, "", None),
    ("abort_shot_file_missing", SMALL, "two_layer", 1, None, None),
    # explicit-geometry aborts (read_receiver.cpp:14-306)
    ("abort_geom_unsorted", SMALL, "two_layer", 0, "500,20,100,20\n400,20,100,20", None),
    ("abort_geom_blank_line", SMALL, "two_layer", 0, "500,20,100,20\n\n600,20,10,20", None),
    ("abort_geom_five_fields", SMALL, "two_layer", 0, "500,20,100,20,5", None),
    ("abort_geom_rec_outside_x", SMALL, "two_layer", 0, "500,20,100,20\n500,20,2000,20", None),
    ("abort_geom_rec_outside_z", SMALL, "two_layer", 0, "500,20,100,20\n500,20,100,2000", None),
    ("abort_geom_rec_offgrid_x", SMALL, "two_layer", 0, "500,20,100,20\n500,20,105,20", None),
    ("abort_geom_rec_offgrid_z", SMALL, "two_layer", 0, "500,20,100,20\n500,20,100,25", None),
    ("abort_geom_src_offgrid", SMALL, "two_layer", 0, "505,20,100,20", None),
    ("abort_geom_empty", SMALL, "two_layer", 0, "", None),
    ("abort_geom_missing", SMALL, "two_layer", 0, None, None),
    # job card (json_parser.cpp:40-265)
    ("abort_card_missing_key", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1",
     ['"Frequency"', '"Freq"']),
    ("abort_card_datadt", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1",
     ['"DATA DT" : "0.001"', '"DATA DT" : "0.0015"']),
    ("abort_card_jobtype", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1",
     ['"New"', '"new"']),
    ("abort_card_flag", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1",
     ['"Read source receiver flag" : "1"', '"Read source receiver flag" : "2"']),
    ("abort_card_free_surface", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1",
     ['"Free surface" : "0"', '"Free surface" : "3"']),
    ("abort_card_missing", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1",
     ["card.json", None]),
    ("card_first_key_wins", SMALL, "two_layer", 1, "500.000000,20.000000,40,0,10.000000,1",
     ['"NX" : "100"', '"NX" : "100"\n    "NX" : "7"\n    "NZ comment" : "1"']),
    # models (model.cpp:12-45) and stability (check_stability.cpp:8-72)
    ("abort_model_missing", SMALL, "missing", 1, "500.000000,20.000000,10,0,10.000000,1", None),
    ("abort_model_short", SMALL, "short", 1, "500.000000,20.000000,10,0,10.000000,1", None),
    ("abort_cfl", dict(SMALL, fd_dt=0.0025, real_dt=0.0025, T=0.005), "two_layer", 1,
     "500.000000,20.000000,10,0,10.000000,1", None),
    ("abort_dispersion", dict(SMALL, fm=16.0), "two_layer", 1,
     "500.000000,20.000000,10,0,10.000000,1", None),
    ("dispersion_at_limit", dict(SMALL, fm=15.0), "two_layer", 1,
     "500.000000,20.000000,10,0,10.000000,1", None),
    # modelling_module's job status (cpumodelling.cpp:299-321): a Restart job
    # without the <job>checkpoint_rank0.txt of an earlier run
    ("abort_restart_without_file", SMALL, "two_layer", 1, "500.000000,20.000000,10,0,10.000000,1",
     ['"New"', '"Restart"']),
]


def models(g, kind):
    nx, nz = g["nx"], g["nz"]
    if kind == "uniform":
        return np.full((nx, nz), 2000, np.float32), np.full((nx, nz), 2000, np.float32)
    vp = np.full((nx, nz), 1500, np.float32); vp[:, nz // 2:] = 3000
    rho = np.full((nx, nz), 1000, np.float32); rho[:, nz // 2:] = 2000
    return vp, rho


def write_inputs(d, g, model, flag, text, card_edit):
    """The card and its files in d, named as refrun.CARD names them."""
    vp, rho = models(g, model)
    if model != "missing":
        (vp if model != "short" else vp.ravel()[:-1]).tofile(os.path.join(d, "vp.bin"))
    rho.tofile(os.path.join(d, "rho.bin"))
    if text is not None:
        open(os.path.join(d, "shots.txt" if flag else "geom.txt"), "w").write(
            text + ("\n" if text else ""))
    card = refrun.CARD.format(nx=g["nx"], nz=g["nz"], dx=g["dx"], dz=g["dz"], T=g["T"],
                              fd_dt=g["fd_dt"], real_dt=g["real_dt"], fm=g["fm"], flag=flag)
    if card_edit and card_edit[1] is None:      # no job card at all
        return
    if card_edit:
        assert card_edit[0] in card, card_edit
        card = card.replace(card_edit[0], card_edit[1])
    open(os.path.join(d, "card.json"), "w").write(card)
    os.makedirs(os.path.join(d, "out"), exist_ok=True)


def run(name, g, model, flag, text, card_edit):
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, g, model, flag, text, card_edit)
        r = subprocess.run([refrun.EXE, "card.json"], cwd=d, capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS="4"), timeout=600)
        m = re.search(r"MPI_Abort\(MPI_COMM_WORLD, (-?\d+)\)", r.stderr + r.stdout)
        gpath = os.path.join(d, "test_geometry.txt")
        geom = open(gpath).read() if os.path.exists(gpath) else None
    rec = dict(name=name, grid=g, model=model, flag=flag, text=text, card_edit=card_edit,
               exit_status=r.returncode, abort_code=int(m.group(1)) if m else None, geometry=geom)
    print(f"{name:28s} exit {r.returncode:3d}  abort {rec['abort_code']}  "
          f"geometry {'yes' if geom else 'no'}", flush=True)
    return rec


def main():
    if not refrun.available():
        sys.exit(f"{refrun.EXE} missing: make -C oracle ref")
    out = [run(*c) for c in CASES]
    path = os.path.join(HERE, "reference_cards.json")
    json.dump({"reference": "src/src_cpu (compiled by oracle/Makefile.ref)", "cases": out},
              open(path, "w"), indent=1)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main()
