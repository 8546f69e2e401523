"""Writes the committed golden fixtures of tests/golden from the reference tree.

Run in a container that has the reference mounted (default /root/reference)
and the reference checker built (make -C oracle):

    python tests/golden/make_fixtures.py [--ref /root/reference]

Fixtures (inputs and expected outputs only):
  t1_curvature.npz   weighted-mean curvature of digitised circles R = 10..60,
                     sigma = 2R/15, from the reference's committed
                     paper/data/validate_theory_foundations/t1_curvature_error.csv
                     (demos/validate_theory_foundations.cpp:107-158)
  g13.npz g14.npz g15.npz
                     curved-bicrystal label fields before/after 800 steps and the
                     (time, area) trajectory, from the reference's committed
                     paper/data/validate_extended_base_models/g1{3,4,5}_*.csv
                     (demos/validate_extended_base_models.cpp:757-830)
  voronoi_*.npz      initial labels and orientations written by the reference's
                     own CLI (demos/gpk.cpp voronoi_init, 51-189) for small decks,
                     run from oracle/_ref/gpk
  cli_*.npz          a deck (text, output dir templated as {out}) and every
                     snapshot file the reference CLI wrote for it (CSV, VTK and
                     PPM bytes), demos/gpk.cpp run_grain_growth (195-282),
                     run_solidification (288-391), run_recrystallisation (477-555),
                     run_precipitation (703-797); cli_precipitate_breakup_3d.npz
                     (run_precipitate_breakup, 1065-1275) keeps the names of all
                     files and the bytes of the summary, the label slices and
                     the first, displayed and last full-grid dumps
"""
import argparse
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def read_grid_csv(path):
    """Label field of write_grid_csv (two header lines, one label per line)."""
    with open(path) as f:
        head = f.readline()
        f.readline()
        vals = np.array([int(line) for line in f], dtype=np.int32)
    ny = int(head.split("Ny=")[1].split()[0])
    nx = int(head.split("Nx=")[1].split()[0])
    return vals.reshape(ny, nx)


def t1(ref):
    src = os.path.join(ref, "paper/data/validate_theory_foundations/t1_curvature_error.csv")
    rows = [line.strip().split(",") for line in open(src).read().splitlines()[1:] if line.strip()]
    R = np.array([int(float(r[0])) for r in rows], np.int32)
    # keep the printed text: the fixture holds 9 significant digits (%.9g)
    kappa_text = np.array([r[1] for r in rows])
    np.savez_compressed(os.path.join(HERE, "t1_curvature.npz"), R=R, kappa_text=kappa_text)


def bicrystal(ref, key):
    d = os.path.join(ref, "paper/data/validate_extended_base_models")
    before = read_grid_csv(os.path.join(d, key + "_grid_before.csv"))
    after = read_grid_csv(os.path.join(d, key + "_grid_after.csv"))
    rows = [line.strip().split(",") for line in
            open(os.path.join(d, key + "_area_t.csv")).read().splitlines()[1:] if line.strip()]
    assert before.max() <= 1 and after.max() <= 1
    np.savez_compressed(os.path.join(HERE, key + ".npz"),
                        shape=np.array(before.shape, np.int32),
                        before=np.packbits(before.astype(np.uint8).ravel()),
                        after=np.packbits(after.astype(np.uint8).ravel()),
                        t_text=np.array([r[0] for r in rows]),
                        area_text=np.array([r[1] for r in rows]))


DECK = """[run]
model = grain_growth
[grid]
Nz = {nz}
Ny = {ny}
Nx = {nx}
dx = 1.0e-6
[time]
t_end = 1e-12
n_snaps = 1
[thermo]
M0 = 1.0e-8
sigma0 = 0.5
[grain_growth]
n_grains = {ng}
seed = {seed}
[geometry]
sigma_cells = 1.5
[output]
format = csv
dir = {out}
write_initial = true
"""


def voronoi(cli, nz, ny, nx, ng, seed):
    with tempfile.TemporaryDirectory() as tmp:
        deck = os.path.join(tmp, "deck.ini")
        out = os.path.join(tmp, "out")
        with open(deck, "w") as f:
            f.write(DECK.format(nz=nz, ny=ny, nx=nx, ng=ng, seed=seed, out=out))
        subprocess.check_call([cli, deck], stdout=subprocess.DEVNULL, cwd=tmp)
        rows = np.loadtxt(os.path.join(out, "snap_0000.csv"), comments="#")
    # CSV rows are written z-major, then y, then x (utils/csv_writer.hpp:44-56)
    phi = rows[:, 3].astype(np.int32).reshape(nz, ny, nx)
    g_text = np.array(["%.6f" % v for v in rows[:, 7]])
    name = "voronoi_%dx%dx%d_%d_%d.npz" % (nz, ny, nx, ng, seed)
    np.savez_compressed(os.path.join(HERE, name), phi=phi, g_text=g_text,
                        dims=np.array([nz, ny, nx, ng, seed], np.int64))


CLI_DECKS = {
    "cli_grain_growth_2d": """[run]
model = grain_growth
[grid]
Nz = 1
Ny = 64
Nx = 48
dx = 1.0e-6
[time]
t_end = 2.0
n_snaps = 2
[thermo]
M0 = 1.0e-10
sigma0 = 0.5
[grain_growth]
n_grains = 30
seed = 7
[geometry]
sigma_cells = 1.5
[output]
format = all
dir = {out}
write_initial = true
""",
    "cli_solidification_2d": """[run]
model = solidification
[grid]
Nz = 1
Ny = 64
Nx = 64
dx = 0.5e-6
[time]
t_end = 2.0e-4
n_snaps = 2
[thermo]
M0 = 5.0e-8
sigma0 = 0.20
eps4 = 0.04
D = 3.0e-9
m_liq = -500.0
k_part = 0.3
T_melt = 933.0
L = 1.0e8
stochastic = false
solute_rejection = true
[initial]
c0 = 0.02
undercooling = 5.0
seed_radius = 4
[output]
format = all
dir = {out}
write_initial = true
""",
    "cli_solidification_3d": """[run]
model = solidification
[grid]
Nz = 24
Ny = 24
Nx = 24
dx = 0.5e-6
[time]
t_end = 1.0e-4
n_snaps = 2
[thermo]
M0 = 5.0e-8
sigma0 = 0.20
eps4 = 0.04
D = 3.0e-9
m_liq = -500.0
k_part = 0.3
T_melt = 933.0
L = 1.0e8
solute_rejection = true
[initial]
c0 = 0.02
undercooling = 5.0
seed_radius = 4
[output]
format = all
dir = {out}
prefix = s3
""",
    "cli_recrystallisation_2d": """[run]
model = recrystallisation
[grid]
Nz = 1
Ny = 64
Nx = 64
dx = 1.0e-6
[time]
t_end = 30.0
n_snaps = 3
[thermo]
M0 = 1.0e-12
sigma0 = 0.3
E_stored = 4.0e5
[nucleation]
n_nuclei = 5
nucleus_r = 3
Jdot = 5.0e14
seed = 42
[geometry]
sigma_cells = 1.5
[output]
format = all
dir = {out}
write_initial = true
""",
    "cli_precipitation_2d": """[run]
model = precipitation
[grid]
Nz = 1
Ny = 64
Nx = 64
dx = 0.5e-6
[time]
t_end = 3.0e-5
n_snaps = 2
[thermo]
M0 = 2.0e-8
sigma0 = 0.18
D = 2.0e-9
n_phases = 3
T_ref = 900.0
a = 0.0, -8.0e6, -7.5e6
b = 0.0,  2.0e7, -1.0e7
c2 = 0.0,  2.5e8,  1.8e8
dT = 0.0,  1.0e4,  5.0e3
[initial]
c0 = 0.06
T0 = 880.0
n_seeds = 12
seed_radius = 2
seed = 123
[geometry]
sigma_cells = 1.5
[output]
format = all
dir = {out}
write_initial = true
""",
    "cli_solidification_elastic_2d": """[run]
model = solidification
[grid]
Nz = 1
Ny = 64
Nx = 64
dx = 0.5e-6
[time]
t_end = 2.0e-4
n_snaps = 2
[thermo]
M0 = 5.0e-8
sigma0 = 0.20
eps4 = 0.0
D = 3.0e-9
m_liq = -500.0
k_part = 0.3
T_melt = 933.0
L = 1.0e8
elastic_enabled = true
elastic_modulus = 5.0e9
eigenstrain_coeff = 0.05
c_ref = 0.02
elastic_smooth_iters = 12
solute_rejection = true
[initial]
c0 = 0.02
undercooling = 10.0
seed_radius = 4
[output]
format = all
dir = {out}
write_initial = true
""",
    "cli_grain_growth_3d": """[run]
model = grain_growth
[grid]
Nz = 16
Ny = 20
Nx = 24
dx = 1.0e-6
[time]
t_end = 0.3
n_snaps = 2
[thermo]
M0 = 1.0e-10
sigma0 = 0.5
[grain_growth]
n_grains = 40
seed = 3
[output]
format = csv
dir = {out}
prefix = g3
""",
}


# the breakup sweep writes a full-grid dump every n_steps / 10 steps; the
# fixture keeps the summary, every label slice and the initial, displayed and
# final full dumps byte for byte, and the names of all files
CLI_BREAKUP = ("cli_precipitate_breakup_3d", """[run]
model = precipitate_breakup
[grid]
Ny = 12
Nz = 20
dx = 5.0e-9
[time]
n_snaps = 10
[thermo]
M0 = 1.0e-9
sigma_broad = 0.091
sigma_side = 0.811
c_sat = 0.15
Omega = 0.0
[breakup]
h_cells = 4
n_steps = 200
min_piece_frac = 0.25
d_list = 4,16
[geometry]
sigma_cells = 1.5
[output]
dir = {out}
""")


def breakup_kept(fn):
    return fn == "summary.csv" or "_xz_" in fn or "_xy_" in fn or fn.endswith(("3d_t0.csv", "3d_tb.csv",
                                                                               "3d_ts0200.csv"))


def cli_snapshots(cli, name, deck, keep=None):
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "out")
        path = os.path.join(tmp, "deck.ini")
        with open(path, "w") as f:
            f.write(deck.format(out=out))
        subprocess.check_call([cli, path], stdout=subprocess.DEVNULL, cwd=tmp)
        files = sorted(os.listdir(out))
        blobs = [np.frombuffer(open(os.path.join(out, fn), "rb").read(), np.uint8)
                 if keep is None or keep(fn) else None for fn in files]
    arrays = {"deck": np.array(deck), "files": np.array(files)}
    for k, b in enumerate(blobs):
        if b is not None:
            arrays["f%d" % k] = b
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    t1(args.ref)
    for key in ("g13", "g14", "g15"):
        bicrystal(args.ref, key)
    cli = os.path.join(REPO, "oracle", "_ref", "gpk")
    for dims in ((1, 64, 48, 30, 7), (1, 96, 96, 200, 42), (16, 20, 24, 40, 3), (24, 24, 24, 60, 42)):
        voronoi(cli, *dims)
    for name, deck in CLI_DECKS.items():
        cli_snapshots(cli, name, deck)
    cli_snapshots(cli, *CLI_BREAKUP, keep=breakup_kept)


if __name__ == "__main__":
    main()
