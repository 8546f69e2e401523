"""The `gpk` deck driver (gpk_amd/gpk, built on the C++ facade) against the
reference CLI (demos/gpk.cpp): the same deck must write the same snapshot
files in the default strict32 precision.

Byte for byte (CSV, VTK and PPM) wherever the model evaluates no libm
transcendental on per-cell data.  The 2-D solidification deck does: its
four-fold anisotropy is sigma0 (1 + eps4 cos(4 atan2(n_y, n_x)))
(src/gpk_thermo.cpp:9-30), and the device's cosf/atan2f may differ from
glibc's in the last place.  That deck is compared field by field instead:
grain ids exactly, floats to 1e-4 relative to each field's scale."""
import os
import subprocess

import numpy as np
import pytest

from tests.util import GOLDEN, REPO

pytestmark = pytest.mark.gpu

CLI = os.path.join(REPO, "gpk_amd", "gpk")
EXACT = ("cli_grain_growth_2d", "cli_grain_growth_3d", "cli_solidification_3d", "cli_recrystallisation_2d",
         "cli_precipitation_2d", "cli_solidification_elastic_2d")


def first_difference(a, b):
    la, lb = a.split(b"\n"), b.split(b"\n")
    for k, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return "line %d: %r != %r" % (k + 1, x[:120], y[:120])
    return "length %d != %d lines" % (len(la), len(lb))


def run_deck(name, tmp_path):
    assert os.path.exists(CLI), "gpk CLI not built (make -C gpk_amd/csrc)"
    fx = np.load(os.path.join(GOLDEN, name + ".npz"))
    out = tmp_path / "out"
    deck = tmp_path / "deck.ini"
    deck.write_text(str(fx["deck"]).format(out=str(out)))
    print("deck", name, flush=True)
    r = subprocess.run([CLI, str(deck)], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    files = list(fx["files"])
    assert sorted(os.listdir(out)) == files
    return fx, out, files


@pytest.mark.parametrize("name", EXACT)
def test_cli_snapshots_match_reference(name, tmp_path):
    fx, out, files = run_deck(name, tmp_path)
    for k, fn in enumerate(files):
        want = fx["f%d" % k].tobytes()
        got = (out / fn).read_bytes()
        assert got == want, (fn, first_difference(got, want))


def csv_rows(blob):
    lines = [ln for ln in blob.decode().splitlines() if not ln.startswith("#")]
    return np.array([[float(v) for v in ln.split()] for ln in lines])


def test_cli_solidification_2d_anisotropic(tmp_path):
    fx, out, files = run_deck("cli_solidification_2d", tmp_path)
    for k, fn in enumerate(files):
        if not fn.endswith(".csv"):
            continue
        want = csv_rows(fx["f%d" % k].tobytes())
        got = csv_rows((out / fn).read_bytes())
        assert want.shape == got.shape, fn
        assert np.array_equal(got[:, :4], want[:, :4]), fn  # x, y, z, grain_id
        for col in range(4, 8):
            scale = max(np.abs(want[:, col]).max(), 1e-30)
            err = np.abs(got[:, col] - want[:, col]).max() / scale
            assert err <= 1e-4, (fn, col, err)


def test_cli_precipitate_breakup_matches_reference(tmp_path):
    # the sweep reads the labels between steps (volume every 4th step, pieces
    # every n_steps / n_snaps), so this also checks the facade's host sync;
    # d = 4 stays one piece, d = 16 severs (summary: stable, BREAK)
    fx, out, files = run_deck("cli_precipitate_breakup_3d", tmp_path)
    kept = 0
    for k, fn in enumerate(files):
        if "f%d" % k not in fx.files:
            continue
        want = fx["f%d" % k].tobytes()
        got = (out / fn).read_bytes()
        assert got == want, (fn, first_difference(got, want))
        kept += 1
    assert kept >= 15
    assert (out / "summary.csv").read_text().splitlines()[1:] == ["4,4,1.000,1,stable", "16,4,4.000,2,BREAK"]
