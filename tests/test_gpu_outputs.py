"""Drop-in outputs: the CLI shim run in-process against the reference's files.

For each deck of tests/golden/dat/ the shim (``alpineem_amd.cli.main``, the
stand-in for ``./FDTD inputs.txt``) runs on the device in a scratch
directory, and its files are compared with what the flang-built reference
wrote for the same deck (tests/golden/make_golden.py):

- data.dat: the same lines; every non-numeric token (labels, header text)
  verbatim; numeric tokens equal as FP32 values on decks without a plane
  wave (whose device series are the reference's bit for bit), and within
  1e-5 rel-L2 per section on plane-wave decks (last-bit exp() differences of
  the analytic incident field); header numbers equal everywhere.
- _geometry.bin, _video_*.bin, _full_*.bin: byte for byte (SHA-256) on
  decks without a plane wave; the plane-wave deck's video frames within the
  field tolerance, with the Fortran record markers checked.

The CLI is called in-process: this process already holds the GPU, so no
child process is started.
"""
import hashlib
import os
import re
import shutil

import numpy as np
import pytest

from alpineem_amd import cli, decks
from alpineem_amd._lib import AemError
from alpineem_amd.solver import Batch, Plan, Simulation

pytestmark = pytest.mark.gpu

EXACT = ["dipole36", "export_small"]            # no plane wave: bit-exact series
CLOSE = ["sphere_oblique", "master_ports", "drude_slab", "c3_mini"]
NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?$")


def _sections(lines):
    """Split data.dat into (label words, label numbers, [numeric rows]) in
    file order: a line with any non-numeric token starts a section (the 17
    header lines are one section each), all-numeric lines are its rows."""
    out = []
    for ln in lines:
        toks = ln.split()
        if toks and all(NUM.match(t) for t in toks):
            if not out:
                out.append(("", [], []))
            out[-1][2].append(toks)
        else:
            words = [t for t in toks if not NUM.match(t)]
            nums = [t for t in toks if NUM.match(t)]
            out.append((" ".join(words), nums, []))
    return out


def _run_cli(golden_dir, name, tmp_path):
    shutil.copy(os.path.join(golden_dir, "decks", name + ".txt"), tmp_path / "inputs.txt")
    assert cli.main([str(tmp_path / "inputs.txt")]) == 0
    return tmp_path


def _sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


@pytest.mark.parametrize("name", EXACT + CLOSE)
def test_cli_data_dat_matches_reference(golden_dir, name, tmp_path, capsys):
    out = _run_cli(golden_dir, name, tmp_path)
    printed = capsys.readouterr().out
    assert "Average of" in printed and "Mcells per second" in printed
    got = open(out / "data.dat").read().splitlines()
    ref = open(os.path.join(golden_dir, "dat", name + ".dat")).read().splitlines()
    assert len(got) == len(ref)
    sg, sr = _sections(got), _sections(ref)
    assert [s[0] for s in sg] == [s[0] for s in sr], "labels differ"
    worst = 0.0
    for q, ((label, la, a), (_, lb, b)) in enumerate(zip(sg, sr)):
        np.testing.assert_array_equal(np.array(la, np.float32), np.array(lb, np.float32), err_msg=label)
        assert [len(r) for r in a] == [len(r) for r in b], label
        x = np.array([float(t) for r in a for t in r], np.float32)
        y = np.array([float(t) for r in b for t in r], np.float32)
        if name in EXACT or x.size <= 4:
            np.testing.assert_array_equal(x, y, err_msg=f"{name}: section '{label}'")
        else:
            nr = np.linalg.norm(y.astype(np.float64))
            r = np.linalg.norm(x.astype(np.float64) - y) / (nr if nr > 0 else 1.0)
            worst = max(worst, r)
            assert r <= 1e-5, (label, r)
    print(f"\n  {name}: {len(got)} lines, {len(sg)} sections, worst section rel-L2 {worst:.2e}")


@pytest.mark.parametrize("name", EXACT + CLOSE)
def test_cli_binary_files_match_reference(golden_dir, name, tmp_path):
    out = _run_cli(golden_dir, name, tmp_path)
    g = np.load(os.path.join(golden_dir, name + ".npz"))
    files = list(g["bin_files"])
    assert sorted(f for f in os.listdir(out) if f.endswith(".bin")) == sorted(files)
    for f, h in zip(files, g["bin_sha256"]):
        if name in EXACT or f.endswith("_geometry.bin"):
            assert _sha(out / f) == h, f
    if name in CLOSE and "video" in g:
        stride = int(g["video_stride"])
        T, a2, a1 = int(g["time_steps"]), g["video"].shape[2], g["video"].shape[3]
        for c, comp in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
            raw = np.fromfile(out / f"data_video_{comp}.bin", np.int32)
            assert raw[0] == raw[-1] == T * a1 * a2 * 4
            v = raw[1:-1].view(np.float32).reshape(T, a2, a1)[stride - 1::stride]
            ref = g["video"][c]
            r = np.linalg.norm(v.astype(np.float64) - ref) / np.linalg.norm(ref)
            assert r <= 5e-5, (comp, r)


def test_video_frames_match_oracle_bit_exact(golden_dir):
    """The device's per-step video slices (read back with aem_get_video)
    equal the reference's frames every 25 steps on the antenna deck."""
    plan = Plan.from_inputs(os.path.join(golden_dir, "decks", "export_small.txt"))
    g = np.load(os.path.join(golden_dir, "export_small.npz"))
    sim = Simulation(plan)
    sim.run()
    stride = int(g["video_stride"])
    for c, comp in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
        frames = sim.video(comp)[stride - 1::stride]
        np.testing.assert_array_equal(frames, g["video"][c], err_msg=comp)


def test_watchdog_stops_unstable_deck(tmp_path, capsys):
    """A Courant factor of 0.5 (twice the stable time step) makes the update
    grow without bound: the run returns AEM_ERR_NUMERIC instead of NaN
    outputs, and the CLI prints the reason and exits with status 1."""
    d = decks.dipole(n=24, steps=300, arm=4, angles=0, factor=0.5)
    sim = Simulation(Plan.from_text(d.text()))
    with pytest.raises(AemError) as e:
        sim.run()
    assert e.value.status == 6 and "non-finite" in str(e.value)
    (tmp_path / "inputs.txt").write_text(d.text())
    assert cli.main([str(tmp_path / "inputs.txt")]) == 1
    assert "non-finite" in capsys.readouterr().out


def test_batch_refuses_full_field_export(golden_dir):
    a = Simulation(Plan.from_inputs(os.path.join(golden_dir, "decks", "export_small.txt")))
    with pytest.raises(AemError) as e:
        Batch([a])
    assert e.value.status == 5
