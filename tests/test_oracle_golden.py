"""The CPU oracle against the reference solver's own outputs.

Fixtures under tests/golden/ were produced by the flang-built reference
(oracle/_ref/fdtd_ref, see tests/golden/make_golden.py) on the decks in
tests/golden/decks/.  The oracle restates the reference's FP32 arithmetic in
the same order, so every series must match bit for bit: incident wave, port
voltages, far-field E_theta/E_phi (including the FP32 retarded-time binning),
periodic-slab reflection/transmission, the phase-centre incident and the
cell-centred video slices of all six field components.
"""
import hashlib
import os

import numpy as np
import pytest

from alpineem_amd.solver import Plan
from tests.oracle_ffi import Oracle

DECKS = ["dipole36", "master_ports", "drude_slab", "sphere_oblique", "antenna_x", "c1_mini", "c2_mini", "c3_mini",
         "sheets_xyz", "pop4_m0", "pop4_m1", "pop4_m2", "pop4_m3", "export_small", "igp_mz", "igp_px"]


def _run(golden_dir, name, fp64=False):
    deck = os.path.join(golden_dir, "decks", name + ".txt")
    plan = Plan.from_inputs(deck)
    orc = Oracle(plan, fp64=fp64)
    orc.step(plan.time_steps)
    orc.finalize()
    return plan, orc, np.load(os.path.join(golden_dir, name + ".npz"))


@pytest.mark.parametrize("name", DECKS)
def test_oracle_bit_exact_vs_reference(golden_dir, name):
    plan, orc, g = _run(golden_dir, name)
    v = plan.view
    assert np.float32(v.dt) == g["dt"], "time step differs from the reference header"
    np.testing.assert_array_equal(orc.series("incident"), g["incident"])
    if "port_v" in g:
        for q in range(v.n_ports):
            np.testing.assert_array_equal(orc.series("port_v", q), g["port_v"][q], err_msg=f"port {q}")
    if "e_theta" in g:
        for a in range(v.n_angles):
            np.testing.assert_array_equal(orc.series("e_theta", a), g["e_theta"][a], err_msg=f"E_theta angle {a}")
            np.testing.assert_array_equal(orc.series("e_phi", a), g["e_phi"][a], err_msg=f"E_phi angle {a}")
        np.testing.assert_array_equal(orc.ff_time(), g["ff_time"])
    if "reflected" in g:
        np.testing.assert_array_equal(orc.series("reflected"), g["reflected"])
        np.testing.assert_array_equal(orc.series("transmitted"), g["transmitted"])
    if "ff_pc" in g:
        np.testing.assert_array_equal(orc.series("ff_pc"), g["ff_pc"])
    if "video" in g:
        stride = int(g["video_stride"])
        for c in range(6):
            for f in range(g["video"].shape[1]):
                np.testing.assert_array_equal(orc.video(c, (f + 1) * stride), g["video"][c, f],
                                              err_msg=f"video comp {c} step {(f + 1) * stride}")


@pytest.mark.parametrize("name", DECKS)
def test_geometry_file_matches_reference(golden_dir, name, tmp_path):
    """_geometry.bin equals the reference's file byte for byte (SHA-256)."""
    plan = Plan.from_inputs(os.path.join(golden_dir, "decks", name + ".txt"))
    g = np.load(os.path.join(golden_dir, name + ".npz"))
    out = tmp_path / "geo.bin"
    plan.write_geometry(str(out))
    geo = np.fromfile(out, np.float32)
    v = plan.view
    assert geo.size == 4 * v.nx * v.ny * v.nz
    assert geo.astype(np.float64).sum() == g["geometry_sum"]
    assert np.count_nonzero(geo) == g["geometry_nnz"]
    ref = dict(zip(g["bin_files"], g["bin_sha256"]))
    assert hashlib.sha256(out.read_bytes()).hexdigest() == ref["data_geometry.bin"]


@pytest.mark.parametrize("name", ["dipole36", "drude_slab", "c2_mini"])
def test_fp64_oracle_close_to_fp32_reference(golden_dir, name):
    """The FP64 restatement on the same FP32 inputs agrees with the FP32
    reference to the expected FP32 rounding level; a dropped or mis-signed
    term would show up at O(1)."""
    plan, orc, g = _run(golden_dir, name, fp64=True)
    key = "port_v" if "port_v" in g else ("reflected" if "reflected" in g else "e_theta")
    got = orc.series({"port_v": "port_v", "reflected": "reflected", "e_theta": "e_theta"}[key], 0)
    ref = g[key][0] if g[key].ndim == 2 else g[key]
    rel = np.linalg.norm(got.astype(np.float64) - ref) / np.linalg.norm(ref)
    assert rel < 1e-4, rel
