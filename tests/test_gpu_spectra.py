"""Device DFT (aem_get_dft) and the post-processor formulas on top of it.

* The DFT at the post-processor's zero-padded FFT frequencies equals
  numpy.fft.fft of the padded series, bin for bin (the reference computes its
  spectra that way, utility_scripts/post_processor.py:191-215), and at
  arbitrary frequencies it equals a direct FP64 sum.
* Full-size regression against data the reference ships: its 150^3,
  15000-step radiating-dipole example and the S11 / realized-gain CSVs its
  post-processor wrote (tests/golden/dipole150_example.npz, every 10th row).
"""
import os

import numpy as np
import pytest

from alpineem_amd import spectra
from alpineem_amd.solver import Plan, Simulation

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(golden_dir):
    plan = Plan.from_inputs(os.path.join(golden_dir, "decks", "c1_mini.txt"))
    sim = Simulation(plan)
    sim.run()
    sim.finalize()
    return plan, sim


def _series64(sim, kind, idx=0):
    return sim.series(kind, idx).astype(np.float64)


@pytest.mark.parametrize("kind,idx", [("port_v", 0), ("incident", 0), ("e_theta", 1), ("e_phi", 3)])
def test_dft_equals_padded_fft_bins(c1, kind, idx):
    plan, sim = c1
    T, dt = plan.time_steps, float(np.float32(plan.view.dt))
    x = _series64(sim, kind, idx)
    pad = 10
    ref = np.fft.fft(np.concatenate([x, np.zeros(T * pad)]))
    bins = np.arange(0, T * (pad + 1) // 2, 37)
    freqs = np.fft.fftfreq(T * (pad + 1), dt)[bins]
    got = sim.dft(kind, freqs, idx)
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got - ref[bins])) / scale
    print(f"\n  {kind}[{idx}] max |DFT - FFT| / max|FFT| = {err:.2e}")
    assert err < 1e-12


def test_dft_arbitrary_frequencies(c1):
    plan, sim = c1
    T, dt = plan.time_steps, float(np.float32(plan.view.dt))
    x = _series64(sim, "port_v")
    rng = np.random.default_rng(3)
    freqs = np.sort(rng.uniform(0.0, 0.5 / dt, 64))
    n = np.arange(T)
    ref = np.exp(-2j * np.pi * np.outer(freqs, n) * dt) @ x
    got = sim.dft("port_v", freqs)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


def test_dft_requires_finalize_for_post_loop_series(golden_dir):
    sim = Simulation(Plan.from_inputs(os.path.join(golden_dir, "decks", "dipole36.txt")))
    sim.run(10)
    with pytest.raises(RuntimeError):
        sim.dft("incident", [1e9])
    assert sim.dft("port_v", [1e9]).shape == (1,)


def test_dipole150_example_matches_reference_csv(golden_dir):
    """The reference's shipped dipole results, recomputed from a full run on
    the device: S11 and the co-polar realized gain at the example's three
    angles."""
    fx = np.load(os.path.join(golden_dir, "dipole150_example.npz"))
    plan = Plan.from_inputs(os.path.join(golden_dir, "decks", "dipole150_example.txt"))
    sim = Simulation(plan)
    sim.run()
    sim.finalize()
    freqs = fx["freq_ghz"] * 1e9
    rep = spectra.antenna_report(sim, freqs)
    s11 = rep["s"][0]
    d_db = np.abs(20 * np.log10(np.abs(s11)) - fx["s11_db"])
    d_rad = np.abs(np.angle(np.exp(1j * (np.angle(s11) - fx["s11_rad"]))))
    print(f"\n  S11: max |d dB| = {d_db.max():.2e}, max |d rad| = {d_rad.max():.2e}")
    assert d_db.max() < 1e-3 and d_rad.max() < 1e-3
    gain = fx["gain"]
    for a in range(3):
        g_db, g_rad = rep["gain_theta"][a]
        ref_db, ref_rad = gain[:, 4 * a], gain[:, 4 * a + 1]
        band = (freqs > 0) & (ref_db > -40)         # co-polar lobe, away from DC
        e_db = np.abs(g_db - ref_db)[band].max()
        e_rad = np.abs(np.angle(np.exp(1j * (g_rad - ref_rad))))[band].max()
        print(f"  gain theta, angle {a}: max |d dB| = {e_db:.2e}, max |d rad| = {e_rad:.2e} over {band.sum()} bins")
        assert e_db < 1e-2 and e_rad < 1e-2
