"""Fixtures from output data shipped with the reference (no code is run).

examples/radiating_dipole_antenna/{S_parameters,Realized_antenna_gain}.csv
hold the reference's post-processed results of its 150^3, 15000-step dipole
deck (the deck written by that example's master.py, reproduced in
tests/golden/decks/dipole150_example.txt).  This script keeps every 10th
frequency row of both files in tests/golden/dipole150_example.npz:
frequency (GHz), S11 (dB, rad) and realized gain (dB, rad) for theta/phi
polarisations at the example's three far-field angles.

examples/radiating_dipole_antenna_w_SPICE_source/S_parameters.csv is the
same dipole fed through a SPICE port by the netlist that example's master.py
writes (a 1 mV Gaussian-modulated 10 GHz sine behind 50 ohm and the grid
capacitance); its deck is tests/golden/decks/dipole150_spice_example.txt.
Every 10th row goes to tests/golden/dipole150_spice_example.npz, with the
source parameters, after checking that the example's incident.npy is half
the source voltage at the solver's steps (source.py of that example).

examples/metal_sphere_monostatic_scattering/Scattering_Far_Field{,-MP,-ACC}.csv
hold the monostatic RCS of a 50-cell-radius sphere (170^3, 10000 steps,
object and clear decks: tests/golden/decks/sphere170{,_clear}_example.txt)
from the reference's serial, OpenMP and OpenACC builds; all 603 rows of the
serial file go to tests/golden/sphere170_example.npz with the largest
serial-to-other-build difference (the reference's own spread).
examples/trans_and_refl_infinite_Drude(plasma)_slab/S_parameters.csv holds
the reflection and transmission of a periodic 3-pole Drude slab (object,
clear and metal decks: tests/golden/decks/drude_slab{,_clear,_metal}_example.txt);
every 4th row goes to tests/golden/drude_slab_example.npz.

Usage:  python tests/golden/make_example_fixtures.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/examples/radiating_dipole_antenna"
STRIDE = 10


def main():
    s = np.loadtxt(os.path.join(SRC, "S_parameters.csv"), delimiter=",", skiprows=1)
    g = np.loadtxt(os.path.join(SRC, "Realized_antenna_gain.csv"), delimiter=",", skiprows=1)
    assert np.array_equal(s[:, 0], g[:, 0])
    rows = np.arange(0, s.shape[0], STRIDE)
    np.savez_compressed(os.path.join(HERE, "dipole150_example.npz"), freq_ghz=s[rows, 0], s11_db=s[rows, 1],
                        s11_rad=s[rows, 2], gain=g[rows, 1:], stride=np.int32(STRIDE), rows=np.int32(s.shape[0]))
    print(f"kept {rows.size} of {s.shape[0]} rows")

    spice = SRC + "_w_SPICE_source"
    s = np.loadtxt(os.path.join(spice, "S_parameters.csv"), delimiter=",", skiprows=1)
    rows = np.arange(0, s.shape[0], STRIDE)
    amp, f0, t0, width, dt = 0.001, 10e9, 150e-12, 50e-12, 1.8873167e-12
    t = np.arange(1, 15001) * dt
    vs = amp * np.sin(2 * np.pi * f0 * t) * np.exp(-np.log(2) * ((t - t0) / width) ** 2)
    inc = np.load(os.path.join(spice, "incident.npy"), allow_pickle=False)
    assert inc.shape == (15000,) and np.max(np.abs(inc - 0.5 * vs)) < 1e-12 * amp
    np.savez_compressed(os.path.join(HERE, "dipole150_spice_example.npz"), freq_ghz=s[rows, 0], s11_db=s[rows, 1],
                        s11_rad=s[rows, 2], source=np.array([amp, f0, t0, width, dt]), stride=np.int32(STRIDE),
                        rows=np.int32(s.shape[0]))
    print(f"SPICE example: kept {rows.size} of {s.shape[0]} rows")

    ex = os.path.dirname(SRC)
    rcs = [np.loadtxt(os.path.join(ex, "metal_sphere_monostatic_scattering", f"Scattering_Far_Field{x}.csv"),
                      delimiter=",", skiprows=1) for x in ("", "-MP", "-ACC")]
    spread_db = max(np.abs(rcs[0][:, 1] - r[:, 1]).max() for r in rcs[1:])
    spread_rad = max(np.abs(np.angle(np.exp(1j * (rcs[0][:, 2] - r[:, 2])))).max() for r in rcs[1:])
    np.savez_compressed(os.path.join(HERE, "sphere170_example.npz"), freq_ghz=rcs[0][:, 0], theta_db=rcs[0][:, 1],
                        theta_rad=rcs[0][:, 2], phi_db=rcs[0][:, 3], phi_rad=rcs[0][:, 4],
                        build_spread=np.array([spread_db, spread_rad]))
    print(f"sphere RCS: {rcs[0].shape[0]} rows, build spread {spread_db:.2e} dB, {spread_rad:.2e} rad")

    d = np.loadtxt(os.path.join(ex, "trans_and_refl_infinite_Drude(plasma)_slab", "S_parameters.csv"), delimiter=",",
                   skiprows=1)
    rows = np.arange(0, d.shape[0], 4)
    np.savez_compressed(os.path.join(HERE, "drude_slab_example.npz"), freq_ghz=d[rows, 0], refl_db=d[rows, 1],
                        refl_rad=d[rows, 2], trans_db=d[rows, 3], trans_rad=d[rows, 4], stride=np.int32(4),
                        rows=np.int32(d.shape[0]))
    print(f"Drude slab: kept {rows.size} of {d.shape[0]} rows")


if __name__ == "__main__":
    main()
