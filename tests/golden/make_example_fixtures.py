"""Fixtures from output data shipped with the reference (no code is run).

examples/radiating_dipole_antenna/{S_parameters,Realized_antenna_gain}.csv
hold the reference's post-processed results of its 150^3, 15000-step dipole
deck (the deck written by that example's master.py, reproduced in
tests/golden/decks/dipole150_example.txt).  This script keeps every 10th
frequency row of both files in tests/golden/dipole150_example.npz:
frequency (GHz), S11 (dB, rad) and realized gain (dB, rad) for theta/phi
polarisations at the example's three far-field angles.

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


if __name__ == "__main__":
    main()
