"""Command-line drop-in for the reference executable: ``python -m alpineem_amd.cli inputs.txt``.

Mirrors ``./FDTD inputs.txt`` (main_fdtd/master.py:438-441): parses the deck,
runs every time step on the GPU in the deck's directory and writes
``data.dat``, ``<name>_geometry.bin`` and, when the deck asks for them, the
video and full-field files there, printing the reference's progress and
timing lines (fdtd_solver.f90:724, 4366-4368, 8313-8316).  Deck errors print
the reference's ``ERROR: ...`` message and exit with status 1, as does a run
stopped by the non-finite field check.  ``--kmax`` reads the deck as the
reference's kmax executable does (its -Duse_kmax_version build) and writes
that build's data.dat.
"""
from __future__ import annotations

import os
import sys
import time

from ._lib import AemError
from .solver import Plan, Simulation


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    kmax = "--kmax" in argv
    argv = [a for a in argv if a != "--kmax"]
    if len(argv) != 1:
        print("usage: python -m alpineem_amd.cli [--kmax] <inputs.txt>", file=sys.stderr)
        return 2
    deck = os.path.abspath(argv[0])
    cwd = os.getcwd()
    # the reference runs in the deck's directory and writes its files there
    os.chdir(os.path.dirname(deck))
    try:
        return _run(deck, kmax)
    except AemError as e:
        print(str(e).split(": ", 1)[-1])
        return 1
    finally:
        os.chdir(cwd)


def _run(deck: str, kmax: bool = False) -> int:
    t0 = time.time()
    print("Reading in the inputs file and setting up the simulation...")
    plan = Plan.from_inputs(deck, kmax=1) if kmax else Plan.from_inputs(deck)
    sim = Simulation(plan)
    print(f"Setup is finished. Time to complete was {time.time() - t0:.3f} seconds.")
    print("Starting main FDTD algorithm...")
    T = plan.time_steps
    t1 = time.time()
    done = 0
    while done < T:
        chunk = min(100, T - done)
        sim.run(chunk)
        done += chunk
        if done % 100 == 0:
            print(f"{done} of {T} time steps")
    el = time.time() - t1
    v = plan.view
    cells = (v.NX - 1) * (v.NY - 1) * (v.NZ - 1)
    print(f"Finished main FDTD algorithm. Time to complete was {el:.3f} seconds.")
    print(f"Average of {1e-6 * cells * T / max(el, 1e-9):.2f} Mcells per second.")
    print("Post processing some data...")
    t2 = time.time()
    sim.finalize()
    sim.write_outputs()
    print(f"Finished post processing. Time to complete was {time.time() - t2:.3f} seconds.")
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
