"""Batched stepping (aem_batch_*): every member of a batch computes what it
computes when run alone.

Members mix every path of a step: fused and separate CPML, thin sheets,
TF/SF plane waves, multi-pole Drude media, periodic boundaries, plane sums,
lumped ports, different grid sizes and planes per thread, and far-field
binning that is batched (LDS histogram, ordered contributor lists) or run
per member (atomic).
Fields and port series must be bit-identical to solo runs; far-field series
binned through FP64 histograms (whose LDS additions have no fixed order)
must agree to FP64 summation noise.
"""
import os

import numpy as np
import pytest

from alpineem_amd import decks
from alpineem_amd.solver import Batch, Plan, Simulation

pytestmark = pytest.mark.gpu

MEMBERS = [("dipole36", "histogram", 0), ("c1_mini", "ordered", 0), ("antenna_x", "histogram", 2),
           ("sheets_xyz", "atomic", 0), ("master_ports", "ordered", 3), ("drude_slab", "ordered", 0),
           ("sphere_oblique", "histogram", 0), ("c3_mini", "ordered", 1)]


def _make(golden_dir, name, ntff, planes):
    plan = Plan.from_inputs(os.path.join(golden_dir, "decks", name + ".txt"))
    return Simulation(plan, ntff=ntff, planes=planes)


def _compare(name, ntff, a, b):
    v = a.plan.view
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(a.field(comp), b.field(comp), err_msg=f"{name} {comp}")
    for q in range(v.n_ports):
        np.testing.assert_array_equal(a.series("port_v", q), b.series("port_v", q), err_msg=f"{name} V{q}")
    for kind in ("incident", "reflected", "transmitted", "ff_pc"):
        np.testing.assert_array_equal(a.series(kind), b.series(kind), err_msg=f"{name} {kind}")
    for ang in range(v.n_angles):
        for kind in ("e_theta", "e_phi"):
            x, y = a.series(kind, ang), b.series(kind, ang)
            if ntff == "histogram":
                np.testing.assert_allclose(x, y, rtol=1e-6, atol=1e-6 * np.abs(y).max(), err_msg=f"{name} {kind}{ang}")
            else:
                np.testing.assert_array_equal(x, y, err_msg=f"{name} {kind}{ang}")


def test_batch_members_equal_solo_runs(golden_dir):
    batched = [_make(golden_dir, *m) for m in MEMBERS]
    solo = [_make(golden_dir, *m) for m in MEMBERS]
    steps = min(s.plan.time_steps for s in batched)
    b = Batch(batched)
    b.run(steps // 2)
    b.run(steps - steps // 2, graph=False)          # eager batched launches continue the same state
    for s in solo:
        s.run(steps)
    for s in batched + solo:
        s.finalize()
    for (name, ntff, _), a, s in zip(MEMBERS, batched, solo):
        print(f"\n  {name:16s} ntff={ntff:9s} {steps} steps")
        _compare(name, ntff, a, s)


def test_batch_population_matches_solo(golden_dir):
    """A small GA-style population (uneven cubes, random arms and dielectric
    blocks, 19 far-field angles, histogram binning) batched and solo."""
    pop = decks.population(count=6, seed=3, steps=160, edges=range(48, 81, 16))
    plans = [Plan.from_text(d.text()) for d in pop]
    batched = [Simulation(p, ntff="histogram") for p in plans]
    solo = [Simulation(Plan.from_text(d.text()), ntff="histogram") for d in pop]
    Batch(batched).run(160)
    for s in solo:
        s.run(160)
    for q, (a, s) in enumerate(zip(batched, solo)):
        a.finalize()
        s.finalize()
        _compare(f"member{q}", "histogram", a, s)


def test_batch_rejects_overrun_and_duplicates(golden_dir):
    a = _make(golden_dir, "dipole36", "ordered", 0)
    with pytest.raises(RuntimeError):
        Batch([a, a])
    b = Batch([a])
    with pytest.raises(RuntimeError):
        b.run(a.plan.time_steps + 1)


def test_batch_after_single_pass_steps():
    """A simulation that took single-pass steps first (its fields in the
    second buffer set, after an odd number of steps) joins a batch and
    continues with the batch's in-place launches, then alone again: its
    state equals a run of the separate H/E launches throughout."""
    from tests.test_gpu_parity import single_pass_deck
    text = single_pass_deck(steps=140).text()
    a = Simulation(Plan.from_text(text))
    other = Simulation(Plan.from_text(single_pass_deck((40, 38, 36), steps=140).text()))
    ref = Simulation(Plan.from_text(text), single_pass=False)
    assert a.single_pass
    a.run(31)
    b = Batch([a, other])
    assert not a.single_pass
    b.run(40)
    a.run(29)
    ref.run(100)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(a.field(comp), ref.field(comp), err_msg=comp)


def _design_sim(steps, eps):
    d, lo, hi = decks.adjoint_deck(n=48, steps=steps, box=(6, 6, 4))
    s = Simulation(Plan.from_text(d.text()), ntff="off")
    s.design_box(lo, hi)
    s.set_design_eps(eps)
    return s, lo, hi


def test_batched_member_refuses_late_setup(golden_dir):
    """A batch captures its members' kernel arguments (coefficient table,
    record and objective state) when it is created, so a design box, a
    record, an objective or a far-field DFT asked for afterwards is refused
    with a state error instead of stepping with stale arguments."""
    d, lo, hi = decks.adjoint_deck(n=48, steps=40, box=(6, 6, 4))
    a = Simulation(Plan.from_text(d.text()), ntff="off")
    b = Batch([a, _make(golden_dir, "dipole36", "ordered", 0)])
    for call in (lambda: a.design_box(lo, hi), lambda: a.record_design(lo, hi),
                 lambda: a.set_objective(30.0, 0.0, 10e9), lambda: a.set_farfield_dft([(30.0, 0.0)], 1e9, 1e9, 2)):
        with pytest.raises(RuntimeError, match="batch"):
            call()
    b.run(5)


def test_design_box_member_batched_equals_solo(golden_dir):
    """A member whose design box (per-cell permittivity) and objective were
    set before the batch was created steps exactly as it does alone."""
    eps = decks.design_permittivity((4, 6, 6), seed=5)
    a, _, _ = _design_sim(60, eps)
    a.set_objective(30.0, 0.0, 10e9)
    s, _, _ = _design_sim(60, eps)
    s.set_objective(30.0, 0.0, 10e9)
    Batch([a, _make(golden_dir, "dipole36", "ordered", 0)]).run(60)
    s.run(60)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(a.field(comp), s.field(comp), err_msg=comp)
    assert a.objective() == s.objective()
