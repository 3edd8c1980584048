"""SPICE co-simulation on the device (aem_port_exchange) against the oracle,
the lumped port and the reference's shipped SPICE example.

- A host circuit answering every step with the gap voltage of a lumped-port
  run reproduces that device run bit for bit (fields and gap currents).
- With the example netlist's circuit (source, 50 ohm, grid capacitance) in
  the loop, the device and the oracle hand the circuit the same currents and
  end with the same fields, port voltages and far field, bit for bit.
- data.dat carries the SPICE voltages as the SPICE build writes them.
- Full size: the reference's radiating_dipole_antenna_w_SPICE_source example
  (150^3, 15000 steps, its deck and netlist reproduced) against the S11 its
  post-processor wrote (tests/golden/dipole150_spice_example.npz).  ngspice
  is absent, so the circuit is integrated by alpineem_amd.cosim and the
  comparison is at the level of the circuit integrators' difference.
"""
import os

import numpy as np
import pytest

from alpineem_amd import cosim, datfile, decks
from alpineem_amd._lib import AemError
from alpineem_amd.solver import Plan, Simulation
from tests.oracle_ffi import Oracle
from tests.test_spice import Replay, _decks, _oracle_cosim

pytestmark = pytest.mark.gpu


def test_replayed_gap_voltage_reproduces_the_lumped_port_on_device():
    base, sp = _decks(n=44, steps=200)
    pb, ps = Plan.from_text(base.text()), Plan.from_text(sp.text())
    a = Simulation(pb)
    a.run(pb.time_steps)
    rep = Replay(a.series("port_v"))
    b = Simulation(ps)
    assert not b.single_pass
    cosim.run_cosim(b, rep)
    assert b.steps_done == pb.time_steps
    np.testing.assert_array_equal(np.concatenate(rep.currents), a.series("port_i"))
    np.testing.assert_array_equal(b.series("spice_i"), a.series("port_i"))
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(b.field(comp), a.field(comp), err_msg=comp)


def test_circuit_in_the_loop_bit_exact_vs_oracle():
    base, sp = _decks(n=40, steps=240, angles=3)
    ps = Plan.from_text(sp.text())
    dt = float(ps.view.dt)
    cap = ps.view.spice[0].cap
    src = cosim.gaussian_sine(1e-3, 10e9, 60 * dt, 20 * dt)
    dev = Simulation(ps)
    c1 = cosim.SeriesSourceCircuit(dt, [50.0], [cap], [src])
    cosim.run_cosim(dev, c1)
    orc = _oracle_cosim(ps, cosim.SeriesSourceCircuit(dt, [50.0], [cap], [src]))
    np.testing.assert_array_equal(dev.series("spice_i"), orc.series("spice_i"))
    np.testing.assert_array_equal(dev.series("spice_v"), orc.series("spice_v"))
    assert np.any(dev.series("spice_v") != 0)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(dev.field(comp), orc.field(comp), err_msg=comp)
    dev.finalize()
    orc.finalize()
    for q in range(ps.view.n_angles):
        np.testing.assert_array_equal(dev.series("e_theta", q), orc.series("e_theta", q))


def test_spice_deck_refuses_plain_stepping_and_writes_its_voltages(tmp_path):
    _, sp = _decks(n=36, steps=60, angles=1)
    sp.spice_extra_nodes = ["src_in"]
    ps = Plan.from_text(sp.text())
    sim = Simulation(ps)
    with pytest.raises(AemError) as e:
        sim.run(1)
    assert "aem_port_exchange" in str(e.value)
    dt = float(ps.view.dt)
    src = cosim.gaussian_sine(1e-3, 10e9, 20 * dt, 8 * dt)

    class WithExtra(cosim.SeriesSourceCircuit):
        def step(self, n, currents):
            v = super().step(n, currents)
            return np.concatenate([v, [self.vs(0, n * self.dt)]])

    circ = WithExtra(dt, [50.0], [ps.view.spice[0].cap], [src])
    cosim.run_cosim(sim, circ, steps=25)
    assert sim.steps_done == 25
    cur = sim.port_exchange(None)                 # step 26 waits at the exchange
    with pytest.raises(AemError):
        sim.finalize()
    with pytest.raises(AemError):
        sim.port_exchange(None)                   # it needs node voltages
    for n in range(26, 61):
        cur = sim.port_exchange(circ.step(n, cur), current=n < 60)
    assert sim.steps_done == 60
    with pytest.raises(AemError):
        sim.port_exchange(None)                   # no step left
    sim.write_outputs(str(tmp_path / "data.dat"))
    d = datfile.read_dat(str(tmp_path / "data.dat"))
    assert d.header["number of spice ports"] == "2"
    assert len(d.spice_v) == 2 and len(d.port_v) == 0
    np.testing.assert_array_equal(d.spice_v[0], sim.series("spice_v", 0))
    np.testing.assert_array_equal(d.spice_v[1], sim.series("spice_v", 1))
    np.testing.assert_allclose(sim.series("spice_v", 1), [src(n * dt) for n in range(1, 61)], rtol=1e-6,
                               atol=1e-12)


def test_dipole150_spice_example_matches_reference_csv(golden_dir):
    """The reference's SPICE-fed dipole example recomputed on the device with
    its netlist's circuit on the host: S11 = (V - V_inc)/V_inc, V_inc half the
    source (the example's source.py), against its shipped S_parameters.csv
    over the band the example plots (2.5-22 GHz)."""
    fx = np.load(os.path.join(golden_dir, "dipole150_spice_example.npz"))
    amp, f0, t0, width, dt_typed = fx["source"]
    plan = Plan.from_inputs(os.path.join(golden_dir, "decks", "dipole150_spice_example.txt"))
    v = plan.view
    src = cosim.gaussian_sine(amp, f0, t0, width)
    sim = Simulation(plan)
    circ = cosim.SeriesSourceCircuit(float(v.dt), [50.0], [v.spice[0].cap], [src])
    cosim.run_cosim(sim, circ)
    T = plan.time_steps
    freqs = fx["freq_ghz"] * 1e9
    dt = float(np.float32(v.dt))
    inc = 0.5 * np.array([src(n * dt_typed) for n in range(1, T + 1)])
    ph = np.exp(-2j * np.pi * np.outer(freqs, np.arange(T)) * dt)
    inc_f = ph @ inc
    vf = sim.dft("spice_v", freqs)
    s11 = (vf - inc_f) / inc_f
    band = (fx["freq_ghz"] >= 2.5) & (fx["freq_ghz"] <= 22.0)
    d_db = np.abs(20 * np.log10(np.abs(s11)) - fx["s11_db"])[band]
    d_rad = np.abs(np.angle(np.exp(1j * (np.angle(s11) - fx["s11_rad"]))))[band]
    print(f"\n  SPICE example S11 over {band.sum()} bins: max |d dB| = {d_db.max():.3e} (median {np.median(d_db):.2e}),"
          f" max |d rad| = {d_rad.max():.3e}")
    assert d_db.max() < 0.5 and np.median(d_db) < 0.05 and d_rad.max() < 0.1
