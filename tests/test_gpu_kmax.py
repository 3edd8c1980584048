"""The kmax (constant k-vector, complex-field) build on the device.

A kmax simulation holds the real and the imaginary part of every field as
two device simulations of one plan stepped together (AEM_IMAG selects the
imaginary part).  On the kmax decks of tests/golden/decks/ (TE plane wave
through a dielectric and a Drude slab with k = (300, 150) rad/m; TM plane
wave launched in -x with k = (0, 250); a lumped-port dipole in an x/z unit
cell with k = (120, 40)) the device must equal the oracle bit for bit in
every field of both parts, the incident wave and the port voltage, and its
TE/TM plane sums (FP64 on the device, FP32 in the reference) must equal the
flang-built kmax reference's within 1e-5 relative L2.
"""
import os

import numpy as np
import pytest

from alpineem_amd.solver import Plan, Simulation
from tests.oracle_ffi import Oracle
from tests.test_kmax import DECKS, kmax_plan, labelled

pytestmark = pytest.mark.gpu

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
SUMS = ("reflected", "reflected_tm", "transmitted", "transmitted_tm")


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", DECKS)
def test_kmax_device_equals_oracle_and_reference(golden_dir, name):
    plan = kmax_plan(golden_dir, name)
    v = plan.view
    T = plan.time_steps
    sim = Simulation(plan)
    orc = Oracle(plan)
    half = T // 2
    sim.run(half)
    orc.step(half)
    for c in COMPS:
        for im in (False, True):
            np.testing.assert_array_equal(sim.field(c, imag=im), orc.field(c, imag=im),
                                          err_msg=f"{name} {c} {'imag' if im else 'real'} at step {half}")
    sim.run(T - half, graph=False)
    orc.step(T - half)
    sim.finalize()
    orc.finalize()
    for c in COMPS:
        for im in (False, True):
            np.testing.assert_array_equal(sim.field(c, imag=im), orc.field(c, imag=im),
                                          err_msg=f"{name} {c} {'imag' if im else 'real'} at step {T}")
    np.testing.assert_array_equal(sim.cseries("incident"), orc.cseries("incident"))
    for q in range(v.n_ports):
        np.testing.assert_array_equal(sim.cseries("port_v", q), orc.cseries("port_v", q))
    g = np.load(os.path.join(golden_dir, name + ".npz"))
    got = labelled({k: sim.cseries(k) for k in SUMS}, v.k_direction)
    worst = 0.0
    for k, s in got.items():
        r = rel_l2(s, g[k])
        worst = max(worst, r)
        assert r <= 1e-5, (name, k, r)
    print(f"\n  {name}: fields and port/incident series equal the oracle in both parts; TE/TM sums within "
          f"{worst:.1e} of the kmax reference")


def test_kmax_zero_k_is_two_ordinary_runs(golden_dir):
    """At k = 0 with no plane wave the Bloch factors are 1: a pair started
    from a real field and a separate imaginary field evolves as two
    ordinary periodic simulations of the same deck (linearity)."""
    text = open(os.path.join(golden_dir, "decks", "kmax_te.txt")).read()
    text = text.replace("300.0,150.0", "0.0,0.0").replace("\n1,0\nTE", "\n0,0\nTE")
    kplan = Plan.from_text(text, kmax=1)
    rng = np.random.default_rng(11)
    shape = kplan.shape
    init = {(c, im): (rng.standard_normal(shape) * 1e-3).astype(np.float32) for c in COMPS for im in (False, True)}
    k = Simulation(kplan)
    orc = Oracle(kplan)
    for (c, im), a in init.items():
        k.set_field(c, a, imag=im)
        orc.set_field(c, a, imag=im)
    k.run(60)
    orc.step(60)
    for c in COMPS:
        for im in (False, True):
            np.testing.assert_array_equal(k.field(c, imag=im), orc.field(c, imag=im), err_msg=c)
    assert np.any(k.field("Ez", imag=True))


def test_kmax_refusals(golden_dir):
    from alpineem_amd._lib import AemError
    from alpineem_amd.solver import Batch
    plan = kmax_plan(golden_dir, "kmax_te")
    sim = Simulation(plan)
    with pytest.raises(AemError):
        Batch([sim])
    with pytest.raises(AemError):
        sim.time_phases(2)
    plain = Simulation(Plan.from_inputs(os.path.join(golden_dir, "decks", "drude_slab.txt")))
    with pytest.raises(AemError):
        plain.series("incident", imag=True)


def test_kmax_cli_writes_the_kmax_data_dat(golden_dir, tmp_path):
    """The CLI with --kmax runs a kmax deck and writes the kmax build's
    data.dat: header records, complex incident, TE/TM sections and port
    voltage equal to the reference's values (FP32 tokens)."""
    import shutil

    from alpineem_amd import cli
    from alpineem_amd.datfile import read_dat
    shutil.copy(os.path.join(golden_dir, "decks", "kmax_port.txt"), tmp_path / "inputs.txt")
    assert cli.main(["--kmax", str(tmp_path / "inputs.txt")]) == 0
    d = read_dat(str(tmp_path / "data.dat"))
    g = np.load(os.path.join(golden_dir, "kmax_port.npz"))
    assert d.kmax and d.k_vec == (120.0, 40.0) and d.mode_type == 0
    assert d.header["number of internal ports"] == "1"
    np.testing.assert_array_equal(d.incident, g["incident"])
    np.testing.assert_array_equal(d.port_v[0], g["port_v"][0])
    for k in ("te_refl", "tm_refl", "te_trans", "tm_trans"):
        assert rel_l2(getattr(d, k), g[k]) <= 1e-5, k
