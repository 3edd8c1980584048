"""The kmax (constant k-vector, complex-field) build: setup, oracle, pinning.

The reference's kmax build (fdtd_solver.f90 compiled with -Duse_kmax_version)
makes every field complex, the periodic walls Bloch-periodic and launches a
modal plane wave from one plane.  Its outputs on tests/golden/decks/kmax_*.txt
come from the flang-built kmax reference (tests/golden/make_golden.py):
the complex incident wave, the TE/TM reflected and transmitted plane sums as
data.dat labels them, and the complex port voltage.  The oracle (two real
states, one per part) must reproduce them bit for bit.
"""
import os

import numpy as np
import pytest

from alpineem_amd import _lib
from alpineem_amd.solver import Plan
from tests.oracle_ffi import Oracle

DECKS = ["kmax_te", "kmax_tm", "kmax_port"]


def kmax_plan(golden_dir, name):
    return Plan.from_inputs(os.path.join(golden_dir, "decks", name + ".txt"), kmax=1)


def labelled(series, k_direction):
    """The four plane-sum series in the writer's labels (f90:8690-8730): with
    k_direction 0 the 'Refl' records hold the transmitted sums."""
    r_te, r_tm, t_te, t_tm = (series[k] for k in ("reflected", "reflected_tm", "transmitted", "transmitted_tm"))
    if k_direction == 1:
        return {"te_refl": r_te, "tm_refl": r_tm, "te_trans": t_te, "tm_trans": t_tm}
    return {"te_refl": t_te, "tm_refl": t_tm, "te_trans": r_te, "tm_trans": r_tm}


def test_kmax_setup_matches_reference_formulas(golden_dir):
    """Source planes, mode weights, exceptions and Bloch phases of the three
    decks from the reference's setup formulas (f90:1506-1517, 1696-1806)."""
    c0 = np.float32(299792458.0)
    for name in DECKS:
        plan = kmax_plan(golden_dir, name)
        v = plan.view
        assert v.kmax == 1 and v.has_tfsf == 0
        a = v.k_axis
        assert v.pbc[a] == 0 and sum(v.pbc) == 2
        if v.k_direction == 1:
            assert (v.kpl_H, v.kpl_E) == (v.tf_lo[a] - 1, v.tf_lo[a])
        else:
            assert (v.kpl_H, v.kpl_E) == (v.tf_hi[a], v.tf_hi[a])
        p, q = [b for b in range(3) if b != a]
        kp, kq = np.float32(v.k_vec[p]), np.float32(v.k_vec[q])
        nrm = np.sqrt(np.float32(kq * kq + kp * kp))
        want = {(True, True): (0.0, 1.0), (True, False): (0.0, kq / nrm), (False, True): (kp / nrm, 0.0),
                (False, False): (kp / nrm, kq / nrm)}[(kp == 0, kq == 0)]
        assert (v.k_exc[p], v.k_exc[q]) == pytest.approx(want, rel=1e-7) and v.k_exc[a] == 0
        kmag = np.sqrt(np.float32(v.k_vec[0]) ** 2 + np.float32(v.k_vec[1]) ** 2 + np.float32(v.k_vec[2]) ** 2)
        assert v.f_adj == pytest.approx(kmag * c0 / (2 * np.pi) + v.f_center, rel=1e-6)
        N = (v.NX, v.NY, v.NZ)
        d = (v.dx, v.dy, v.dz)
        for b in range(3):
            th = np.float32(v.k_vec[b]) * np.float32(N[b] - 2) * np.float32(d[b])
            assert v.bloch_c[b] == pytest.approx(np.cos(np.float64(th)), abs=1e-7)
            assert v.bloch_s[b] == pytest.approx(np.sin(np.float64(th)), abs=1e-7)
        if v.mode_type == 0:
            assert list(v.WE) == [2.0 * v.plane_wave_amp] * 3 and list(v.WH) == [0, 0, 0]


@pytest.mark.parametrize("name", DECKS)
def test_kmax_oracle_bit_exact_vs_reference(golden_dir, name):
    plan = kmax_plan(golden_dir, name)
    orc = Oracle(plan)
    orc.step(plan.time_steps)
    orc.finalize()
    g = np.load(os.path.join(golden_dir, name + ".npz"))
    v = plan.view
    np.testing.assert_array_equal(orc.cseries("incident"), g["incident"])
    got = labelled({k: orc.cseries(k) for k in ("reflected", "reflected_tm", "transmitted", "transmitted_tm")},
                   v.k_direction)
    for k, s in got.items():
        np.testing.assert_array_equal(s, g[k], err_msg=k)
    for q in range(v.n_ports):
        np.testing.assert_array_equal(orc.cseries("port_v", q), g["port_v"][q], err_msg=f"port {q}")
    peak = max(np.abs(g[k]).max() for k in got)
    assert peak > 0.05, "the deck launches nothing"


def test_kmax_zero_k_reduces_to_real_periodic_run(golden_dir):
    """With k = 0 every Bloch factor is 1 and every source phase real at t=0;
    the imaginary part then evolves from the imaginary source alone, and
    real + i imag is the complex run: check linearity by running the oracle
    with the imaginary part's source only (set through the state) -- here the
    simpler invariant: at k = 0 the Bloch copies equal plain periodic copies,
    so a kmax pair started from a real field with no source keeps a zero
    imaginary part and its real part equals the ordinary build's."""
    text = open(os.path.join(golden_dir, "decks", "kmax_te.txt")).read()
    text = text.replace("300.0,150.0", "0.0,0.0").replace("\n1,0\nTE", "\n0,0\nTE")
    plan = Plan.from_text(text, kmax=1)
    assert plan.view.plane_wave_amp == 0
    orc = Oracle(plan)
    rng = np.random.default_rng(7)
    shape = plan.shape
    for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        orc.set_field(c, rng.standard_normal(shape).astype(np.float32) * 1e-3)
    orc.step(40)
    for c in ("Ex", "Hz"):
        assert not np.any(orc.field(c, imag=True))
    assert np.any(orc.field("Ez"))


@pytest.mark.parametrize("name", DECKS)
def test_kmax_plan_waveforms_equal_reference_incident(golden_dir, name):
    """The plan's complex waveform tables (what the device's port and
    incident-wave kernels read) hold the reference's incident series: row 2
    (plane wave) or the clear-line input rows of an antenna deck."""
    plan = kmax_plan(golden_dir, name)
    v = plan.view
    T = plan.time_steps
    g = np.load(os.path.join(golden_dir, name + ".npz"))
    re = np.ctypeslib.as_array(v.src_table, shape=(3 * T,))
    im = np.ctypeslib.as_array(v.src_table_im, shape=(3 * T,))
    if v.plane_wave_amp > 0:
        np.testing.assert_array_equal(re[2 * T:] + 1j * im[2 * T:].astype(np.complex64), g["incident"])
    else:
        # antenna: the first current-source sample is the waveform at n = 1
        assert re[0] != 0 and np.all(re[2 * T:] == 0) and np.all(im[2 * T:] == 0)
