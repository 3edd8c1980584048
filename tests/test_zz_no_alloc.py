"""aem_get_dft_noalloc (static device scratch, no device allocation) against
aem_get_dft: the same (re, im) pairs bit for bit, on two calls in a row, the
second with fewer frequencies and another series, and a clear error past
the scratch's frequency limit.  Called through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from alpineem_amd import decks
from alpineem_amd._lib import SERIES, lib
from alpineem_amd.solver import Plan, Simulation

pytestmark = pytest.mark.gpu

MAX_FREQS = 16384


def _dft(fn, sim, kind, idx, freqs):
    f = np.ascontiguousarray(freqs, np.float64)
    out = np.full(2 * f.size, np.nan, np.float64)
    st = fn(sim._h, SERIES[kind], idx, f.ctypes.data, f.size, out.ctypes.data)
    return st, out


def test_dft_noalloc_equals_dft():
    plan = Plan.from_text(decks.dipole(n=36, steps=200, angles=2, arm=7).text())
    sim = Simulation(plan)
    sim.run(plan.time_steps)
    sim.finalize()
    L = lib()
    dt = float(plan.view.dt)
    f1 = np.linspace(0.0, 0.5 / dt, 3001)
    st_a, a = _dft(L.aem_get_dft, sim, "port_v", 0, f1)
    st_b, b = _dft(L.aem_get_dft_noalloc, sim, "port_v", 0, f1)
    assert st_a == 0 and st_b == 0, L.aem_last_error()
    assert np.any(a != 0)
    np.testing.assert_array_equal(b, a)
    # second call: fewer frequencies, another series; the scratch keeps the
    # first call's contents beyond them
    f2 = np.linspace(1e8, 8e9, 37)
    st_a, a = _dft(L.aem_get_dft, sim, "e_theta", 1, f2)
    st_b, b = _dft(L.aem_get_dft_noalloc, sim, "e_theta", 1, f2)
    assert st_a == 0 and st_b == 0, L.aem_last_error()
    assert np.any(a != 0)
    np.testing.assert_array_equal(b, a)
    # past the limit: an error, nothing written
    f3 = np.zeros(MAX_FREQS + 1)
    st_b, b = _dft(L.aem_get_dft_noalloc, sim, "port_v", 0, f3)
    assert st_b == 4
    assert "static scratch" in L.aem_last_error().decode()
    assert np.all(np.isnan(b))
    sim.close()
