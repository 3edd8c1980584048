"""Frequency-domain products of a run: S-parameters, realized gain, RCS.

The reference computes these after the run with numpy FFTs of the
``data.dat`` series, zero-padded by a factor ``padding+1``
(``utility_scripts/post_processor.py:191-266, 351-388``).  Here every
spectrum comes from :meth:`Simulation.dft`, an FP64 Kahan-compensated DFT on
the device at exactly the requested frequencies.  Zero padding adds nothing
to a DFT sum, so at the post-processor's frequencies ``k/((padding+1) T dt)``
the values equal its FFT bins; :func:`fft_frequencies` gives that grid and
:func:`pulse_band` the band it keeps.

The formulas below follow the post-processor's, with its constants.
"""
from __future__ import annotations

import numpy as np

IMPED_FREE = 376.730313          # post_processor.py:58


def fft_frequencies(time_steps: int, dt: float, padding: int = 10) -> np.ndarray:
    """Non-negative frequencies of numpy.fft.fftfreq(T*(padding+1), dt)
    (post_processor.py:192 after the padding of :162-179)."""
    n = time_steps * (padding + 1)
    return np.fft.fftfreq(n, dt)[: (n - 1) // 2 + 1]


def pulse_band(pulse_type: int, f_center: float):
    """Frequency band the post-processor writes out (post_processor.py:30-33, 273-281)."""
    lo, hi = (0.0, 2.146) if pulse_type == 1 else (0.183, 3.087)
    return lo * f_center, hi * f_center


def antenna_s_parameters(v_ports, incident, impedances, excited: int):
    """S_{q,exc} for every lumped port of an antenna run (post_processor.py:255-262).

    ``v_ports[q]`` and ``incident`` are complex spectra, ``excited`` is the
    0-based excited port.  Returns a list of complex arrays."""
    out = []
    z_exc = np.real(impedances[excited])
    for q, vq in enumerate(v_ports):
        if q == excited:
            out.append((vq - incident) / incident)
        else:
            out.append((vq / np.sqrt(np.real(impedances[q]))) / (incident / np.sqrt(z_exc)))
    return out


def realized_gain(e_far, incident, z_excited: float, freqs, t_corr: float):
    """Realized gain (dB) and phase of one far-field polarisation of an
    antenna run (post_processor.py:263-265, 364-371)."""
    p_ant = np.abs(incident) ** 2 / (2 * np.real(z_excited))
    gain_db = 10 * np.log10(4 * np.pi * np.abs(e_far) ** 2 / (2 * IMPED_FREE * p_ant))
    phase = np.angle(e_far * np.exp(1j * 2 * np.pi * freqs * t_corr - 1j * np.pi / 2) / incident)
    return gain_db, phase


def scattered_far_field(e_obj, e_clear, t_obj: float, t_clear: float, inc_pc, freqs):
    """rE_scat/E_inc of a plane-wave run, object minus clear run, each
    phase-centred by its far-field time correction (post_processor.py:245-250)."""
    normal = np.exp(1j * 2 * np.pi * freqs * t_obj - 1j * np.pi / 2)
    clear = np.exp(1j * 2 * np.pi * freqs * t_clear - 1j * np.pi / 2)
    return (e_obj * normal - e_clear * clear) / inc_pc


def rcs_db(scatter):
    """RCS in dBsm from rE_scat/E_inc (post_processor.py:351-353)."""
    return 20 * np.log10(np.abs(scatter)) + 10 * np.log10(4 * np.pi)


def wave_port_r_t(obj, clear, metal):
    """Periodic-slab reflection and transmission from the plane-sum spectra
    of object, clear and metal runs (post_processor.py:223-229):
    [reflection, transmission]."""
    refl = -1 * (obj[0] - clear[0]) / (metal[0] - clear[0])
    trans = -1 * (obj[1] - metal[1]) / (clear[1] - metal[1])
    return [refl, trans]


def antenna_report(sim, freqs):
    """S-parameters of every port and realized gain at every far-field angle
    of a finalized antenna run, from device DFTs at ``freqs``.

    Returns ``{"s": [complex per port], "gain_theta": [(dB, rad) per angle],
    "gain_phi": [...]}``."""
    v = sim.plan.view
    freqs = np.asarray(freqs, np.float64)
    exc = int(v.excitation_port) - 1
    inc = sim.dft("incident", freqs)
    vp = [sim.dft("port_v", freqs, q) for q in range(v.n_ports)]
    z = [float(v.ports[q].Z) for q in range(v.n_ports)]
    rep = {"s": antenna_s_parameters(vp, inc, z, exc), "gain_theta": [], "gain_phi": []}
    if v.n_angles > 0:
        tc = sim.ff_time()
        for a in range(v.n_angles):
            for key, kind in (("gain_theta", "e_theta"), ("gain_phi", "e_phi")):
                rep[key].append(realized_gain(sim.dft(kind, freqs, a), inc, z[exc], freqs, float(tc[a])))
    return rep
