"""The planes tuner (alpineem_amd/tune.py).

CPU: bad arguments are refused on the host, before any simulation is made.
GPU: the tuner returns one of its candidates with a time for each, and a run
with the chosen value gives the same results, bit for bit, as a run with the
library's default chunking.
"""
import os

import numpy as np
import pytest

from alpineem_amd.solver import Plan, Simulation
from alpineem_amd.tune import tune_batch_planes, tune_planes


def _plan(golden_dir, name):
    return Plan.from_inputs(os.path.join(golden_dir, "decks", name + ".txt"))


@pytest.mark.parametrize("kw, text", [
    ({"steps": 0}, "at least 1"),
    ({"warmup": 0}, "at least 1"),
    ({"candidates": ()}, "non-empty"),
    ({"candidates": (2, -1)}, ">= 0"),
    ({"steps": 190, "warmup": 20}, "time_steps"),       # dipole36 has 200 steps
])
def test_bad_arguments_are_refused_on_the_host(golden_dir, kw, text):
    plan = _plan(golden_dir, "dipole36")
    with pytest.raises(ValueError, match=text):
        tune_planes(plan, **kw)
    with pytest.raises(ValueError, match=text):
        tune_batch_planes([plan], **kw)


def test_empty_batch_is_refused():
    with pytest.raises(ValueError, match="no plans"):
        tune_batch_planes([])


@pytest.mark.gpu
def test_tuned_planes_change_no_result(golden_dir):
    plan = _plan(golden_dir, "dipole36")
    best, ms = tune_planes(plan, candidates=(0, 1, 3, 8), steps=10, warmup=5)
    print(f"\n[tune dipole36] ms per step {ms}, best {best}")
    assert best in (0, 1, 3, 8) and sorted(ms) == [0, 1, 3, 8]
    assert all(v > 0 for v in ms.values())
    tuned, default = Simulation(plan, planes=best), Simulation(plan)
    tuned.run()
    default.run()
    np.testing.assert_array_equal(tuned.series("port_v"), default.series("port_v"))
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(tuned.field(comp), default.field(comp), err_msg=comp)
    tuned.close()
    default.close()


@pytest.mark.gpu
def test_batch_tuner_times_every_candidate(golden_dir):
    plans = [_plan(golden_dir, "dipole36"), _plan(golden_dir, "antenna_x"), _plan(golden_dir, "c1_mini")]
    best, ms = tune_batch_planes(plans, candidates=(0, 2, 5), steps=10, warmup=5, ntff="histogram")
    print(f"\n[tune batch of 3] ms per step {ms}, best {best}")
    assert best in (0, 2, 5) and sorted(ms) == [0, 2, 5]
    assert all(v > 0 for v in ms.values())
