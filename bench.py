"""Throughput benchmark of the FDTD leapfrog loop on one MI355X.

Workload (the reference metric's configuration): a 400^3 FP32 grid, dx = 1 mm,
dt = 0.99 x the 3-D Courant limit, 10-cell CPML on all faces, vacuum, and a
z-directed 50 ohm lumped port at the centre driven by a Gaussian of spread
40 dt and t_spread 160 dt.  A "step" is one full leapfrog step (H phase,
E phase, CPML, port, per-step outputs).  Inputs are resident in HBM before
the timed region; W warm-up steps are run first, then exactly K steps are
timed (wall clock around a device-synchronised replay of K captured step
graphs).  value = 400^3 x K / seconds, in Gcell-updates/s.

The JSON line also carries
  roofline      the dominant kernel's algorithmic HBM bytes per launch over its
                average duration, timed with HIP events on the sim's stream;
  cpu_baseline  the reference solver (oracle/_ref/fdtd_ref_omp, built from
                the reference source by oracle/Makefile) on the same deck for a
                bounded number of steps, read from its own Mcells/s line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
BYTES_H = 36          # read Hx,Hy,Hz,Ex,Ey,Ez + write Hx,Hy,Hz (FP32)
BYTES_E = 37          # read Ex,Ey,Ez,Hx,Hy,Hz + 1-byte coefficient index, write Ex,Ey,Ez
BYTES_STEP = BYTES_H + BYTES_E
BYTES_SP = 49         # single pass: read Ex..Hz + 1-byte index, write Ex..Hz once per step


def cpu_baseline(n: int, steps: int) -> dict:
    """Time the reference OpenMP build (or, if it is absent, the single-thread
    C oracle) on a bounded sample of the same workload."""
    from alpineem_amd import decks
    ref = os.path.join(ROOT, "oracle", "_ref", "fdtd_ref_omp")
    cores = min(16, os.cpu_count() or 1)
    deck = decks.roofline_deck(n, steps).text()
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as tmp:
            open(os.path.join(tmp, "inputs.txt"), "w").write(deck)
            env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
            r = subprocess.run([ref, "inputs.txt"], cwd=tmp, env=env, capture_output=True, text=True, timeout=600)
            m = re.search(r"Average of\s+([0-9.Ee+-]+)\s+Mcells per second", r.stdout)
            if r.returncode == 0 and m:
                return {"value": round(float(m.group(1)) / 1000.0, 5), "unit": "Gcell/s", "cores": cores,
                        "kind": "reference",
                        "sample": f"reference OpenMP build, {n}^3 deck, {steps} steps, its own Mcells/s meter"}
    # fallback: the C restatement, one thread, a few steps
    from alpineem_amd.solver import Plan
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from tests.oracle_ffi import Oracle
    plan = Plan.from_text(deck)
    orc = Oracle(plan)
    t0 = time.perf_counter()
    orc.step(steps)
    el = time.perf_counter() - t0
    cells = n ** 3
    return {"value": round(cells * steps / el / 1e9, 5), "unit": "Gcell/s", "cores": 1, "kind": "port",
            "sample": f"C oracle, {n}^3 deck, {steps} steps, 1 thread"}


def copy_probe(mib: int = 256, reps: int = 3) -> None:
    """Device-to-device copies of a known size (``mib`` MiB), for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0        # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy probe: {reps} x {mib} MiB device-to-device", file=sys.stderr, flush=True)


COPY_KIB = 256 * 1024      # --copy-probe: each copy reads and writes 256 MiB


def population_bench(steps: int, count: int = 64, ntff: str = "ordered") -> dict:
    """The batched-population workload: ``count`` independent dipole antennas
    with cube edges drawn from {80, 96, ..., 160} (seed 3), random arms and up
    to four dielectric blocks, 10-cell CPML, 50 ohm port, 19 far-field
    angles (decks.population).  Timed for ``steps`` steps: once as 64
    sequential solo runs, once as one batch; the same far-field binning in
    both: "ordered" (the reference's FP32 sums in its order, from contributor
    lists; deterministic and bit-exact) or "histogram" (FP64 LDS histograms,
    whose atomic additions have no fixed order).  The batch's z chunk is the
    autotuner's choice (alpineem_amd.tune) over throwaway members, untimed.
    Only stepping is timed (each solo run and the batch include their one
    graph capture); setup is not."""
    from alpineem_amd import decks
    from alpineem_amd.solver import Batch, Plan, Simulation

    pop = decks.population(count=count, seed=3, steps=steps)
    plans = [Plan.from_text(d.text()) for d in pop]
    cells = sum(p.view.nx * p.view.ny * p.view.nz for p in plans)
    seq = 0.0
    for q, p in enumerate(plans):
        s = Simulation(p, ntff=ntff)
        t0 = time.perf_counter()
        s.run(steps)
        seq += time.perf_counter() - t0
        s.close()
        if q % 16 == 15:
            print(f"[bench] population ({ntff}): {q + 1} of {len(plans)} solo runs", file=sys.stderr, flush=True)
    # the batch's z chunk chosen by the autotuner on throwaway members (not timed)
    from alpineem_amd.tune import tune_batch_planes
    t0 = time.perf_counter()
    planes, tuned = tune_batch_planes(plans, candidates=(0, 2, 4, 8), steps=10, warmup=3, ntff=ntff)
    t_tune = time.perf_counter() - t0
    print(f"[bench] population ({ntff}): tuner chose planes={planes} in {t_tune:.1f} s", file=sys.stderr, flush=True)
    sims = [Simulation(p, ntff=ntff, planes=planes) for p in plans]
    b = Batch(sims)
    t0 = time.perf_counter()
    b.run(steps)
    bat = time.perf_counter() - t0
    res = {"members": count, "cells_per_step": cells, "steps": steps, "ntff": f"{ntff}, 19 angles",
           "sequential_gcells": round(cells * steps / seq / 1e9, 3), "batched_gcells": round(cells * steps / bat / 1e9, 3),
           "speedup": round(seq / bat, 3), "sequential_s": round(seq, 3), "batched_s": round(bat, 3),
           "tuned_planes": planes, "tuner_ms_per_step": {str(k): round(v, 4) for k, v in tuned.items()},
           "tuner_s": round(t_tune, 2)}
    b.close()
    for s in sims:
        s.close()
    return res


def adjoint_bench(steps: int) -> dict:
    """The differentiable workload (decks.adjoint_deck): 160^3, 12-cell CPML,
    x-polarised TF/SF differentiated Gaussian, a 40 x 40 x 20 design box of
    eps_r U[1, 4] (seed 5), objective |r E_theta(30 deg, 0, 10 GHz)|^2 by
    the frequency-domain transform (the time-domain far-field series are off
    in both runs).  Times one plain forward run (no objective, no design box), the recorded
    forward run and the adjoint sweep of the gradient (wall clock, device
    synchronised; each run includes its one graph capture)."""
    from alpineem_amd import decks
    from alpineem_amd.solver import Plan, Simulation

    d, lo, hi = decks.adjoint_deck(steps=steps)
    plan = Plan.from_text(d.text(), cpml_cells=12)
    shape = tuple(hi[a] - lo[a] + 1 for a in (2, 1, 0))
    eps = decks.design_permittivity(shape, seed=5)
    plain = Simulation(plan, ntff="off")
    t0 = time.perf_counter()
    plain.run()
    t_plain = time.perf_counter() - t0
    plain.close()
    sim = Simulation(plan, ntff="off")
    sim.set_objective(30.0, 0.0, 10e9, "theta")
    sim.design_box(lo, hi)
    sim.set_design_eps(eps)
    t0 = time.perf_counter()
    sim.run()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = sim.design_gradient()
    t_bwd = time.perf_counter() - t0
    J = abs(sim.objective()) ** 2
    sim.close()
    # the same gradient with checkpoints every 200 steps (a 200-step record,
    # each interval recomputed from its snapshot in the backward sweep)
    ck = Simulation(plan, ntff="off")
    ck.checkpoint(200)
    ck.set_objective(30.0, 0.0, 10e9, "theta")
    ck.design_box(lo, hi)
    ck.set_design_eps(eps)
    t0 = time.perf_counter()
    ck.run()
    t_fwd_ck = time.perf_counter() - t0
    t0 = time.perf_counter()
    g_ck = ck.design_gradient()
    t_bwd_ck = time.perf_counter() - t0
    ck.close()
    n = plan.view.nx * plan.view.ny * plan.view.nz
    return {"workload": f"{plan.view.nx}^3, {steps} steps, design box {shape[2]}x{shape[1]}x{shape[0]}",
            "forward_s": round(t_plain, 4), "recorded_forward_s": round(t_fwd, 4), "adjoint_s": round(t_bwd, 4),
            "gradient_over_forward": round((t_fwd + t_bwd) / t_plain, 3),
            "forward_gcells": round(n * steps / t_plain / 1e9, 2), "J": J,
            "grad_norm": float(np.linalg.norm(g)),
            "checkpoint_every": 200, "checkpointed_forward_s": round(t_fwd_ck, 4),
            "checkpointed_adjoint_s": round(t_bwd_ck, 4),
            "checkpointed_gradient_over_forward": round((t_fwd_ck + t_bwd_ck) / t_plain, 3),
            "checkpointed_equal": bool(np.array_equal(g, g_ck))}


def configs_bench(only: str = None) -> dict:
    """The benchmark configurations with every output of each, timed end to
    end on the device (wall clock around the stepping; setup excluded):

    C1  100^3 radiating dipole (arms 20 cells, 1-cell gap), 2000 steps,
        10-cell CPML with kappa_max 5 / alpha_max 0.05, 50 ohm port; port
        V/I series, S11 at 101 frequencies over 0.5-8 GHz (device DFT) and
        the far field at 37 theta angles (phi = 0) with the reference's
        ordered FP32 binning (its outputs bit for bit).
    C2  200^3 TF/SF PEC sphere of radius 40 cells, dx 2 mm, 4000 steps,
        12-cell CPML, TF/SF box 15 cells inside the CPML, far-field box 5
        cells outside it; monostatic far field (ordered binning) and the
        running FP64/Kahan DFT far field at 19 bistatic angles x 201
        frequencies.
    C3  32 x 32 x 1200 periodic layered slab, five lossy anisotropic layers,
        a 6-pole Drude slab and a 100 ohm/sq sheet, 16-cell CPML in z, 12000
        steps; reflection/transmission plane sums and their spectra at 401
        frequencies.
    R5  the 400^3 throughput deck with the survey's performance CPML
        (kappa_max 5, alpha_max 0.05), 200 steps.
    ``only`` runs one of them ("C1", "C2", "C3", "R5")."""
    from alpineem_amd import decks
    from alpineem_amd.solver import Plan, Simulation

    out = {}

    def timed(name, plan, sim, prep=None, post=None):
        t0 = time.perf_counter()
        sim.run()
        step_s = time.perf_counter() - t0
        t1 = time.perf_counter()
        if post:
            post(sim)
        post_s = time.perf_counter() - t1
        v = plan.view
        cells = v.nx * v.ny * v.nz
        out[name] = {"cells": cells, "steps": plan.time_steps, "step_s": round(step_s, 4),
                     "gcells": round(cells * plan.time_steps / step_s / 1e9, 3), "outputs_s": round(post_s, 4)}
        print(f"[bench] {name}: {v.nx}x{v.ny}x{v.nz}, {plan.time_steps} steps, {out[name]['gcells']} Gcell/s "
              f"(stepping {step_s:.3f} s, outputs {post_s:.3f} s)", file=sys.stderr, flush=True)
        sim.close()

    def want(name):
        return only is None or only == name

    if want("C1"):
        plan = Plan.from_text(decks.dipole(n=100, steps=2000, angles=37, arm=20).text(), cpml_kappa_max=5.0,
                              cpml_alpha_max=0.05)
        sim = Simulation(plan)

        def c1_post(s):
            s.finalize()
            f = np.linspace(0.5e9, 8e9, 101)
            _ = (s.dft("port_v", f) - s.dft("incident", f)) / s.dft("incident", f)
            for a in range(plan.view.n_angles):
                s.series("e_theta", a)
        timed("C1", plan, sim, post=c1_post)
        out["C1"]["outputs"] = "port V/I, S11 x101, far field 37 angles (ordered FP32 binning)"

    if want("C2"):
        plan = Plan.from_text(decks.pec_sphere(n=200, steps=4000, radius=40, angles=[(180.0, 0.0)]).text(),
                              cpml_cells=12, tfsf_inset=15, ntff_gap=5)
        sim = Simulation(plan)
        sim.set_farfield_dft([(float(t), 0.0) for t in np.linspace(0.0, 180.0, 19)], 0.5e9, 0.05e9, 201)

        def c2_post(s):
            s.finalize()
            s.series("e_theta", 0)
            s.farfield_dft()
        timed("C2", plan, sim, post=c2_post)
        out["C2"]["outputs"] = ("monostatic far field (ordered FP32 binning), running FP64 DFT far field 19 angles "
                                "x 201")

    if want("C3"):
        plan = Plan.from_text(decks.layered_slab().text(), cpml_cells=16)
        sim = Simulation(plan)

        def c3_post(s):
            s.finalize()
            f = np.linspace(1e9, 100e9, 401)
            s.dft("reflected", f)
            s.dft("transmitted", f)
        timed("C3", plan, sim, post=c3_post)
        out["C3"]["outputs"] = "reflected/transmitted plane sums, spectra x401"

    if want("R5"):   # performance CPML at 400^3
        plan = Plan.from_text(decks.roofline_deck(400, 220).text(), cpml_kappa_max=5.0, cpml_alpha_max=0.05)
        sim = Simulation(plan)
        sim.run(20)
        t0 = time.perf_counter()
        sim.run(200)
        el = time.perf_counter() - t0
        out["R400_kappa5"] = {"cells": 400 ** 3, "steps": 200, "gcells": round(400 ** 3 * 200 / el / 1e9, 3),
                              "single_pass": sim.single_pass}
        print(f"[bench] 400^3 with kappa_max 5, alpha_max 0.05: {out['R400_kappa5']['gcells']} Gcell/s",
              file=sys.stderr, flush=True)
        sim.close()
    return out


def _short_name(kernel: str) -> str:
    """'void aemk::k_e_bulk<unsigned char, 1>(aemk::Grid, ...)' -> 'k_e_bulk<unsigned char, 1>'."""
    s = kernel.split("(")[0].replace("void ", "").strip()
    return s.split("::")[-1] if "<" not in s else s.split("::", 1)[-1]


def pmc_summary(fetch_csv: str, write_csv: str, out_json: str) -> dict:
    """HBM bytes per launch from two rocprofv3 --pmc passes over the same
    bench command with --copy-probe (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch).  Each counter is scaled by its reading for the known-size
    device copies of the same pass; on gfx950 FETCH_SIZE reads about half
    the copied bytes and WRITE_SIZE reads them exactly."""
    import csv

    def load(path):
        per, copies = {}, []
        for r in csv.DictReader(open(path)):
            v = float(r["Counter_Value"])
            if "copyBuffer" in r["Kernel_Name"]:
                if v > COPY_KIB / 4:
                    copies.append(v)
                continue
            per.setdefault(_short_name(r["Kernel_Name"]), []).append(v)
        if not copies:
            raise SystemExit(f"{path}: no known-size copy found (run the pass with --copy-probe)")
        return {k: sum(v) / len(v) for k, v in per.items()}, COPY_KIB / (sum(copies) / len(copies))

    fetch, sf = load(fetch_csv)
    write, sw = load(write_csv)
    out = {"_scale": {"FETCH_SIZE": round(sf, 4), "WRITE_SIZE": round(sw, 4)}}
    for k in fetch:
        if k in write:
            out[k] = {"fetch_bytes": round(fetch[k] * sf * 1024), "write_bytes": round(write[k] * sw * 1024),
                      "bytes": round((fetch[k] * sf + write[k] * sw) * 1024)}
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return out


def main() -> int:
    if len(sys.argv) == 5 and sys.argv[1] == "--pmc-summary":
        res = pmc_summary(sys.argv[2], sys.argv[3], sys.argv[4])
        for k, v in res.items():
            print(k, v)
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=400)
    ap.add_argument("--phase-steps", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="also run a known-size device copy (counter scaling)")
    ap.add_argument("--no-single-pass", action="store_true",
                    help="separate H and E bulk launches instead of the single-pass core launch")
    ap.add_argument("--configs", type=int, default=1, help="1: also time the other configurations with full outputs")
    ap.add_argument("--population-steps", type=int, default=3000,
                    help="steps of the 64-member batched-population comparison (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="steps of the 160^3 design-gradient timing (0: skip)")
    ap.add_argument("--leg", choices=["C1", "C2", "C3", "R5"], default=None,
                    help="run only this configuration leg, as the first and only simulation of the process, and "
                         "print its figures (one workload per profiled process)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this solver runs on one GPU; runs over several GPUs (z-slab sharding with one "
              "halo plane per half-step) are not built", file=sys.stderr)
        return 2

    if args.leg:
        print(json.dumps({"leg": args.leg, "configs": configs_bench(args.leg)}), flush=True)
        return 0

    from alpineem_amd import decks
    from alpineem_amd.solver import Plan, Simulation

    n, K, W, P = args.n, args.steps, args.warmup, args.phase_steps
    deck = decks.roofline_deck(n, W + K + P)
    t0 = time.perf_counter()
    plan = Plan.from_text(deck.text())
    sim = Simulation(plan, single_pass=not args.no_single_pass)
    setup_s = time.perf_counter() - t0
    print(f"[bench] setup {setup_s:.1f}s, grid {n}^3, tuples {plan.view.n_tuple}", file=sys.stderr, flush=True)

    sim.run(W)                                    # warm-up (also captures the step graph)
    t0 = time.perf_counter()
    sim.run(K)                                    # returns after the device is synchronised
    wall = time.perf_counter() - t0
    cells = n ** 3
    value = cells * K / wall / 1e9
    ms_step = wall / K * 1e3
    print(f"[bench] timed {K} steps: {ms_step:.4f} ms/step, {value:.2f} Gcell/s", file=sys.stderr, flush=True)

    single = sim.single_pass                      # core box as one H+E launch (k_step_sp)
    ph = sim.time_phases(P)                       # ms per step per phase, HIP events on the sim stream
    names = (["core_step", "rest_h", "core_e", "rest_e", "outputs"] if single
             else ["h_core", "h_rest", "e_core", "e_rest", "outputs"])
    print("[bench] phases (ms): " + ", ".join(f"{a}={b:.4f}" for a, b in zip(names, ph)), file=sys.stderr, flush=True)
    core = sim.core_cells                         # cells of the core launch timed in phase 0 (and 2 if split)
    c = 1 if core < cells else 0                  # fused CPML: the core kernel carries the x-slab test (kCx)
    if single:
        dom, bpc, t_ms = f"k_step_sp<unsigned char, {c}>", BYTES_SP, ph[0]
    elif ph[2] >= ph[0]:
        dom, bpc, t_ms = f"k_e_bulk<unsigned char, {c}, false, unsigned int>", BYTES_E, ph[2]
    else:
        dom, bpc, t_ms = f"k_h_bulk<{c}, false, unsigned int>", BYTES_H, ph[0]
    achieved = bpc * core / (t_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None, "kernel": dom,
                "bytes_per_cell": bpc, "cells_per_launch": core, "launch_ms": round(float(t_ms), 5),
                "step_frac": round(value * (BYTES_SP if single else BYTES_STEP) / PEAK_HBM_GBS, 4)}
    traffic_file = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(traffic_file):
        entry = json.load(open(traffic_file)).get(dom)
        if entry:
            # HBM bytes per launch (copy-scaled FETCH_SIZE + WRITE_SIZE) against the algorithmic bytes
            roofline["traffic"] = entry["bytes"]
            roofline["traffic_over_algorithmic"] = round(entry["bytes"] / (bpc * core), 3)

    if args.copy_probe:
        copy_probe()
    sim.close()
    pop = pop_h = None
    if args.population_steps > 0:
        pop = population_bench(args.population_steps, ntff="ordered")
        pop_h = population_bench(args.population_steps, ntff="histogram")
        for r in (pop, pop_h):
            print(f"[bench] population ({r['ntff']}): 64 members, {r['cells_per_step'] / 1e6:.1f} Mcells per step, "
                  f"sequential {r['sequential_gcells']} Gcell/s, batched {r['batched_gcells']} Gcell/s, "
                  f"speedup {r['speedup']}x (tuned z chunk {r['tuned_planes']}: {r['tuner_ms_per_step']} ms, "
                  f"{r['tuner_s']} s)", file=sys.stderr, flush=True)
    adj = None
    if args.adjoint_steps > 0:
        adj = adjoint_bench(args.adjoint_steps)
        print(f"[bench] adjoint: {adj['workload']}: forward {adj['forward_s']} s, recorded forward "
              f"{adj['recorded_forward_s']} s, adjoint {adj['adjoint_s']} s, gradient/forward "
              f"{adj['gradient_over_forward']}; with checkpoints every {adj['checkpoint_every']} steps: "
              f"{adj['checkpointed_forward_s']} + {adj['checkpointed_adjoint_s']} s, gradient/forward "
              f"{adj['checkpointed_gradient_over_forward']}, equal {adj['checkpointed_equal']}",
              file=sys.stderr, flush=True)
    cfg = configs_bench() if args.configs else None
    cpu = None if args.no_cpu else cpu_baseline(n, args.cpu_steps)
    line = {
        "metric": "Gcell-updates/s at 400^3 FP32 (CPML, lumped port)",
        "value": round(value, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 vacuum, dx=1mm, dt=0.99 CFL, 10-cell CPML, z-directed 50 ohm lumped port, "
                               f"Gaussian spread 40dt / t_spread 160dt", "cells": cells,
                   "phase_ms": {a: round(float(b), 5) for a, b in zip(names, ph)}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "population": pop,
        "population_histogram": pop_h,
        "adjoint": adj,
        "configs": cfg,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
