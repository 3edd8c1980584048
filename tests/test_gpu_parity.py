"""HIP path (libaem on gfx950) against the oracle and the reference fixtures.

For every golden deck the device run must reproduce the reference outputs
(tests/golden/*.npz) to a relative L2 of 1e-5 or better -- the FP32 parity
bar of this path -- and the fields left in device memory must match the
oracle's.  Bulk, CPML, sheet, Drude and port arithmetic is order-identical
FP32, so port voltages and fields of decks without plane-wave sources are
expected bit-exact; the residual differences elsewhere come from the device
expf() (TF/SF incident field) and from FP64 instead of FP32 accumulation of
the far-field bins and the periodic-slab plane sums.
"""
import os

import numpy as np
import pytest

from alpineem_amd.solver import Plan, Simulation
from tests.oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

DECKS = ["dipole36", "master_ports", "drude_slab", "sphere_oblique", "antenna_x", "c1_mini", "c2_mini", "c3_mini",
         "sheets_xyz", "igp_mz", "igp_px"]
TOL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def runs(golden_dir):
    out = {}
    for name in DECKS:
        plan = Plan.from_inputs(os.path.join(golden_dir, "decks", name + ".txt"))
        sim = Simulation(plan)
        sim.run(plan.time_steps)
        sim.finalize()
        out[name] = (plan, sim, np.load(os.path.join(golden_dir, name + ".npz")))
    return out


def _series_sets(src, g, v):
    """[(label, got, ref)] for every output series of a deck; ``src`` is a
    Simulation or an Oracle (same series names)."""
    checks = [("incident", src.series("incident"), g["incident"])]
    if "port_v" in g:
        checks += [(f"port_v[{q}]", src.series("port_v", q), g["port_v"][q]) for q in range(v.n_ports)]
    if "e_theta" in g:
        # one far-field vector (E_theta, E_phi) per angle: the cross-polar
        # component alone is numerical noise (|E_phi| ~ 1e-7 |E_theta| for a
        # z-dipole at phi=0) and has no meaningful relative error of its own
        for a in range(v.n_angles):
            got = np.concatenate([src.series("e_theta", a), src.series("e_phi", a)])
            ref = np.concatenate([g["e_theta"][a], g["e_phi"][a]])
            checks.append((f"far[{a}]", got, ref))
    if "reflected" in g:
        checks.append(("reflected", src.series("reflected"), g["reflected"]))
        checks.append(("transmitted", src.series("transmitted"), g["transmitted"]))
    if "ff_pc" in g:
        checks.append(("ff_pc", src.series("ff_pc"), g["ff_pc"]))
    return checks


@pytest.mark.parametrize("name", DECKS)
def test_series_match_reference(runs, name):
    """Every series within rel-L2 1e-5 of the reference, including signals
    that nearly cancel (dipole pattern nulls, fields behind metal), where the
    reference itself sits far further from FP64 arithmetic.  The distances
    from the FP64 oracle are printed beside each error."""
    plan, sim, g = runs[name]
    v = plan.view
    print(f"\n[{name}] grid {v.nx}x{v.ny}x{v.nz}, {plan.time_steps} steps")
    o64 = Oracle(plan, fp64=True)
    o64.step(plan.time_steps)
    o64.finalize()
    sets64 = _series_sets(o64, g, v)
    floors = {label: rel_l2(got, ref) for label, got, ref in sets64}
    fp64 = {label: got for label, got, _ in sets64}
    bad = []
    for label, got, ref in _series_sets(sim, g, v):
        r = rel_l2(got, ref)
        exact = int(np.sum(got == ref))
        print(f"  {label:14s} relL2={r:.2e}  (reference vs FP64: {floors[label]:.2e}, device vs FP64: "
              f"{rel_l2(got, fp64[label]):.2e})  exact {exact}/{ref.size}")
        if not r <= TOL:
            bad.append((label, r))
    if "ff_time" in g:
        np.testing.assert_array_equal(sim.ff_time(), g["ff_time"])
    assert not bad, bad


@pytest.mark.parametrize("name", ["dipole36", "antenna_x", "c1_mini", "sheets_xyz"])
def test_fields_bit_exact_vs_oracle(runs, name):
    """Antenna decks (no TF/SF plane wave): every field value left in device
    memory equals the oracle's, which equals the serial reference's."""
    plan, sim, _ = runs[name]
    orc = Oracle(plan)
    orc.step(plan.time_steps)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(sim.field(comp), orc.field(comp), err_msg=comp)


@pytest.mark.parametrize("name", ["drude_slab", "master_ports", "sphere_oblique", "c2_mini", "c3_mini"])
def test_fields_close_to_oracle(runs, name):
    """Plane-wave decks: the only difference from the oracle is the last-bit
    rounding of exp() in the analytic incident field."""
    plan, sim, _ = runs[name]
    orc = Oracle(plan)
    orc.step(plan.time_steps)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        a, b = sim.field(comp), orc.field(comp)
        r = rel_l2(a, b)
        print(f"  {name} {comp}: relL2={r:.2e} exact={np.mean(a == b):.6f}")
        assert r <= 5e-5, (comp, r)


@pytest.mark.parametrize("name", ["dipole36", "c1_mini", "antenna_x", "sheets_xyz"])
def test_antenna_decks_bit_exact(runs, name):
    """No plane-wave source: port voltage, incident wave and far field (with
    the reference's FP32 retarded-time binning order) are the reference's,
    bit for bit."""
    plan, sim, g = runs[name]
    np.testing.assert_array_equal(sim.series("incident"), g["incident"])
    for q in range(plan.view.n_ports):
        np.testing.assert_array_equal(sim.series("port_v", q), g["port_v"][q])
    for a in range(plan.view.n_angles):
        np.testing.assert_array_equal(sim.series("e_theta", a), g["e_theta"][a], err_msg=f"E_theta {a}")
        np.testing.assert_array_equal(sim.series("e_phi", a), g["e_phi"][a], err_msg=f"E_phi {a}")


@pytest.mark.parametrize("name", ["c1_mini", "drude_slab", "master_ports", "sphere_oblique", "c2_mini"])
@pytest.mark.parametrize("planes", [0, 1, 7])
def test_cpml_fusion_matches_separate_kernels(golden_dir, name, planes):
    """Folding the CPML correction into the bulk kernels, and the TF/SF face
    pairs into one launch per phase, keeps the reference's per-element add
    order: fields equal the separate-kernel path (one launch per CPML axis
    and per TF/SF face pair), for any number of z planes per bulk thread."""
    deck = os.path.join(golden_dir, "decks", name + ".txt")
    fused = Simulation(Plan.from_inputs(deck), planes=planes)
    plain = Simulation(Plan.from_inputs(deck), fuse_cpml=False)
    fused.run(120)
    plain.run(120)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(fused.field(comp), plain.field(comp), err_msg=comp)


def single_pass_deck(cells=(70, 45, 39), steps=160):
    """Dipole with a port, a dielectric sphere and an anisotropic block in a
    CPML corner on a grid whose x extent spans two 64-cell tiles, the tile
    seam (x = 64) lying in the upper x slab."""
    from alpineem_amd import decks
    nx, ny, nz = cells
    c = (nx // 2, ny // 2, nz // 2)
    d = decks.dipole(n=40, steps=steps, arm=6, angles=3)
    d.cells = cells
    d.phase_centre = c
    d.materials = [decks.Material(4, sigma=(1e8, 1e8, 1e8)), decks.Material(5, eps=(2.5, 3.0, 4.0)),
                   decks.Material(6, eps=(1.5, 1.5, 1.5), sigma=(0.02, 0.01, 0.03))]
    d.objects = [decks.Block(4, (c[0], c[1], c[2] - 6), (1, 1, 6)), decks.Block(4, (c[0], c[1], c[2] + 1), (1, 1, 6)),
                 decks.Sphere(5, (c[0] + 14, c[1], c[2]), 5), decks.Block(6, (3, 3, 3), (12, 8, 6)),
                 decks.Block(5, (nx - 12, ny - 15, nz - 19), (11, 9, 10))]
    d.ports = [decks.Port("z", 50.0, c, (0, 0, 1))]
    return d


@pytest.mark.parametrize("cells", [(70, 45, 39), (150, 37, 29)])
@pytest.mark.parametrize("planes", [0, 1, 5])
def test_single_pass_matches_split_launches(planes, cells):
    """The single-pass core launch (H and E of a step together, halo H
    formed again from the step-start fields) leaves fields, port series and
    far-field state equal, bit for bit, to separate H and E bulk launches --
    across x/y tile seams (in an x slab on the 70-cell grid; slab-free and
    slab tiles as separate launches on the 150-cell grid), z chunks, graph
    and eager steps and odd step counts (the two buffer sets alternate)."""
    from alpineem_amd.solver import Plan as P
    text = single_pass_deck(cells).text()
    sp = Simulation(P.from_text(text), planes=planes)
    ref = Simulation(P.from_text(text), planes=planes, single_pass=False)
    assert sp.single_pass and not ref.single_pass
    for steps, graph in ((37, True), (1, False), (2, False), (60, True), (3, False), (57, True)):
        sp.run(steps, graph=graph)
        ref.run(steps, graph=graph)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(sp.field(comp), ref.field(comp), err_msg=comp)
    sp.finalize()
    ref.finalize()
    np.testing.assert_array_equal(sp.series("port_v", 0), ref.series("port_v", 0))
    np.testing.assert_array_equal(sp.series("port_i", 0), ref.series("port_i", 0))
    for a in range(3):
        np.testing.assert_array_equal(sp.series("e_theta", a), ref.series("e_theta", a))


def edge_deck(cells, sizes, steps=90):
    """A port-fed deck with a dielectric sphere on an arbitrary grid and cell."""
    from alpineem_amd import decks
    nx, ny, nz = cells
    c = (nx // 2, ny // 2, nz // 2)
    d = decks.dipole(n=40, steps=steps, arm=2, angles=2)
    d.cells = cells
    d.cell_size = sizes
    d.phase_centre = c
    d.materials = [decks.Material(4, sigma=(1e8, 1e8, 1e8)), decks.Material(5, eps=(3.0, 2.0, 4.0))]
    d.objects = [decks.Block(4, (c[0], c[1], c[2] - 2), (1, 1, 2)), decks.Block(4, (c[0], c[1], c[2] + 1), (1, 1, 2)),
                 decks.Sphere(5, (c[0] + 2, c[1] - 1, c[2]), 2)]
    d.ports = [decks.Port("z", 50.0, c, (0, 0, 1))]
    return d


@pytest.mark.parametrize("cells,sizes", [
    ((65, 31, 27), (1e-3, 1e-3, 1e-3)),        # a 1-cell second x tile
    ((129, 27, 29), (1e-3, 1.3e-3, 0.8e-3)),   # three x tiles, the last 1 cell; unequal cell sizes
    ((33, 64, 26), (0.9e-3, 1e-3, 1.1e-3)),    # one x tile; rows not a multiple of a work-group's
])
@pytest.mark.parametrize("planes", [0, 3])
def test_single_pass_edge_grids(cells, sizes, planes):
    """The single pass equals the split launches bit for bit on grids whose
    x extent leaves a 1-cell tile, whose rows and planes leave partial work-
    groups and chunks, and on unequal cell sizes."""
    from alpineem_amd.solver import Plan as P
    text = edge_deck(cells, sizes).text()
    sp = Simulation(P.from_text(text), planes=planes)
    ref = Simulation(P.from_text(text), planes=planes, single_pass=False)
    assert sp.single_pass
    sp.run(45)
    ref.run(45)
    sp.run(45, graph=False)
    ref.run(45, graph=False)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(sp.field(comp), ref.field(comp), err_msg=comp)


@pytest.mark.parametrize("name", ["dipole36", "c1_mini", "antenna_x"])
def test_single_pass_golden_decks(golden_dir, name):
    """Decks that qualify run the single-pass step and stay equal to the
    split launches and, through the parity tests above, to the oracle."""
    deck = os.path.join(golden_dir, "decks", name + ".txt")
    sp = Simulation(Plan.from_inputs(deck), planes=3)
    ref = Simulation(Plan.from_inputs(deck), single_pass=False)
    assert sp.single_pass
    sp.run(101)
    ref.run(101)
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        np.testing.assert_array_equal(sp.field(comp), ref.field(comp), err_msg=comp)


@pytest.mark.parametrize("mode", ["atomic", "histogram"])
@pytest.mark.parametrize("name", ["c1_mini", "sphere_oblique"])
def test_ntff_fp64_modes_close(golden_dir, name, mode):
    """The FP64 binnings (atomic, LDS histogram) agree with the ordered FP32
    binning -- the reference's output -- within its FP32 noise: 1e-5 rel-L2,
    or the reference's own distance from FP64 arithmetic where that is
    larger (cross-polar nulls)."""
    deck = os.path.join(golden_dir, "decks", name + ".txt")
    plan = Plan.from_inputs(deck)
    a = Simulation(plan)
    b = Simulation(Plan.from_inputs(deck), ntff=mode)
    for s in (a, b):
        s.run()
        s.finalize()
    o64 = Oracle(plan, fp64=True)
    o64.step(plan.time_steps)
    o64.finalize()
    for ang in range(plan.view.n_angles):
        ea = np.concatenate([a.series("e_theta", ang), a.series("e_phi", ang)])
        eb = np.concatenate([b.series("e_theta", ang), b.series("e_phi", ang)])
        e64 = np.concatenate([o64.series("e_theta", ang), o64.series("e_phi", ang)])
        assert rel_l2(eb, ea) <= max(1e-5, rel_l2(ea, e64)), (ang, rel_l2(eb, ea))


def test_graph_and_eager_paths_agree(golden_dir):
    """Replaying the captured step graph gives the same fields as launching
    every kernel eagerly; several replays are interleaved with host syncs."""
    deck = os.path.join(golden_dir, "decks", "master_ports.txt")
    pa = Plan.from_inputs(deck)
    a = Simulation(pa)
    b = Simulation(Plan.from_inputs(deck))
    for _ in range(4):
        a.run(25, graph=True)
        b.run(25, graph=False)
        for comp in ("Ex", "Hz"):
            np.testing.assert_array_equal(a.field(comp), b.field(comp))


@pytest.mark.parametrize("opts", [dict(cpml_cells=16, cpml_kappa_max=5.0, cpml_alpha_max=0.05),
                                  dict(cpml_cells=12, cpml_grading=4, tfsf_inset=6, ntff_gap=4)])
def test_setup_options_bit_exact_vs_oracle(opts):
    """Setup constants other than the reference's (CPML thickness, grading,
    kappa/alpha maxima, far-field box gap): the device and the oracle run
    the same plan, and an antenna deck's fields, port voltage and ordered
    far field stay equal bit for bit, in the single pass and in the split
    launches."""
    from alpineem_amd import decks
    text = decks.dipole(n=56, steps=140, angles=3, arm=6).text()
    plan = Plan.from_text(text, **opts)
    assert plan.view.npml[0] == opts["cpml_cells"] + 1
    orc = Oracle(plan)
    orc.step(plan.time_steps)
    orc.finalize()
    for sp in (True, False):
        sim = Simulation(plan, single_pass=sp)
        sim.run(plan.time_steps)
        sim.finalize()
        for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            np.testing.assert_array_equal(sim.field(comp), orc.field(comp), err_msg=f"{comp} single_pass={sp}")
        np.testing.assert_array_equal(sim.series("port_v"), orc.series("port_v"))
        for a in range(plan.view.n_angles):
            np.testing.assert_array_equal(sim.series("e_theta", a), orc.series("e_theta", a))
            np.testing.assert_array_equal(sim.series("e_phi", a), orc.series("e_phi", a))
        sim.close()


@pytest.mark.parametrize("name,angles", [("c1_mini", None), ("sphere_oblique", None), ("dipole_37", 37)])
def test_ntff_lists_equal_row_scan(golden_dir, name, angles):
    """Ordered far-field binning: the contributor-list kernel and the
    row-scan kernel (default) give the same FP32 accumulators bit for bit;
    the device time per step with each is printed."""
    from alpineem_amd import decks
    if angles:
        plan = Plan.from_text(decks.dipole(n=64, steps=300, angles=angles, arm=12).text())
    else:
        plan = Plan.from_inputs(os.path.join(golden_dir, "decks", name + ".txt"))
    out, ms = {}, {}
    for scan in (False, True):
        sim = Simulation(plan, ntff_lists=not scan)
        ms[scan] = sim.run(plan.time_steps)
        out[scan] = [sim.ntff_raw(w) for w in range(6)]
        sim.close()
    print(f"\n  {name}: {plan.view.n_angles} angles, ms per step: lists {ms[False]:.4f}, row scan {ms[True]:.4f}")
    assert any(np.any(out[False][w] != 0) for w in range(6))
    for w in range(6):
        np.testing.assert_array_equal(out[False][w], out[True][w], err_msg=f"accumulator {w}")


def test_ground_plane_refusals(golden_dir):
    """An infinite ground plane deck's far field runs only in the ordered
    binning from contributor lists; the other binnings, the frequency-domain
    far field and the adjoint refuse it with AEM_ERR_UNSUPPORTED."""
    from alpineem_amd._lib import AemError
    deck = os.path.join(golden_dir, "decks", "igp_mz.txt")
    for kw in (dict(ntff="histogram"), dict(ntff="atomic"), dict(ntff_lists=False)):
        with pytest.raises(AemError, match="ground plane"):
            Simulation(Plan.from_inputs(deck), **kw)
    sim = Simulation(Plan.from_inputs(deck))
    with pytest.raises(AemError, match="ground plane"):
        sim.set_objective(0.0, 0.0, 10e9, "theta")
    with pytest.raises(AemError, match="ground plane"):
        sim.record_design((20, 20, 24), (21, 21, 25))
    sim.close()


@pytest.mark.parametrize("name", ["igp_mz", "igp_px"])
def test_ground_plane_fields_close_to_oracle(golden_dir, name):
    """Infinite ground planes (igp_mz: -z under a metal half sphere; igp_px:
    +x beside a dielectric block, oblique incidence): one TF/SF face off, a
    mirrored plane wave.  While the pulse enters and crosses the box the
    fields stay within the exp() rounding of the incident and mirrored
    waves from the oracle's (rel-L2 5e-5); after it has
    left, within 2e-6 of the incident amplitude plus 1e-6 of the value (the
    residual of one field is 1e-4 of the other's scale, so a rel-L2 of it
    alone magnifies the same rounding)."""
    plan = Plan.from_inputs(os.path.join(golden_dir, "decks", name + ".txt"))
    sim = Simulation(plan)
    orc = Oracle(plan)
    done = 0
    for n in (5, 120, 200, plan.time_steps):
        sim.run(n - done, graph=n > 5)
        orc.step(n - done)
        done = n
        for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            a, b = sim.field(comp), orc.field(comp)
            if n < plan.time_steps:
                assert rel_l2(a, b) <= 5e-5, (n, comp, rel_l2(a, b))
            else:
                amp = float(plan.view.plane_wave_amp) * (1.0 if comp[0] == "E" else 1.0 / 376.73)
                d = np.abs(a.astype(np.float64) - b)
                assert np.all(d <= 2e-6 * amp + 1e-6 * np.abs(b)), (comp, d.max())
    sim.close()
